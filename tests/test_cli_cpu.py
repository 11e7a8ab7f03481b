"""The C++ driver executables (build/bin): argument handling of the reference
(src/argument_parsing.cpp) and error paths that need no GPU."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "build", "bin")
GOLD = os.path.join(ROOT, "tests", "golden")

pytestmark = pytest.mark.skipif(not os.path.exists(os.path.join(BIN, "sssp")),
                                reason="drivers not built (make -C eegraph_amd/csrc)")


def run(*args):
    p = subprocess.run([os.path.join(BIN, args[0])] + list(args[1:]), capture_output=True, text=True,
                       timeout=60)
    return p.returncode, p.stdout + p.stderr


def test_help_and_usage_errors_exit_zero_like_the_reference():
    rc, out = run("sssp")
    assert rc == 0 and "Required arguments" in out
    rc, out = run("bfs", "--help")
    assert rc == 0 and "--variant" in out and "--k" in out
    rc, out = run("bfs", "--input")  # odd number of tokens
    assert rc == 0 and "error parsing command line arguments" in out
    rc, out = run("bfs", "--input", "g.el", "--bogus", "1")
    assert rc == 0 and "<--bogus>" in out
    rc, out = run("bfs", "--input", "g.el", "--variant", "fast")
    assert rc == 0 and "Variant not recognized" in out
    rc, out = run("bfs", "--source", "3")
    assert rc == 0 and "Input graph file argument is required" in out
    rc, out = run("pr", "--input", "g.el", "--energy", "true")
    assert rc == 0 and "energy file and/or energy stats" in out
    rc, out = run("bfs", "--input", "g.el", "--k", "7")
    assert rc == 0 and "K must be" in out
    rc, out = run("bfs", "--input", "g.el", "--frontier", "bitmap")
    assert rc == 0 and "Frontier must be sorted, compact or flags" in out
    rc, out = run("bfs", "--help")
    assert "sorted, compact or flags (Default: sorted)" in out


def test_unreadable_graph_fails_with_status():
    rc, out = run("sssp", "--input", "/nonexistent/graph.wel")
    assert rc == 1 and "could not be read" in out
    rc, out = run("sssp", "--input", os.path.join(GOLD, "make_golden.py"))
    assert rc == 1 and "not recognized" in out


def test_reading_reports_reference_counts():
    rc, out = run("cc", "--input", os.path.join(GOLD, "loader_zero.el"))
    assert "Number of nodes = 4" in out and "Number of edges = 4" in out


UM_CALLER = r"""
// The unified-memory callers of src/main.cu:45,309-325: a UMGraph read from a file and handed
// to the eegraph_<alg>_um entry points.
#include "eegraph_host.hpp"

int main(int argc, char** argv) {
    ArgumentParser arguments(argc, argv, true, false);
    Result result;
    UMGraph um_graph(arguments.input, true);
    um_graph.ReadGraph();
    result = eegraph_bfs_um(arguments, um_graph);
    result = eegraph_cc_um(arguments, um_graph);
    result = eegraph_pr_um(arguments, um_graph);
    result = eegraph_sssp_um(arguments, um_graph);
    result = eegraph_sswp_um(arguments, um_graph);
    return (int)result.time;
}
"""


def test_unified_memory_callers_compile_against_the_host_surface(tmp_path):
    """The reference's _um call sites build unchanged against host/eegraph_host.hpp and link
    with the host objects and libeegraph.so."""
    src = tmp_path / "um_caller.cpp"
    src.write_text(UM_CALLER)
    obj = os.path.join(ROOT, "build", "obj")
    host_objs = [os.path.join(obj, "host_argument_parsing.o"), os.path.join(obj, "host_eegraph.o")]
    if not all(os.path.exists(o) for o in host_objs):
        pytest.skip("host objects not built (make -C eegraph_amd/csrc)")
    exe = tmp_path / "um_caller"
    p = subprocess.run(["g++", "-std=c++17", "-I", os.path.join(ROOT, "eegraph_amd", "csrc", "host"), str(src)]
                       + host_objs + ["-L", os.path.join(ROOT, "eegraph_amd", "lib"), "-leegraph", "-lpthread",
                                      "-o", str(exe)], capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stderr
