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


def test_unreadable_graph_fails_with_status():
    rc, out = run("sssp", "--input", "/nonexistent/graph.wel")
    assert rc == 1 and "could not be read" in out
    rc, out = run("sssp", "--input", os.path.join(GOLD, "make_golden.py"))
    assert rc == 1 and "not recognized" in out


def test_reading_reports_reference_counts():
    rc, out = run("cc", "--input", os.path.join(GOLD, "loader_zero.el"))
    assert "Number of nodes = 4" in out and "Number of edges = 4" in out
