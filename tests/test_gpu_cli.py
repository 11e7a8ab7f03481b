"""The C++ driver executables end to end on the GPU: file in, labels file out (the reference
output format "<id> <value>" per line, src/utilities.cpp:28-38), checked against the oracle."""
import os
import subprocess

import numpy as np
import pytest

import eegraph_amd as eg
import oracle_py as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "build", "bin")

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def wel(tmp_path_factory):
    s, d, w = eg.gen_rmat(12, 16, seed=1, wmin=1, wmax=64)
    s[0] = 0  # vertex 0 occurs, so the file is read 0-based
    p = tmp_path_factory.mktemp("g") / "rmat12.wel"
    eg.save_edgelist(p, 1 << 12, s, d, w)
    return str(p), s, d, w


def read_labels(path, dtype):
    ids, vals = [], []
    with open(path) as f:
        for line in f:
            a, b = line.split()
            ids.append(int(a))
            vals.append(float(b) if dtype == np.float32 else int(b))
    assert ids == list(range(len(ids)))
    return np.array(vals, dtype=dtype)


@pytest.mark.parametrize("exe,alg", [("bfs", O.BFS), ("sssp", O.SSSP), ("sswp", O.SSWP), ("cc", O.CC)])
@pytest.mark.parametrize("variant", ["sync_push_dd", "async_push_td"])
def test_driver_outputs_match_oracle(wel, tmp_path, exe, alg, variant):
    path, s, d, w = wel
    n = int(max(s.max(), d.max())) + 1
    src = int(np.argmax(np.bincount(s, minlength=n)))
    out = tmp_path / f"{exe}.txt"
    p = subprocess.run([os.path.join(BIN, exe), "--input", path, "--source", str(src), "--variant", variant,
                        "--output", str(out), "--debug", "true"], capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stdout + p.stderr
    assert "Number of iterations" in p.stdout and "CPU check: identical" in p.stdout
    lab = read_labels(out, np.uint32)
    assert np.array_equal(lab, O.reference_labels(alg, n, s, d, w, src))


def test_pagerank_driver(wel, tmp_path):
    path, s, d, w = wel
    n = int(max(s.max(), d.max())) + 1
    out = tmp_path / "pr.txt"
    p = subprocess.run([os.path.join(BIN, "pr"), "--input", path, "--variant", "sync_push_dd",
                        "--output", str(out)], capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stdout + p.stderr
    x = read_labels(out, np.float32)
    row, col, _ = O.csr(n, s, d)
    xf, _, _ = O.pagerank(n, row, col, 0.85, 1e-13)
    assert np.abs(x.astype(np.float64) - xf).sum() < 1e-6


def test_source_out_of_range_is_an_error(wel):
    path, s, d, w = wel
    p = subprocess.run([os.path.join(BIN, "bfs"), "--input", path, "--source", "99999999"],
                       capture_output=True, text=True, timeout=120)
    assert p.returncode == 1 and "out of range" in p.stdout
