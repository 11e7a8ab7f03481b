"""Per-iteration activity (the reference's sumLabels vector, src/gpu_utils.cu:28-34, collected
with --activity, src/main.cu:171-174) against what BFS levels fix exactly.

In sync_push_dd with compact queues, iteration k processes the vertices at BFS level k-1 that
have out-edges, so the activity vector is the level histogram of those vertices.  With flag
scans (and in the other modes) an entry counts the label updates of the iteration, which for
BFS are the vertices of level k.  The CLI's "Max active pct" and "Pct over threshold" lines are
utilities::maxActivePct / pctIterOverThreshold (src/utilities.cpp:62-81) of that vector."""
import os
import re
import subprocess

import numpy as np
import pytest

import eegraph_amd as eg
import oracle_py as O

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "build", "bin")


@pytest.fixture(scope="module")
def rmat14():
    s, d, w = eg.gen_rmat(14, 16, seed=1, wmin=1, wmax=64)
    n = 1 << 14
    g = eg.Graph.from_edges(s, d, w, n=n)
    row, col, _ = O.csr(n, s, d, w)
    src = g.info()["argmax_out_degree"]
    lev = O.bfs(n, row, col, src)
    outdeg = np.diff(row)
    yield n, s, d, w, g, src, lev, outdeg
    g.close()


def _trim(a):
    a = list(int(x) for x in a)
    while a and a[-1] == 0:
        a.pop()
    return a


def _level_hist(lev, mask):
    reached = lev != O.INF
    top = int(lev[reached].max())
    return [int(np.count_nonzero((lev == k) & mask)) for k in range(top + 1)]


@pytest.mark.parametrize("K", [4, 8, 32])
def test_queue_activity_is_the_level_histogram_of_expanding_vertices(rmat14, K):
    n, s, d, w, g, src, lev, outdeg = rmat14
    lab, st = g.run("bfs", src, variant="sync_push_dd", K=K, activity_cap=4096)
    assert np.array_equal(lab, lev)
    want = _trim(_level_hist(lev, outdeg > 0))
    assert _trim(st["activity"]) == want
    assert st["activity_len"] >= len(want)


@pytest.mark.parametrize("K", [4, 32])
def test_sorted_queue_activity_is_the_level_histogram_too(rmat14, K):
    n, s, d, w, g, src, lev, outdeg = rmat14
    lab, st = g.run("bfs", src, variant="sync_push_dd", frontier="sorted", K=K, activity_cap=4096)
    assert np.array_equal(lab, lev)
    assert _trim(st["activity"]) == _trim(_level_hist(lev, outdeg > 0))


def test_flag_scan_activity_counts_each_level_once(rmat14):
    n, s, d, w, g, src, lev, outdeg = rmat14
    lab, st = g.run("bfs", src, variant="sync_push_dd", frontier="flags", K=8, activity_cap=4096)
    assert np.array_equal(lab, lev)
    hist = _level_hist(lev, np.ones(n, bool))
    assert _trim(st["activity"]) == hist[1:]  # iteration k updates the vertices of level k
    assert sum(st["activity"]) == st["label_updates"]


def test_activity_of_other_algorithms_sums_to_the_updates(rmat14):
    n, s, d, w, g, src, lev, outdeg = rmat14
    for alg, var in (("sssp", "sync_push_td"), ("sswp", "async_push_dd"), ("cc", "sync_push_dd")):
        lab, st = g.run(alg, src, variant=var, K=8, activity_cap=1 << 16)
        ref = O.reference_labels({"sssp": O.SSSP, "sswp": O.SSWP, "cc": O.CC}[alg], n, s, d, w, src)
        assert np.array_equal(lab, ref), alg
        assert st["activity_len"] >= 1
        if var != "sync_push_dd":  # label updates per iteration
            assert sum(st["activity"]) == st["label_updates"], alg


def test_cli_activity_summary(rmat14, tmp_path):
    n, s, d, w, g, src, lev, outdeg = rmat14
    path = tmp_path / "rmat14.wel"
    eg.save_edgelist(path, n, s, d, w)
    nn = int(max(s.max(), d.max())) + 1
    p = subprocess.run([os.path.join(BIN, "bfs"), "--input", str(path), "--source", str(src), "--variant",
                        "sync_push_dd", "--activity", "true"], capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stdout + p.stderr
    hist = _level_hist(lev, outdeg > 0)
    mx = float(re.search(r"Max active pct: ([0-9.eE+-]+)", p.stdout).group(1))
    over = float(re.search(r"Pct over threshold: ([0-9.eE+-]+)", p.stdout).group(1))
    act = _trim(hist)  # the queue-mode vector: one entry per iteration that had a frontier
    assert abs(mx - max(act) / nn) < 1e-5
    frac = sum(1 for a in act if a / nn > 0.5)
    assert abs(over - frac / len(act)) < 1e-5
