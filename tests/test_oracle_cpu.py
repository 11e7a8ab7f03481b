"""Pins the CPU oracle before it is trusted as the checker (no GPU needed).

- the restated reference sweeps (seq_cpu over the Tigr layout) equal the textbook oracles;
- the textbook oracles equal Floyd-Warshall closures on tiny graphs (committed vectors);
- closed forms: paths, stars, cycles, disconnected pieces, PageRank on a cycle / two nodes;
- sensitivity: a dropped weight, a reversed edge or a wrong damping changes the answer, so the
  checks above would catch those mistakes.
"""
import json
import os

import numpy as np
import pytest

import eegraph_amd as eg
import oracle_py as O

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
INF = O.INF


@pytest.fixture(scope="module")
def rmat12():
    s, d, w = eg.gen_rmat(12, 16, seed=1, wmin=1, wmax=64)
    return 1 << 12, s, d, w


@pytest.mark.parametrize("alg", [O.BFS, O.SSSP, O.SSWP, O.CC])
@pytest.mark.parametrize("K", [4, 8, 32])
def test_restated_sweep_equals_textbook(rmat12, alg, K):
    n, s, d, w = rmat12
    src = int(np.argmax(np.bincount(s, minlength=n)))
    if alg == O.CC:
        lay = O.tigr_layout(n, np.concatenate([s, d]), np.concatenate([d, s]), None, K)
    else:
        lay = O.tigr_layout(n, s, d, w, K)
    lab = O.init_labels(alg, n, src)
    passes, conv = O.seq_sweep(alg, n, lay, K, lab)
    assert conv
    assert np.array_equal(lab, O.reference_labels(alg, n, s, d, w, src))


def test_restated_sweep_on_grid():
    s, d, w = eg.gen_grid(40, 30, 0.01, seed=3, wmin=1, wmax=1000)
    n = 1200
    for alg in (O.BFS, O.SSSP, O.SSWP, O.CC):
        lay = O.tigr_layout(n, s, d, w, 4)
        lab = O.init_labels(alg, n, 7)
        O.seq_sweep(alg, n, lay, 4, lab)
        assert np.array_equal(lab, O.reference_labels(alg, n, s, d, w, 7)), alg


def test_textbook_equals_committed_bruteforce():
    with open(os.path.join(GOLD, "tiny_bruteforce.json")) as f:
        cases = json.load(f)
    for c in cases:
        s, d, w = (np.array(c[k], np.uint32) for k in ("src", "dst", "w"))
        for name, code in (("bfs", O.BFS), ("sssp", O.SSSP), ("sswp", O.SSWP), ("cc", O.CC)):
            assert O.reference_labels(code, c["n"], s, d, w, c["source"]).tolist() == c[name]
            assert O.brute(code, c["n"], s, d, w, c["source"]).tolist() == c[name]


def test_closed_forms_path_star_components():
    # path 0 -> 1 -> ... -> 9 with weight i+1 on edge i; vertices 10, 11 isolated
    n = 12
    s = np.arange(9, dtype=np.uint32)
    d = s + 1
    w = s + 1
    assert O.reference_labels(O.BFS, n, s, d, w, 0).tolist() == list(range(10)) + [INF, INF]
    sssp = [int(k * (k + 1) // 2) for k in range(10)] + [INF, INF]
    assert O.reference_labels(O.SSSP, n, s, d, w, 0).tolist() == sssp
    sswp = [INF] + [1] * 9 + [0, 0]  # the first edge (weight 1) bounds every path
    assert O.reference_labels(O.SSWP, n, s, d, w, 0).tolist() == sswp
    assert O.reference_labels(O.CC, n, s, d, w, 0).tolist() == [0] * 10 + [10, 11]
    # star with centre 5: widest path to leaf i is its edge weight
    s = np.full(5, 5, np.uint32)
    d = np.array([0, 1, 2, 3, 4], np.uint32)
    w = np.array([9, 3, 7, 1, 4], np.uint32)
    assert O.reference_labels(O.SSWP, 6, s, d, w, 5).tolist() == [9, 3, 7, 1, 4, INF]
    assert O.reference_labels(O.SSSP, 6, s, d, w, 5).tolist() == [9, 3, 7, 1, 4, 0]
    # min-label components: {0,3}, {1,2,4}, {5}
    s = np.array([3, 4, 2], np.uint32)
    d = np.array([0, 2, 1], np.uint32)
    assert O.reference_labels(O.CC, 6, s, d, None, 0).tolist() == [0, 1, 1, 0, 1, 5]


def test_sensitivity_of_the_checks(rmat12):
    """A dropped weight term, a transposed edge list, or min/max swapped changes the answer."""
    n, s, d, w = rmat12
    src = int(np.argmax(np.bincount(s, minlength=n)))
    sssp = O.reference_labels(O.SSSP, n, s, d, w, src)
    assert not np.array_equal(sssp, O.reference_labels(O.SSSP, n, s, d, np.ones_like(w), src))
    assert not np.array_equal(sssp, O.reference_labels(O.SSSP, n, d, s, w, src))
    sswp = O.reference_labels(O.SSWP, n, s, d, w, src)
    assert not np.array_equal(sswp, O.reference_labels(O.SSWP, n, d, s, w, src))
    bfs = O.reference_labels(O.BFS, n, s, d, w, src)
    assert not np.array_equal(bfs, sssp)


def test_tigr_layout_invariants(rmat12):
    n, s, d, w = rmat12
    K = 8
    np_, el, parts = O.tigr_layout(n, s, d, w, K)
    deg = np.bincount(s, minlength=n).astype(np.uint64)
    assert np.array_equal(np_, np.concatenate([[0], np.cumsum(2 * deg + 1)[:-1]]).astype(np.uint64))
    assert np.array_equal(el[np_.astype(np.int64)], deg.astype(np.uint32))
    nparts = (deg + K - 1) // K
    assert len(parts) == int(nparts.sum())
    # parts ascending by (v, p) and covering p < ceil(deg/K)
    expect = np.concatenate([np.stack([np.full(k, v), np.arange(k)], 1)
                             for v, k in enumerate(nparts) if k]).astype(np.uint32)
    assert np.array_equal(parts, expect)
    # every edge sits once, in input order within its source, at np+1+2j (dst) / np+2+2j (w)
    row, col, wo = O.csr(n, s, d, w)
    for v in (0, 1, 17, n - 1):
        k = int(deg[v])
        base = int(np_[v])
        assert el[base + 1: base + 1 + 2 * k: 2].tolist() == col[row[v]:row[v + 1]].tolist()
        assert el[base + 2: base + 2 + 2 * k: 2].tolist() == wo[row[v]:row[v + 1]].tolist()
    # the strided part ownership j = p + i*ceil(deg/K) covers each edge exactly once
    v = int(np.argmax(deg))
    k = int(deg[v])
    npv = (k + K - 1) // K
    owned = sorted(p + i * npv for p in range(npv) for i in range(K) if p + i * npv < k)
    assert owned == list(range(k))


def test_pagerank_closed_forms():
    # directed cycle: uniform 1/N
    n = 50
    s = np.arange(n, dtype=np.uint32)
    row, col, _ = O.csr(n, s, (s + 1) % n)
    x, it, res = O.pagerank(n, row, col, 0.85, 1e-14)
    assert np.allclose(x, 1.0 / n, rtol=0, atol=1e-14)
    # two nodes 0 -> 1, 1 dangling: x0 = (1-d)/2, x1 = (1-d)/2 + d*x0
    row, col, _ = O.csr(2, np.array([0], np.uint32), np.array([1], np.uint32))
    x, _, _ = O.pagerank(2, row, col, 0.85, 1e-15)
    assert abs(x[0] - 0.075) < 1e-14 and abs(x[1] - (0.075 + 0.85 * 0.075)) < 1e-14
    # wrong damping is detected
    x2, _, _ = O.pagerank(2, row, col, 0.80, 1e-15)
    assert abs(x2[1] - x[1]) > 1e-3
    # personalised on the same graph: restart mass at 0 flows to 1
    x, _, _ = O.pagerank(2, row, col, 0.85, 1e-15, personalised=True, source=0)
    assert abs(x[0] - 0.15) < 1e-14 and abs(x[1] - 0.85 * 0.15) < 1e-14


def test_delta_push_matches_jacobi(rmat12):
    """The restated delta-push PageRank (the CPU baseline for PR) reaches the Jacobi fixpoint."""
    n, s, d, w = rmat12
    row, col, _ = O.csr(n, s, d)
    xj, _, _ = O.pagerank(n, row, col, 0.85, 1e-13)
    xp, passes, conv = O.pr_delta_push(n, row, col, 0.85, 1e-9 * 0.15 / n)
    assert conv
    assert np.abs(xp.astype(np.float64) - xj).sum() < 1e-6


def test_rmat14_fixture_and_oracle_labels():
    """The committed config-1 fixture: the generator still yields the same edges and the
    oracles the same labels."""
    import hashlib

    with open(os.path.join(GOLD, "rmat14_seed1.json")) as f:
        g = json.load(f)
    s, d, w = eg.gen_rmat(g["scale"], g["edge_factor"], seed=g["seed"], wmin=g["wmin"], wmax=g["wmax"])
    h = hashlib.sha256()
    for a in (s, d, w):
        h.update(np.ascontiguousarray(a).tobytes())
    assert h.hexdigest() == g["sha256_src_dst_w"]
    assert [[int(a), int(b), int(c)] for a, b, c in zip(s[:64], d[:64], w[:64])] == g["head"]
    n = 1 << g["scale"]
    src = int(np.argmax(np.bincount(s, minlength=n)))
    assert src == g["argmax_source"]
    for name, code in (("bfs", O.BFS), ("sssp", O.SSSP), ("sswp", O.SSWP), ("cc", O.CC)):
        lab = O.reference_labels(code, n, s, d, w, src)
        assert hashlib.sha256(lab.tobytes()).hexdigest() == g["labels_sha256"][name]


def test_pagerank_pull_equals_push_form():
    """The parallel pull form of the FP64 oracle equals the serial push form (itself pinned by
    closed forms) on an RMAT graph with dangling vertices, for any thread count."""
    import eegraph_amd as eg
    s, d, w = eg.gen_rmat(12, 8, seed=3)
    n = 1 << 12
    row, col, _ = O.csr(n, s, d)
    irow, icol, _ = O.csr(n, d, s)
    outdeg = np.diff(row).astype(np.uint32)
    xp, itp, _ = O.pagerank(n, row, col, 0.85, 1e-12)
    for threads in (1, 3, 8):
        xq, itq, _ = O.pagerank_pull(n, irow, icol, outdeg, 0.85, 1e-12, threads=threads)
        assert itq == itp
        assert np.abs(xq - xp).max() < 1e-15 and np.abs(xq - xp).sum() < 1e-12


def test_personalised_pull_equals_push_form():
    """The personalised FP64 pull (one-hot restart) equals the serial personalised push form for
    several restart vertices and thread counts; a different restart vertex gives other ranks."""
    s, d, w = eg.gen_rmat(12, 8, seed=3)
    n = 1 << 12
    row, col, _ = O.csr(n, s, d)
    irow, icol, _ = O.csr(n, d, s)
    outdeg = np.diff(row).astype(np.uint32)
    last = None
    for src in (0, 77, 1000):
        xp, itp, _ = O.pagerank(n, row, col, 0.85, 1e-12, personalised=True, source=src)
        for threads in (1, 4):
            xq, itq, _ = O.pagerank_pull(n, irow, icol, outdeg, 0.85, 1e-12, threads=threads,
                                         personalised=True, source=src)
            assert itq == itp
            assert np.abs(xq - xp).max() < 1e-15 and np.abs(xq - xp).sum() < 1e-12
        if last is not None:
            assert np.abs(xq - last).sum() > 1e-3
        last = xq


@pytest.mark.parametrize("shape", ["rmat", "grid", "zero_weights"])
def test_bucket_queue_oracles_equal_heap_oracles(shape):
    """Dial's bucket queues give Dijkstra's and max-min Dijkstra's answers (RMAT with weights
    1..64, a grid with shortcuts and weights 1..1000, a graph with zero-weight edges), from
    several sources; swapping the two widest rules changes the answer."""
    if shape == "rmat":
        s, d, w = eg.gen_rmat(13, 16, seed=9, wmin=1, wmax=64)
        n = 1 << 13
    elif shape == "grid":
        s, d, w = eg.gen_grid(60, 50, 0.01, seed=3, wmin=1, wmax=1000)
        n = 3000
    else:
        s, d, w = eg.gen_rmat(11, 8, seed=4, wmin=0, wmax=3)
        n = 1 << 11
    row, col, wo = O.csr(n, s, d, w)
    outdeg = np.diff(row)
    for src in (int(np.argmax(outdeg)), 0, 5):
        assert np.array_equal(O.dial(n, row, col, wo, src), O.dijkstra(n, row, col, wo, src))
        assert np.array_equal(O.dial_widest(n, row, col, wo, src), O.widest(n, row, col, wo, src))
    with pytest.raises(ValueError):
        O.dial(n, row, col, wo, 0, wmax=int(wo.max()) - 1)


def test_bucket_queue_oracles_equal_bruteforce():
    with open(os.path.join(GOLD, "tiny_bruteforce.json")) as f:
        cases = json.load(f)
    for c in cases:
        s, d, w = (np.array(c[k], np.uint32) for k in ("src", "dst", "w"))
        row, col, wo = O.csr(c["n"], s, d, w)
        assert O.dial(c["n"], row, col, wo, c["source"]).tolist() == c["sssp"]
        assert O.dial_widest(c["n"], row, col, wo, c["source"]).tolist() == c["sswp"]


def test_csr_build_is_a_stable_counting_sort():
    """The parallel CSR build keeps every row in input order (equal to a stable argsort by
    source), with hubs, empty rows and duplicate edges."""
    s, d, w = eg.gen_rmat(14, 16, seed=12, wmin=1, wmax=64)
    n = 1 << 14
    row, col, wo = O.csr(n, s, d, w)
    order = np.argsort(s, kind="stable")
    assert np.array_equal(row, np.concatenate([[0], np.cumsum(np.bincount(s, minlength=n))]).astype(np.uint64))
    assert np.array_equal(col, d[order]) and np.array_equal(wo, w[order])
