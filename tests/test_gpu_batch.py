"""Batched traversals: every row of a batched answer must equal the single-query run (and the
oracle) bit for bit — the fixpoints are unique."""
import numpy as np
import pytest

import eegraph_amd as eg
import oracle_py as O

pytestmark = pytest.mark.gpu

ALG = {"bfs": O.BFS, "sssp": O.SSSP, "sswp": O.SSWP}


@pytest.fixture(scope="module")
def rmat13():
    s, d, w = eg.gen_rmat(13, 16, seed=4, wmin=1, wmax=64)
    n = 1 << 13
    g = eg.Graph.from_edges(s, d, w, n=n)
    yield n, s, d, w, g
    g.close()


@pytest.fixture(scope="module")
def rmat22():
    g = eg.Graph.rmat(22, 16, seed=2, wmin=1, wmax=64)
    yield g
    g.close()


@pytest.mark.parametrize("alg", list(ALG))
def test_batch_full_width_rmat22(rmat22, alg):
    """256 sources (four 64-bit mask words per vertex) on RMAT scale 22, EF 16: columns from
    every mask word equal their single-query runs bit for bit."""
    g = rmat22
    n = g.info()["n"]
    srcs = np.random.default_rng(22).integers(0, n, 256).astype(np.uint32)
    srcs[0] = g.info()["argmax_out_degree"]
    out, st = g.run_batch(alg, srcs, K=8)
    assert out.shape == (256, n)
    for q in (0, 1, 63, 64, 127, 128, 200, 255):
        single, _ = g.run(alg, int(srcs[q]), variant="sync_push_dd", K=8)
        assert np.array_equal(out[q], single), (alg, q)
    del out


@pytest.mark.parametrize("alg", list(ALG))
@pytest.mark.parametrize("nsrc", [1, 37, 64, 100, 150, 200, 256, 300])
def test_batch_equals_single_and_oracle(rmat13, alg, nsrc):
    n, s, d, w, g = rmat13
    rng = np.random.default_rng(nsrc)
    srcs = rng.integers(0, n, nsrc).astype(np.uint32)
    srcs[0] = int(np.argmax(np.bincount(s, minlength=n)))
    if nsrc > 3:
        srcs[3] = srcs[2]  # duplicate sources share a frontier entry
    out, st = g.run_batch(alg, srcs, K=8)
    assert out.shape == (nsrc, n)
    for q in sorted(set([0, 1, 2, 3, nsrc - 1, nsrc // 2]) & set(range(nsrc))):
        single, _ = g.run(alg, int(srcs[q]), variant="sync_push_dd", K=8)
        assert np.array_equal(out[q], single), (alg, q)
        assert np.array_equal(out[q], O.reference_labels(ALG[alg], n, s, d, w, int(srcs[q]))), (alg, q)


@pytest.mark.parametrize("alg", list(ALG))
def test_batch_all_rows_vs_oracle(rmat13, alg):
    n, s, d, w, g = rmat13
    srcs = np.arange(0, n, n // 64, dtype=np.uint32)[:64]
    out, _ = g.run_batch(alg, srcs, K=4)
    row, col, wo = O.csr(n, s, d, w)
    for q, src in enumerate(srcs):
        ref = {"bfs": lambda: O.bfs(n, row, col, int(src)),
               "sssp": lambda: O.dijkstra(n, row, col, wo, int(src)),
               "sswp": lambda: O.widest(n, row, col, wo, int(src))}[alg]()
        assert np.array_equal(out[q], ref), (alg, q)


def test_batch_sources_without_edges():
    n = 50
    s = np.array([1, 2, 3], np.uint32)
    d = np.array([2, 3, 4], np.uint32)
    w = np.array([5, 6, 7], np.uint32)
    g = eg.Graph.from_edges(s, d, w, n=n)
    srcs = np.array([0, 1, 49, 1], np.uint32)
    for alg in ALG:
        out, _ = g.run_batch(alg, srcs, K=4)
        for q, src in enumerate(srcs):
            assert np.array_equal(out[q], O.brute(ALG[alg], n, s, d, w, int(src))), (alg, q)


@pytest.mark.parametrize("alg", list(ALG))
@pytest.mark.parametrize("nsrc", [5, 130])
def test_replayed_chunks_match_eager_batches(rmat13, alg, nsrc):
    """Batched runs replay a graph of six iterations unless per-launch timing is asked for; the
    two paths alternate with other work in between and must give the same answers.  5 sources
    use the two-pass plane kernel, 130 the dense one."""
    n, s, d, w, g = rmat13
    rng = np.random.default_rng(nsrc)
    for rep in range(3):
        srcs = rng.integers(0, n, nsrc).astype(np.uint32)
        replayed, st_r = g.run_batch(alg, srcs, K=8)
        g.run("sssp", int(srcs[0]), variant="sync_push_dd", K=8)  # other work in between
        eager, st_e = g.run_batch(alg, srcs, K=8, timing=True)
        assert np.array_equal(replayed, eager), (alg, nsrc, rep)
        # iteration counts may differ: labels are updated in place, so a vertex can see a
        # neighbour's improvement from the same iteration
        for q in (0, nsrc - 1):
            single, _ = g.run(alg, int(srcs[q]), variant="sync_push_dd", K=8)
            assert np.array_equal(replayed[q], single), (alg, nsrc, rep, q)


@pytest.mark.parametrize("alg", list(ALG))
@pytest.mark.parametrize("nsrc", [12, 256])
def test_batch_grid_size_does_not_change_answers(rmat13, alg, nsrc):
    """The same batched answers for one, 48 (the default) and 64 workgroups per CU, through
    both plane kernels (12 queries: the per-edge pass; 256: the wavefront-per-row pass)."""
    n, s, d, w, g = rmat13
    srcs = np.random.default_rng(7 + nsrc).integers(0, n, nsrc).astype(np.uint32)
    ref, _ = g.run_batch(alg, srcs, K=8)
    for wpc in (1, 48, 64):
        out, _ = g.run_batch(alg, srcs, K=8, wgs_per_cu=wpc)
        assert np.array_equal(out, ref), (alg, nsrc, wpc)
    single, _ = g.run(alg, int(srcs[-1]), variant="sync_push_dd", K=8)
    assert np.array_equal(ref[-1], single)


@pytest.mark.parametrize("nsrc", [5, 16, 40])
def test_batched_ppr_planes_against_single_and_fp64(rmat13, nsrc):
    """Personalised PageRank for many restart vertices in planes of 16 or 32 columns (one
    gather pass per iteration for all of them): every row within L1 1e-6 / max 1e-7 of the FP64
    oracle and of its single-query run.  Sources include the hub, a vertex without out-edges, a
    vertex without any edge, and repeats."""
    n, s, d, w, g = rmat13
    row, col, _ = O.csr(n, s, d, w)
    outdeg = np.diff(row)
    indeg = np.bincount(d, minlength=n)
    rng = np.random.default_rng(nsrc)
    srcs = rng.integers(0, n, nsrc).astype(np.uint32)
    srcs[0] = g.info()["argmax_out_degree"]
    srcs[1] = int(np.flatnonzero((outdeg == 0) & (indeg > 0))[0])
    iso = np.flatnonzero((outdeg == 0) & (indeg == 0))
    if len(iso):
        srcs[2] = int(iso[0])
    srcs[-1] = srcs[0]
    lab, st = g.run_batch("ppr", srcs)
    assert lab.shape == (nsrc, n) and lab.dtype == np.float32
    for q, x in enumerate(srcs):
        ref, _, _ = O.pagerank(n, row, col, 0.85, 1e-13, personalised=True, source=int(x))
        err = np.abs(lab[q].astype(np.float64) - ref)
        assert err.sum() < 1e-6 and err.max() < 1e-7, (q, int(x), err.sum(), err.max())
        if q < 4:
            single, _ = g.run("ppr", int(x), variant="sync_push_dd")
            assert np.abs(lab[q].astype(np.float64) - single.astype(np.float64)).sum() < 1e-6


@pytest.mark.parametrize("nsrc", [3, 40])
def test_batched_ppr_hub_chunks(nsrc):
    """Vertices with more than 64 segments of 64 in-edges (in-degree > 4096) add their partial
    sums in chunks of 64 segments first: two hubs (in-degrees 9000 and 4500, just past one
    chunk) and one at exactly 64 segments (4096, no chunks) against the FP64 oracle, with
    restart vertices on and next to the hubs.  Run to an L1 change below 1e-10, so the distance
    left to the fixed point (about 5.7x the last change) stays far below the tolerance; with
    the default 1e-7 a restart on vertex 7 stops ~1.1e-7 short there."""
    n = 10000
    rng = np.random.default_rng(11)
    s = np.concatenate([np.arange(1, 9001), np.arange(5000, 9500), np.arange(2000, 6096),
                        rng.integers(0, n, 30000)]).astype(np.uint32)
    d = np.concatenate([np.zeros(9000), np.full(4500, 7), np.full(4096, 11),
                        rng.integers(12, n, 30000)]).astype(np.uint32)
    g = eg.Graph.from_edges(s, d, np.ones_like(s), n=n)
    try:
        row, col, _ = O.csr(n, s, d)
        srcs = np.array([0, 7, 11, 5, 9000, 9999, 1234][:nsrc] + list(rng.integers(0, n, max(0, nsrc - 7))), np.uint32)
        lab, st = g.run_batch("ppr", srcs, epsilon=1e-10)
        for q, x in enumerate(srcs):
            ref, _, _ = O.pagerank(n, row, col, 0.85, 1e-13, personalised=True, source=int(x))
            err = np.abs(lab[q].astype(np.float64) - ref)
            assert err.sum() < 1e-6 and err.max() < 1e-7, (q, int(x), err.sum(), err.max())
    finally:
        g.close()


@pytest.mark.parametrize("alg,bits", [("sssp", 8), ("sswp", 8)])
@pytest.mark.parametrize("nsrc", [17, 37, 256])
@pytest.mark.parametrize("K", [4, 8, 16, 32])
def test_batch_narrow_labels_equal_32_bit(rmat13, alg, bits, nsrc, K):
    """More than 16 sources take narrow label planes (every weight here <= 64: 8-bit distances,
    queries past 254 answered again wider; 8-bit widest-path labels); their answers are the
    32-bit planes' bit for bit."""
    n, s, d, w, g = rmat13
    srcs = np.random.default_rng(nsrc + K).integers(0, n, nsrc).astype(np.uint32)
    srcs[0] = int(np.argmax(np.bincount(s, minlength=n)))
    out, st = g.run_batch(alg, srcs, K=K)
    assert st["label_bits"] == bits
    wide, stw = g.run_batch(alg, srcs, K=K, label_bits=32)
    assert stw["label_bits"] == 32
    assert np.array_equal(out, wide)
    for q in (0, nsrc - 1):
        assert np.array_equal(out[q], O.reference_labels(ALG[alg], n, s, d, w, int(srcs[q]))), (alg, q)


def _chain(total):
    """A path 0 -> 1 -> ... -> 9 whose weights sum to `total`, plus a side edge."""
    n = 12
    s = np.array(list(range(9)) + [10], np.uint32)
    d = np.array(list(range(1, 10)) + [11], np.uint32)
    w = np.full(10, total // 9, np.uint32)
    w[8] += total - int(w[:9].sum())
    w[9] = 3
    return n, s, d, w


@pytest.mark.parametrize("total,bits", [(65534, 16), (65535, 32), (200000, 32)])
def test_batch_distances_past_16_bits_rerun_with_32(total, bits):
    """The largest 16-bit distance is 65534 (65535 means unreached); a pass whose distances reach
    65535 is answered again with 32-bit labels, and every answer equals the oracle's."""
    n, s, d, w = _chain(total)
    g = eg.Graph.from_edges(s, d, w, n=n)
    try:
        srcs = np.array([0] * 10 + [1, 2, 3, 4, 5, 6, 7, 8, 10, 11], np.uint32)
        out, st = g.run_batch("sssp", srcs)
        assert st["label_bits"] == bits
        assert out[0][9] == total
        for q, x in enumerate(srcs):
            assert np.array_equal(out[q], O.reference_labels(O.SSSP, n, s, d, w, int(x))), q
    finally:
        g.close()


@pytest.mark.parametrize("wmax,bits", [(254, 8), (255, 32)])
def test_batch_widest_labels_need_weights_below_255(wmax, bits):
    """8-bit widest-path labels hold weights up to 254 (255 stands for the source's INF); a
    graph with a heavier edge takes 32-bit labels.  Answers equal the oracle's either way."""
    n = 64
    rng = np.random.default_rng(wmax)
    s = rng.integers(0, n, 400).astype(np.uint32)
    d = rng.integers(0, n, 400).astype(np.uint32)
    w = rng.integers(1, wmax + 1, 400).astype(np.uint32)
    w[0] = wmax
    g = eg.Graph.from_edges(s, d, w, n=n)
    try:
        srcs = np.arange(40, dtype=np.uint32)
        out, st = g.run_batch("sswp", srcs)
        assert st["label_bits"] == bits
        for q, x in enumerate(srcs):
            ref = O.reference_labels(O.SSWP, n, s, d, w, int(x))
            assert np.array_equal(out[q], ref), q
            assert out[q][x] == ref[x]  # the source keeps INF
    finally:
        g.close()


def _heavy_chain(step):
    """A path 0 -> 1 -> ... -> 9 with every weight `step`, a side edge 10 -> 11 and a shortcut
    0 -> 5 of weight 1."""
    n = 12
    s = np.array(list(range(9)) + [10, 0], np.uint32)
    d = np.array(list(range(1, 10)) + [11, 5], np.uint32)
    w = np.array([step] * 9 + [3, 1], np.uint32)
    return n, s, d, w


@pytest.mark.parametrize("step", [31, 50, 100, 127])
def test_batch_8_bit_distances_rerun_the_queries_past_254(step):
    """Weights <= 127 take 8-bit distances; a query with a finite distance past 254 is flagged
    and answered again on wider planes, the others keep their 8-bit answers.  From source x the
    farthest vertex is 9, at (9 - x) * step, or (4 * step + 1) through the shortcut from 0."""
    n, s, d, w = _heavy_chain(step)
    g = eg.Graph.from_edges(s, d, w, n=n)
    try:
        srcs = np.array([0] * 6 + [1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 0, 3, 7], np.uint32)
        out, st = g.run_batch("sssp", srcs)
        assert st["label_bits"] == 8
        far = [(4 * step + 1) if x == 0 else ((9 - x) * step if x <= 9 else 0) for x in srcs]
        assert st["narrow_reruns"] == sum(1 for f in far if f > 254)
        for q, x in enumerate(srcs):
            assert np.array_equal(out[q], O.reference_labels(O.SSSP, n, s, d, w, int(x))), q
        wide, _ = g.run_batch("sssp", srcs, label_bits=32)
        assert np.array_equal(out, wide)
    finally:
        g.close()


def test_batch_8_bit_reruns_on_16_bit_planes():
    """More than 16 flagged queries are answered again together on 16-bit planes: 40 sources on
    the heavy chain (weight 100), 30 of them with a distance past 254."""
    n, s, d, w = _heavy_chain(100)
    g = eg.Graph.from_edges(s, d, w, n=n)
    try:
        srcs = np.array([0] * 14 + [1, 2, 3, 4, 5, 6] * 2 + [7, 8, 9, 10, 11] * 2 + [1, 2, 3, 4], np.uint32)
        assert len(srcs) == 40
        far = [(4 * 100 + 1) if x == 0 else ((9 - x) * 100 if x <= 9 else 0) for x in srcs]
        out, st = g.run_batch("sssp", srcs)
        assert st["label_bits"] == 8
        assert st["narrow_reruns"] == sum(1 for f in far if f > 254) == 30
        for q, x in enumerate(srcs):
            assert np.array_equal(out[q], O.reference_labels(O.SSSP, n, s, d, w, int(x))), q
    finally:
        g.close()


def test_batch_label_bits_must_be_0_or_32(rmat13):
    n, s, d, w, g = rmat13
    with pytest.raises(eg.EEGError):
        g.run_batch("sssp", np.arange(20, dtype=np.uint32), label_bits=16)


@pytest.mark.parametrize("nsrc", [17, 64])
def test_batch_16_bit_distances_heavy_weights(nsrc):
    """Weights above 127 take 16-bit distance planes: RMAT-13 with weights 1..1000, answers equal
    the 32-bit planes' and the oracle's."""
    s, d, w = eg.gen_rmat(13, 16, seed=9, wmin=1, wmax=1000)
    n = 1 << 13
    g = eg.Graph.from_edges(s, d, w, n=n)
    try:
        srcs = np.random.default_rng(nsrc).integers(0, n, nsrc).astype(np.uint32)
        srcs[0] = int(np.argmax(np.bincount(s, minlength=n)))
        out, st = g.run_batch("sssp", srcs)
        assert st["label_bits"] == 16
        wide, _ = g.run_batch("sssp", srcs, label_bits=32)
        assert np.array_equal(out, wide)
        for q in (0, 1, nsrc - 1):
            assert np.array_equal(out[q], O.reference_labels(O.SSSP, n, s, d, w, int(srcs[q]))), q
    finally:
        g.close()
