"""GPU parity at the full sizes of every workload, against the CPU oracle.

Each module fixture generates its graph twice: in HBM through the library and on the host with
the same counter-based generator (host == device is pinned at small scale), and the oracle
runs on the host copy.  Integer labels must match bit for bit (the full-array compare of the
reference's --debug path, src/eegraph.cu:239-257); PageRank within L1 < 1e-6 and max < 1e-7 of
the FP64 fixed point.

  RMAT-14 (EF 16, seed 1) SSSP sync push through the text loader and the driver executable;
  RMAT-22 (EF 16, seed 2): BFS/SSSP/SSWP/CC x 4 variants x K in {4, 8, 16, 32}, flag scans,
      PageRank sync and async;
  4096 x 4096 grid + 1% shortcuts (seed 3, w 1..1000): 4 algorithms, compact queues against
      flag scans, and the async variants;
  RMAT-24 (EF 16, seed 4) x 256 sources: every batched row equals its single-query run, a
      sample of rows equals the oracle;
  64 RMAT graphs (scale 12..20, EF 8..32, seed 5) packed in one traversal, every graph against
      the oracle;
  RMAT-26 (EF 16, seed 6) serving stream: every served answer equals its single-query run, a
      sample equals the oracle (PPR against the FP64 oracle).
"""
import os
import subprocess
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import eegraph_amd as eg
import eegraph_amd._lib as L
import oracle_py as O

pytestmark = [pytest.mark.gpu, pytest.mark.timeout(900)]

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "build", "bin")
VARIANTS = ["async_push_td", "async_push_dd", "sync_push_td", "sync_push_dd"]
PR_L1_TOL, PR_MAX_TOL = 1e-6, 1e-7
POOL = ThreadPoolExecutor(max_workers=16)  # oracle calls release the GIL


def oracle_labels(alg, n, row, col, w, src, s=None, d=None):
    """Textbook answer; bucket queues for the weighted ones (integer weights)."""
    if alg == "bfs":
        return O.bfs(n, row, col, src)
    if alg == "sssp":
        return O.dial(n, row, col, w, src)
    if alg == "sswp":
        return O.dial_widest(n, row, col, w, src)
    return O.wcc(n, s, d)


def _first_diff(a, b):
    bad = np.flatnonzero(a != b)
    return f"{len(bad)} labels differ, first at {bad[0]}: {a[bad[0]]} vs {b[bad[0]]}" if len(bad) else ""


def assert_same(got, ref, what):
    assert got.shape == ref.shape, what
    assert np.array_equal(got, ref), f"{what}: {_first_diff(got, ref)}"


# ------------------------------------------------------------------ RMAT-14 through the loader
def test_rmat14_sssp_sync_push_through_the_driver(tmp_path):
    """The RMAT-14 workload through the text loader, the Tigr build and the sssp executable
    (sync_push_dd, K = 8, source = the maximum out-degree vertex), against Dijkstra."""
    s, d, w = eg.gen_rmat(14, 16, seed=1, wmin=1, wmax=64)
    path = tmp_path / "rmat14.wel"
    eg.save_edgelist(path, 1 << 14, s, d, w)
    n = int(max(s.max(), d.max())) + 1  # id 0 occurs, so the loader keeps ids (src/graph.cpp:102-105)
    assert s.min() == 0 or d.min() == 0
    src = int(np.argmax(np.bincount(s, minlength=n)))
    out = tmp_path / "sssp.txt"
    p = subprocess.run([os.path.join(BIN, "sssp"), "--input", str(path), "--source", str(src), "--variant",
                        "sync_push_dd", "--k", "8", "--output", str(out)], capture_output=True, text=True,
                       timeout=300)
    assert p.returncode == 0, p.stdout + p.stderr
    vals = np.loadtxt(out, dtype=np.uint64)
    assert np.array_equal(vals[:, 0], np.arange(n))
    row, col, wo = O.csr(n, s, d, w)
    assert_same(vals[:, 1].astype(np.uint32), O.dijkstra(n, row, col, wo, src), "rmat14 sssp")


# ------------------------------------------------------------------ RMAT-22
@pytest.fixture(scope="module")
def rmat22():
    g = eg.Graph.rmat(22, 16, seed=2)
    n = g.info()["n"]
    g.prepare(8, L.PREP_TIGR | L.PREP_SYM | L.PREP_TRANSPOSE)
    s, d, w = eg.gen_rmat(22, 16, seed=2, wmin=1, wmax=64, threads=16)
    src = g.info()["argmax_out_degree"]
    assert src == int(np.argmax(np.bincount(s, minlength=n)))
    row, col, wo = O.csr(n, s, d, w)
    futs = {a: POOL.submit(oracle_labels, a, n, row, col, wo, src, s, d) for a in ("bfs", "sssp", "sswp", "cc")}
    refs = {a: f.result() for a, f in futs.items()}
    yield g, n, src, refs, (s, d)
    g.close()


@pytest.mark.parametrize("K", [4, 8, 16, 32])
@pytest.mark.parametrize("variant", VARIANTS)
@pytest.mark.parametrize("alg", ["bfs", "sssp", "sswp", "cc"])
def test_rmat22_all_variants_and_K(rmat22, alg, variant, K):
    g, n, src, refs, _ = rmat22
    lab, st = g.run(alg, src, variant=variant, K=K)
    assert_same(lab, refs[alg], f"rmat22 {alg} {variant} K={K}")


@pytest.mark.parametrize("frontier", ["flags", "sorted"])
@pytest.mark.parametrize("alg", ["bfs", "sssp", "sswp", "cc"])
def test_rmat22_flag_scans(rmat22, alg, frontier):
    g, n, src, refs, _ = rmat22
    lab, _ = g.run(alg, src, variant="sync_push_dd", frontier=frontier, K=8)
    assert_same(lab, refs[alg], f"rmat22 {alg} {frontier}")


@pytest.fixture(scope="module")
def rmat22_pr(rmat22):
    g, n, src, refs, (s, d) = rmat22
    irow, icol, _ = O.csr(n, d, s)
    outdeg = np.bincount(s, minlength=n).astype(np.uint32)
    xf, it, res = O.pagerank_pull(n, irow, icol, outdeg, 0.85, 1e-10, threads=16)
    assert res < 1e-10
    return xf


@pytest.mark.parametrize("K", [4, 8, 16, 32])
@pytest.mark.parametrize("variant", ["sync_push_dd", "async_push_dd"])
def test_rmat22_pagerank_against_fp64(rmat22, rmat22_pr, variant, K):
    g, n, src, refs, _ = rmat22
    x, st = g.run("pr", 0, variant=variant, K=K)
    err = np.abs(x.astype(np.float64) - rmat22_pr)
    assert st["pr_residual"] < 1e-7
    assert err.sum() < PR_L1_TOL and err.max() < PR_MAX_TOL, (variant, K, err.sum(), err.max())


# ------------------------------------------------------------------ 4096^2 grid + 1% shortcuts
@pytest.fixture(scope="module")
def grid4096():
    s, d, w = eg.gen_grid(4096, 4096, 0.01, seed=3, wmin=1, wmax=1000)
    n = 4096 * 4096
    g = eg.Graph.from_edges(s, d, w, n=n)
    g.prepare(8, L.PREP_TIGR | L.PREP_SYM)
    row, col, wo = O.csr(n, s, d, w)
    futs = {a: POOL.submit(oracle_labels, a, n, row, col, wo, 0, s, d) for a in ("bfs", "sssp", "sswp", "cc")}
    refs = {a: f.result() for a, f in futs.items()}
    del row, col, wo, s, d, w
    yield g, refs
    g.close()


GRID_FORMS = {"compact": dict(variant="sync_push_dd", frontier="compact"),
              "sorted": dict(variant="sync_push_dd", frontier="sorted"),
              "flags": dict(variant="sync_push_dd", frontier="flags"),
              "async_push_dd": dict(variant="async_push_dd"),
              "async_push_td": dict(variant="async_push_td")}


@pytest.mark.parametrize("form", list(GRID_FORMS))
@pytest.mark.parametrize("alg", ["bfs", "sssp", "sswp", "cc"])
def test_grid4096_compaction_and_flag_scans(grid4096, alg, form):
    g, refs = grid4096
    lab, _ = g.run(alg, 0, K=8, **GRID_FORMS[form])
    assert_same(lab, refs[alg], f"grid {alg} {form}")


# ------------------------------------------------------------------ RMAT-24 x 256 sources
@pytest.fixture(scope="module")
def rmat24():
    g = eg.Graph.rmat(24, 16, seed=4)
    n = g.info()["n"]
    g.prepare(8, L.PREP_TIGR)
    srcs = np.random.default_rng(4).integers(0, n, 256).astype(np.uint32)
    s, d, w = eg.gen_rmat(24, 16, seed=4, wmin=1, wmax=64, threads=16)
    row, col, wo = O.csr(n, s, d, w)
    del s, d, w
    yield g, n, srcs, (row, col, wo)
    g.close()


@pytest.mark.parametrize("alg", ["bfs", "sssp", "sswp"])
def test_rmat24_256_sources_batched(rmat24, alg):
    """Every row of the batched pass equals its single-query run; rows of 4 sample sources
    (the first sources reaching more than 1000 vertices, or the first ones) equal the oracle."""
    g, n, srcs, (row, col, wo) = rmat24
    lab, st = g.run_batch(alg, srcs)
    assert lab.shape == (256, n)
    reach = []
    for q in range(256):
        single, _ = g.run(alg, int(srcs[q]), variant="sync_push_dd", K=8)
        assert_same(lab[q], single, f"rmat24 batched {alg} row {q}")
        unreached = 0 if alg == "sswp" else L.DIST_INFINITY
        reach.append(int(np.count_nonzero(single != unreached)))
    big = [q for q in range(256) if reach[q] > 1000][:4]
    sample = big + [q for q in range(256) if q not in big][: 4 - len(big)]
    futs = {q: POOL.submit(oracle_labels, alg, n, row, col, wo, int(srcs[q])) for q in sample}
    for q, f in futs.items():
        assert_same(lab[q], f.result(), f"rmat24 batched {alg} row {q} vs oracle")
    assert len(big) > 0


# ------------------------------------------------------------------ 64 packed RMAT graphs
@pytest.fixture(scope="module")
def packed64():
    rng = np.random.default_rng(5)
    shapes = [(int(rng.integers(12, 21)), int(rng.integers(8, 33))) for _ in range(64)]
    gs, hosts = [], []
    for i, (scale, ef) in enumerate(shapes):
        gs.append(eg.Graph.rmat(scale, ef, seed=5000 + i))
        hosts.append(eg.gen_rmat(scale, ef, seed=5000 + i, wmin=1, wmax=64, threads=16))
    srcs = np.array([x.info()["argmax_out_degree"] for x in gs], np.uint32)

    def refs_of(i):
        n = 1 << shapes[i][0]
        s, d, w = hosts[i]
        row, col, wo = O.csr(n, s, d, w)
        return {a: oracle_labels(a, n, row, col, wo, int(srcs[i]), s, d) for a in ("bfs", "sssp", "sswp", "cc")}

    refs = list(POOL.map(refs_of, range(64)))
    del hosts
    yield gs, srcs, refs
    for x in gs:
        x.close()


@pytest.mark.parametrize("alg", ["bfs", "sssp", "sswp", "cc"])
def test_packed_64_graphs_against_oracle(packed64, alg):
    gs, srcs, refs = packed64
    outs, st = eg.run_packed(gs, alg, srcs, K=8)
    for i in range(64):
        assert_same(outs[i], refs[i][alg], f"packed graph {i} {alg}")


# ------------------------------------------------------------------ RMAT-26 serving stream
@pytest.fixture(scope="module")
def rmat26():
    g = eg.Graph.rmat(26, 16, seed=6)
    n = g.info()["n"]
    g.prepare(8, L.PREP_TIGR | L.PREP_TRANSPOSE)
    yield g, n
    g.close()


def test_rmat26_served_answers(rmat26):
    """A 40/30/20/10 BFS/SSSP/SSWP/PPR stream (sources ~ U[0, N), seed 6, plus the maximum
    out-degree vertex for each algorithm) through the default server: every answer equals its
    single-query run; the answers from the maximum out-degree vertex equal the oracle (BFS,
    SSSP, SSWP bit for bit, PPR within L1 1e-6 / max 1e-7 of FP64)."""
    g, n = rmat26
    hub = g.info()["argmax_out_degree"]
    rng = np.random.default_rng(6)
    mix = list(rng.choice(["bfs", "sssp", "sswp", "ppr"], size=24, p=[0.4, 0.3, 0.2, 0.1]))
    srcs = [int(x) for x in rng.integers(0, n, 24)]
    mix += ["bfs", "sssp", "sswp", "ppr"]
    srcs += [hub] * 4
    answers = []
    with eg.Server(g) as srv:
        tk = [srv.submit(a, x) for a, x in zip(mix, srcs)]
        for t, a in zip(tk, mix):
            lab, r = srv.wait(t, algo=a)
            assert r["status"] == L.OK
            answers.append(lab)
    for a, x, lab in zip(mix, srcs, answers):
        single, _ = g.run(a, x, variant="sync_push_dd", K=8)
        if a == "ppr":
            assert np.abs(lab.astype(np.float64) - single.astype(np.float64)).sum() < PR_L1_TOL
        else:
            assert_same(lab, single, f"served {a} from {x}")
    # the oracle on the host copy of the graph
    s, d, w = eg.gen_rmat(26, 16, seed=6, wmin=1, wmax=64, threads=16)
    row, col, wo = O.csr(n, s, d, w)
    futs = {a: POOL.submit(oracle_labels, a, n, row, col, wo, hub) for a in ("bfs", "sssp", "sswp")}
    outdeg = np.diff(row).astype(np.uint32)
    del row, col, wo, w
    irow, icol, _ = O.csr(n, d, s)
    del s, d
    xf, _, res = O.pagerank_pull(n, irow, icol, outdeg, 0.85, 1e-10, threads=16, personalised=True, source=hub)
    for a, f in futs.items():
        assert_same(answers[mix.index(a, 24)], f.result(), f"served {a} from the hub vs oracle")
    err = np.abs(answers[27].astype(np.float64) - xf)
    assert err.sum() < PR_L1_TOL and err.max() < PR_MAX_TOL, (err.sum(), err.max())
