"""Launch geometry does not change answers, on a graph large enough that every workgroup count
tried is a different grid: RMAT-20 has more traversal tiles (256 virtual nodes each at K = 8)
and more PageRank tiles (512 virtual nodes each) than 256 CUs x the largest workgroups-per-CU
value used below, so grid = CUs x wgs_per_cu in every case (on a 256-CU MI355X)."""
import numpy as np
import pytest

import eegraph_amd as eg
import oracle_py as O

pytestmark = pytest.mark.gpu
CUS = 256


@pytest.fixture(scope="module")
def rmat20():
    s, d, w = eg.gen_rmat(20, 16, seed=8, wmin=1, wmax=64, threads=16)
    n = 1 << 20
    g = eg.Graph.from_edges(s, d, w, n=n)
    src = g.info()["argmax_out_degree"]
    yield n, s, d, w, g, src
    g.close()


def _tiles(deg, K, tv):
    return int(np.sum((deg + K - 1) // K) + tv - 1) // tv


@pytest.mark.parametrize("alg", ["bfs", "sssp", "sswp", "cc"])
def test_static_sweep_grids_do_not_change_labels(rmat20, alg):
    n, s, d, w, g, src = rmat20
    deg = np.bincount(s, minlength=n) if alg != "cc" else np.bincount(np.concatenate([s, d]), minlength=n)
    wpcs = (1, 4, 16)
    assert _tiles(deg, 8, 2048 // 8) > CUS * max(wpcs)  # every grid below is CUs x wpc, all different
    ref = O.reference_labels({"bfs": O.BFS, "sssp": O.SSSP, "sswp": O.SSWP, "cc": O.CC}[alg], n, s, d, w, src)
    for variant, frontier in (("sync_push_dd", "flags"), ("async_push_dd", "compact"),
                              ("async_push_td", "compact"), ("sync_push_dd", "compact")):
        for wpc in wpcs:
            lab, _ = g.run(alg, src, variant=variant, frontier=frontier, K=8, wgs_per_cu=wpc)
            assert np.array_equal(lab, ref), f"{alg} {variant} {frontier} wgs_per_cu={wpc}"


def test_pagerank_grids_do_not_change_bits(rmat20):
    n, s, d, w, g, src = rmat20
    wpcs = (1, 2, 8)
    assert _tiles(np.bincount(d, minlength=n), 8, 4096 // 8) > CUS * max(wpcs)
    ref, st0 = g.run("pr", 0, variant="sync_push_dd", K=8, wgs_per_cu=wpcs[0])
    for wpc in wpcs[1:]:
        x, st = g.run("pr", 0, variant="sync_push_dd", K=8, wgs_per_cu=wpc)
        assert st["iterations"] == st0["iterations"]
        assert np.array_equal(x.view(np.uint32), ref.view(np.uint32)), f"wgs_per_cu={wpc}"
    row, col, _ = O.csr(n, s, d)
    irow, icol, _ = O.csr(n, d, s)
    xf, _, _ = O.pagerank_pull(n, irow, icol, np.diff(row).astype(np.uint32), 0.85, 1e-10, threads=16)
    err = np.abs(ref.astype(np.float64) - xf)
    assert err.sum() < 1e-6 and err.max() < 1e-7
    for wpc in wpcs:  # in place: within the tolerance for every grid
        xa, _ = g.run("pr", 0, variant="async_push_dd", K=8, wgs_per_cu=wpc)
        assert np.abs(xa.astype(np.float64) - xf).sum() < 1e-6, wpc
