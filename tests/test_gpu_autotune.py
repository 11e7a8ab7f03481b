"""Autotuner: the tuned options must be valid, reproducible from the shape cache, and give
answers identical to the default options (the integer fixpoints are unique)."""
import numpy as np
import pytest

import eegraph_amd as eg
import oracle_py as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def rmat14():
    s, d, w = eg.gen_rmat(14, 16, seed=1, wmin=1, wmax=64)
    n = 1 << 14
    g = eg.Graph.from_edges(s, d, w, n=n)
    yield n, s, d, w, g
    g.close()


@pytest.mark.parametrize("alg", ["bfs", "sssp", "sswp", "cc"])
def test_autotune_integer(rmat14, alg):
    n, s, d, w, g = rmat14
    src = g.info()["argmax_out_degree"]
    o, ms = g.autotune(alg, [src, 7, 100])
    assert ms > 0 and np.isfinite(ms)
    assert o.K in (4, 8, 16, 32)
    assert o.block in (0, 256, 512, 1024)  # the workgroup size is one of the searched choices
    per = max(o.block, 256) // 256
    assert o.wgs_per_cu in (0,) + tuple(max(1, x // per) for x in (8, 12, 16, 24, 32, 48))
    assert o.variant in (eg.VARIANTS["async_push_td"], eg.VARIANTS["async_push_dd"],
                         eg.VARIANTS["sync_push_dd"])
    tuned, _ = g.run(alg, src, opts=o)
    default, _ = g.run(alg, src, variant="sync_push_dd", K=8)
    assert np.array_equal(tuned, default)
    ref = O.reference_labels({"bfs": O.BFS, "sssp": O.SSSP, "sswp": O.SSWP, "cc": O.CC}[alg],
                             n, s, d, w, src)
    assert np.array_equal(tuned, ref)
    # a second call on the same shape is answered from the cache: same choice
    o2, ms2 = g.autotune(alg, [src, 7, 100])
    assert (o2.variant, o2.frontier, o2.K, o2.block, o2.wgs_per_cu) == (o.variant, o.frontier, o.K, o.block,
                                                                        o.wgs_per_cu)
    assert ms2 > 0


def test_autotune_pagerank_keeps_numeric_settings(rmat14):
    n, s, d, w, g = rmat14
    base = eg.make_opts(damping=0.85, epsilon=1e-7)
    o, ms = g.autotune("pr", opts=base)
    assert abs(o.damping - 0.85) < 1e-7 and o.epsilon == 1e-7
    x, st = g.run("pr", opts=o)
    row, col, _ = O.csr(n, s, d, w)
    ref, _, _ = O.pagerank(n, row, col, 0.85, 1e-13)
    assert np.abs(x.astype(np.float64) - ref).sum() < 1e-6


def test_autotune_rejects_bad_source(rmat14):
    n, s, d, w, g = rmat14
    with pytest.raises(eg.EEGError):
        g.autotune("bfs", [n + 5])
