"""Packed graphs: many independent graphs in one traversal.  Every graph's answer must equal its
own single-query run and the oracle bit for bit."""
import numpy as np
import pytest

import eegraph_amd as eg
import eegraph_amd._lib as L
import oracle_py as O

pytestmark = pytest.mark.gpu

ALG = {"bfs": O.BFS, "sssp": O.SSSP, "sswp": O.SSWP, "cc": O.CC}
VARIANTS = [dict(variant="sync_push_dd"), dict(variant="sync_push_dd", frontier="flags"),
            dict(variant="async_push_dd"), dict(variant="async_push_td")]


@pytest.fixture(scope="module")
def family():
    """12 RMAT graphs of scales 5..11 and edge factors 2..24, plus a graph whose source has
    no out-edges and a graph with isolated vertices."""
    rng = np.random.default_rng(5)
    gs = []
    for i in range(12):
        scale = int(rng.integers(5, 12))
        ef = int(rng.integers(2, 25))
        s, d, w = eg.gen_rmat(scale, ef, seed=100 + i)
        n = 1 << scale
        gs.append((n, s, d, w))
    gs.append((40, np.array([1, 2, 3], np.uint32), np.array([2, 3, 4], np.uint32),
               np.array([5, 6, 7], np.uint32)))
    gs.append((30, np.array([5, 6, 6, 29], np.uint32), np.array([6, 5, 29, 0], np.uint32),
               np.array([1, 2, 3, 4], np.uint32)))
    handles = [eg.Graph.from_edges(s, d, w, n=n) for n, s, d, w in gs]
    srcs = []
    for (n, s, d, w), g in zip(gs, handles):
        srcs.append(g.info()["argmax_out_degree"] if len(srcs) % 2 == 0 else int(rng.integers(0, n)))
    srcs[12] = 0  # no out-edges from vertex 0 in that graph
    yield gs, handles, np.array(srcs, np.uint32)
    for g in handles:
        g.close()


@pytest.mark.parametrize("alg", list(ALG))
@pytest.mark.parametrize("vi", range(len(VARIANTS)))
def test_packed_equals_single_and_oracle(family, alg, vi):
    gs, handles, srcs = family
    outs, st = eg.run_packed(handles, alg, srcs, K=8, **VARIANTS[vi])
    assert st["iterations"] >= 1
    for i, ((n, s, d, w), g) in enumerate(zip(gs, handles)):
        single, _ = g.run(alg, int(srcs[i]), K=8, **VARIANTS[vi])
        assert np.array_equal(outs[i], single), (alg, i)
        ref = O.reference_labels(ALG[alg], n, s, d, w, int(srcs[i]))
        assert np.array_equal(outs[i], ref), (alg, i)


@pytest.mark.parametrize("K", [4, 32])
def test_packed_repeat_and_reorder(family, K):
    gs, handles, srcs = family
    a, _ = eg.run_packed(handles, "sssp", srcs, K=K)
    b, _ = eg.run_packed(handles, "sssp", srcs, K=K)  # cached union
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
    order = list(range(len(handles)))[::-1]
    c, _ = eg.run_packed([handles[i] for i in order], "sssp", srcs[order], K=K)  # new union
    for j, i in enumerate(order):
        assert np.array_equal(c[j], a[i])
    # the same graph twice in one pack
    dup, _ = eg.run_packed([handles[0], handles[0]], "bfs", srcs[[0, 0]])
    single, _ = handles[0].run("bfs", int(srcs[0]))
    assert np.array_equal(dup[0], single) and np.array_equal(dup[1], single)


def test_packed_errors(family):
    gs, handles, srcs = family
    with pytest.raises(eg.EEGError) as e:
        eg.run_packed(handles, "pr", srcs)
    assert e.value.status == L.ERR_UNSUPPORTED
    bad = srcs.copy()
    bad[0] = gs[0][0]
    with pytest.raises(eg.EEGError) as e:
        eg.run_packed(handles, "bfs", bad)
    assert e.value.status == L.ERR_INVALID_ARG
    s = np.array([0, 1], np.uint32)
    gu = eg.Graph.from_edges(s, s[::-1].copy(), None, n=2)
    with pytest.raises(eg.EEGError) as e:
        eg.run_packed([handles[0], gu], "bfs", np.array([0, 0], np.uint32))
    assert e.value.status == L.ERR_INVALID_ARG
    gu.close()
