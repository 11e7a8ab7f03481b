"""The allocation-free entry point (eeg_run_device_noalloc) against eeg_run_device, through the
C-ABI: identical labels on two calls in a row, the second on a smaller graph, and a clear
rejection of graphs above its vertex limit."""
import ctypes as C

import numpy as np
import pytest

import eegraph_amd as eg
import eegraph_amd._lib as L

pytestmark = pytest.mark.gpu

CASES = [(alg, var) for alg in ("bfs", "sssp", "sswp", "cc") for var in ("sync_push_dd", "async_push_td")]

# Device buffers of the test come from the HIP runtime libeegraph.so is linked against (torch
# ships its own runtime, which must not be mixed with it in one process).
_hip = C.CDLL("libamdhip64.so.7")
_hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
_hip.hipFree.argtypes = [C.c_void_p]
_hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
_hip.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
D2H = 2


class DevBuf:
    def __init__(self, n, fill):
        self.n = n
        self.p = C.c_void_p()
        assert _hip.hipMalloc(C.byref(self.p), n * 4) == 0
        assert _hip.hipMemset(self.p, fill, n * 4) == 0

    def host(self):
        out = np.empty(self.n, np.uint32)
        assert _hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), self.p, self.n * 4, D2H) == 0
        return out

    def free(self):
        _hip.hipFree(self.p)


def _call(fn, g, alg, src, variant, out):
    o = eg.make_opts(variant=variant, K=8)
    st = L.Stats()
    rc = fn(g.handle, L.ALGOS[alg], int(src), C.byref(o), out.p, None, C.byref(st))
    L.check(rc, fn.__name__)
    return out.host(), st


def test_noalloc_equals_run_device_on_two_calls_second_smaller():
    for scale, seed in ((17, 21), (12, 22)):  # the second call on a smaller graph
        g = eg.Graph.rmat(scale, 16, seed=seed)
        g.prepare(8, L.PREP_TIGR | L.PREP_SYM)
        n = g.info()["n"]
        src = g.info()["argmax_out_degree"]
        a, b = DevBuf(n, 7), DevBuf(n, 9)
        for alg, var in CASES:
            ref, _ = _call(L.lib.eeg_run_device, g, alg, src, var, a)
            got, st = _call(L.lib.eeg_run_device_noalloc, g, alg, src, var, b)
            assert np.array_equal(got, ref), (scale, alg, var)
            assert st.iterations > 0
        a.free()
        b.free()
        g.close()


def test_noalloc_rejects_graphs_above_its_limit_and_unsupported_variants():
    nmax = 2097152
    s = np.array([0], np.uint32)
    d = np.array([1], np.uint32)
    g = eg.Graph.from_edges(s, d, None, n=nmax + 1)
    out = DevBuf(nmax + 1, 0)
    o = eg.make_opts(variant="sync_push_dd", K=8)
    rc = L.lib.eeg_run_device_noalloc(g.handle, L.BFS, 0, C.byref(o), out.p, None, None)
    assert rc == L.ERR_INVALID_ARG and b"at most" in L.lib.eeg_last_error()
    out.free()
    g.close()
    g = eg.Graph.rmat(10, 8, seed=3)
    out = DevBuf(1024, 0)
    o = eg.make_opts(variant="sync_push_dd", K=8)
    rc = L.lib.eeg_run_device_noalloc(g.handle, L.BFS, 0, C.byref(o), out.p, None, None)
    assert rc == L.ERR_INVALID_ARG and b"eeg_graph_prepare" in L.lib.eeg_last_error()  # tables not built
    g.prepare(8, L.PREP_TIGR)
    o = eg.make_opts(variant="async_push_dd", K=8)
    rc = L.lib.eeg_run_device_noalloc(g.handle, L.BFS, 0, C.byref(o), out.p, None, None)
    assert rc == L.ERR_UNSUPPORTED
    o = eg.make_opts(variant="sync_push_dd", frontier="flags", K=8)  # flags per vertex
    rc = L.lib.eeg_run_device_noalloc(g.handle, L.BFS, 0, C.byref(o), out.p, None, None)
    assert rc == L.ERR_UNSUPPORTED
    out.free()
    g.close()
