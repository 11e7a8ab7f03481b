"""ctypes wrapper of the CPU oracle (oracle/build/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's CPU
baseline leg.  Nothing in eegraph_amd/ imports it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "build", "liboracle.so")

BFS, SSSP, SSWP, CC = 0, 1, 2, 3
INF = 4294967294


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE])


if not os.path.exists(LIB_PATH):
    build()
_lib = C.CDLL(LIB_PATH)
_vp = C.c_void_p


def _sig(name, res, *args):
    f = getattr(_lib, name)
    f.restype = res
    f.argtypes = list(args)


_sig("orc_csr_build", None, C.c_uint32, C.c_uint64, _vp, _vp, _vp, _vp, _vp, _vp)
_sig("orc_tigr_layout", C.c_uint64, C.c_uint32, C.c_uint64, _vp, _vp, _vp, C.c_uint32, _vp, _vp, _vp)
_sig("orc_seq_sweep", C.c_uint32, C.c_int, C.c_uint32, C.c_uint64, _vp, _vp, _vp, C.c_uint32, _vp,
     C.c_double, C.POINTER(C.c_int))
_sig("orc_init_labels", None, C.c_int, C.c_uint32, C.c_uint32, _vp)
_sig("orc_bfs", None, C.c_uint32, _vp, _vp, C.c_uint32, _vp)
_sig("orc_dijkstra", None, C.c_uint32, _vp, _vp, _vp, C.c_uint32, _vp)
_sig("orc_widest", None, C.c_uint32, _vp, _vp, _vp, C.c_uint32, _vp)
_sig("orc_wcc", None, C.c_uint32, C.c_uint64, _vp, _vp, _vp)
_sig("orc_brute", C.c_int, C.c_int, C.c_uint32, C.c_uint64, _vp, _vp, _vp, C.c_uint32, _vp)
_sig("orc_pagerank", C.c_uint32, C.c_uint32, _vp, _vp, C.c_double, C.c_double, C.c_uint32, C.c_int,
     C.c_uint32, _vp, C.POINTER(C.c_double))
_sig("orc_pagerank_pull", C.c_uint32, C.c_uint32, _vp, _vp, _vp, C.c_double, C.c_double, C.c_uint32,
     C.c_int, C.c_int, C.c_uint32, _vp, C.POINTER(C.c_double))
_sig("orc_dial", C.c_int, C.c_uint32, _vp, _vp, _vp, C.c_uint32, C.c_uint32, _vp)
_sig("orc_dial_widest", C.c_int, C.c_uint32, _vp, _vp, _vp, C.c_uint32, C.c_uint32, _vp)
_sig("orc_pr_delta_push", C.c_uint32, C.c_uint32, _vp, _vp, C.c_float, C.c_float, _vp, C.c_double,
     C.POINTER(C.c_int))


def _p(a):
    return None if a is None else a.ctypes.data_as(_vp)


def _u32(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.uint32)


def csr(n, src, dst, w=None):
    src, dst, w = _u32(src), _u32(dst), _u32(w)
    m = len(src)
    row = np.empty(n + 1, np.uint64)
    col = np.empty(m, np.uint32)
    wo = np.empty(m, np.uint32) if w is not None else None
    _lib.orc_csr_build(n, m, _p(src), _p(dst), _p(w), _p(row), _p(col), _p(wo))
    return row, col, wo


def tigr_layout(n, src, dst, w, K):
    src, dst, w = _u32(src), _u32(dst), _u32(w)
    m = len(src)
    P = _lib.orc_tigr_layout(n, m, _p(src), _p(dst), _p(w), K, None, None, None)
    np_ = np.empty(n, np.uint64)
    el = np.empty(2 * m + n, np.uint32)
    parts = np.empty(2 * P, np.uint32)
    _lib.orc_tigr_layout(n, m, _p(src), _p(dst), _p(w), K, _p(np_), _p(el), _p(parts))
    return np_, el, parts.reshape(-1, 2)


def init_labels(alg, n, source):
    lab = np.empty(n, np.uint32)
    _lib.orc_init_labels(alg, n, source, _p(lab))
    return lab


def seq_sweep(alg, n, layout, K, labels, time_cap_s=0.0):
    """Runs the restated seq_cpu sweep in place; returns (passes, converged)."""
    np_, el, parts = layout
    conv = C.c_int(0)
    passes = _lib.orc_seq_sweep(alg, n, len(parts), _p(np_), _p(el),
                                _p(np.ascontiguousarray(parts)), K, _p(labels), time_cap_s,
                                C.byref(conv))
    return passes, bool(conv.value)


def bfs(n, row, col, source):
    out = np.empty(n, np.uint32)
    _lib.orc_bfs(n, _p(row), _p(col), source, _p(out))
    return out


def dijkstra(n, row, col, w, source):
    out = np.empty(n, np.uint32)
    _lib.orc_dijkstra(n, _p(row), _p(col), _p(w), source, _p(out))
    return out


def widest(n, row, col, w, source):
    out = np.empty(n, np.uint32)
    _lib.orc_widest(n, _p(row), _p(col), _p(w), source, _p(out))
    return out


def wcc(n, src, dst):
    src, dst = _u32(src), _u32(dst)
    out = np.empty(n, np.uint32)
    _lib.orc_wcc(n, len(src), _p(src), _p(dst), _p(out))
    return out


def brute(alg, n, src, dst, w, source):
    src, dst, w = _u32(src), _u32(dst), _u32(w)
    out = np.empty(n, np.uint32)
    rc = _lib.orc_brute(alg, n, len(src), _p(src), _p(dst), _p(w), source, _p(out))
    if rc != 0:
        raise ValueError("brute force handles n <= 64")
    return out


def pagerank(n, row, col, d=0.85, eps=1e-12, max_iter=10000, personalised=False, source=0):
    x = np.empty(n, np.float64)
    res = C.c_double(0)
    it = _lib.orc_pagerank(n, _p(row), _p(col), d, eps, max_iter, 1 if personalised else 0,
                           source, _p(x), C.byref(res))
    return x, it, res.value


def pagerank_pull(n, in_row, in_col, outdeg, d=0.85, eps=1e-12, max_iter=10000, threads=0,
                  personalised=False, source=0):
    """FP64 Jacobi PageRank over the in-edge CSR, OpenMP over destinations."""
    x = np.empty(n, np.float64)
    res = C.c_double(0)
    it = _lib.orc_pagerank_pull(n, _p(in_row), _p(in_col), _p(outdeg), d, eps, max_iter, threads,
                                1 if personalised else 0, source, _p(x), C.byref(res))
    return x, it, res.value


def dial(n, row, col, w, source, wmax=None):
    """Dijkstra's answer with Dial's bucket queue (integer weights <= wmax)."""
    if wmax is None:
        wmax = int(w.max()) if w is not None and len(w) else 1
    out = np.empty(n, np.uint32)
    if _lib.orc_dial(n, _p(row), _p(col), _p(w), wmax, source, _p(out)) != 0:
        raise ValueError("a weight exceeds wmax")
    return out


def dial_widest(n, row, col, w, source, wmax=None):
    """Max-min Dijkstra's answer with a bucket per label value (integer weights <= wmax)."""
    if wmax is None:
        wmax = int(w.max()) if w is not None and len(w) else 1
    out = np.empty(n, np.uint32)
    if _lib.orc_dial_widest(n, _p(row), _p(col), _p(w), wmax, source, _p(out)) != 0:
        raise ValueError("a weight exceeds wmax")
    return out


def pr_delta_push(n, row, col, d=0.85, thr=1e-10, time_cap_s=0.0):
    x = np.empty(n, np.float32)
    conv = C.c_int(0)
    passes = _lib.orc_pr_delta_push(n, _p(row), _p(col), d, thr, _p(x), time_cap_s, C.byref(conv))
    return x, passes, bool(conv.value)


def reference_labels(alg, n, src, dst, w, source):
    """Textbook answer for BFS/SSSP/SSWP/CC on an edge list."""
    if alg == CC:
        return wcc(n, src, dst)
    row, col, wo = csr(n, src, dst, w)
    if alg == BFS:
        return bfs(n, row, col, source)
    if alg == SSSP:
        return dijkstra(n, row, col, wo, source)
    return widest(n, row, col, wo, source)
