"""Python view of the C-ABI: device graphs, queries, generators and edge-list files.

Host arrays are numpy; results come back as numpy arrays (u32 labels or f32 ranks).
This layer only marshals arguments: every computation runs in libeegraph.so.
"""
import ctypes as C

import numpy as np

from . import _lib as L


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _u32(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.uint32)


def device_count():
    c = C.c_int(0)
    L.check(L.lib.eeg_device_count(C.byref(c)), "eeg_device_count")
    return c.value


def gen_rmat(scale, edge_factor, a=0.57, b=0.19, c=0.19, seed=1, wmin=1, wmax=64,
             weights=True, threads=0):
    """RMAT edge list (src, dst, w) of edge_factor << scale edges, counter-based."""
    m = edge_factor << scale
    src = np.empty(m, np.uint32)
    dst = np.empty(m, np.uint32)
    w = np.empty(m, np.uint32) if weights else None
    L.check(L.lib.eeg_gen_rmat(scale, edge_factor, a, b, c, seed, wmin, wmax, _ptr(src),
                               _ptr(dst), _ptr(w), threads), "eeg_gen_rmat")
    return src, dst, w


def gen_grid(width, height, shortcut_frac=0.01, seed=3, wmin=1, wmax=1000, weights=True,
             threads=0):
    """4-neighbour grid (both directions) plus random shortcut pairs."""
    m = C.c_uint64(0)
    L.check(L.lib.eeg_gen_grid_size(width, height, shortcut_frac, C.byref(m)), "eeg_gen_grid_size")
    src = np.empty(m.value, np.uint32)
    dst = np.empty(m.value, np.uint32)
    w = np.empty(m.value, np.uint32) if weights else None
    L.check(L.lib.eeg_gen_grid(width, height, shortcut_frac, seed, wmin, wmax, _ptr(src),
                               _ptr(dst), _ptr(w), threads), "eeg_gen_grid")
    return src, dst, w


def load_edgelist(path, weighted=True):
    """Reads a text (.el/.wel/.txt/.edges/.edgelist) or Subway binary (.bcsr/.bwcsr) file.
    Returns (n, src, dst, w, has_zero_id) with ids already 0-based."""
    el = L.EdgeList()
    L.check(L.lib.eeg_load_edgelist(str(path).encode(), 1 if weighted else 0, C.byref(el)),
            "eeg_load_edgelist")
    try:
        m = el.m
        src = np.ctypeslib.as_array(el.src, shape=(m,)).copy() if m else np.empty(0, np.uint32)
        dst = np.ctypeslib.as_array(el.dst, shape=(m,)).copy() if m else np.empty(0, np.uint32)
        w = None
        if bool(el.w):
            w = np.ctypeslib.as_array(el.w, shape=(m,)).copy() if m else np.empty(0, np.uint32)
        return el.n, src, dst, w, bool(el.has_zero_id)
    finally:
        L.lib.eeg_edgelist_free(C.byref(el))


def save_edgelist(path, n, src, dst, w=None):
    src, dst, w = _u32(src), _u32(dst), _u32(w)
    el = L.EdgeList(n=n, has_zero_id=1, m=len(src),
                    src=src.ctypes.data_as(C.POINTER(C.c_uint32)),
                    dst=dst.ctypes.data_as(C.POINTER(C.c_uint32)),
                    w=w.ctypes.data_as(C.POINTER(C.c_uint32)) if w is not None else None)
    L.check(L.lib.eeg_save_edgelist(str(path).encode(), C.byref(el)), "eeg_save_edgelist")


def make_opts(variant="sync_push_dd", K=8, frontier="sorted", max_iters=0, timing=False,
              count_work=False, collect_activity=False, damping=0.85, epsilon=0.0, acc=0.0,
              wgs_per_cu=0, energy=False, block=0, label_bits=0):
    o = L.Opts()
    o.variant = L.VARIANTS[variant] if isinstance(variant, str) else int(variant)
    o.K = K
    o.block = block
    o.frontier = {"compact": L.FRONTIER_COMPACT, "flags": L.FRONTIER_FLAGS,
                  "sorted": L.FRONTIER_SORTED}[frontier] \
        if isinstance(frontier, str) else int(frontier)
    o.max_iters = max_iters
    o.timing = 1 if timing else 0
    o.count_work = 1 if count_work else 0
    o.collect_activity = 1 if collect_activity else 0
    o.damping = damping
    o.epsilon = epsilon
    o.acc = acc
    o.wgs_per_cu = wgs_per_cu
    o.energy = 1 if energy else 0
    o.label_bits = label_bits
    return o


class Graph:
    """A graph resident in HBM (out-edge CSR plus lazily built Tigr / symmetrised /
    in-edge tables).  Mirrors the reference's Graph + VirtualGraph pair, built once and
    reused by every query instead of per eegraph_<alg> call."""

    def __init__(self, handle):
        self._h = C.c_void_p(handle)

    @classmethod
    def from_edges(cls, src, dst, w=None, n=None, device=0):
        src, dst, w = _u32(src), _u32(dst), _u32(w)
        if n is None:
            n = int(max(src.max(initial=0), dst.max(initial=0))) + 1 if len(src) else 0
        h = C.c_void_p()
        L.check(L.lib.eeg_graph_create(_ptr(src), _ptr(dst), _ptr(w), len(src), n, device,
                                       C.byref(h)), "eeg_graph_create")
        return cls(h.value)

    @classmethod
    def from_device_edges(cls, d_src, d_dst, d_w, m, n, device=0):
        """d_* are device addresses (e.g. torch tensor data_ptr())."""
        h = C.c_void_p()
        L.check(L.lib.eeg_graph_create_device(C.c_void_p(d_src), C.c_void_p(d_dst),
                                              C.c_void_p(d_w) if d_w else None, m, n, device,
                                              C.byref(h)), "eeg_graph_create_device")
        return cls(h.value)

    @classmethod
    def rmat(cls, scale, edge_factor, a=0.57, b=0.19, c=0.19, seed=1, wmin=1, wmax=64,
             device=0):
        """RMAT generated directly in HBM (same edges as gen_rmat)."""
        h = C.c_void_p()
        L.check(L.lib.eeg_graph_create_rmat(scale, edge_factor, a, b, c, seed, wmin, wmax,
                                            device, C.byref(h)), "eeg_graph_create_rmat")
        return cls(h.value)

    def close(self):
        if self._h and self._h.value:
            L.lib.eeg_graph_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def handle(self):
        return self._h

    def info(self):
        i = L.GraphInfo()
        L.check(L.lib.eeg_graph_get_info(self._h, C.byref(i)), "eeg_graph_get_info")
        return {k: getattr(i, k) for k, _ in i._fields_}

    def prepare(self, K=8, what=L.PREP_TIGR):
        L.check(L.lib.eeg_graph_prepare(self._h, K, what), "eeg_graph_prepare")

    def tigr_export(self, K=8):
        info = self.info()
        n, m = info["n"], info["m"]
        P = C.c_uint64(0)
        L.check(L.lib.eeg_tigr_export(self._h, K, None, None, None, C.byref(P)), "eeg_tigr_export")
        np_ = np.empty(n, np.uint64)
        el = np.empty(2 * m + n, np.uint32)
        parts = np.empty(2 * P.value, np.uint32)
        L.check(L.lib.eeg_tigr_export(self._h, K, _ptr(np_), _ptr(el), _ptr(parts), C.byref(P)),
                "eeg_tigr_export")
        return np_, el, parts.reshape(-1, 2)

    def csr_export(self):
        info = self.info()
        row = np.empty(info["n"] + 1, np.uint64)
        col = np.empty(info["m"], np.uint32)
        w = np.empty(info["m"], np.uint32) if info["weighted"] else None
        L.check(L.lib.eeg_csr_export(self._h, _ptr(row), _ptr(col), _ptr(w)), "eeg_csr_export")
        return row, col, w

    def csr_export_derived(self, which):
        """which: "sym" (both directions, used by CC) or "in" (in-edges, used by PageRank)."""
        code = {"sym": L.PREP_SYM, "in": L.PREP_TRANSPOSE}[which]
        m = C.c_uint64(0)
        L.check(L.lib.eeg_csr_export_derived(self._h, code, None, None, C.byref(m)),
                "eeg_csr_export_derived")
        row = np.empty(self.info()["n"] + 1, np.uint64)
        col = np.empty(m.value, np.uint32)
        L.check(L.lib.eeg_csr_export_derived(self._h, code, _ptr(row), _ptr(col), C.byref(m)),
                "eeg_csr_export_derived")
        return row, col

    def run(self, algo, source=0, opts=None, want_labels=True, activity_cap=0, **kw):
        """One query; returns (labels or None, stats dict)."""
        a = L.ALGOS[algo] if isinstance(algo, str) else int(algo)
        o = opts if opts is not None else make_opts(**kw)
        n = self.info()["n"]
        out = None
        if want_labels:
            out = np.empty(n, np.float32 if a in (L.PR, L.PPR) else np.uint32)
        st = L.Stats()
        act = None
        if activity_cap:  # per-iteration activity (the reference's --activity) into act
            act = np.zeros(activity_cap, np.uint32)
            st.activity = act.ctypes.data_as(C.POINTER(C.c_uint32))
            st.activity_cap = activity_cap
            o.collect_activity = 1
        rc = L.lib.eeg_run(self._h, a, source, C.byref(o), _ptr(out), C.byref(st))
        L.check(rc, f"eeg_run({algo})")
        d = st.to_dict()
        if act is not None:
            d["activity"] = act[: st.activity_len].copy()
        return out, d

    def autotune(self, algo, sources=None, opts=None, **kw):
        """Times short runs and returns (tuned Opts, median traversal ms)."""
        a = L.ALGOS[algo] if isinstance(algo, str) else int(algo)
        o = opts if opts is not None else make_opts(**kw)
        srcs = _u32(np.asarray(sources if sources is not None else [], dtype=np.uint32))
        ms = C.c_double(0)
        L.check(L.lib.eeg_autotune(self._h, a, _ptr(srcs) if len(srcs) else None, len(srcs),
                                   C.byref(o), C.byref(ms)), f"eeg_autotune({algo})")
        return o, ms.value

    def run_batch(self, algo, sources, opts=None, want_labels=True, **kw):
        """Many sources in one traversal; returns (labels[nsrc, n] or None, stats dict)."""
        a = L.ALGOS[algo] if isinstance(algo, str) else int(algo)
        o = opts if opts is not None else make_opts(**kw)
        srcs = _u32(np.asarray(sources))
        n = self.info()["n"]
        out = (np.empty((len(srcs), n), np.float32 if a == L.PPR else np.uint32)
               if want_labels else None)
        st = L.Stats()
        L.check(L.lib.eeg_run_batch(self._h, a, _ptr(srcs), len(srcs), C.byref(o), _ptr(out),
                                    C.byref(st)), f"eeg_run_batch({algo})")
        return out, st.to_dict()


def run_packed(graphs, algo, sources, opts=None, want_labels=True, **kw):
    """Several graphs in one launch sequence; returns (list of label arrays or None, stats)."""
    a = L.ALGOS[algo] if isinstance(algo, str) else int(algo)
    o = opts if opts is not None else make_opts(**kw)
    hs = (C.c_void_p * len(graphs))(*[g.handle.value for g in graphs])
    srcs = _u32(np.asarray(sources))
    outs = [np.empty(g.info()["n"], np.uint32) for g in graphs] if want_labels else None
    ptrs = (C.c_void_p * len(graphs))(*[o_.ctypes.data for o_ in outs]) if want_labels else None
    st = L.Stats()
    L.check(L.lib.eeg_run_packed(hs, len(graphs), a, _ptr(srcs), C.byref(o), ptrs, C.byref(st)),
            "eeg_run_packed")
    return outs, st.to_dict()


class Server:
    """Continuous-batching request loop over one graph (eeg_server_*)."""

    def __init__(self, graph, label_budget_bytes=0, max_batch=0, K=0, damping=0.0,
                 ppr_epsilon=0.0, discard_labels=False, lane_wgs_per_cu=0, lanes=0,
                 plane_engines=False, bfs_engine=True, first_iteration=0, ppr_engine=True,
                 ppr_columns=0, lane_batch=True):
        cfg = L.ServerCfg()
        cfg.label_budget_bytes = label_budget_bytes
        cfg.max_batch = max_batch
        cfg.K = K
        cfg.damping = damping
        cfg.ppr_epsilon = ppr_epsilon
        cfg.flags = ((L.SERVER_DISCARD_LABELS if discard_labels else 0) |
                     (L.SERVER_PLANE_ENGINES if plane_engines else 0) |
                     (0 if bfs_engine else L.SERVER_NO_BFS_ENGINE) |
                     (0 if ppr_engine else L.SERVER_NO_PPR_ENGINE) |
                     (0 if lane_batch else L.SERVER_NO_LANE_BATCH))
        cfg.lane_wgs_per_cu = lane_wgs_per_cu
        cfg.lanes = lanes
        cfg.first_iteration = first_iteration
        cfg.ppr_columns = ppr_columns
        self.graph = graph  # keeps the graph alive while the server runs
        self.n = graph.info()["n"]
        self.discard = discard_labels
        h = C.c_void_p()
        L.check(L.lib.eeg_server_create(graph.handle, C.byref(cfg), C.byref(h)), "eeg_server_create")
        self._h = h

    def submit(self, algo, source=0):
        a = L.ALGOS[algo] if isinstance(algo, str) else int(algo)
        t = C.c_uint64(0)
        L.check(L.lib.eeg_server_submit(self._h, a, int(source), C.byref(t)), "eeg_server_submit")
        return t.value

    def wait(self, ticket, want_labels=True, timeout_ms=-1.0, algo=None):
        """Returns (labels or None, result dict)."""
        out = None
        if want_labels and not self.discard:
            out = np.empty(self.n, np.float32 if algo in ("pr", "ppr", L.PR, L.PPR) else np.uint32)
        r = L.QueryResult()
        rc = L.lib.eeg_server_wait(self._h, ticket, _ptr(out), timeout_ms, C.byref(r))
        L.check(rc, "eeg_server_wait")
        return out, {"algo": r.algo, "source": r.source, "batch_size": r.batch_size,
                     "status": r.status, "t_submit_ms": r.t_submit_ms,
                     "t_start_ms": r.t_start_ms, "t_done_ms": r.t_done_ms}

    def cancel(self, ticket):
        L.check(L.lib.eeg_server_cancel(self._h, ticket), "eeg_server_cancel")

    def close(self):
        if self._h is not None and self._h.value:
            L.check(L.lib.eeg_server_destroy(self._h), "eeg_server_destroy")
        self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def energy_read(device=0):
    """Device energy accumulator in microjoules, or None if amd-smi is unavailable."""
    r = energy_read_ts(device)
    return None if r is None else r[0]


def energy_read_ts(device=0):
    """(microjoules, counter timestamp in ns) of the device energy accumulator, or None.  The
    counter advances in steps (about once per second on MI355X); the timestamp is that of the
    last step."""
    uj = C.c_double(0)
    ts = C.c_uint64(0)
    rc = L.lib.eeg_energy_read(device, C.byref(uj), C.byref(ts))
    return (uj.value, ts.value) if rc == L.OK else None
