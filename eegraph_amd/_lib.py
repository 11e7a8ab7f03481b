"""ctypes bindings for libeegraph.so (the C-ABI declared in include/eeg_capi.h).

The shared library is built in-tree by ``make -C eegraph_amd/csrc`` (``__graft_entry__.build``)
into ``eegraph_amd/lib/libeegraph.so``.  There is no fallback: if the library is missing,
importing the package raises, so a missing native build can never be masked.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libeegraph.so")

if not os.path.exists(LIB_PATH):
    raise ImportError(
        f"libeegraph.so not found at {LIB_PATH}; build it with `make -C eegraph_amd/csrc`")

lib = C.CDLL(LIB_PATH)

u32 = C.c_uint32
u64 = C.c_uint64
P = C.POINTER

# status codes (eeg_status)
OK = 0
ERR_INVALID_ARG = 1
ERR_OOM = 2
ERR_DEVICE = 3
ERR_NOT_CONVERGED = 4
ERR_REJECTED = 5
ERR_NO_DEVICE = 6
ERR_IO = 7
ERR_UNSUPPORTED = 8
ERR_CANCELLED = 9
ERR_TIMEOUT = 10

# algorithms (eeg_algo)
BFS, SSSP, SSWP, CC, PR, PPR = range(6)
ALGOS = {"bfs": BFS, "sssp": SSSP, "sswp": SSWP, "cc": CC, "pr": PR, "ppr": PPR}

# variants (eeg_variant), same order as the reference enum Variant
ASYNC_PUSH_TD, ASYNC_PUSH_DD, SYNC_PUSH_TD, SYNC_PUSH_DD = range(4)
VARIANTS = {"async_push_td": ASYNC_PUSH_TD, "async_push_dd": ASYNC_PUSH_DD,
            "sync_push_td": SYNC_PUSH_TD, "sync_push_dd": SYNC_PUSH_DD}
FRONTIER_SORTED, FRONTIER_FLAGS, FRONTIER_COMPACT = 0, 1, 2

PREP_TIGR, PREP_SYM, PREP_TRANSPOSE = 1, 2, 4
DIST_INFINITY = 4294967294


class Opts(C.Structure):
    _fields_ = [("variant", u32), ("K", u32), ("block", u32), ("frontier", u32),
                ("max_iters", u32), ("collect_activity", u32), ("timing", u32),
                ("count_work", u32), ("damping", C.c_float), ("acc", C.c_float),
                ("epsilon", C.c_double), ("wgs_per_cu", u32), ("energy", u32),
                ("label_bits", u32), ("reserved", u32 * 5)]


class Stats(C.Structure):
    _fields_ = [("iterations", u32), ("launches", u32), ("ms_traversal", C.c_double),
                ("ms_total", C.c_double), ("ms_kernel_sum", C.c_double),
                ("kernels_timed", u32), ("activity_len", u32), ("edges_relaxed", u64),
                ("vnodes_processed", u64), ("vertices_touched", u64), ("label_updates", u64),
                ("bytes_alg", C.c_double), ("pr_residual", C.c_double),
                ("energy_j", C.c_double), ("activity", P(u32)), ("activity_cap", u32),
                ("label_bits", u32), ("narrow_reruns", u32), ("reserved", u32 * 5)]

    def to_dict(self):
        d = {}
        for name, _ in self._fields_:
            if name in ("reserved", "activity"):
                continue
            d[name] = getattr(self, name)
        return d


class GraphInfo(C.Structure):
    _fields_ = [("n", u32), ("weighted", u32), ("m", u64), ("max_out_degree", u32),
                ("argmax_out_degree", u32), ("nonzero_out", u32), ("device", u32),
                ("device_bytes", u64)]


class EdgeList(C.Structure):
    _fields_ = [("n", u32), ("has_zero_id", u32), ("m", u64), ("src", P(u32)),
                ("dst", P(u32)), ("w", P(u32))]


class ServerCfg(C.Structure):
    _fields_ = [("label_budget_bytes", u64), ("max_batch", u32), ("K", u32),
                ("damping", C.c_float), ("ppr_epsilon", C.c_double), ("flags", u32),
                ("lane_wgs_per_cu", u32), ("lanes", u32), ("first_iteration", u32),
                ("ppr_columns", u32), ("reserved", u32 * 3)]


SERVER_DISCARD_LABELS, SERVER_PLANE_ENGINES, SERVER_NO_BFS_ENGINE, SERVER_NO_PPR_ENGINE = 1, 2, 4, 8
SERVER_NO_LANE_BATCH = 16


class QueryResult(C.Structure):
    _fields_ = [("algo", u32), ("source", u32), ("batch_size", u32), ("status", u32),
                ("t_submit_ms", C.c_double), ("t_start_ms", C.c_double),
                ("t_done_ms", C.c_double)]


def _sig(name, res, *args):
    f = getattr(lib, name)
    f.restype = res
    f.argtypes = list(args)
    return f


_sig("eeg_strerror", C.c_char_p, C.c_int)
_sig("eeg_last_error", C.c_char_p)
_sig("eeg_device_count", C.c_int, P(C.c_int))
_sig("eeg_gen_rmat", C.c_int, u32, u32, C.c_double, C.c_double, C.c_double, u64, u32, u32,
     C.c_void_p, C.c_void_p, C.c_void_p, C.c_int)
_sig("eeg_gen_grid_size", C.c_int, u32, u32, C.c_double, P(u64))
_sig("eeg_gen_grid", C.c_int, u32, u32, C.c_double, u64, u32, u32, C.c_void_p, C.c_void_p,
     C.c_void_p, C.c_int)
_sig("eeg_load_edgelist", C.c_int, C.c_char_p, C.c_int, P(EdgeList))
_sig("eeg_save_edgelist", C.c_int, C.c_char_p, P(EdgeList))
_sig("eeg_edgelist_free", None, P(EdgeList))
_sig("eeg_graph_create", C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, u64, u32, C.c_int,
     P(C.c_void_p))
_sig("eeg_graph_create_device", C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, u64, u32, C.c_int,
     P(C.c_void_p))
_sig("eeg_graph_create_rmat", C.c_int, u32, u32, C.c_double, C.c_double, C.c_double, u64, u32,
     u32, C.c_int, P(C.c_void_p))
_sig("eeg_graph_destroy", C.c_int, C.c_void_p)
_sig("eeg_graph_get_info", C.c_int, C.c_void_p, P(GraphInfo))
_sig("eeg_graph_prepare", C.c_int, C.c_void_p, u32, u32)
_sig("eeg_tigr_export", C.c_int, C.c_void_p, u32, C.c_void_p, C.c_void_p, C.c_void_p, P(u64))
_sig("eeg_csr_export", C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p)
_sig("eeg_csr_export_derived", C.c_int, C.c_void_p, u32, C.c_void_p, C.c_void_p, P(u64))
_sig("eeg_run", C.c_int, C.c_void_p, u32, u32, P(Opts), C.c_void_p, P(Stats))
_sig("eeg_run_device", C.c_int, C.c_void_p, u32, u32, P(Opts), C.c_void_p, C.c_void_p,
     P(Stats))
_sig("eeg_run_device_noalloc", C.c_int, C.c_void_p, u32, u32, P(Opts), C.c_void_p, C.c_void_p,
     P(Stats))
_sig("eeg_run_batch", C.c_int, C.c_void_p, u32, C.c_void_p, u32, P(Opts), C.c_void_p, P(Stats))
_sig("eeg_run_packed", C.c_int, C.c_void_p, u32, u32, C.c_void_p, P(Opts), C.c_void_p,
     P(Stats))
_sig("eeg_autotune", C.c_int, C.c_void_p, u32, C.c_void_p, u32, P(Opts), P(C.c_double))
_sig("eeg_server_create", C.c_int, C.c_void_p, P(ServerCfg), P(C.c_void_p))
_sig("eeg_server_submit", C.c_int, C.c_void_p, u32, u32, P(u64))
_sig("eeg_server_wait", C.c_int, C.c_void_p, u64, C.c_void_p, C.c_double, P(QueryResult))
_sig("eeg_server_cancel", C.c_int, C.c_void_p, u64)
_sig("eeg_server_destroy", C.c_int, C.c_void_p)
_sig("eeg_energy_read", C.c_int, C.c_int, P(C.c_double), P(u64))

# Every entry point include/eeg_capi.h declares (checked by the CPU test suite).
EXPORTS = [
    "eeg_strerror", "eeg_last_error", "eeg_device_count", "eeg_gen_rmat", "eeg_gen_grid_size",
    "eeg_gen_grid", "eeg_load_edgelist", "eeg_save_edgelist", "eeg_edgelist_free",
    "eeg_graph_create", "eeg_graph_create_device", "eeg_graph_create_rmat", "eeg_graph_destroy",
    "eeg_graph_get_info", "eeg_graph_prepare", "eeg_tigr_export", "eeg_csr_export",
    "eeg_csr_export_derived", "eeg_run",
    "eeg_run_device", "eeg_run_device_noalloc", "eeg_run_batch", "eeg_run_packed", "eeg_autotune", "eeg_server_create",
    "eeg_server_submit", "eeg_server_wait", "eeg_server_cancel", "eeg_server_destroy",
    "eeg_energy_read",
]


class EEGError(RuntimeError):
    def __init__(self, status, where):
        self.status = status
        detail = lib.eeg_last_error().decode(errors="replace")
        msg = lib.eeg_strerror(status).decode()
        super().__init__(f"{where}: {msg} (status {status}){': ' + detail if detail else ''}")


def check(status, where):
    if status != OK:
        raise EEGError(status, where)
    return status
