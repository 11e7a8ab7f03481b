"""MI355X-native push-style graph analytics (BFS, SSSP, SSWP, CC, PageRank) over Tigr
virtual graphs resident in HBM.

The computation lives in libeegraph.so (hand-written HIP kernels for gfx950 behind the C-ABI
of include/eeg_capi.h); this package is a thin ctypes layer over it.
"""
from . import _lib
from ._lib import EEGError, DIST_INFINITY, ALGOS, VARIANTS
from .api import (Graph, Server, device_count, energy_read, energy_read_ts, gen_grid, gen_rmat,
                  load_edgelist, make_opts, run_packed, save_edgelist)

__all__ = ["Graph", "Server", "EEGError", "DIST_INFINITY", "ALGOS", "VARIANTS", "device_count",
           "energy_read", "energy_read_ts", "gen_grid", "gen_rmat", "load_edgelist", "make_opts", "run_packed",
           "save_edgelist"]
