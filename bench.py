#!/usr/bin/env python3
"""Benchmark of the push-style traversal path on one MI355X.

Workload (one "step"): RMAT scale 22, edge factor 16 (67,108,864 edges), a/b/c/d =
0.57/0.19/0.19/0.05, seed 2, weights U[1,64], generated directly in HBM; BFS, SSSP and SSWP
from the maximum out-degree vertex, CC (minimum-id labels on the symmetrised list) and
PageRank (d = 0.85, L1 residual < 1e-7), each in its sync and async push variant, K = 8.

value = input edges traversed by the step's queries / wall time of the step (GTEPS), with the
graph already resident in HBM.  Per-algorithm GTEPS use the device-event traversal time.

    python bench.py [--steps K] [--warmup W] [--gpus 1]

Prints one JSON line on stdout; progress goes to stderr.
"""
import argparse
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "GTEPS & joules/run per algorithm; batched queries/s + p99 vs one-at-a-time"
ALGS = ["bfs", "sssp", "sswp", "cc", "pr"]
VARIANT_OF = {"sync": "sync_push_dd", "async": "async_push_dd"}
HBM_PEAK_GBS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def parse():
    p = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=5)
    p.add_argument("--warmup", type=int, default=2)
    p.add_argument("--scale", type=int, default=22)
    p.add_argument("--edge-factor", type=int, default=16)
    p.add_argument("--seed", type=int, default=2)
    p.add_argument("--K", type=int, default=8)
    p.add_argument("--variants", default="sync,async")
    p.add_argument("--algs", default=",".join(ALGS))
    p.add_argument("--energy-seconds", type=float, default=3.5,
                   help="loop length per job for the energy reading (the counter steps about once "
                        "per second, so at least 3 s)")
    p.add_argument("--cpu-cap", type=float, default=12.0,
                   help="seconds per algorithm for the CPU baseline (extrapolated past it)")
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--no-energy", action="store_true")
    p.add_argument("--no-batch", action="store_true", help="skip the batched and packed section")
    p.add_argument("--no-serve", action="store_true", help="skip the serving section")
    p.add_argument("--no-grid", action="store_true", help="skip the grid section")
    p.add_argument("--no-sweep", action="store_true", help="skip the K x variant sweep and the autotuner")
    p.add_argument("--batch-scale", type=int, default=24)
    p.add_argument("--serve-scale", type=int, default=26)
    p.add_argument("--serve-qps", default="25,50,100,200,400,1000,5000",
                   help="offered loads (queries/s) of the serving sweep")
    p.add_argument("--serve-queries", type=int, default=200, help="queries per serving point")
    p.add_argument("--serve-ppr-columns", type=int, default=0,
                   help="PPR engine width of the serving section (8, 16 or 32; 0 = the server's default)")
    p.add_argument("--serve-lanes", type=int, default=0,
                   help="single-query lanes of the serving section (0 = the server's default, 2)")
    p.add_argument("--serve-queries-top", type=int, default=1500,
                   help="queries at the highest offered load (the saturated throughput point; long enough "
                        "that the last queries' latency does not set the rate: 600 measured 2.3-2.45x, "
                        "1500 2.6x on the same server)")
    p.add_argument("--batched-only", choices=("bfs", "sssp", "sswp", "ppr"),
                   help="only run two multi-source passes (the second timed) of this algorithm on the "
                        "batched section's graph and sources (ppr: the first 64, as in the batched "
                        "section), for profiling the batched kernels")
    p.add_argument("--calibrate-copy", action="store_true",
                   help="only run a plain 512 MiB device-to-device copy kernel three times "
                        "(scale reference for rocprofv3 FETCH_SIZE / WRITE_SIZE)")
    return p.parse_args()


def calibrate_copy():
    import torch

    n = 512 * 1024 * 1024 // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda:0")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"calibration": "device copy", "bytes_read": n * 4, "bytes_written": n * 4,
                      "launches": 3}), flush=True)


def traffic_entry(kernel, workload):
    """HBM bytes per launch measured with rocprofv3 --pmc (FETCH_SIZE, WRITE_SIZE), committed
    under profiles/ as data; None when no measurement for this kernel and workload exists."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(path):
        return None
    try:
        with open(path) as f:
            table = json.load(f)
        e = table.get(f"{workload}:{kernel}")
        return None if e is None else float(e["hbm_bytes_per_launch"])
    except Exception:
        return None


def cpu_baseline(scale, ef, seed, K, algs, cap):
    """The oracle's restatement of the reference CPU path (seq_cpu sweeps over the Tigr layout,
    serial delta-push PageRank), one thread, on the same graph."""
    import numpy as np

    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import eegraph_amd as eg
    import oracle_py as O

    n = 1 << scale
    s, d, w = eg.gen_rmat(scale, ef, seed=seed, wmin=1, wmax=64, threads=16)
    m = len(s)
    src = int(np.argmax(np.bincount(s, minlength=n)))
    lay = O.tigr_layout(n, s, d, w, K)
    total_edges, total_t, parts, extrapolated = 0, 0.0, {}, False
    codes = {"bfs": O.BFS, "sssp": O.SSSP, "sswp": O.SSWP, "cc": O.CC}
    for alg in algs:
        if alg == "pr":
            row, col, _ = O.csr(n, s, d)
            t = time.perf_counter()
            x, passes, conv = O.pr_delta_push(n, row, col, 0.85, 1e-7 * 0.15 / n, cap)
            dt = time.perf_counter() - t
        else:
            L = lay
            if alg == "cc":
                L = O.tigr_layout(n, np.concatenate([s, d]), np.concatenate([d, s]), None, K)
            lab = O.init_labels(codes[alg], n, src)
            t = time.perf_counter()
            passes, conv = O.seq_sweep(codes[alg], n, L, K, lab, cap)
            dt = time.perf_counter() - t
            del L
        if not conv:
            # passes needed is unknown; the GPU's own count of sweeps is a lower bound, so the
            # rate below is an upper bound on the CPU's throughput
            extrapolated = True
        parts[alg] = {"seconds": round(dt, 3), "passes": passes, "converged": conv,
                      "gteps": m / dt / 1e9}
        total_edges += m
        total_t += dt
        log(f"cpu {alg}: {passes} passes in {dt:.2f} s (converged={conv})")
    return {"value": total_edges / total_t / 1e9, "unit": "GTEPS", "cores": 1, "kind": "port",
            "sample": (f"RMAT scale {scale} EF{ef} seed {seed}, K={K}: restated seq_cpu sweeps "
                       f"({', '.join(a for a in algs if a != 'pr')}) and serial delta-push PageRank, "
                       f"one query each, 1 thread, cap {cap:.0f} s per algorithm"
                       + (" (capped runs: lower bound on time)" if extrapolated else "")),
            "per_algorithm": parts, "host_cpu": cpu_model(), "nproc": os.cpu_count()}


class EnergyMeter:
    """Samples the amd-smi energy accumulator in a thread.  The counter advances in steps (about
    once per second on MI355X), so power over a region is taken between the first and the last
    counter step inside it, from the counter's own values and timestamps."""

    def __init__(self, eg):
        import threading

        self.eg = eg
        self.samples = []  # (wall s, microjoules, counter ns) at every counter step
        self.stop = False
        first = eg.energy_read_ts(0)
        self.ok = first is not None
        if self.ok:
            self.samples.append((time.perf_counter(), first[0], first[1]))
            self.th = threading.Thread(target=self._loop, daemon=True)
            self.th.start()

    def _loop(self):
        while not self.stop:
            r = self.eg.energy_read_ts(0)
            if r is not None and r[1] != self.samples[-1][2]:
                self.samples.append((time.perf_counter(), r[0], r[1]))
            time.sleep(0.002)

    def close(self):
        self.stop = True
        if self.ok:
            self.th.join()

    def power(self, t0, t1):
        """Average watts between the first and last counter steps observed in [t0, t1] (wall
        clock), with the seconds covered; None with fewer than two steps."""
        inside = [x for x in self.samples if t0 <= x[0] <= t1]
        if len(inside) < 2:
            return None
        a, b = inside[0], inside[-1]
        dt = (b[2] - a[2]) * 1e-9
        return ((b[1] - a[1]) * 1e-6 / dt, dt) if dt > 0 else None

    def period_ms(self):
        import numpy as np

        ts = [x[2] for x in self.samples]
        return float(np.median(np.diff(ts))) * 1e-6 if len(ts) > 2 else None


def energy_section(eg, run, jobs, seconds):
    """Idle power after a settle period, then each job looped for `seconds`: average watts
    between counter steps inside the loop, J/run = watts x seconds per run, and the same above
    idle (idle watts subtracted)."""
    meter = EnergyMeter(eg)
    if not meter.ok:
        return None
    time.sleep(2.0)  # settle: the previous GPU work is past the counter's next step
    t = time.perf_counter()
    time.sleep(max(3.5, seconds))
    idle = meter.power(t, time.perf_counter())
    out = {"counter_period_ms": None, "idle_w": round(idle[0], 1) if idle else None,
           "method": "counter steps inside each loop (the accumulator advances about once per second)"}
    for alg, var in jobs:
        runs, t = 0, time.perf_counter()
        while time.perf_counter() - t < seconds or runs < 3:
            run(alg, var)
            runs += 1
        t1 = time.perf_counter()
        p = meter.power(t, t1)
        per_run = (t1 - t) / runs
        e = {"runs": runs, "s_per_run": round(per_run, 6)}
        if p:
            e.update({"avg_w": round(p[0], 1), "j_per_run": round(p[0] * per_run, 5), "window_s": round(p[1], 2)})
            if idle:
                e["j_per_run_above_idle"] = round((p[0] - idle[0]) * per_run, 5)
        out[f"{alg}/{var}"] = e
    out["counter_period_ms"] = meter.period_ms()
    meter.close()
    return out


def cpu_model():
    """The host CPU's model string (first "model name" of /proc/cpuinfo)."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or platform.machine()


def sweep_section(g, a, src, m):
    """The RMAT-22 workload over K in {4, 8, 16, 32} and all four variants (BFS/SSSP/SSWP/CC),
    PageRank sync and async per K: device traversal ms (median of 3 after one untimed run)."""
    from eegraph_amd import _lib as L

    out = {}
    for K in (4, 8, 16, 32):
        g.prepare(K, L.PREP_TIGR | L.PREP_SYM | L.PREP_TRANSPOSE)
        row = {}
        for alg in ("bfs", "sssp", "sswp", "cc", "pr"):
            variants = ("sync_push_dd", "async_push_dd") if alg == "pr" else \
                ("async_push_td", "async_push_dd", "sync_push_td", "sync_push_dd")
            r = {}
            for var in variants:
                g.run(alg, src, variant=var, K=K, want_labels=False)
                ms = sorted(g.run(alg, src, variant=var, K=K, want_labels=False)[1]["ms_traversal"]
                            for _ in range(3))[1]
                r[var] = {"ms": round(ms, 4), "gteps": round(m / ms / 1e6, 2)}
            row[alg] = r
        out[f"K{K}"] = row
        log(f"sweep K={K}: " + ", ".join(f"{alg} {min(v['ms'] for v in r.values()):.3f} ms" for alg, r in row.items()))
    return out


def autotune_section(g, src, default_K):
    """eeg_autotune per algorithm against the default options (sync_push_dd, sorted queues,
    K = 8, default grid): median device traversal ms of 5 runs each, same source."""
    import eegraph_amd as eg

    res = {}
    for alg in ("bfs", "sssp", "sswp", "cc", "pr"):
        o, tuned_ms = g.autotune(alg, [src])
        d = eg.make_opts(variant="sync_push_dd", K=default_K)

        def med(opts):
            g.run(alg, src, opts=opts, want_labels=False)
            return sorted(g.run(alg, src, opts=opts, want_labels=False)[1]["ms_traversal"] for _ in range(5))[2]

        dm, tm = med(d), med(o)
        res[alg] = {"default_ms": round(dm, 4), "tuned_ms": round(tm, 4), "speedup": round(dm / tm, 3),
                    "tuned": {"variant": int(o.variant), "K": int(o.K), "block": int(o.block or 256),
                              "frontier": int(o.frontier), "wgs_per_cu": int(o.wgs_per_cu)}}
        log(f"autotune {alg}: {res[alg]}")
    return res


def pct(xs, q):
    import numpy as np
    return round(float(np.percentile(np.asarray(xs), q)), 3)


def batched_section(eg, a):
    """Many queries per traversal against one-at-a-time launches of the same queries.
    Multi-source: RMAT scale 24 (EF 16, seed 4), 256 sources ~ U[0, N) with seed 4, BFS with
    64-bit visit masks and SSSP/SSWP label planes.  Packed: 64 RMAT graphs, scale ~ U{12..20},
    edge factor ~ U{8..32} (seed 5), one source each (max out-degree vertex), in one traversal.
    Latency of a query: from the start of the pass to its answer (a batched answer is ready when
    the batch ends; a one-at-a-time answer after the queries before it)."""
    import numpy as np

    res = {}
    g = eg.Graph.rmat(a.batch_scale, 16, seed=4)
    n = g.info()["n"]
    srcs = np.random.default_rng(4).integers(0, n, 256).astype(np.uint32)
    res[f"rmat{a.batch_scale}_single_query"] = single_query_section(g, a.K)
    multi = {}
    for alg in ("bfs", "sssp", "sswp"):
        g.run(alg, int(srcs[0]), want_labels=False)
        g.run_batch(alg, srcs, want_labels=False)  # same size as the timed pass
        lat = []
        t = time.perf_counter()
        for x in srcs:
            g.run(alg, int(x), want_labels=False)
            lat.append((time.perf_counter() - t) * 1e3)
        one = time.perf_counter() - t
        t = time.perf_counter()
        _, st = g.run_batch(alg, srcs, want_labels=False)
        bat = time.perf_counter() - t
        multi[alg] = {"one_at_a_time_qps": round(256 / one, 1), "batched_qps": round(256 / bat, 1),
                      "speedup": round(one / bat, 2), "one_at_a_time_p99_ms": pct(lat, 99),
                      "batched_p99_ms": round(bat * 1e3, 3), "iterations": st["iterations"]}
        if alg != "bfs":
            multi[alg]["label_bits"] = st["label_bits"]
            multi[alg]["narrow_reruns"] = st["narrow_reruns"]
        log(f"batched {alg}: {multi[alg]}")
    res[f"rmat{a.batch_scale}_256_sources"] = multi
    # personalised PageRank: 64 restart vertices (the first 64 sources) as planes of 32
    psrc = srcs[:64]
    g.run("ppr", int(psrc[0]), want_labels=False)
    g.run_batch("ppr", psrc[:32], want_labels=False)
    lat = []
    t = time.perf_counter()
    for x in psrc:
        g.run("ppr", int(x), want_labels=False)
        lat.append((time.perf_counter() - t) * 1e3)
    one = time.perf_counter() - t
    t = time.perf_counter()
    _, st = g.run_batch("ppr", psrc, want_labels=False)
    bat = time.perf_counter() - t
    res[f"rmat{a.batch_scale}_64_ppr_sources"] = {
        "one_at_a_time_qps": round(64 / one, 1), "batched_qps": round(64 / bat, 1), "speedup": round(one / bat, 2),
        "one_at_a_time_p99_ms": pct(lat, 99), "batched_p99_ms": round(bat * 1e3, 3),
        "iterations": st["iterations"], "plane_iterations": st["launches"]}
    log(f"batched ppr: {res[f'rmat{a.batch_scale}_64_ppr_sources']}")
    g.close()

    rng = np.random.default_rng(5)
    gs = [eg.Graph.rmat(int(rng.integers(12, 21)), int(rng.integers(8, 33)), seed=5000 + i) for i in range(64)]
    psrc = np.array([x.info()["argmax_out_degree"] for x in gs], np.uint32)
    packed = {"graphs": 64, "edges": int(sum(x.info()["m"] for x in gs))}
    for alg in ("bfs", "sssp", "sswp", "cc"):
        for x, y in zip(gs, psrc):
            x.run(alg, int(y), want_labels=False)
        eg.run_packed(gs, alg, psrc, want_labels=False)
        lat = []
        t = time.perf_counter()
        for x, y in zip(gs, psrc):
            x.run(alg, int(y), want_labels=False)
            lat.append((time.perf_counter() - t) * 1e3)
        one = time.perf_counter() - t
        t = time.perf_counter()
        _, st = eg.run_packed(gs, alg, psrc, want_labels=False)
        bat = time.perf_counter() - t
        packed[alg] = {"one_at_a_time_qps": round(64 / one, 1), "batched_qps": round(64 / bat, 1),
                       "speedup": round(one / bat, 2), "one_at_a_time_p99_ms": pct(lat, 99),
                       "batched_p99_ms": round(bat * 1e3, 3), "iterations": st["iterations"]}
        log(f"packed {alg}: {packed[alg]}")
    res["packed_64_rmat_graphs"] = packed
    res["packed_64_rmat_graphs_tuned"] = packed_tuned_section(eg, gs, psrc)
    for x in gs:
        x.close()
    return res


def single_query_section(g, K):
    """One query per algorithm and variant (sync_push_dd, async_push_dd) from the maximum
    out-degree vertex of the batched section's graph: device traversal ms (median of 3 after
    one untimed run), GTEPS on the input edges, and the achieved fraction of the 8 TB/s HBM peak
    on the algorithmic bytes counted by a separate run (count_work)."""
    from eegraph_amd import _lib as L

    info = g.info()
    src, m = info["argmax_out_degree"], info["m"]
    g.prepare(K, L.PREP_TIGR | L.PREP_SYM | L.PREP_TRANSPOSE)
    out = {"source": src, "edges": m}
    for alg in ("bfs", "sssp", "sswp", "cc", "pr"):
        for var in ("sync_push_dd", "async_push_dd"):
            c = g.run(alg, src, variant=var, K=K, want_labels=False, count_work=True)[1]
            g.run(alg, src, variant=var, K=K, want_labels=False)
            ms = sorted(g.run(alg, src, variant=var, K=K, want_labels=False)[1]["ms_traversal"]
                        for _ in range(3))[1]
            gbs = c["bytes_alg"] / (ms * 1e-3) / 1e9
            out[f"{alg}/{var}"] = {"ms": round(ms, 4), "gteps": round(m / ms / 1e6, 2),
                                   "iterations": c["iterations"], "alg_bytes": c["bytes_alg"],
                                   "alg_gbs": round(gbs, 1), "hbm_frac": round(gbs / HBM_PEAK_GBS, 4)}
            log(f"single query rmat24 {alg}/{var}: {out[f'{alg}/{var}']}")
    return out


def packed_tuned_section(eg, gs, psrc):
    """The packed workload with tuned options on both sides: every graph's one-at-a-time query
    runs with eeg_autotune's choice for that graph (choices cached per graph shape), the packed
    pass with the fastest of the variant / frontier / workgroup-size configurations tried on
    it.  Both sides timed as in the default comparison."""
    import eegraph_amd as eg_

    out = {}
    forms = [dict(variant="sync_push_dd"), dict(variant="sync_push_dd", frontier="compact"),
             dict(variant="sync_push_dd", frontier="flags"), dict(variant="async_push_dd"),
             dict(variant="async_push_td")]
    for alg in ("bfs", "sssp", "sswp", "cc"):
        tuned = [x.autotune(alg, [int(y)])[0] for x, y in zip(gs, psrc)]
        for x, y, o in zip(gs, psrc, tuned):
            x.run(alg, int(y), opts=o, want_labels=False)
        t = time.perf_counter()
        for x, y, o in zip(gs, psrc, tuned):
            x.run(alg, int(y), opts=o, want_labels=False)
        one = time.perf_counter() - t
        best = None
        for f in forms:
            for block in (256, 512):
                o = eg_.make_opts(K=8, **f)
                o.block = block
                eg.run_packed(gs, alg, psrc, opts=o, want_labels=False)
                t = time.perf_counter()
                eg.run_packed(gs, alg, psrc, opts=o, want_labels=False)
                dt = time.perf_counter() - t
                if best is None or dt < best[0]:
                    best = (dt, dict(f, block=block))
        out[alg] = {"one_at_a_time_qps": round(64 / one, 1), "batched_qps": round(64 / best[0], 1),
                    "speedup": round(one / best[0], 2), "packed_options": best[1]}
        log(f"packed tuned {alg}: {out[alg]}")
    return out


def batched_only(eg, a):
    """Two multi-source passes of one algorithm (RMAT scale --batch-scale, EF 16, seed 4, the
    same 256 sources as the batched section); prints the timed pass as one JSON line."""
    import numpy as np

    g = eg.Graph.rmat(a.batch_scale, 16, seed=4)
    n = g.info()["n"]
    srcs = np.random.default_rng(4).integers(0, n, 256).astype(np.uint32)
    if a.batched_only == "ppr":
        srcs = srcs[:64]
    g.run_batch(a.batched_only, srcs, want_labels=False)
    t = time.perf_counter()
    _, st = g.run_batch(a.batched_only, srcs, want_labels=False)
    bat = time.perf_counter() - t
    print(json.dumps({"batched_only": a.batched_only,
                      "workload": f"rmat{a.batch_scale}_ef16_seed4_{len(srcs)}_sources",
                      "ms": round(bat * 1e3, 3), "ms_traversal": round(st["ms_traversal"], 3),
                      "queries_per_s": round(len(srcs) / bat, 1), "iterations": st["iterations"],
                      "plane_iterations": st["launches"], "label_bits": st["label_bits"],
                      "narrow_reruns": st["narrow_reruns"]}), flush=True)
    g.close()


def grid_section(eg, a):
    """High-diameter case: 4096 x 4096 grid, 4-neighbour edges both ways plus 1% random
    shortcut pairs (seed 3), w ~ U{1..1000}, source 0.  BFS/SSSP/SSWP/CC with frontier queues
    (sorted: rebuilt from a bitmap in vertex order; compact: in improvement order) against
    flag scans (sync_push_dd), and the async variants; device traversal time of one query
    each (after one untimed run)."""
    t = time.perf_counter()
    s, d, w = eg.gen_grid(4096, 4096, 0.01, seed=3, wmin=1, wmax=1000)
    g = eg.Graph.from_edges(s, d, w, n=4096 * 4096)
    del s, d, w
    m = g.info()["m"]
    log(f"grid graph m={m} built in {time.perf_counter() - t:.2f} s")
    res = {"workload": "grid4096x4096 + 1% shortcuts, seed 3, w 1..1000, source 0", "edges": m}
    for alg in ("bfs", "sssp", "sswp", "cc"):
        r = {}
        for name, var, fr in (("sync_sorted", "sync_push_dd", "sorted"), ("sync_compact", "sync_push_dd", "compact"),
                              ("sync_flags", "sync_push_dd", "flags"),
                              ("async_push_dd", "async_push_dd", "compact"), ("async_push_td", "async_push_td", "compact")):
            c = g.run(alg, 0, variant=var, frontier=fr, K=a.K, want_labels=False, count_work=True)[1]
            _, st = g.run(alg, 0, variant=var, frontier=fr, K=a.K, want_labels=False)
            r[name] = {"ms": round(st["ms_traversal"], 3), "gteps": round(m / st["ms_traversal"] / 1e6, 2),
                       "iterations": st["iterations"], "alg_bytes": c["bytes_alg"],
                       "hbm_frac": round(c["bytes_alg"] / (st["ms_traversal"] * 1e-3) / 8e12, 4)}
        res[alg] = r
        log(f"grid {alg}: {r}")
    g.close()
    return res


def serving_section(eg, a):
    """Serving loop on RMAT scale 26 (EF 16, seed 6, w ~ U{1..64}), offered loads swept over
    --serve-qps (Poisson arrivals, seed 6), each point a stream of --serve-queries queries mixed
    40% BFS / 30% SSSP / 20% SSWP / 10% PPR (damping 0.85, one-hot restart), sources ~ U[0, N).
    Server: default configuration (BFS and PPR engines with continuous batching, SSSP/SSWP on
    two single-query lanes), answers not kept.  Baseline: the same queries run one at a time
    (eeg_run each, serially); its latencies are those of a FIFO queue with the measured service
    times on the same arrival times.  Latency = answer time - arrival."""
    import threading
    import numpy as np

    t = time.perf_counter()
    g = eg.Graph.rmat(a.serve_scale, 16, seed=6)
    n = g.info()["n"]
    g.prepare(8, 1 | 4)  # Tigr of the out- and in-edge CSRs (PPR gathers over in-edges)
    log(f"serving graph n={n} m={g.info()['m']} built in {time.perf_counter() - t:.2f} s")
    rng = np.random.default_rng(6)
    nmax = max(a.serve_queries, a.serve_queries_top)
    mix = rng.choice(["bfs", "sssp", "sswp", "ppr"], size=nmax, p=[0.4, 0.3, 0.2, 0.1])
    srcs = rng.integers(0, n, nmax)
    gaps = rng.exponential(1.0, nmax)  # unit-rate gaps, scaled per offered load

    # one at a time: service time of every query (eeg_run each, serially)
    o = eg.make_opts(variant="sync_push_dd", K=8)
    for alg in ("bfs", "sssp", "sswp", "ppr"):
        g.run(alg, 0, opts=o, want_labels=False)
    meter = EnergyMeter(eg)
    svc = []
    t1 = time.perf_counter()
    for i in range(nmax):
        t = time.perf_counter()
        g.run(mix[i], int(srcs[i]), opts=o, want_labels=False)
        svc.append(time.perf_counter() - t)
    t2 = time.perf_counter()
    p1 = meter.power(t1, t2)
    one_j = round(p1[0] * (t2 - t1) / nmax, 4) if p1 else None
    log(f"one at a time: {nmax} queries in {t2 - t1:.2f} s ({nmax / (t2 - t1):.1f} q/s)")

    srv = eg.Server(g, discard_labels=True, ppr_columns=a.serve_ppr_columns, lanes=a.serve_lanes)
    for alg in ("bfs", "sssp", "sswp", "ppr"):  # warm every path once
        srv.wait(srv.submit(alg, 0))
    points = []
    rates = [float(x) for x in a.serve_qps.split(",")]
    for rate in rates:
        nq = a.serve_queries_top if rate == rates[-1] else a.serve_queries
        arrive = np.cumsum(gaps[:nq]) / rate
        tickets = [None] * nq

        def producer():
            t0 = time.perf_counter()
            for i in range(nq):
                d = arrive[i] - (time.perf_counter() - t0)
                if d > 0:
                    time.sleep(d)
                tickets[i] = srv.submit(mix[i], int(srcs[i]))

        th = threading.Thread(target=producer)
        tw0 = time.perf_counter()
        th.start()
        lat, per, first, last, tprint = [], {}, None, 0.0, time.time()
        for i in range(nq):
            while tickets[i] is None:
                time.sleep(0.0002)
            _, r = srv.wait(tickets[i])
            lat.append(r["t_done_ms"] - r["t_submit_ms"])
            per.setdefault(str(mix[i]), []).append(lat[-1])
            first = r["t_submit_ms"] if first is None else min(first, r["t_submit_ms"])
            last = max(last, r["t_done_ms"])
            if time.time() - tprint > 5:
                tprint = time.time()
                log(f"serving {rate:g} q/s: {i + 1}/{nq} answered")
        th.join()
        tw1 = time.perf_counter()
        ps = meter.power(tw0, tw1)
        served_s = (last - first) / 1e3
        fin, lat1 = 0.0, []
        for i in range(nq):
            fin = max(arrive[i], fin) + svc[i]
            lat1.append((fin - arrive[i]) * 1e3)
        one_s = fin - arrive[0]
        pt = {"offered_qps": rate, "queries": nq,
              "served": {"qps": round(nq / served_s, 1), "p50_ms": pct(lat, 50), "p99_ms": pct(lat, 99),
                         "per_algorithm_p50_p99_ms": {k: [pct(v, 50), pct(v, 99)] for k, v in per.items()},
                         "j_per_query": round(ps[0] * (tw1 - tw0) / nq, 4) if ps else None},
              "one_at_a_time": {"qps": round(nq / one_s, 1), "p50_ms": pct(lat1, 50), "p99_ms": pct(lat1, 99)}}
        pt["qps_ratio"] = round(pt["served"]["qps"] / pt["one_at_a_time"]["qps"], 2)
        points.append(pt)
        log(f"serving point: {pt}")
    srv.close()
    meter.close()
    g.close()
    out = {"workload": f"rmat{a.serve_scale}_ef16_seed6, {a.serve_queries} queries per point "
                       f"({a.serve_queries_top} at the highest load), Poisson arrivals, "
                       f"40/30/20/10 bfs/sssp/sswp/ppr",
           "queries": a.serve_queries,
           "one_at_a_time_capacity_qps": round(nmax / sum(svc), 1),
           "one_at_a_time_j_per_query": one_j,
           "points": points,
           "qps_ratio": points[-1]["qps_ratio"] if points else None}
    return out


def main():
    a = parse()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if a.gpus != 1 or world != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path (asked for "
              f"--gpus {a.gpus}, WORLD_SIZE {world}); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    if a.calibrate_copy:
        calibrate_copy()
        return

    import eegraph_amd as eg
    from eegraph_amd import _lib as L

    if a.batched_only:
        batched_only(eg, a)
        return

    algs = [x for x in a.algs.split(",") if x]
    variants = [VARIANT_OF[v] for v in a.variants.split(",") if v]
    jobs = [(alg, var) for alg in algs for var in variants]
    workload = f"rmat{a.scale}_ef{a.edge_factor}_seed{a.seed}"

    t = time.perf_counter()
    g = eg.Graph.rmat(a.scale, a.edge_factor, seed=a.seed, wmin=1, wmax=64)
    info = g.info()
    g.prepare(a.K, L.PREP_TIGR | L.PREP_SYM | L.PREP_TRANSPOSE)
    src, m = info["argmax_out_degree"], info["m"]
    log(f"graph n={info['n']} m={m} built+prepared in {time.perf_counter() - t:.2f} s; source {src}")

    def run(alg, var, **kw):
        return g.run(alg, src, variant=var, K=a.K, want_labels=False, **kw)[1]

    # counted run: algorithmic bytes and work per job (device counters, not in the timed region)
    counted = {}
    for alg, var in jobs:
        counted[(alg, var)] = run(alg, var, count_work=True)
    log("counted run done")

    for i in range(a.warmup):
        for alg, var in jobs:
            run(alg, var)
    log(f"warmup {a.warmup} done")

    # timed region: the default launch path; each run reports its device time from HIP events
    # the library records on its own stream (first kernel to the end of the last working
    # iteration), so a launch (one iteration) averages device time / working iterations
    per_job = {j: [] for j in jobs}
    t0 = time.perf_counter()
    for k in range(a.steps):
        for alg, var in jobs:
            per_job[(alg, var)].append(run(alg, var))
        log(f"step {k + 1}/{a.steps}")
    wall = time.perf_counter() - t0
    edges_per_step = m * len(jobs)
    value = edges_per_step * a.steps / wall / 1e9

    per_alg = {}
    best = None  # dominant kernel: most device time in the timed region
    for (alg, var), sts in per_job.items():
        ms = sorted(s["ms_traversal"] for s in sts)[len(sts) // 2]
        ksum = sum(s["ms_traversal"] for s in sts)
        kn = sum(s["iterations"] for s in sts)
        c = counted[(alg, var)]
        its = max(1, c["iterations"])
        bytes_per_launch = c["bytes_alg"] / its
        avg_launch_ms = ksum / kn if kn else float("nan")
        key = f"{alg}/{var}"
        per_alg[key] = {"gteps": m / ms / 1e6, "ms": round(ms, 4), "iterations": sts[0]["iterations"],
                        "edges_relaxed": c["edges_relaxed"], "label_updates": c["label_updates"],
                        "alg_bytes": c["bytes_alg"],
                        "kernel_ms_avg": round(avg_launch_ms, 5),
                        "alg_gbs": bytes_per_launch / avg_launch_ms / 1e6 if kn else None}
        if alg == "pr":
            per_alg[key]["residual"] = sts[0]["pr_residual"]
        if kn and (best is None or ksum > best[0]):
            best = (ksum, alg, var, bytes_per_launch, avg_launch_ms, kn)

    kernel_names = {"bfs": "k_relax<BFS>", "sssp": "k_relax<SSSP>", "sswp": "k_relax<SSWP>",
                    "cc": "k_relax<CC>", "pr": "k_pr_pull+k_pr_fix+k_pr_reduce"}
    roofline = None
    if best:
        _, alg, var, bpl, avg_ms, kn = best
        kname = kernel_names[alg]
        achieved = bpl / avg_ms / 1e6
        traffic = traffic_entry(f"{alg}/{var}", workload)
        roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                    "kernel": f"{kname} ({alg} {var}, K={a.K})", "launches_timed": kn,
                    "alg_bytes_per_launch": round(bpl), "avg_launch_ms": round(avg_ms, 5)}

    energy = None
    if not a.no_energy:
        energy = energy_section(eg, run, jobs, a.energy_seconds)
        log(f"energy done: idle {energy and energy.get('idle_w')} W")

    sweep = None if a.no_sweep else sweep_section(g, a, src, m)
    tuned = None if a.no_sweep else autotune_section(g, src, a.K)
    cpu = None
    if not a.no_cpu:
        log("cpu baseline ...")
        cpu = cpu_baseline(a.scale, a.edge_factor, a.seed, a.K, algs, a.cpu_cap)
    g.close()
    grid = None if a.no_grid else grid_section(eg, a)
    batched = None if a.no_batch else batched_section(eg, a)
    serving = None if a.no_serve else serving_section(eg, a)

    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "GTEPS",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(wall / a.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32 labels; PR f32 compensated pairs, f64 residual",
        "data": "synthetic (RMAT generated in HBM, counter-based RNG)",
        "config": {"workload": f"{workload}: {','.join(algs)} x {','.join(variants)}, K={a.K}, "
                               f"source=argmax out-degree ({src})",
                   "scale": a.scale, "edge_factor": a.edge_factor, "seed": a.seed, "K": a.K,
                   "edges": m, "queries_per_step": len(jobs)},
        "timing": "value: wall time of the step's queries with the graph resident in HBM, labels "
                  "left on the device (no copy-out or un-permute, which add n x 4 bytes over PCIe "
                  "per query); per-algorithm ms: device events from the first kernel to the end "
                  "of the last working iteration",
        "roofline": roofline,
        "cpu_baseline": cpu,
        "per_algorithm": per_alg,
        "sweep_rmat22": sweep,
        "autotune_rmat22": tuned,
        "energy": energy,
        "grid": grid,
        "batched": batched,
        "serving": serving,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
