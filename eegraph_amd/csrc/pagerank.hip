// PageRank and personalised PageRank.
//
// The reference pushes float deltas with atomicAdd (src/pr.cu:3-288): the order of additions
// changes from run to run and every part of a multi-part vertex consumes the same delta
// (D6).  Here:
//   sync variants  - Jacobi iteration x' = (1-d)/N + d * sum_{u->v} x_u/outdeg(u), x0 = 1/N,
//                    gathered per destination over the in-edge CSR with Tigr virtual nodes
//                    (K in-edges each, K lanes, one edge per lane).  Each virtual node reduces
//                    its K contributions with a fixed TwoSum tree, each vertex adds its virtual
//                    nodes of a tile in order with compensation (TwoSum; runs longer than
//                    PR_LONG as 64 in-order chunks joined by a fixed tree), vertices that span
//                    tiles are finished by a fix-up pass in tile order.  Results are
//                    deterministic.
//                    The L1 residual is accumulated in FP64 and reduced in a fixed order.
//   async variants - the same gather with ranks updated in place (Gauss-Seidel in tile order):
//                    fewer sweeps, no atomics, order of reads not fixed (result within the
//                    tolerance, not bit-reproducible).
//   personalised   - the same iterations with a one-hot restart: x0 = e_s and
//                    x' = (1-d) e_s + d * sum_{u->v} x_u/outdeg(u), stopped at the same L1
//                    residual.  A residual push with a per-vertex threshold small enough for
//                    an L1 bound of 1e-7 at N = 2^26 touches every vertex many times with float
//                    atomics; the gather does each iteration in one pass without atomics.
//   Mass of vertices without out-edges is dropped (not redistributed) in both forms.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <vector>

#include "compensated.hpp"
#include "engine.hpp"

namespace eeg {
namespace {

struct PullArgs {
    const uint64_t* row;     // in-edge CSR
    const uint32_t* col;     // sources of in-edges
    const uint32_t* vlist;   // vertices with in-edges
    const uint32_t* vscan;   // their virtual-node prefix
    uint32_t nz;
    uint32_t P;
    const float* c_old;      // contributions x_u / outdeg(u)
    const float* x_old;      // ranks as compensated pairs: x = hi + lo
    const float* xl_old;
    float* x_new;
    float* xl_new;
    float* c_new;
    const uint32_t* outdeg;  // out-degrees for the new contributions
    float2* spill;           // per tile: head and tail partials
    double* res_part;        // per workgroup FP64 residual
    Ctl* ctl;
    float base;              // teleport term of every vertex: (1-d)/N, or 0 when personalised
    float damping;
    uint32_t it;
    uint32_t pers;           // personalised: restart vertex (internal id), else 0xffffffff
    float pers_base;         // its teleport term, 1-d
    uint32_t strided;        // in place: workgroup b takes tiles b, b + grid, ... (see k_pr_pull)
    const uint32_t* tstart;  // first entry of each tile (vlist position)
};

// New rank of v from its gathered sum; xo = old rank (hi + lo), d = out-degree.
__device__ __forceinline__ void finish_vertex(const PullArgs& a, uint32_t v, Comp sum, double xo, uint32_t d,
                                              double& res) {
    float hi, lo;
    rank_pair(v == a.pers ? a.pers_base : a.base, a.damping, sum, &hi, &lo);
    a.x_new[v] = hi;
    a.xl_new[v] = lo;
    a.c_new[v] = d ? (hi + lo) / (float)d : 0.0f;
    res += fabs(((double)hi + (double)lo) - xo);
}

// Edge slots per pull tile: twice the traversal tile, so each lane keeps 16 independent
// gathers in flight and the per-tile staging is amortised over more edges.
constexpr int PR_TILE = 2 * TILE_EDGES;
// Virtual nodes of one entry in one tile beyond which a wave, not a thread, sums them.
constexpr int PR_LONG = 32;

template <int K>
// At most 64 VGPRs, 8 waves per SIMD (K = 4: LDS holds 4 workgroups per CU, so 128): the
// gathers are latency-bound, so the extra resident waves are worth more than the registers
// (-4% per iteration against 71 VGPRs, 7 waves, at K = 8).
__global__ __launch_bounds__(WG, K >= 8 ? 8 : 4) void k_pr_pull(PullArgs a) {
    constexpr int VPR = WG / K;
    constexpr int TV = PR_TILE / K;
    constexpr int ROUNDS = PR_TILE / WG;
    constexpr int EMAX = TV + 2;
    __shared__ uint32_t sS[EMAX];
    __shared__ uint64_t sRow[EMAX];
    __shared__ uint32_t sDeg[EMAX];
    __shared__ uint16_t sOwn[TV];
    __shared__ float2 sPart[TV];
    __shared__ float2 sXo[EMAX];     // old rank (hi, lo) of each entry, loaded with its row
    __shared__ uint32_t sOd[EMAX];   // out-degree of each entry
    __shared__ double sRed[WG / WAVE];

    if (a.ctl->done) return;
    double res = 0.0;
    const uint32_t T = a.P, nE = a.nz;
    const uint32_t ntiles = (T + TV - 1) / TV;
    const uint32_t c0 = (uint32_t)((uint64_t)blockIdx.x * ntiles / gridDim.x);
    const uint32_t c1 = (uint32_t)((uint64_t)(blockIdx.x + 1) * ntiles / gridDim.x);
    const int grp = threadIdx.x / K, lj = threadIdx.x % K;
    // Tiles: a contiguous run per workgroup, or (in place) strided over the grid, so the first
    // tiles of every workgroup - the highest-degree vertices in degree order, whose
    // contributions most edges gather - are finished in the first round and the later rounds
    // already read their new values (Gauss-Seidel by rounds of `grid` tiles).
    const uint32_t tb = a.strided ? blockIdx.x : c0, te = a.strided ? ntiles : c1;
    const uint32_t tstep = a.strided ? gridDim.x : 1u;
    if (tb < te) {
        uint32_t e_start = a.tstart[tb];
        for (uint32_t tile = tb; tile < te; tile += tstep) {
            const uint32_t t0 = tile * TV;
            const uint32_t t1 = min(t0 + (uint32_t)TV, T);
            // the entry holding vnode t1 (the next tile's first entry; nE past the last tile):
            // at most TV entries start in (t0, t1], so it lies within the staged window
            const uint32_t e_next = (tile + 1 < ntiles) ? a.tstart[tile + 1] : nE;
            for (int k = threadIdx.x; k < EMAX; k += WG) {
                const uint64_t i = (uint64_t)e_start + k;
                sS[k] = (i <= nE) ? a.vscan[i] : 0xffffffffu;  // vscan[nE] = P closes the last
            }
            __syncthreads();
            // entries in the tile: up to e_next, and e_next itself when it starts before t1
            const int nent = (int)(e_next - e_start) + (sS[e_next - e_start] < t1 ? 1 : 0);
            // entry data, including what finishing the vertex needs, so those loads overlap
            // with the row loads instead of trailing the gathers
            for (int k = threadIdx.x; k < nent; k += WG) {
                const uint32_t v = a.vlist[e_start + k];
                const uint64_t r0 = a.row[v];
                sRow[k] = r0;
                sDeg[k] = (uint32_t)(a.row[v + 1] - r0);
                sXo[k] = make_float2(a.x_old[v], a.xl_old[v]);
                sOd[k] = a.outdeg[v];
            }
            for (int j = threadIdx.x; j < (int)(t1 - t0); j += WG) {
                const uint32_t t = t0 + j;
                int lo = 0, hi = nent;
                while (hi - lo > 1) {
                    const int mid = (lo + hi) >> 1;
                    if (sS[mid] <= t) lo = mid; else hi = mid;
                }
                sOwn[j] = (uint16_t)lo;
            }
            __syncthreads();
            // one in-edge per lane, K lanes per virtual node, TwoSum tree per virtual node;
            // the source indices of all rounds are loaded first, then all contributions, so a
            // lane keeps ROUNDS independent requests in flight
            // Branch-free rounds: a slot past the tile or past its virtual node's edges reads
            // edge 0 and its contribution (always in range, and one shared address per wave)
            // and drops the value, so no exec-mask branches surround the loads.
            uint32_t pidx[ROUNDS];
            bool pv[ROUNDS];
#pragma unroll
            for (int r = 0; r < ROUNDS; ++r) {
                const int slot = r * VPR + grp;  // < TV: sOwn[slot] is in range
                const uint32_t t = t0 + slot;
                const bool tin = t < t1;
                const int k = tin ? (int)sOwn[slot] : 0;
                const uint32_t eoff = (t - sS[k]) * K + lj;
                pv[r] = tin && eoff < sDeg[k];
                const uint64_t e = pv[r] ? sRow[k] + eoff : 0ull;
                // streamed once per iteration: keep it from evicting the gathered contributions
                pidx[r] = __builtin_nontemporal_load(&a.col[e]);
            }
            float pc[ROUNDS];
#pragma unroll
            for (int r = 0; r < ROUNDS; ++r) {
                const float c = a.c_old[pidx[r]];
                pc[r] = pv[r] ? c : 0.0f;
            }
#pragma unroll
            for (int r = 0; r < ROUNDS; ++r) {
                const int slot = r * VPR + grp;
                const Comp g = group_sum_first_lane<K>(pc[r]);
                if (lj == 0 && t0 + slot < t1) sPart[slot] = make_float2(g.s, g.e);
            }
            __syncthreads();
            // Per entry: its virtual nodes of this tile added in order, one thread per entry.  A
            // run longer than PR_LONG (the hubs: up to a whole tile) is summed by one wave
            // instead, each lane a contiguous chunk in order and the lanes by a fixed tree, so a
            // single thread does not serialise the tile.
            auto finish_entry = [&](int k, Comp acc) {
                const bool started_before = sS[k] < t0;
                const bool continues = sS[k + 1] > t1;
                if (!started_before && !continues) {
                    const float2 xo = sXo[k];
                    finish_vertex(a, a.vlist[e_start + k], acc, (double)xo.x + (double)xo.y, sOd[k], res);
                } else if (started_before) {
                    a.spill[2 * tile] = make_float2(acc.s, acc.e);      // head
                } else {
                    a.spill[2 * tile + 1] = make_float2(acc.s, acc.e);  // tail
                }
            };
            for (int k = threadIdx.x; k < nent; k += WG) {
                const uint32_t s0 = max(sS[k], t0), s1 = min(sS[k + 1], t1);
                if (s1 - s0 > (uint32_t)PR_LONG) continue;
                Comp acc{0.0f, 0.0f};
                for (uint32_t t = s0; t < s1; ++t) {
                    const float2 p = sPart[t - t0];
                    acc = comp_add(acc, Comp{p.x, p.y});
                }
                finish_entry(k, acc);
            }
            {
                const int wv = threadIdx.x / WAVE, lane = (int)lane_id();
                int nlong = 0;  // long entries seen so far; the j-th goes to wave j % waves
                for (int base = 0; base < nent; base += WAVE) {
                    const int k = base + lane;
                    bool lng = false;
                    if (k < nent) lng = min(sS[k + 1], t1) - max(sS[k], t0) > (uint32_t)PR_LONG;
                    uint64_t mask = __ballot(lng);
                    while (mask) {
                        const int kk = base + (int)(__ffsll((unsigned long long)mask) - 1);
                        mask &= mask - 1;
                        if (nlong++ % (WG / WAVE) != wv) continue;
                        const uint32_t s0 = max(sS[kk], t0), s1 = min(sS[kk + 1], t1);
                        const uint32_t chunk = (s1 - s0 + WAVE - 1) / WAVE;
                        const uint32_t q0 = min(s0 + (uint32_t)lane * chunk, s1), q1 = min(q0 + chunk, s1);
                        Comp acc{0.0f, 0.0f};
                        for (uint32_t t = q0; t < q1; ++t) {
                            const float2 p = sPart[t - t0];
                            acc = comp_add(acc, Comp{p.x, p.y});
                        }
#pragma unroll
                        for (int off = WAVE / 2; off > 0; off >>= 1) {
                            Comp o;
                            o.s = __shfl_down(acc.s, off, WAVE);
                            o.e = __shfl_down(acc.e, off, WAVE);
                            acc = comp_add(acc, o);
                        }
                        if (lane == 0) finish_entry(kk, acc);
                    }
                }
            }
            __syncthreads();
            if (!a.strided) e_start = e_next;
            else if (tile + tstep < te) e_start = a.tstart[tile + tstep];
        }
    }
    // fixed-order block reduction of the FP64 residual
    res = wave_sum(res);
    if (lane_id() == 0) sRed[threadIdx.x / WAVE] = res;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0;
        for (int i = 0; i < WG / WAVE; ++i) t += sRed[i];
        a.res_part[blockIdx.x] = t;
    }
}

// Tiles crossed by a vertex beyond which a wave, not a thread, combines its partial sums.
constexpr uint32_t FIX_LONG = 8;

// Vertices whose virtual nodes span tiles: tail of the first tile + heads of the following
// tiles, in tile order.  `span` lists their vlist positions: one thread per vertex, and one
// wave per vertex in `span_long` (the hubs, crossing more than FIX_LONG tiles), whose lanes
// load the partial sums at once and whose lane 0 adds them in the same order.
__global__ __launch_bounds__(WG) void k_pr_fix(PullArgs a, const uint32_t* __restrict__ span, uint32_t nspan,
                                               const uint32_t* __restrict__ span_long, uint32_t nlong,
                                               uint32_t TV, double* res_part) {
    __shared__ double sRed[WG / WAVE];
    if (a.ctl->done) return;
    double res = 0.0;
    for (uint32_t i = blockIdx.x * WG + threadIdx.x; i < nspan; i += gridDim.x * WG) {
        const uint32_t pos = span[i];
        const uint32_t s0 = a.vscan[pos], s1 = a.vscan[pos + 1];
        const uint32_t ta = s0 / TV, tb = (s1 - 1) / TV;
        const float2 tl = a.spill[2 * ta + 1];
        Comp acc{tl.x, tl.y};
        for (uint32_t t = ta + 1; t <= tb; ++t) {
            const float2 h = a.spill[2 * t];
            acc = comp_add(acc, Comp{h.x, h.y});
        }
        const uint32_t v = a.vlist[pos];
        finish_vertex(a, v, acc, (double)a.x_old[v] + (double)a.xl_old[v], a.outdeg[v], res);
    }
    const int lane = (int)lane_id();
    const uint32_t nw = gridDim.x * (WG / WAVE);
    for (uint32_t j = blockIdx.x * (WG / WAVE) + threadIdx.x / WAVE; j < nlong; j += nw) {
        const uint32_t pos = span_long[j];
        const uint32_t s0 = a.vscan[pos], s1 = a.vscan[pos + 1];
        const uint32_t ta = s0 / TV, np = (s1 - 1) / TV - ta + 1;  // parts: tail of ta, heads after
        // starting from 0 + tail gives the tail's bits exactly, as the thread form does
        Comp acc{0.0f, 0.0f};
        for (uint32_t base = 0; base < np; base += WAVE) {
            const uint32_t p = base + (uint32_t)lane;
            float2 h = make_float2(0.0f, 0.0f);
            if (p < np) h = a.spill[p == 0 ? 2 * ta + 1 : 2 * (ta + p)];
            const uint32_t cnt = min(np - base, (uint32_t)WAVE);
            for (uint32_t l = 0; l < cnt; ++l) {
                const float hs = __shfl(h.x, (int)l, WAVE), he = __shfl(h.y, (int)l, WAVE);
                acc = comp_add(acc, Comp{hs, he});
            }
        }
        if (lane == 0) {
            const uint32_t v = a.vlist[pos];
            finish_vertex(a, v, acc, (double)a.x_old[v] + (double)a.xl_old[v], a.outdeg[v], res);
        }
    }
    res = wave_sum(res);
    if (lane_id() == 0) sRed[threadIdx.x / WAVE] = res;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0;
        for (int i = 0; i < WG / WAVE; ++i) t += sRed[i];
        res_part[blockIdx.x] = t;
    }
}

// Vertices without in-edges hold their teleport term from the first iteration on.
__global__ void k_pr_isolated(const uint64_t* __restrict__ in_row, const uint32_t* __restrict__ outdeg,
                              uint32_t n, const float* __restrict__ x_old, const float* __restrict__ xl_old,
                              float* __restrict__ x_new, float* __restrict__ xl_new,
                              float* __restrict__ c_new, float base0, uint32_t pers, float pers_base,
                              double* res_part, const Ctl* ctl) {
    __shared__ double sRed[WG / WAVE];
    if (ctl->done) return;
    double res = 0;
    for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n;
         v += (uint64_t)gridDim.x * blockDim.x) {
        if (in_row[v + 1] != in_row[v]) continue;
        const float base = (v == pers) ? pers_base : base0;
        x_new[v] = base;
        xl_new[v] = 0.0f;
        const uint32_t d = outdeg[v];
        c_new[v] = d ? base / (float)d : 0.0f;
        res += fabs((double)base - ((double)x_old[v] + (double)xl_old[v]));
    }
    res = wave_sum(res);
    if (lane_id() == 0) sRed[threadIdx.x / WAVE] = res;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0;
        for (int i = 0; i < WG / WAVE; ++i) t += sRed[i];
        res_part[blockIdx.x] = t;
    }
}

// Sums the per-workgroup residuals in a fixed order and decides convergence.
// One workgroup of RED_WG threads sums the partials in a fixed order (strided per thread, then
// a tree), so the residual is deterministic.
constexpr int RED_WG = 1024;
__global__ __launch_bounds__(RED_WG) void k_pr_reduce(const double* __restrict__ part, uint32_t nparts, Ctl* ctl,
                                                      uint32_t it, double eps) {
    __shared__ double sRed[RED_WG];
    if (ctl->done) return;
    // strided per thread, four loads in flight, added in index order (the same sum as one at
    // a time)
    double t = 0;
    uint32_t i = threadIdx.x;
    for (; i + 3 * RED_WG < nparts; i += 4 * RED_WG) {
        const double p0 = part[i], p1 = part[i + RED_WG], p2 = part[i + 2 * RED_WG], p3 = part[i + 3 * RED_WG];
        t += p0;
        t += p1;
        t += p2;
        t += p3;
    }
    for (; i < nparts; i += RED_WG) t += part[i];
    sRed[threadIdx.x] = t;
    __syncthreads();
    for (int off = RED_WG / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) sRed[threadIdx.x] += sRed[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        ctl->residual[it & 3] = sRed[0];
        ctl->iters = it;
        if (sRed[0] < eps) ctl->done = 1;
    }
}

// x0 = 1/N everywhere, or e_s when personalised (pers = s)
__global__ void k_pr_init(const uint32_t* __restrict__ outdeg, uint32_t n, float x0, uint32_t pers,
                          float* __restrict__ x, float* __restrict__ xl, float* __restrict__ c) {
    for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n;
         v += (uint64_t)gridDim.x * blockDim.x) {
        const float xv = (pers == 0xffffffffu) ? x0 : (v == pers ? 1.0f : 0.0f);
        x[v] = xv;
        xl[v] = 0.0f;
        const uint32_t d = outdeg[v];
        c[v] = d ? xv / (float)d : 0.0f;
    }
}

__global__ void k_pr_output(const float* __restrict__ hi, const float* __restrict__ lo, uint32_t n,
                            float* __restrict__ out) {
    for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n;
         v += (uint64_t)gridDim.x * blockDim.x)
        out[v] = hi[v] + lo[v];
}

__global__ void k_outdeg(const uint64_t* __restrict__ row, uint32_t n, uint32_t* __restrict__ deg) {
    for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n;
         v += (uint64_t)gridDim.x * blockDim.x)
        deg[v] = (uint32_t)(row[v + 1] - row[v]);
}

// First entry of every tile: the largest i < nz with vscan[i] <= tile * TV.
__global__ void k_tile_starts(const uint32_t* __restrict__ vscan, uint32_t nz, uint32_t TV, uint32_t ntiles,
                              uint32_t* __restrict__ tstart) {
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < ntiles; t += gridDim.x * blockDim.x) {
        const uint64_t first = (uint64_t)t * TV;
        uint32_t lo = 0, hi = nz;
        while (hi - lo > 1) {
            const uint32_t mid = lo + (hi - lo) / 2;
            if (vscan[mid] <= first) lo = mid; else hi = mid;
        }
        tstart[t] = lo;
    }
}

// Vertices crossing a tile boundary: those crossing at most FIX_LONG tiles from the front of
// `span`, the others from its back (cap entries: a boundary is crossed by one vertex at most).
__global__ void k_list_spanning(const uint32_t* __restrict__ vscan, uint32_t nz, uint32_t TV, uint32_t cap,
                                uint32_t* __restrict__ span, uint32_t* __restrict__ count) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nz;
         i += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t s0 = vscan[i], s1 = vscan[i + 1];
        const uint32_t ta = s0 / TV, tb = (s1 - 1) / TV;
        if (ta == tb) continue;
        if (tb - ta + 1 > FIX_LONG) span[cap - 1 - atomicAdd(&count[1], 1u)] = (uint32_t)i;
        else span[atomicAdd(&count[0], 1u)] = (uint32_t)i;
    }
}

}  // namespace

// Vertices of the in-edge Tigr table whose virtual nodes cross a pull tile, listed once per
// graph and K (their partial sums are combined by k_pr_fix).  Called with the graph's lock held
// or before concurrent queries start.
eeg_status ensure_pr_span(eeg_graph* g, uint32_t K) {
    Tigr& t = g->tin[k_index(K)];
    if (t.span_built) return EEG_OK;
    if (!g->outdeg.p) {  // out-degrees as u32, read by every PageRank iteration
        EEG_TRY(g->outdeg.alloc((size_t)g->n * 4 + 4));
        if (g->n) {
            k_outdeg<<<std::max(1u, std::min(8192u, (g->n + WG - 1) / WG)), WG, 0, g->stream>>>(
                g->out.rowp(), g->n, g->outdeg.as<uint32_t>());
            EEG_HIP(hipGetLastError());
        }
    }
    const uint32_t TV = PR_TILE / K;
    const uint64_t ntiles = (t.P + TV - 1) / TV;
    hipStream_t s = g->stream;
    DevMem cnt;
    const uint32_t cap = (uint32_t)ntiles + 1;
    EEG_TRY(t.span.alloc((size_t)cap * 4));
    EEG_TRY(cnt.alloc(8));
    EEG_HIP(hipMemsetAsync(cnt.p, 0, 8, s));
    EEG_TRY(t.prstart.alloc((size_t)ntiles * 4 + 4));
    if (t.nz) {
        k_list_spanning<<<std::max(1u, std::min(8192u, (t.nz + WG - 1) / WG)), WG, 0, s>>>(
            t.vscan.as<uint32_t>(), t.nz, TV, cap, t.span.as<uint32_t>(), cnt.as<uint32_t>());
        EEG_HIP(hipGetLastError());
        k_tile_starts<<<std::max<uint64_t>(1, std::min<uint64_t>(4096, (ntiles + WG - 1) / WG)), WG, 0, s>>>(
            t.vscan.as<uint32_t>(), t.nz, TV, (uint32_t)ntiles, t.prstart.as<uint32_t>());
        EEG_HIP(hipGetLastError());
    }
    uint32_t ns[2] = {0, 0};
    EEG_HIP(hipMemcpyAsync(ns, cnt.p, 8, hipMemcpyDeviceToHost, s));
    EEG_HIP(hipStreamSynchronize(s));
    if ((uint64_t)ns[0] + ns[1] > cap) return fail(EEG_ERR_DEVICE, "PageRank span list overflow");
    t.nspan = ns[0];
    t.nspan_long = ns[1];
    t.span_cap = cap;
    t.span_built = true;
    return EEG_OK;
}

eeg_status run_pagerank(eeg_graph* g, uint32_t algo, uint32_t source, const eeg_opts& o,
                        eeg_stats* st, Work* wk, hipStream_t sk) {
    const auto t_host0 = std::chrono::steady_clock::now();
    const bool personalised = (algo == EEG_PPR);
    // async variants update ranks in place (Gauss-Seidel in tile order): a gather may see a
    // neighbour's rank of this sweep or of the last one, so the sum order is not fixed, but the
    // iteration converges in fewer sweeps and needs no atomics
    const bool inplace = (o.variant == EEG_ASYNC_PUSH_TD || o.variant == EEG_ASYNC_PUSH_DD);
    const uint32_t K = o.K ? o.K : 8;
    if (k_index(K) < 0) return fail(EEG_ERR_INVALID_ARG, "K must be 4, 8, 16 or 32 (got %u)", K);
    const float d = o.damping > 0 ? o.damping : 0.85f;
    if (!(d > 0 && d < 1)) return fail(EEG_ERR_INVALID_ARG, "damping must be in (0,1)");
    const double eps = o.epsilon > 0 ? o.epsilon : 1e-7;
    const uint32_t max_iters = o.max_iters ? o.max_iters : 10000;
    const uint32_t n = g->n;
    hipStream_t s = wk ? sk : g->stream;
    Work& w = wk ? *wk : g->work;
    EEG_TRY(ensure_in(g));
    EEG_TRY(ensure_tigr(g, g->in, g->tin, K));
    const Tigr& t = g->tin[k_index(K)];
    const uint32_t TV = PR_TILE / K;
    const uint64_t ntiles = (t.P + TV - 1) / TV;

    EEG_TRY(w.ctl.alloc(sizeof(Ctl)));
    if (!w.host_ctl) EEG_HIP(hipHostMalloc((void**)&w.host_ctl, 2 * sizeof(Ctl), hipHostMallocDefault));
    for (int i = 0; i < 2; ++i) {
        EEG_TRY(w.fx[i].alloc((size_t)n * 4 + 4));
        EEG_TRY(w.fc[i].alloc((size_t)n * 4 + 4));
        EEG_TRY(w.fl[i].alloc((size_t)n * 4 + 4));
    }
    EEG_TRY(w.spill.alloc((size_t)(ntiles + 1) * 2 * sizeof(float2)));
    // vertices whose virtual nodes cross a tile boundary (finished by k_pr_fix)
    EEG_TRY(ensure_pr_span(g, K));
    const DevMem& span = t.span;
    const uint32_t nshort = t.nspan, nlong = t.nspan_long;
    const uint32_t nspan = nshort + nlong;
    const uint32_t* span_long = span.as<uint32_t>() + (t.span_cap - nlong);
    // the fix-up order only matters within a vertex, so the list order is irrelevant

    int cus = 256;
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, g->device);
    // workgroups per CU: more than stay resident (8), so tiles are handed out as workgroups
    // finish; 48 for the Jacobi gather, 24 in place (whose sweep count also depends on it)
    const uint32_t wpc = o.wgs_per_cu ? o.wgs_per_cu : (inplace ? 24u : 48u);
    const unsigned grid_pull = (unsigned)std::max<uint64_t>(1, std::min<uint64_t>(ntiles, (uint64_t)cus * wpc));
    const unsigned grid_fix = std::max(1u, std::min(1024u, std::max((nshort + WG - 1) / WG, (nlong + 3) / 4)));
    const unsigned grid_iso = std::max(1u, std::min(4096u, (n + WG - 1) / WG));
    const uint32_t nparts = grid_pull + grid_fix + grid_iso;
    EEG_TRY(w.red.alloc((size_t)nparts * sizeof(double)));
    EEG_HIP(hipMemsetAsync(w.red.p, 0, (size_t)nparts * sizeof(double), s));

    Ctl* ctl = w.ctl.as<Ctl>();
    EEG_HIP(hipMemsetAsync(ctl, 0, sizeof(Ctl), s));
    const float x0 = 1.0f / (float)n;
    const float base = personalised ? 0.0f : (1.0f - d) / (float)n;
    const uint32_t pers = personalised ? source : 0xffffffffu;
    k_pr_init<<<grid_iso, WG, 0, s>>>(g->outdeg.as<uint32_t>(), n, x0, pers, w.fx[0].as<float>(), w.fl[0].as<float>(),
                                      w.fc[0].as<float>());
    EEG_HIP(hipGetLastError());

    PullArgs pa{};
    pa.row = g->in.rowp();
    pa.col = g->in.colp();
    pa.vlist = t.vlist.as<uint32_t>();
    pa.vscan = t.vscan.as<uint32_t>();
    pa.nz = t.nz;
    pa.P = (uint32_t)t.P;
    pa.outdeg = g->outdeg.as<uint32_t>();
    pa.spill = w.spill.as<float2>();
    pa.ctl = ctl;
    pa.base = base;
    pa.damping = d;
    pa.pers = pers;
    pa.pers_base = 1.0f - d;
    pa.strided = inplace ? 1u : 0u;
    pa.tstart = t.prstart.as<uint32_t>();
    double* red = w.red.as<double>();
    auto pull_fn = (K == 4) ? k_pr_pull<4> : (K == 8) ? k_pr_pull<8> : (K == 16) ? k_pr_pull<16> : k_pr_pull<32>;

    hipEvent_t ev_start;
    EEG_TRY(w.event(0, &ev_start));
    // a timing event after iteration `it` is recorded only where the run may end (the last
    // iteration of every chunk and the iterations near the predicted end): each record costs
    // device time between the launches
    std::vector<uint8_t> recorded;
    uint32_t predicted_end = 0xffffffffu;
    auto launch = [&](uint32_t it, bool rec) -> eeg_status {
        pa.it = it;
        const int ob = inplace ? 0 : (int)((it - 1) & 1), nb = inplace ? 0 : (int)(it & 1);
        pa.x_old = w.fx[ob].as<float>();
        pa.xl_old = w.fl[ob].as<float>();
        pa.c_old = w.fc[ob].as<float>();
        pa.x_new = w.fx[nb].as<float>();
        pa.xl_new = w.fl[nb].as<float>();
        pa.c_new = w.fc[nb].as<float>();
        pa.res_part = red;
        hipEvent_t eb, ea;
        if (o.timing) {
            EEG_TRY(w.event(2 * (size_t)it - 1, &eb));
            EEG_HIP(hipEventRecord(eb, s));
        }
        if (t.P) {
            void* args[] = {&pa};
            EEG_HIP(hipLaunchKernel((const void*)pull_fn, dim3(grid_pull), dim3(WG), args, 0, s));
        } else {
            EEG_HIP(hipMemsetAsync(red, 0, grid_pull * sizeof(double), s));
        }
        if (nspan) {
            k_pr_fix<<<grid_fix, WG, 0, s>>>(pa, span.as<uint32_t>(), nshort, span_long, nlong, TV,
                                             red + grid_pull);
            EEG_HIP(hipGetLastError());
        }
        const uint32_t iso_iters = inplace ? 1 : 2;  // vertices without in-edges settle at once
        if (it <= iso_iters) {
            k_pr_isolated<<<grid_iso, WG, 0, s>>>(g->in.rowp(), g->outdeg.as<uint32_t>(), n, pa.x_old, pa.xl_old,
                                                  pa.x_new, pa.xl_new, pa.c_new, base, pers, 1.0f - d,
                                                  red + grid_pull + grid_fix, ctl);
            EEG_HIP(hipGetLastError());
        } else if (it == iso_iters + 1) {
            EEG_HIP(hipMemsetAsync(red + grid_pull + grid_fix, 0, grid_iso * sizeof(double), s));
        }
        k_pr_reduce<<<1, RED_WG, 0, s>>>(red, nparts, ctl, it, eps);
        EEG_HIP(hipGetLastError());
        if (recorded.size() <= it) recorded.resize((size_t)it + 64, 0);
        recorded[it] = (rec || o.timing) ? 1 : 0;
        if (recorded[it]) {
            hipEvent_t ea2;
            EEG_TRY(w.event(2 * (size_t)it, &ea2));
            EEG_HIP(hipEventRecord(ea2, s));
        }
        (void)ea;
        return EEG_OK;
    };
    if (!nspan) EEG_HIP(hipMemsetAsync(red + grid_pull, 0, grid_fix * sizeof(double), s));

    EEG_HIP(hipEventRecord(ev_start, s));
    uint32_t it = 1;
    eeg_status rc = EEG_OK;
    {
        uint32_t chunk = 4;
        EEG_TRY(w.poll_events());
        hipEvent_t* poll = w.poll;
        bool pending[2] = {false, false};  // a copy of this run is in flight in the slot
        // queues up to `chunk` iterations, never past max_iters (callers check it <= max_iters)
        auto submit = [&](int slot) -> eeg_status {
            const uint32_t cnt = std::min<uint32_t>(chunk, max_iters + 1 - it);
            for (uint32_t k = 0; k < cnt; ++k, ++it)
                EEG_TRY(launch(it, k + 1 == cnt || (predicted_end != 0xffffffffu && it + 2 >= predicted_end)));
            EEG_HIP(hipMemcpyAsync(&w.host_ctl[slot], ctl, sizeof(Ctl), hipMemcpyDeviceToHost, s));
            EEG_HIP(hipEventRecord(poll[slot], s));
            pending[slot] = true;
            chunk = std::min<uint32_t>(chunk * 2, 16);
            return EEG_OK;
        };
        // Iterations still needed after the polled one: the residual decays geometrically, so
        // log(eps / r_k) / log(r_k / r_k-1) (16 when the last two residuals do not show it).
        auto remaining = [&](const Ctl& h) -> uint32_t {
            const uint32_t k = h.iters;
            if (k < 2) return 16;
            const double r1 = h.residual[k & 3], r0 = h.residual[(k - 1) & 3];
            if (!(r1 > 0.0) || !(r0 > r1)) return 16;
            const double need = std::log(eps / r1) / std::log(r1 / r0);
            return need < 1.0 ? 1u : (uint32_t)std::min(16.0, std::ceil(need));
        };
        int cur = 0;
        rc = submit(0);
        while (rc == EEG_OK) {
            if (it <= max_iters) rc = submit(cur ^ 1);
            if (rc != EEG_OK) break;
            EEG_HIP(hipEventSynchronize(poll[cur]));
            pending[cur] = false;
            const Ctl& h = w.host_ctl[cur];
            if (h.done) break;
            if (!pending[cur ^ 1]) {  // max_iters launched and none of them converged
                rc = fail(EEG_ERR_NOT_CONVERGED, "PageRank residual above %g after %u iterations", eps, max_iters);
                break;
            }
            // size the next chunk to the predicted remainder beyond what is in flight (+1), so
            // few launches run after convergence; a short prediction costs one more poll
            const uint32_t inflight = (it - 1) - h.iters, rem = remaining(h);
            chunk = rem > inflight ? std::min<uint32_t>(rem - inflight + 1, 16) : 1u;
            predicted_end = h.iters + rem;
            cur ^= 1;
        }
    }
    EEG_HIP(hipStreamSynchronize(s));
    EEG_HIP(hipMemcpy(w.host_ctl, ctl, sizeof(Ctl), hipMemcpyDeviceToHost));
    const Ctl hc = w.host_ctl[0];
    const uint32_t iters = hc.iters;
    if (rc == EEG_OK && !hc.done)
        rc = fail(EEG_ERR_NOT_CONVERGED, "PageRank residual above %g after %u iterations", eps, max_iters);
    // result: the buffer written by the last iteration, copied into the label slot
    EEG_TRY(w.lab.alloc((size_t)n * 4 + 4));
    const int fb = inplace ? 0 : (int)(iters & 1);
    k_pr_output<<<grid_iso, WG, 0, s>>>(w.fx[fb].as<float>(), w.fl[fb].as<float>(), n, w.lab.as<float>());
    EEG_HIP(hipGetLastError());
    EEG_HIP(hipStreamSynchronize(s));
    if (st) {
        // the first recorded event at or after the last working iteration (every chunk ends
        // with one, so at worst a few no-op launches past it)
        uint32_t ie = std::max<uint32_t>(iters, 1);
        while (ie < recorded.size() && !recorded[ie]) ++ie;
        if (ie >= recorded.size()) ie = (uint32_t)recorded.size() - 1;
        hipEvent_t ev_end;
        EEG_TRY(w.event(2 * (size_t)ie, &ev_end));
        float ms = 0;
        EEG_HIP(hipEventElapsedTime(&ms, ev_start, ev_end));
        st->iterations = iters;
        st->launches = it - 1;
        st->ms_traversal = ms;
        st->pr_residual = hc.residual[iters & 3];
        st->edges_relaxed = (uint64_t)iters * g->m;
        st->vertices_touched = (uint64_t)iters * n;
        // per iteration: 4 B source index + 4 B contribution per in-edge and 8 B per vertex
        // (rank read and write); the row offsets, out-degrees, second rank word and
        // contributions this layout also moves (32 B more per vertex) are not counted
        st->bytes_alg = (double)iters * (8.0 * (double)g->m + 8.0 * (double)n);
        double ksum = 0;
        uint32_t kn = 0;
        if (o.timing) {
            for (uint32_t i = 1; i <= iters; ++i) {
                hipEvent_t eb, ea;
                EEG_TRY(w.event(2 * (size_t)i - 1, &eb));
                EEG_TRY(w.event(2 * (size_t)i, &ea));
                float km = 0;
                EEG_HIP(hipEventElapsedTime(&km, eb, ea));
                ksum += km;
                ++kn;
            }
        }
        st->ms_kernel_sum = ksum;
        st->kernels_timed = kn;
        st->ms_total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_host0).count();
    }
    return rc;
}


}  // namespace eeg
