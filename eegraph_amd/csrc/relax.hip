// Push relaxation over Tigr virtual nodes for BFS / SSSP / SSWP / CC, and the host loop that
// drives it.  Replaces the reference kernels {bfs,sssp,sswp,cc}::{async,sync}_push_{td,dd}
// (src/bfs.cu:27-273, src/sssp.cu:26-299, src/sswp.cu:26-271, src/cc.cu:26-271) and the
// iteration loops of src/eegraph.cu (e.g. :2113-2236).
//
// Work decomposition.  A frontier is a list of entries (vertices with out-edges) plus the
// exclusive prefix S of their virtual-node counts ceil(deg/K).  The grid walks the virtual
// nodes in tiles of TILE_EDGES/K; every workgroup owns a contiguous run of tiles, stages the
// tile's entries (vertex, row offset, degree, label) in LDS, maps each virtual node to its
// entry by an LDS binary search, and then relaxes one edge per lane with K lanes per
// virtual node, so a wavefront reads 64 consecutive destination slots of a hub.
//
// Frontier modes (eeg_variant):
//   M_TD    every virtual node every sweep; stop after a sweep without improvement.
//   M_FLAGS per-vertex iteration stamps: active iff flag[v] == iter; improvements stamp iter+1.
//   M_ASYNC per-virtual-node "last label pushed": a virtual node runs when its vertex label is
//           better than what it pushed last, labels in place (no sibling clears, unlike D5).
//   M_QUEUE compacted queue: improved vertices are deduplicated with a stamp, gathered in LDS
//           with wavefront ballots and appended with one packed 64-bit atomic per tile, which
//           reserves queue slots and virtual-node prefix together.
#include <algorithm>
#include <chrono>
#include <cstring>
#include <vector>

#include "engine.hpp"
#include "frontier.hpp"

namespace eeg {
namespace {

enum Mode : int { M_TD = 0, M_FLAGS = 1, M_ASYNC = 2, M_QUEUE = 3 };

struct RelaxArgs {
    const uint64_t* row;
    const uint32_t* col;
    const uint32_t* wgt;
    const uint32_t* vnc;  // per vertex: number of virtual nodes, ceil(deg/K)
    uint32_t* lab;
    const uint32_t* qin;  // frontier entries (vertex ids)
    const uint32_t* sin;  // their virtual-node prefix
    uint32_t* qout;
    uint32_t* sout;
    uint32_t* stamp;
    uint32_t* flag;
    uint32_t* vnpush;
    uint32_t* bitmap;   // sorted frontiers (M_QUEUE): improved vertices set their bit here and
                        // k_compact_frontier builds the next queue; null: queued by the kernel
    Ctl* ctl;
    uint32_t n_static;  // entries / vnodes when qslot_in < 0
    uint32_t P_static;
    int qslot_in;       // Ctl::qpack slot holding the input size, or -1
    int qslot_out;
    uint32_t cap_out;
    uint32_t it;        // 1-based iteration of this traversal
    uint32_t stampv;    // stamp value of this iteration
    uint32_t count_stats;
    // Replayed graph (M_QUEUE): when dev_it is set, dev_it[0] is the iteration and dev_it[1]
    // the stamp base, read from device memory; it, stampv, the control slots and the queue
    // ping-pong are derived from them, so one captured chunk serves every later chunk and run.
    const uint32_t* dev_it;
    uint32_t it_off;              // replayed graph: this node's iteration is dev_it[0] + it_off
    uint32_t *qa, *sa, *qb, *sb;  // queue buffers 0 and 1
    // Tile skipping of the static sweeps (M_ASYNC; M_FLAGS except CC, whose flag scan measured
    // slower with it): an improvement of vertex u in iteration it stamps u's 64-vertex block
    // with tbase + it, and iteration it runs only the tiles whose entries cover a block stamped
    // in iteration it-1 or it (stamps only grow, so nothing is cleared; stamps of earlier
    // traversals are below tbase).  Iteration 1 runs the blocks stamped tbase + 0 at the start
    // (the seeds), or every tile with all_first.
    const uint32_t* tstart;  // first entry of each tile
    const uint64_t* tblk;    // first and last block of each tile's entries
    uint32_t* vdirty;
    uint32_t tbase;
    uint32_t all_first;      // iteration 1 runs every tile (CC: every vertex starts active)
};

// BS threads per workgroup (256, 512 or 1024): a tile keeps TILE_EDGES edge slots, so a larger
// workgroup runs fewer rounds of one edge per lane.
template <int ALG, int K, int MODE, int BS>
__global__ __launch_bounds__(BS) void k_relax(RelaxArgs args) {
    using R = Rule<ALG>;
    RelaxArgs a = args;
    if (MODE == M_QUEUE && a.dev_it) {
        const uint32_t it = *a.dev_it + a.it_off;
        const bool odd = ((it - 1) & 1) != 0;  // input queue = buffer (it-1)&1
        a.it = it;
        a.stampv = a.dev_it[1] + it;
        a.qslot_in = (int)((it - 1) % 3);
        a.qslot_out = (int)(it % 3);
        a.qin = odd ? a.qb : a.qa;
        a.sin = odd ? a.sb : a.sa;
        a.qout = odd ? a.qa : a.qb;
        a.sout = odd ? a.sa : a.sb;
    }
    constexpr int VPR = BS / K;             // virtual nodes per round
    constexpr int TV = TILE_EDGES / K;      // virtual nodes per tile
    constexpr int ROUNDS = TILE_EDGES / BS;
    constexpr int EMAX = TV + 2;            // staged entries: every entry has >= 1 vnode
    constexpr int QCAP = (MODE == M_QUEUE) ? TILE_EDGES : 1;

    __shared__ uint32_t sS[EMAX];
    __shared__ uint64_t sRow[EMAX];
    __shared__ uint32_t sDeg[EMAX];
    __shared__ uint32_t sLab[EMAX];
    __shared__ uint16_t sOwn[TV];
    __shared__ uint32_t sQ[QCAP];
    __shared__ uint32_t sQn[QCAP];
    __shared__ uint32_t sMisc[8];
    __shared__ uint64_t sScan[BS / WAVE + 1];

    Ctl* ctl = a.ctl;
    if (ctl->done) return;
    uint32_t nE, T;
    if (a.qslot_in >= 0) {
        const unsigned long long pk = ctl->qpack[a.qslot_in];
        nE = (uint32_t)(pk & 0xffffffffull);
        T = (uint32_t)(pk >> 32);
    } else {
        nE = a.n_static;
        T = a.P_static;
    }
    const bool lead = (blockIdx.x == 0 && threadIdx.x == 0);
    if (MODE == M_QUEUE) {
        if (nE == 0 || T == 0) {
            if (lead) ctl->done = 1;
            return;
        }
    } else if (a.it > 1 && ctl->changed[(a.it - 1) & 3] == 0) {
        if (lead) ctl->done = 1;
        return;
    }
    if (lead) {
        ctl->iters = a.it;
        if (MODE == M_QUEUE) ctl->qpack[(a.it + 1) % 3] = 0;
        else ctl->changed[(a.it + 1) & 3] = 0;
    }
    if (T == 0) return;

    uint64_t st_ent = 0, st_vn = 0, st_edge = 0, st_upd = 0;
    bool changed = false;
    const uint32_t ntiles = (uint32_t)(((uint64_t)T + TV - 1) / TV);
    const uint32_t c0 = (uint32_t)((uint64_t)blockIdx.x * ntiles / gridDim.x);
    const uint32_t c1 = (uint32_t)((uint64_t)(blockIdx.x + 1) * ntiles / gridDim.x);
    const int grp = threadIdx.x / K, lj = threadIdx.x % K;

    constexpr bool SKIP = MODE == M_ASYNC || (MODE == M_FLAGS && ALG != A_CC);
    if (c0 < c1) {
        uint32_t e_start = SKIP ? 0u : wg_search<BS>(a.sin, nE, c0 * TV);
        for (uint32_t tile = c0; tile < c1; ++tile) {
            if constexpr (SKIP) {
                const uint64_t bb = a.tblk[tile];
                const uint32_t blo = (uint32_t)bb, bhi = (uint32_t)(bb >> 32);
                bool hot = a.it == 1 && a.all_first;
                if (!hot)
                    for (uint32_t b = blo + threadIdx.x; b <= bhi; b += BS) hot |= a.vdirty[b] + 1u >= a.tbase + a.it;
                if (!__syncthreads_or(hot)) continue;  // the same answer in every thread
                e_start = a.tstart[tile];
            }
            const uint32_t t0 = tile * TV;
            const uint32_t t1 = min(t0 + (uint32_t)TV, T);
            for (int k = threadIdx.x; k < EMAX; k += BS) {
                const uint64_t i = (uint64_t)e_start + k;
                sS[k] = (i < nE) ? a.sin[i] : 0xffffffffu;
            }
            if (threadIdx.x == 0) {
                sMisc[0] = 1;
                sMisc[1] = 1;
                sMisc[2] = 0;
            }
            __syncthreads();
            for (int k = threadIdx.x; k < EMAX; k += BS) {
                const uint32_t cur = sS[k];
                const uint32_t nxt = (k + 1 < EMAX) ? sS[k + 1] : 0xffffffffu;
                if (cur <= t1 - 1 && nxt > t1 - 1) sMisc[0] = k + 1;  // entries in the tile
                if (cur <= t1 && nxt > t1) sMisc[1] = k + 1;          // ... and up to t1
            }
            __syncthreads();
            const int nent = (int)sMisc[0];
            const int nnext = (int)sMisc[1];
            for (int k = threadIdx.x; k < nent; k += BS) {
                const uint32_t v = a.qin[e_start + k];
                const uint64_t r0 = a.row[v], r1 = a.row[v + 1];
                const uint32_t lab = a.lab[v];
                uint32_t deg = (uint32_t)(r1 - r0);
                if (MODE == M_FLAGS && a.flag[v] != a.stampv) deg = 0;
                if (!R::usable(lab)) deg = 0;
                sRow[k] = r0;
                sDeg[k] = deg;
                sLab[k] = lab;
                st_ent += 1;
            }
            for (int j = threadIdx.x; j < (int)(t1 - t0); j += BS) {
                const uint32_t t = t0 + j;
                int lo = 0, hi = nent;
                while (hi - lo > 1) {
                    const int mid = (lo + hi) >> 1;
                    if (sS[mid] <= t) lo = mid; else hi = mid;
                }
                sOwn[j] = (uint16_t)lo;
            }
            __syncthreads();

            // Phase 1: each lane issues the edge loads of all its rounds before waiting on
            // any of them; phases 2 and 3 likewise issue every label gather, then every
            // atomic, so a tile keeps ROUNDS independent requests in flight per lane.
            uint32_t ed[ROUNDS], ew[ROUNDS], el[ROUNDS];
            bool ev[ROUNDS];
            // Unweighted frontier queues (BFS, CC) issue the edge and label loads from every
            // lane without exec-mask branches: a slot past the tile or past its virtual node's
            // edges reads edge 0 and that edge's destination label (in range, one shared
            // address per wave) and ev[] drops them.  Measured faster there; the sweeps (many
            // inactive virtual nodes) and the weighted queues keep the branches.
            constexpr bool kFlat = (MODE == M_QUEUE) && !R::weighted;
            if constexpr (kFlat) {
#pragma unroll
                for (int r = 0; r < ROUNDS; ++r) {
                    const int slot = r * VPR + grp;  // < TV: sOwn[slot] is in range
                    const uint32_t t = t0 + slot;
                    const bool tin = t < t1;
                    const int k = tin ? (int)sOwn[slot] : 0;
                    const uint32_t eoff = (t - sS[k]) * K + lj;
                    const uint32_t deg = sDeg[k];
                    const bool act = tin && deg > 0;
                    if (act && lj == 0) ++st_vn;
                    ev[r] = act && eoff < deg;
                    const uint64_t e = ev[r] ? sRow[k] + eoff : 0ull;
                    ed[r] = __builtin_nontemporal_load(&a.col[e]);
                    ew[r] = 1u;
                    el[r] = sLab[k];
                    st_edge += ev[r];
                }
            } else {
#pragma unroll
            for (int r = 0; r < ROUNDS; ++r) {
                const int slot = r * VPR + grp;
                const uint32_t t = t0 + slot;
                ev[r] = false;
                ed[r] = 0;
                ew[r] = 1u;
                el[r] = 0;
                if (t < t1) {
                    const int k = sOwn[slot];
                    const uint32_t eoff = (t - sS[k]) * K + lj;
                    const uint32_t deg = sDeg[k];
                    const uint32_t lab = sLab[k];
                    bool act = deg > 0;
                    if (MODE == M_ASYNC && act) {
                        act = R::better(lab, a.vnpush[t]);
                        if (act && lj == 0) a.vnpush[t] = lab;
                    }
                    if (act && lj == 0) ++st_vn;
                    if (act && eoff < deg) {
                        const uint64_t e = sRow[k] + eoff;
                        ev[r] = true;
                        // edge data is read once per visit: keep it from evicting labels
                        ed[r] = __builtin_nontemporal_load(&a.col[e]);
                        if constexpr (R::weighted) ew[r] = __builtin_nontemporal_load(&a.wgt[e]);
                        el[r] = lab;
                        ++st_edge;
                    }
                }
            }
            }
            bool eu[ROUNDS];
            {
                uint32_t ecur[ROUNDS];
#pragma unroll
                for (int r = 0; r < ROUNDS; ++r) {
                    ew[r] = R::cand(el[r], ew[r]);  // label offered to the destination
                    if constexpr (kFlat) {
                        const uint32_t cur = a.lab[ed[r]];
                        ecur[r] = ev[r] ? cur : 0u;
                    } else {
                        ecur[r] = ev[r] ? a.lab[ed[r]] : 0u;
                    }
                }
                uint32_t eo[ROUNDS];
                bool ea[ROUNDS];
#pragma unroll
                for (int r = 0; r < ROUNDS; ++r) {
                    ea[r] = ev[r] && R::better(ew[r], ecur[r]);
                    eo[r] = ecur[r];
                    if (ea[r]) eo[r] = R::apply(&a.lab[ed[r]], ew[r]);
                }
#pragma unroll
                for (int r = 0; r < ROUNDS; ++r) eu[r] = ea[r] && R::better(ew[r], eo[r]);
            }
            // Phase 4: frontier bookkeeping.  The queue needs each improved vertex's
            // virtual-node count and a dedup stamp; both are issued for all rounds first.
            // Sorted frontiers only set the vertex's bit (an OR whose old value nobody waits
            // for); counts, dedup and order come from k_compact_frontier.
            uint32_t env[ROUNDS];
            if constexpr (MODE == M_QUEUE) {
                if (a.bitmap) {
#pragma unroll
                    for (int r = 0; r < ROUNDS; ++r) {
                        env[r] = 0u;
                        if (eu[r]) atomicOr(&a.bitmap[ed[r] >> 5], 1u << (ed[r] & 31u));
                    }
                } else {
#pragma unroll
                    for (int r = 0; r < ROUNDS; ++r) env[r] = eu[r] ? a.vnc[ed[r]] : 0u;
                    uint32_t eso[ROUNDS];
#pragma unroll
                    for (int r = 0; r < ROUNDS; ++r) {
                        eso[r] = a.stampv;
                        if (env[r]) eso[r] = atomicExch(&a.stamp[ed[r]], a.stampv);
                    }
#pragma unroll
                    for (int r = 0; r < ROUNDS; ++r)
                        if (eso[r] == a.stampv) env[r] = 0;
                }
            }
#pragma unroll
            for (int r = 0; r < ROUNDS; ++r) {
                const bool upd = eu[r];
                const uint32_t dv = ed[r];
                st_upd += upd;
                if (MODE == M_QUEUE) {
                    const bool cand = env[r] != 0;
                    const uint64_t mask = __ballot(cand);
                    if (mask) {
                        const int leader = __ffsll((unsigned long long)mask) - 1;
                        uint32_t base = 0;
                        if (lane_id() == leader) base = atomicAdd(&sMisc[2], (uint32_t)__popcll(mask));
                        base = __shfl(base, leader, WAVE);
                        if (cand) {
                            const uint32_t pos = base + (uint32_t)__popcll(mask & lanemask_lt());
                            sQ[pos] = dv;
                            sQn[pos] = env[r];
                        }
                    }
                } else {
                    if (upd) {
                        changed = true;
                        if (MODE == M_FLAGS) a.flag[dv] = a.stampv + 1;
                        if constexpr (SKIP) a.vdirty[dv >> 6] = a.tbase + a.it;
                    }
                }
            }

            if (MODE == M_QUEUE) {
                __syncthreads();
                const uint32_t cnt = sMisc[2];
                if (cnt > 0) {
                    const uint32_t per = (cnt + BS - 1) / BS;
                    const uint32_t b0 = threadIdx.x * per;
                    uint64_t loc = 0;
                    for (uint32_t i = 0; i < per; ++i)
                        if (b0 + i < cnt) loc += sQn[b0 + i];
                    uint64_t tot;
                    uint64_t pre = block_excl_scan<uint64_t, BS>(loc, sScan, &tot);
                    for (uint32_t i = 0; i < per; ++i)
                        if (b0 + i < cnt) {
                            const uint32_t x = sQn[b0 + i];
                            sQn[b0 + i] = (uint32_t)pre;
                            pre += x;
                        }
                    if (threadIdx.x == 0) {
                        const unsigned long long add = ((unsigned long long)tot << 32) | cnt;
                        const unsigned long long old = atomicAdd(&ctl->qpack[a.qslot_out], add);
                        sMisc[3] = (uint32_t)(old & 0xffffffffull);
                        sMisc[4] = (uint32_t)(old >> 32);
                    }
                    __syncthreads();
                    const uint32_t bq = sMisc[3], bv = sMisc[4];
                    if ((uint64_t)bq + cnt <= a.cap_out) {
                        for (uint32_t i = threadIdx.x; i < cnt; i += BS) {
                            a.qout[bq + i] = sQ[i];
                            a.sout[bq + i] = bv + sQn[i];
                        }
                    } else if (threadIdx.x == 0) {
                        atomicOr(&ctl->err, 1u);
                    }
                }
            }
            __syncthreads();
            if constexpr (!SKIP) e_start += (uint32_t)(nnext - 1);
        }
    }
    if (MODE != M_QUEUE) {
        if (__ballot(changed) && lane_id() == 0) ctl->changed[a.it & 3] = 1;
    }
    if (a.count_stats) {
        st_ent = wave_sum(st_ent);
        st_vn = wave_sum(st_vn);
        st_edge = wave_sum(st_edge);
        st_upd = wave_sum(st_upd);
        if (lane_id() == 0 && (st_ent | st_vn | st_edge | st_upd)) {
            atomicAdd(&ctl->stat[0], (unsigned long long)st_ent);
            atomicAdd(&ctl->stat[1], (unsigned long long)st_vn);
            atomicAdd(&ctl->stat[2], (unsigned long long)st_edge);
            atomicAdd(&ctl->stat[3], (unsigned long long)st_upd);
        }
    }
}

// ---------------------------------------------------------------- initialisation
// Initial labels (src/eegraph.cu:38-49 BFS/SSSP, :1549-1560 SSWP, :544-553 CC): the host
// picks the values; CC passes the original ids, so components end labelled by their minimum
// original id although vertices are stored in degree order.  Written as a select rather than
// a switch on the algorithm: hipcc 7.2 lowered that switch for gfx950 with the default arm
// unreachable (the CC label was never written).
__global__ void k_init_labels(uint32_t* __restrict__ lab, uint32_t n, uint32_t fill,
                              uint32_t at_source, uint32_t source, const uint32_t* __restrict__ ids) {
    for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n;
         v += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t x = (v == source) ? at_source : fill;
        lab[v] = ids ? ids[v] : x;  // CC: the original id of internal vertex v
    }
}

// Tables of the tile skipping: first entry of every tile (the largest i < nz with
// vscan[i] <= tile * TV) and first virtual node of every vertex with edges.
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
// Blocks of every tile: from its first entry's vertex to the vertex of the entry that holds
// the next tile's first virtual node (a superset of the tile's entries).
__global__ void k_tile_blocks(const uint32_t* __restrict__ vlist, const uint32_t* __restrict__ tstart, uint32_t nz,
                              uint32_t ntiles, uint64_t* __restrict__ tblk) {
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < ntiles; t += gridDim.x * blockDim.x) {
        const uint32_t e0 = tstart[t], e1 = (t + 1 < ntiles) ? tstart[t + 1] : nz - 1;
        tblk[t] = (uint64_t)(vlist[e0] >> 6) | ((uint64_t)(vlist[e1] >> 6) << 32);
    }
}
// Stamps the blocks of the seed vertices (or of `single` when seeds is null) with `val`.
__global__ void k_mark_blocks(const uint32_t* __restrict__ seeds, uint32_t nseeds, uint32_t single,
                              uint32_t* __restrict__ vdirty, uint32_t val) {
    const uint32_t cnt = seeds ? nseeds : 1u;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < cnt; i += gridDim.x * blockDim.x)
        vdirty[(seeds ? seeds[i] : single) >> 6] = val;
}

// Advances the device iteration counter of a replayed chunk (one thread, vector store).
__global__ void k_tick(uint32_t* it, uint32_t step) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *it += step;
}

__global__ void k_fill_u32(uint32_t* __restrict__ a, uint64_t n, uint32_t val) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        a[i] = val;
}

// Clears the control block and seeds the first frontier with `source` (when seed != 0).
__global__ void k_init_ctl(Ctl* ctl, const uint64_t* row, uint32_t K, int seed, uint32_t source,
                           uint32_t* q0, uint32_t* s0) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    for (int i = 0; i < 3; ++i) ctl->qpack[i] = 0;
    for (int i = 0; i < 4; ++i) {
        ctl->changed[i] = 0;
        ctl->stat[i] = 0;
        ctl->residual[i] = 0;
    }
    ctl->done = 0;
    ctl->iters = 0;
    ctl->err = 0;
    ctl->mass = 0;
    if (seed) {
        const uint64_t d = row[source + 1] - row[source];
        if (d > 0) {
            q0[0] = source;
            s0[0] = 0;
            ctl->qpack[0] = ((unsigned long long)((d + K - 1) / K) << 32) | 1ull;
        }
    }
}

// Several sources at once (graphs packed side by side, one source each): labels of the seeds,
// and for frontier queues the first frontier with its virtual-node prefix.  `seeds` are
// distinct internal ids.
__global__ void k_seed_labels(uint32_t* __restrict__ lab, const uint32_t* __restrict__ seeds,
                              uint32_t nseeds, uint32_t val) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nseeds;
         i += (uint64_t)gridDim.x * blockDim.x)
        lab[seeds[i]] = val;
}

// One workgroup: appends the seeds with out-edges to q0/s0 in seed order.
__global__ __launch_bounds__(WG) void k_seed_queue(Ctl* ctl, const uint64_t* __restrict__ row, uint32_t K,
                                                   const uint32_t* __restrict__ seeds, uint32_t nseeds,
                                                   uint32_t* __restrict__ q0, uint32_t* __restrict__ s0) {
    __shared__ uint32_t sc[WG / WAVE + 1];
    __shared__ uint64_t sv[WG / WAVE + 1];
    uint32_t cnt = 0;
    uint64_t vn = 0;
    for (uint32_t b = 0; b < nseeds; b += WG) {
        const uint32_t i = b + threadIdx.x;
        uint32_t v = 0;
        uint64_t d = 0;
        if (i < nseeds) {
            v = seeds[i];
            d = row[v + 1] - row[v];
        }
        const uint32_t keep = d > 0 ? 1u : 0u;
        const uint64_t nv = (d + K - 1) / K;
        uint32_t tc;
        uint64_t tv;
        const uint32_t pc = block_excl_scan<uint32_t>(keep, sc, &tc);
        const uint64_t pv = block_excl_scan<uint64_t>(nv, sv, &tv);
        if (keep) {
            q0[cnt + pc] = v;
            s0[cnt + pc] = (uint32_t)(vn + pv);
        }
        cnt += tc;
        vn += tv;
    }
    if (threadIdx.x == 0) ctl->qpack[0] = ((unsigned long long)vn << 32) | cnt;
}

__global__ void k_seed_flags(uint32_t* __restrict__ flag, const uint32_t* __restrict__ seeds,
                             uint32_t nseeds, uint32_t val) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nseeds;
         i += (uint64_t)gridDim.x * blockDim.x)
        flag[seeds[i]] = val;
}

unsigned grid_for(uint64_t work, unsigned cap) {
    uint64_t g = (work + WG - 1) / WG;
    if (g == 0) g = 1;
    return (unsigned)(g > cap ? cap : g);
}

using KernelFn = void (*)(RelaxArgs);

template <int ALG, int K, int BS>
KernelFn pick_mode(int mode) {
    switch (mode) {
        case M_TD: return k_relax<ALG, K, M_TD, BS>;
        case M_FLAGS: return k_relax<ALG, K, M_FLAGS, BS>;
        case M_ASYNC: return k_relax<ALG, K, M_ASYNC, BS>;
        default: return k_relax<ALG, K, M_QUEUE, BS>;
    }
}
template <int ALG, int BS>
KernelFn pick_k(uint32_t K, int mode) {
    switch (K) {
        case 4: return pick_mode<ALG, 4, BS>(mode);
        case 8: return pick_mode<ALG, 8, BS>(mode);
        case 16: return pick_mode<ALG, 16, BS>(mode);
        default: return pick_mode<ALG, 32, BS>(mode);
    }
}
template <int BS>
KernelFn pick_alg(int alg, uint32_t K, int mode) {
    switch (alg) {
        case A_BFS: return pick_k<A_BFS, BS>(K, mode);
        case A_SSSP: return pick_k<A_SSSP, BS>(K, mode);
        case A_SSWP: return pick_k<A_SSWP, BS>(K, mode);
        default: return pick_k<A_CC, BS>(K, mode);
    }
}
// `block` is 256, 512 or 1024 (checked by the callers)
KernelFn pick_kernel(int alg, uint32_t K, int mode, uint32_t block) {
    switch (block) {
        case 512: return pick_alg<512>(alg, K, mode);
        case 1024: return pick_alg<1024>(alg, K, mode);
        default: return pick_alg<256>(alg, K, mode);
    }
}

bool block_ok(uint32_t block) { return block == 0 || block == 256 || block == 512 || block == 1024; }

int cu_count(int device) {
    static int cached[64] = {0};
    if (device >= 0 && device < 64 && cached[device]) return cached[device];
    int c = 0;
    if (hipDeviceGetAttribute(&c, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess || c <= 0)
        c = 256;
    if (device >= 0 && device < 64) cached[device] = c;
    return c;
}

}  // namespace

// Memory the traversal state needs beyond the graph (used by admission control too).
size_t traversal_state_bytes(uint32_t n, uint64_t P, int mode) {
    size_t b = (size_t)n * 4 + sizeof(Ctl);
    if (mode == M_QUEUE) b += (size_t)n * 4 * 5 + ((size_t)n + 31) / 32 * 4;  // queues, stamps, bitmap
    if (mode == M_FLAGS) b += (size_t)n * 4;
    if (mode == M_ASYNC) b += (size_t)P * 4;
    return b;
}

namespace {

// Tile tables of a Tigr layout for the static sweeps (built once per graph and K).
eeg_status ensure_tiles(Tigr& t, uint32_t K, hipStream_t s) {
    if (t.tiles_built) return EEG_OK;
    const uint32_t TV = TILE_EDGES / K;
    const uint64_t ntiles = (t.P + TV - 1) / TV;
    EEG_TRY(t.tstart.alloc((size_t)ntiles * 4 + 4));
    EEG_TRY(t.tblk.alloc((size_t)ntiles * 8 + 8));
    if (t.nz) {
        const unsigned gt = (unsigned)std::max<uint64_t>(1, std::min<uint64_t>(4096, (ntiles + WG - 1) / WG));
        k_tile_starts<<<gt, WG, 0, s>>>(t.vscan.as<uint32_t>(), t.nz, TV, (uint32_t)ntiles, t.tstart.as<uint32_t>());
        k_tile_blocks<<<gt, WG, 0, s>>>(t.vlist.as<uint32_t>(), t.tstart.as<uint32_t>(), t.nz, (uint32_t)ntiles,
                                        t.tblk.as<uint64_t>());
        EEG_HIP(hipGetLastError());
    }
    t.tiles_built = true;
    return EEG_OK;
}

// Iterations per replayed chunk: a multiple of 2 (queue ping-pong) and 3 (control slots), so
// every replay starts in the same state pattern.
constexpr uint32_t GRAPH_CHUNK = 6;

// The executable graph of one chunk for these kernel arguments (buffers, grid, kernel): built
// on first use and kept in the work state with the last few others, so queries that
// alternate algorithms reuse theirs.  It holds kernel nodes only.
eeg_status chunk_graph(Work& w, KernelFn kfn, unsigned grid, unsigned block, const RelaxArgs& ga, hipStream_t s,
                       hipGraphExec_t* out, const CompactArgs* ca = nullptr, unsigned grid_c = 0) {
    const size_t kc = ca ? sizeof(CompactArgs) + sizeof(grid_c) : 0;
    std::vector<uint8_t> key(sizeof(RelaxArgs) + sizeof(kfn) + sizeof(grid) + sizeof(block) + kc);
    memcpy(key.data(), &ga, sizeof(RelaxArgs));
    memcpy(key.data() + sizeof(RelaxArgs), &kfn, sizeof(kfn));
    memcpy(key.data() + sizeof(RelaxArgs) + sizeof(kfn), &grid, sizeof(grid));
    memcpy(key.data() + sizeof(RelaxArgs) + sizeof(kfn) + sizeof(grid), &block, sizeof(block));
    if (ca) {
        uint8_t* kp = key.data() + sizeof(RelaxArgs) + sizeof(kfn) + sizeof(grid) + sizeof(block);
        memcpy(kp, ca, sizeof(CompactArgs));
        memcpy(kp + sizeof(CompactArgs), &grid_c, sizeof(grid_c));
    }
    for (size_t i = 0; i < w.graphs.size(); ++i) {
        if (w.graphs[i].key == key) {
            Work::GraphEntry hit = std::move(w.graphs[i]);
            w.graphs.erase(w.graphs.begin() + (long)i);
            w.graphs.push_back(std::move(hit));
            *out = w.graphs.back().exec;
            return EEG_OK;
        }
    }
    if (w.graphs.size() >= Work::kGraphCache) {  // every earlier run has finished on its stream
        (void)hipGraphExecDestroy(w.graphs.front().exec);
        w.graphs.erase(w.graphs.begin());
    }
    // built node by node (no stream capture, which would interact with other threads'
    // HIP calls): GRAPH_CHUNK relaxations in one dependency chain, node i running iteration
    // dev_it[0] + i (each followed by its frontier compaction for sorted frontiers), then one
    // tick that advances dev_it[0] by GRAPH_CHUNK
    hipGraph_t graph = nullptr;
    EEG_HIP(hipGraphCreate(&graph, 0));
    RelaxArgs a = ga;
    uint32_t* itp = const_cast<uint32_t*>(ga.dev_it);
    void* rargs[] = {&a};
    hipError_t e = hipSuccess;
    hipGraphNode_t prev = nullptr;
    for (uint32_t i = 0; i < GRAPH_CHUNK && e == hipSuccess; ++i) {
        a.it_off = i;  // the node keeps its own copy of the arguments
        hipKernelNodeParams kp{};
        kp.func = (void*)kfn;
        kp.gridDim = dim3(grid);
        kp.blockDim = dim3(block);
        kp.sharedMemBytes = 0;
        kp.kernelParams = rargs;
        kp.extra = nullptr;
        hipGraphNode_t nr = nullptr;
        e = hipGraphAddKernelNode(&nr, graph, prev ? &prev : nullptr, prev ? 1 : 0, &kp);
        prev = nr;
        if (ca && e == hipSuccess) {
            CompactArgs cx = *ca;
            cx.it_off = i;
            void* cargs[] = {&cx};
            hipKernelNodeParams cp{};
            cp.func = (void*)k_compact_frontier;
            cp.gridDim = dim3(grid_c);
            cp.blockDim = dim3(WG);
            cp.sharedMemBytes = 0;
            cp.kernelParams = cargs;
            cp.extra = nullptr;
            hipGraphNode_t nc = nullptr;
            e = hipGraphAddKernelNode(&nc, graph, &prev, 1, &cp);
            prev = nc;
        }
    }
    uint32_t step = GRAPH_CHUNK;
    void* targs[] = {&itp, &step};
    if (e == hipSuccess) {
        hipKernelNodeParams tp{};
        tp.func = (void*)k_tick;
        tp.gridDim = dim3(1);
        tp.blockDim = dim3(64);
        tp.sharedMemBytes = 0;
        tp.kernelParams = targs;
        tp.extra = nullptr;
        hipGraphNode_t nt = nullptr;
        e = hipGraphAddKernelNode(&nt, graph, &prev, 1, &tp);
    }
    if (e != hipSuccess) {
        (void)hipGraphDestroy(graph);
        return fail(EEG_ERR_DEVICE, "building the iteration chunk failed: %s", hipGetErrorString(e));
    }
    hipGraphExec_t exec = nullptr;
    e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) return fail(EEG_ERR_DEVICE, "hipGraphInstantiate failed: %s", hipGetErrorString(e));
    w.graphs.push_back(Work::GraphEntry{std::move(key), exec});
    *out = exec;
    return EEG_OK;
}

}  // namespace

eeg_status run_traversal(eeg_graph* g, uint32_t algo, uint32_t source, const eeg_opts& o,
                         eeg_stats* st, const uint32_t* d_seeds, uint32_t nseeds, Work* wk, hipStream_t sk) {
    const auto t_host0 = std::chrono::steady_clock::now();
    const uint32_t K = o.K ? o.K : 8;
    if (k_index(K) < 0) return fail(EEG_ERR_INVALID_ARG, "K must be 4, 8, 16 or 32 (got %u)", K);
    if (!block_ok(o.block)) return fail(EEG_ERR_INVALID_ARG, "block must be 256, 512 or 1024 (got %u)", o.block);
    const uint32_t block = o.block ? o.block : (uint32_t)WG;
    if (o.wgs_per_cu > 64) return fail(EEG_ERR_INVALID_ARG, "wgs_per_cu must be <= 64");
    int mode;
    switch (o.variant) {
        case EEG_ASYNC_PUSH_TD:
        case EEG_SYNC_PUSH_TD: mode = M_TD; break;
        case EEG_ASYNC_PUSH_DD: mode = M_ASYNC; break;
        case EEG_SYNC_PUSH_DD: mode = (o.frontier == EEG_FRONTIER_FLAGS) ? M_FLAGS : M_QUEUE; break;
        default: return fail(EEG_ERR_INVALID_ARG, "unknown variant %u", o.variant);
    }
    if (o.frontier > EEG_FRONTIER_COMPACT) return fail(EEG_ERR_INVALID_ARG, "unknown frontier kind %u", o.frontier);
    const bool sorted = mode == M_QUEUE && o.frontier == EEG_FRONTIER_SORTED;
    const int alg = (int)algo;  // EEG_BFS..EEG_CC map onto A_BFS..A_CC
    if ((alg == A_SSSP || alg == A_SSWP) && !g->weighted)
        return fail(EEG_ERR_INVALID_ARG, "SSSP/SSWP need a weighted graph");
    if (alg == A_CC) EEG_TRY(ensure_sym(g));
    Csr& c = (alg == A_CC) ? g->sym : g->out;
    Tigr* tt = (alg == A_CC) ? g->tsym : g->tout;
    EEG_TRY(ensure_tigr(g, c, tt, K));
    const Tigr& t = tt[k_index(K)];
    const uint32_t n = g->n;
    hipStream_t s = wk ? sk : g->stream;
    Work& w = wk ? *wk : g->work;

    EEG_TRY(w.lab.alloc((size_t)n * 4 + 4));
    EEG_TRY(w.ctl.alloc(sizeof(Ctl)));
    if (!w.host_ctl) EEG_HIP(hipHostMalloc((void**)&w.host_ctl, 2 * sizeof(Ctl), hipHostMallocDefault));
    if (mode == M_QUEUE) {
        for (int i = 0; i < 2; ++i) {
            EEG_TRY(w.q[i].alloc((size_t)n * 4 + 4));
            EEG_TRY(w.s[i].alloc((size_t)n * 4 + 4));
        }
    }
    const uint32_t max_iters = o.max_iters ? o.max_iters : (n + 2);
    const uint32_t nwords = (n + 31) / 32;
    if (sorted) {
        EEG_TRY(w.bitmap.alloc((size_t)nwords * 4 + 4));
        // a run that stopped early (error, iteration limit) may have left bits behind
        k_fill_u32<<<grid_for(nwords, 8192), WG, 0, s>>>(w.bitmap.as<uint32_t>(), nwords, 0u);
        EEG_HIP(hipGetLastError());
    }
    if (mode == M_QUEUE && !w.stamp.p) {
        EEG_TRY(w.stamp.alloc((size_t)n * 4 + 4));
        EEG_HIP(hipMemsetAsync(w.stamp.p, 0, w.stamp.bytes, s));
    }
    if (mode == M_FLAGS && !w.flag.p) {
        EEG_TRY(w.flag.alloc((size_t)n * 4 + 4));
        EEG_HIP(hipMemsetAsync(w.flag.p, 0, w.flag.bytes, s));
    }
    // tile skipping of the static sweeps: tables of the layout, stamps per tile
    const bool skip = mode == M_ASYNC || (mode == M_FLAGS && alg != A_CC);
    if (skip) {
        EEG_TRY(ensure_tiles(tt[k_index(K)], K, s));
        const size_t need = (size_t)((n + 63) / 64) * 4 + 4;
        if (w.vdirty.bytes < need) {
            EEG_TRY(w.vdirty.alloc(need));
            EEG_HIP(hipMemsetAsync(w.vdirty.p, 0, w.vdirty.bytes, s));
        }
    }
    if ((uint64_t)w.epoch + max_iters + 200 >= 0xfffffff0ull) {  // stamps would wrap: reset
        if (w.stamp.p) EEG_HIP(hipMemsetAsync(w.stamp.p, 0, w.stamp.bytes, s));
        if (w.flag.p) EEG_HIP(hipMemsetAsync(w.flag.p, 0, w.flag.bytes, s));
        if (w.vdirty.p) EEG_HIP(hipMemsetAsync(w.vdirty.p, 0, w.vdirty.bytes, s));
        w.epoch = 1;
    }
    if (mode == M_ASYNC) EEG_TRY(w.vnpush.alloc((size_t)t.P * 4 + 4));

    // ---- labels and control block
    {
        const uint32_t fill = (alg == A_SSWP) ? 0u : INF;
        const uint32_t at_src = (alg == A_SSWP) ? INF : 0u;
        k_init_labels<<<grid_for(n, 8192), WG, 0, s>>>(w.lab.as<uint32_t>(), n, fill, at_src,
                                                       d_seeds ? 0xffffffffu : source,
                                                       alg == A_CC ? g->inv.as<uint32_t>() : nullptr);
        if (d_seeds && alg != A_CC && nseeds)
            k_seed_labels<<<grid_for(nseeds, 1024), WG, 0, s>>>(w.lab.as<uint32_t>(), d_seeds, nseeds, at_src);
    }
    EEG_HIP(hipGetLastError());
    if (mode == M_ASYNC) {
        const uint32_t never = (alg == A_SSWP) ? 0u : 0xffffffffu;
        k_fill_u32<<<grid_for(t.P, 8192), WG, 0, s>>>(w.vnpush.as<uint32_t>(), t.P, never);
        EEG_HIP(hipGetLastError());
    }
    const uint32_t base = w.epoch;  // iteration it uses stamp base + it
    const bool seeded = (alg != A_CC);
    if (skip && seeded) {  // the first sweep starts from the seeds' blocks
        const uint32_t cnt = d_seeds ? nseeds : 1u;
        k_mark_blocks<<<std::max(1u, std::min(1024u, (cnt + WG - 1) / WG)), WG, 0, s>>>(
            d_seeds, nseeds, source, w.vdirty.as<uint32_t>(), base);
        EEG_HIP(hipGetLastError());
    }
    Ctl* ctl = w.ctl.as<Ctl>();
    k_init_ctl<<<1, 64, 0, s>>>(ctl, c.rowp(), K, (seeded && mode == M_QUEUE && !d_seeds) ? 1 : 0, source,
                                mode == M_QUEUE ? w.q[0].as<uint32_t>() : nullptr,
                                mode == M_QUEUE ? w.s[0].as<uint32_t>() : nullptr);
    EEG_HIP(hipGetLastError());
    if (seeded && mode == M_QUEUE && d_seeds && nseeds) {
        k_seed_queue<<<1, WG, 0, s>>>(ctl, c.rowp(), K, d_seeds, nseeds, w.q[0].as<uint32_t>(),
                                      w.s[0].as<uint32_t>());
        EEG_HIP(hipGetLastError());
    }
    if (mode == M_FLAGS) {
        if (seeded && d_seeds) {
            if (nseeds) k_seed_flags<<<grid_for(nseeds, 1024), WG, 0, s>>>(w.flag.as<uint32_t>(), d_seeds, nseeds, base + 1);
        } else if (seeded) {
            k_fill_u32<<<1, 64, 0, s>>>(w.flag.as<uint32_t>() + source, 1, base + 1);
        } else {
            k_fill_u32<<<grid_for(n, 8192), WG, 0, s>>>(w.flag.as<uint32_t>(), n, base + 1);
        }
        EEG_HIP(hipGetLastError());
    }

    // ---- launch geometry: a resident grid that walks tiles
    const int cus = cu_count(g->device);
    const uint32_t TV = TILE_EDGES / K;
    const uint64_t static_tiles = (t.P + TV - 1) / TV;
    // workgroups per CU of the grid (8 stay resident; more hand out tiles as workgroups finish):
    // measured on RMAT-22 and the 4096^2 grid, flag scans gain from 48, in-place sweeps from 40
    // (32 and 48 each lost on one of the two), queues from 24; topology sweeps stay at 24
    // (defaults for 256-thread workgroups; larger workgroups keep the same threads per CU)
    const uint32_t per = block / WG;
    const uint32_t wpc_static =
        o.wgs_per_cu ? o.wgs_per_cu : std::max(1u, (mode == M_FLAGS ? 48u : mode == M_ASYNC ? 40u : 24u) / per);
    const uint32_t wpc_queue = o.wgs_per_cu ? o.wgs_per_cu : std::max(1u, 24u / per);
    const unsigned grid_static =
        (unsigned)std::max<uint64_t>(1, std::min<uint64_t>(static_tiles, (uint64_t)cus * wpc_static));
    const unsigned grid_queue = (unsigned)cus * wpc_queue;
    KernelFn kfn = pick_kernel(alg, K, mode, block);

    RelaxArgs ra{};
    ra.row = c.rowp();
    ra.col = c.colp();
    ra.wgt = c.wp();
    ra.vnc = t.vnc.as<uint32_t>();
    ra.lab = w.lab.as<uint32_t>();
    ra.stamp = w.stamp.as<uint32_t>();
    ra.flag = w.flag.as<uint32_t>();
    ra.vnpush = w.vnpush.as<uint32_t>();
    ra.bitmap = sorted ? w.bitmap.as<uint32_t>() : nullptr;
    ra.ctl = ctl;
    ra.cap_out = n;
    ra.n_static = t.nz;
    ra.P_static = (uint32_t)t.P;
    ra.count_stats = (o.count_work || o.collect_activity) ? 1u : 0u;
    if (skip) {
        ra.tstart = t.tstart.as<uint32_t>();
        ra.tblk = t.tblk.as<uint64_t>();
        ra.vdirty = w.vdirty.as<uint32_t>();
        ra.tbase = base;
        ra.all_first = seeded ? 0u : 1u;
    }

    CompactArgs ca{};
    unsigned grid_c = 1;
    if (sorted) {
        ca.bitmap = w.bitmap.as<uint32_t>();
        ca.nwords = nwords;
        ca.n = n;
        ca.vnc = t.vnc.as<uint32_t>();
        ca.ctl = ctl;
        ca.qa = w.q[0].as<uint32_t>();
        ca.sa = w.s[0].as<uint32_t>();
        ca.qb = w.q[1].as<uint32_t>();
        ca.sb = w.s[1].as<uint32_t>();
        ca.cap = n;
        grid_c = (unsigned)std::max<uint64_t>(1, std::min<uint64_t>(((uint64_t)nwords + WG - 1) / WG, (uint64_t)cus * 4));
    }

    // Event pool: per launch L (it = L+1) an "after" event at 2L+1 and, with timing, a
    // "before" event at 2L; event 0 of the pool tail marks the traversal start.
    hipEvent_t ev_start;
    EEG_TRY(w.event(0, &ev_start));
    auto launch = [&](uint32_t it) -> eeg_status {
        ra.it = it;
        ra.stampv = base + it;
        unsigned grid = grid_static;
        if (mode == M_QUEUE) {
            if (it == 1 && !seeded) {
                ra.qin = t.vlist.as<uint32_t>();
                ra.sin = t.vscan.as<uint32_t>();
                ra.qslot_in = -1;
            } else {
                ra.qin = w.q[(it - 1) & 1].as<uint32_t>();
                ra.sin = w.s[(it - 1) & 1].as<uint32_t>();
                ra.qslot_in = (int)((it - 1) % 3);
                grid = grid_queue;
            }
            ra.qout = w.q[it & 1].as<uint32_t>();
            ra.sout = w.s[it & 1].as<uint32_t>();
            ra.qslot_out = (int)(it % 3);
        } else {
            ra.qin = t.vlist.as<uint32_t>();
            ra.sin = t.vscan.as<uint32_t>();
            ra.qslot_in = -1;
        }
        hipEvent_t eb, ea;
        if (o.timing) {
            EEG_TRY(w.event(2 * (size_t)it - 1, &eb));
            EEG_HIP(hipEventRecord(eb, s));
        }
        void* args[] = {&ra};
        EEG_HIP(hipLaunchKernel((const void*)kfn, dim3(grid), dim3(block), args, 0, s));
        if (sorted) {
            ca.it = it;
            void* cargs[] = {&ca};
            EEG_HIP(hipLaunchKernel((const void*)k_compact_frontier, dim3(grid_c), dim3(WG), cargs, 0, s));
        }
        EEG_TRY(w.event(2 * (size_t)it, &ea));
        EEG_HIP(hipEventRecord(ea, s));
        return EEG_OK;
    };

    EEG_HIP(hipEventRecord(ev_start, s));
    uint32_t it = 1;
    eeg_status rc = EEG_OK;
    uint32_t graph_launches = 0;  // nonzero: the replayed-graph path ran
    std::vector<uint32_t> activity;
    if (o.collect_activity) {
        // One launch at a time so every iteration's activity can be read back (the reference's
        // sumLabels count of active vertices, src/gpu_utils.cu:28-34): with frontier queues the
        // entries of the frontier the iteration processed (vertices with out-edges), otherwise
        // the label updates the iteration made.
        unsigned long long upd_before = 0;
        while (true) {
            EEG_TRY(launch(it));
            EEG_HIP(hipMemcpyAsync(w.host_ctl, ctl, sizeof(Ctl), hipMemcpyDeviceToHost, s));
            EEG_HIP(hipStreamSynchronize(s));
            const Ctl& h = w.host_ctl[0];
            if (h.err || h.done) break;
            if (mode == M_QUEUE) activity.push_back((uint32_t)(h.qpack[(it - 1) % 3] & 0xffffffffull));
            else activity.push_back((uint32_t)(h.stat[3] - upd_before));
            upd_before = h.stat[3];
            if (++it > max_iters) {
                rc = fail(EEG_ERR_NOT_CONVERGED, "no fixpoint after %u iterations", max_iters);
                break;
            }
        }
    } else if (mode == M_QUEUE && !o.timing) {
        // Iteration 1 eagerly (it may read the static entry list), then a graph of GRAPH_CHUNK
        // iterations (relaxation + counter tick each) replayed until the control block reports
        // done.  The next replay is queued before the previous one's control
        // block is inspected.
        EEG_TRY(launch(1));
        it = 2;
        EEG_TRY(w.devit.alloc(8));
        uint32_t* dit = w.devit.as<uint32_t>();
        k_fill_u32<<<1, 64, 0, s>>>(dit, 1, 2u);
        k_fill_u32<<<1, 64, 0, s>>>(dit + 1, 1, base);
        EEG_HIP(hipGetLastError());
        RelaxArgs ga = ra;
        ga.it = 0;
        ga.stampv = 0;
        ga.qslot_in = 0;
        ga.qslot_out = 0;
        ga.qin = ga.sin = nullptr;
        ga.qout = ga.sout = nullptr;
        ga.dev_it = dit;
        ga.qa = w.q[0].as<uint32_t>();
        ga.sa = w.s[0].as<uint32_t>();
        ga.qb = w.q[1].as<uint32_t>();
        ga.sb = w.s[1].as<uint32_t>();
        hipGraphExec_t exec = nullptr;
        CompactArgs gca = ca;
        gca.it = 0;
        gca.dev_it = dit;
        EEG_TRY(chunk_graph(w, kfn, grid_queue, block, ga, s, &exec, sorted ? &gca : nullptr, grid_c));
        EEG_TRY(w.poll_events());
        hipEvent_t* poll = w.poll;
        uint32_t replays = 0;
        auto submit = [&](int slot) -> eeg_status {
            EEG_HIP(hipGraphLaunch(exec, s));
            ++replays;
            it += GRAPH_CHUNK;
            hipEvent_t er;
            EEG_TRY(w.event(2 + replays, &er));
            EEG_HIP(hipEventRecord(er, s));
            EEG_HIP(hipMemcpyAsync(&w.host_ctl[slot], ctl, sizeof(Ctl), hipMemcpyDeviceToHost, s));
            EEG_HIP(hipEventRecord(poll[slot], s));
            return EEG_OK;
        };
        int cur = 0;
        rc = submit(0);
        while (rc == EEG_OK) {
            if (it <= max_iters + GRAPH_CHUNK) rc = submit(cur ^ 1);
            if (rc != EEG_OK) break;
            EEG_HIP(hipEventSynchronize(poll[cur]));
            const Ctl& h = w.host_ctl[cur];
            if (h.err || h.done) break;
            if (it > max_iters + GRAPH_CHUNK) {
                rc = fail(EEG_ERR_NOT_CONVERGED, "no fixpoint after %u iterations", max_iters);
                break;
            }
            cur ^= 1;
        }
        graph_launches = 1 + replays * GRAPH_CHUNK;
    } else {
        // Chunks of launches; the next chunk is queued before the previous chunk's control
        // block is inspected, so the GPU never waits for the host between chunks.
        uint32_t chunk = 2;
        EEG_TRY(w.poll_events());
        hipEvent_t* poll = w.poll;
        auto submit = [&](int slot) -> eeg_status {
            for (uint32_t k = 0; k < chunk; ++k, ++it) EEG_TRY(launch(it));
            EEG_HIP(hipMemcpyAsync(&w.host_ctl[slot], ctl, sizeof(Ctl), hipMemcpyDeviceToHost, s));
            EEG_HIP(hipEventRecord(poll[slot], s));
            chunk = std::min<uint32_t>(chunk * 2, 8);
            return EEG_OK;
        };
        int cur = 0;
        rc = submit(0);
        while (rc == EEG_OK) {
            if (it <= max_iters + 8) rc = submit(cur ^ 1);
            if (rc != EEG_OK) break;
            EEG_HIP(hipEventSynchronize(poll[cur]));
            const Ctl& h = w.host_ctl[cur];
            if (h.err || h.done) break;
            if (it > max_iters + 8) {
                rc = fail(EEG_ERR_NOT_CONVERGED, "no fixpoint after %u iterations", max_iters);
                break;
            }
            cur ^= 1;
        }
    }
    EEG_HIP(hipStreamSynchronize(s));
    EEG_HIP(hipMemcpy(w.host_ctl, ctl, sizeof(Ctl), hipMemcpyDeviceToHost));
    const Ctl hc = w.host_ctl[0];
    if (rc == EEG_OK && hc.err) rc = fail(EEG_ERR_DEVICE, "frontier queue overflow (err=%u)", hc.err);
    if (rc == EEG_OK && !hc.done) rc = fail(EEG_ERR_NOT_CONVERGED, "no fixpoint after %u iterations", max_iters);
    w.epoch = base + it + 2;

    if (st) {
        // end of the last working iteration: its launch, or the replay that contains it
        hipEvent_t ev_end;
        if (graph_launches && hc.iters > 1)
            EEG_TRY(w.event(2 + (hc.iters - 1 + GRAPH_CHUNK - 1) / GRAPH_CHUNK, &ev_end));
        else
            EEG_TRY(w.event(2 * (size_t)std::max<uint32_t>(hc.iters, 1), &ev_end));
        float ms = 0;
        EEG_HIP(hipEventElapsedTime(&ms, ev_start, ev_end));
        st->iterations = hc.iters;
        st->launches = graph_launches ? graph_launches : it - 1;
        st->ms_traversal = ms;
        st->edges_relaxed = hc.stat[2];
        st->vnodes_processed = hc.stat[1];
        st->vertices_touched = hc.stat[0];
        st->label_updates = hc.stat[3];
        const double ce = (alg == A_SSSP || alg == A_SSWP) ? 12.0 : 8.0;
        double fb = 0;
        if (mode == M_QUEUE) fb = 8.0 * (double)hc.stat[3];
        if (mode == M_FLAGS) fb = 4.0 * (double)hc.stat[0];
        if (mode == M_ASYNC) fb = 4.0 * (double)hc.stat[1];
        st->bytes_alg = 12.0 * hc.stat[0] + ce * hc.stat[2] + 4.0 * hc.stat[3] + fb;
        double ksum = 0;
        uint32_t kn = 0;
        if (o.timing) {
            for (uint32_t i = 1; i <= hc.iters; ++i) {
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
        st->activity_len = 0;
        if (st->activity && st->activity_cap) {
            const uint32_t L = (uint32_t)std::min<size_t>(activity.size(), st->activity_cap);
            for (uint32_t i = 0; i < L; ++i) st->activity[i] = activity[i];
            st->activity_len = L;
        }
        st->ms_total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_host0).count();
    }
    return rc;
}

// A traversal on caller-given device buffers that allocates nothing: no device or pinned
// memory, events or executable graphs.  Frontier queues (sync_push_dd, compact) or the
// topology-driven sweep; the host reads the control block after each chunk of launches.
// The dedup stamps are cleared at the start, so nothing carries over between runs.
eeg_status run_traversal_fixed(eeg_graph* g, uint32_t algo, uint32_t source, const eeg_opts& o,
                               const FixedState& fs, hipStream_t s, eeg_stats* st) {
    const auto t_host0 = std::chrono::steady_clock::now();
    const uint32_t K = o.K ? o.K : 8;
    if (k_index(K) < 0) return fail(EEG_ERR_INVALID_ARG, "K must be 4, 8, 16 or 32 (got %u)", K);
    if (!block_ok(o.block)) return fail(EEG_ERR_INVALID_ARG, "block must be 256, 512 or 1024 (got %u)", o.block);
    const uint32_t block = o.block ? o.block : (uint32_t)WG;
    if (o.wgs_per_cu > 64) return fail(EEG_ERR_INVALID_ARG, "wgs_per_cu must be <= 64");
    if (o.collect_activity || o.timing)
        return fail(EEG_ERR_UNSUPPORTED, "activity and per-launch timing need allocations");
    int mode;
    switch (o.variant) {
        case EEG_ASYNC_PUSH_TD:
        case EEG_SYNC_PUSH_TD: mode = M_TD; break;
        case EEG_SYNC_PUSH_DD:
            // sorted frontiers keep a bitmap: the compact queue stands in (same answers)
            if (o.frontier == EEG_FRONTIER_FLAGS)
                return fail(EEG_ERR_UNSUPPORTED, "flag-scan frontiers are not available without allocation");
            if (o.frontier > EEG_FRONTIER_COMPACT)
                return fail(EEG_ERR_INVALID_ARG, "unknown frontier kind %u", o.frontier);
            mode = M_QUEUE;
            break;
        case EEG_ASYNC_PUSH_DD:
            return fail(EEG_ERR_UNSUPPORTED, "async_push_dd keeps per-virtual-node state; not available without allocation");
        default: return fail(EEG_ERR_INVALID_ARG, "unknown variant %u", o.variant);
    }
    const int alg = (int)algo;
    if (alg > A_CC) return fail(EEG_ERR_UNSUPPORTED, "only BFS, SSSP, SSWP and CC run without allocation");
    if ((alg == A_SSSP || alg == A_SSWP) && !g->weighted)
        return fail(EEG_ERR_INVALID_ARG, "SSSP/SSWP need a weighted graph");
    if (alg == A_CC && !g->sym.built)
        return fail(EEG_ERR_INVALID_ARG, "CC needs the symmetrised CSR: call eeg_graph_prepare first");
    Csr& c = (alg == A_CC) ? g->sym : g->out;
    Tigr* tt = (alg == A_CC) ? g->tsym : g->tout;
    if (!tt[k_index(K)].built)
        return fail(EEG_ERR_INVALID_ARG, "Tigr tables for K = %u not built: call eeg_graph_prepare first", K);
    const Tigr& t = tt[k_index(K)];
    const uint32_t n = g->n;
    const uint32_t max_iters = o.max_iters ? o.max_iters : (n + 2);

    const uint32_t fill = (alg == A_SSWP) ? 0u : INF;
    const uint32_t at_src = (alg == A_SSWP) ? INF : 0u;
    k_init_labels<<<grid_for(n, 8192), WG, 0, s>>>(fs.lab, n, fill, at_src, source,
                                                   alg == A_CC ? g->inv.as<uint32_t>() : nullptr);
    if (mode == M_QUEUE) k_fill_u32<<<grid_for(n, 8192), WG, 0, s>>>(fs.stamp, n, 0u);
    const bool seeded = (alg != A_CC);
    k_init_ctl<<<1, 64, 0, s>>>(fs.ctl, c.rowp(), K, (seeded && mode == M_QUEUE) ? 1 : 0, source, fs.q[0], fs.s[0]);
    EEG_HIP(hipGetLastError());

    const int cus = cu_count(g->device);
    const uint32_t TV = TILE_EDGES / K;
    const uint64_t static_tiles = (t.P + TV - 1) / TV;
    const uint32_t wpc = o.wgs_per_cu ? o.wgs_per_cu : std::max(1u, 24u * (uint32_t)WG / block);
    const unsigned grid_static = (unsigned)std::max<uint64_t>(1, std::min<uint64_t>(static_tiles, (uint64_t)cus * wpc));
    const unsigned grid_queue = (unsigned)cus * wpc;
    KernelFn kfn = pick_kernel(alg, K, mode, block);
    RelaxArgs ra{};
    ra.row = c.rowp();
    ra.col = c.colp();
    ra.wgt = c.wp();
    ra.vnc = t.vnc.as<uint32_t>();
    ra.lab = fs.lab;
    ra.stamp = fs.stamp;
    ra.ctl = fs.ctl;
    ra.cap_out = n;
    ra.n_static = t.nz;
    ra.P_static = (uint32_t)t.P;
    ra.count_stats = o.count_work ? 1u : 0u;
    const uint32_t base = 1;  // stamps were cleared to 0: iteration it stamps 1 + it
    auto launch = [&](uint32_t it) -> eeg_status {
        ra.it = it;
        ra.stampv = base + it;
        unsigned grid = grid_static;
        if (mode == M_QUEUE && !(it == 1 && !seeded)) {
            ra.qin = fs.q[(it - 1) & 1];
            ra.sin = fs.s[(it - 1) & 1];
            ra.qslot_in = (int)((it - 1) % 3);
            grid = grid_queue;
        } else {
            ra.qin = t.vlist.as<uint32_t>();
            ra.sin = t.vscan.as<uint32_t>();
            ra.qslot_in = -1;
        }
        if (mode == M_QUEUE) {
            ra.qout = fs.q[it & 1];
            ra.sout = fs.s[it & 1];
            ra.qslot_out = (int)(it % 3);
        }
        void* args[] = {&ra};
        EEG_HIP(hipLaunchKernel((const void*)kfn, dim3(grid), dim3(block), args, 0, s));
        return EEG_OK;
    };
    uint32_t it = 1, chunk = 2;
    Ctl hc;
    memset(&hc, 0, sizeof(hc));
    eeg_status rc = EEG_OK;
    while (true) {
        for (uint32_t k = 0; k < chunk && it <= max_iters; ++k, ++it) EEG_TRY(launch(it));
        EEG_HIP(hipMemcpyAsync(&hc, fs.ctl, sizeof(Ctl), hipMemcpyDeviceToHost, s));
        EEG_HIP(hipStreamSynchronize(s));
        if (hc.err || hc.done) break;
        if (it > max_iters) {
            rc = fail(EEG_ERR_NOT_CONVERGED, "no fixpoint after %u iterations", max_iters);
            break;
        }
        chunk = std::min<uint32_t>(chunk * 2, 16);
    }
    if (rc == EEG_OK && hc.err) rc = fail(EEG_ERR_DEVICE, "frontier queue overflow (err=%u)", hc.err);
    if (st) {
        st->iterations = hc.iters;
        st->launches = it - 1;
        st->edges_relaxed = hc.stat[2];
        st->vnodes_processed = hc.stat[1];
        st->vertices_touched = hc.stat[0];
        st->label_updates = hc.stat[3];
        st->ms_total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_host0).count();
        st->ms_traversal = st->ms_total;
    }
    return rc;
}

}  // namespace eeg
