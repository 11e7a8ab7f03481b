// Batched traversals: many sources over one HBM-resident graph in one launch sequence.
//
//  BFS        multi-source BFS with 64-bit visit masks (bit-parallel BFS, Then et al., "The More
//             the Merrier: Efficient Multi-Source Graph Traversal", VLDB 2015), push form:
//             every frontier vertex ORs its new bits into each neighbour's `next` mask where the
//             neighbour has not seen them; a finalize pass turns next & ~seen into the next
//             frontier and writes the level of every newly seen (vertex, query) pair.
//  SSSP/SSWP  label planes stored vertex-major (B labels of a vertex are contiguous): a
//             wavefront takes one edge and 64 queries, so the source and destination label
//             reads and the atomics are contiguous 256-byte accesses.  A per-vertex change mask
//             selects the queries a vertex still has to push.
// Frontiers are compacted vertex queues with the Tigr virtual-node prefix, as in relax.hip,
// and the same three-slot control rotation lets launches run ahead of the host.
// Every answer equals the single-query run: the fixpoints are unique.
#include <algorithm>
#include <chrono>
#include <cstring>
#include <vector>

#include "batch.hpp"
#include "bytes8.hpp"
#include "frontier.hpp"

namespace eeg {
namespace {

constexpr uint32_t MAXW = BATCH_MAXW;

unsigned grid_n(uint64_t work, unsigned cap = 16384) {
    uint64_t g = (work + WG - 1) / WG;
    if (g == 0) g = 1;
    return (unsigned)(g > cap ? cap : g);
}


// Replayed graphs: the per-iteration arguments come from device memory (see MsArgs).
__device__ __forceinline__ void ms_resolve(MsArgs& a, bool finalize) {
    if (!a.dev_it) return;
    const uint32_t lvl = a.dev_it[0];
    const bool odd = ((lvl - 1) & 1) != 0;  // input frontier = buffer (level-1)&1
    a.level = lvl;
    a.stampv = a.dev_it[1] + lvl;
    a.launch = finalize ? 2 * lvl : 2 * lvl - 1;
    a.qin = odd ? a.qb : a.qa;
    a.sin = odd ? a.sb : a.sa;
    a.qout = odd ? a.qa : a.qb;
    a.sout = odd ? a.sa : a.sb;
}
__device__ __forceinline__ void plane_resolve(PlaneArgs& a) {
    if (!a.dev_it) return;
    const uint32_t it = a.dev_it[0];
    const bool odd = ((it - 1) & 1) != 0;
    a.it = it;
    a.stampv = a.dev_it[1] + it;
    a.qin = odd ? a.qb : a.qa;
    a.sin = odd ? a.sb : a.sa;
    a.qout = odd ? a.qa : a.qb;
    a.sout = odd ? a.sa : a.sb;
}

// ---------------------------------------------------------------- multi-source BFS
template <int K>
// At most 168 VGPRs (3 waves per SIMD, 8 bytes of spill per lane): -9% per pass against 200
// VGPRs and 2 waves on RMAT-24 x 256; 4 waves spilled 136 bytes.
__global__ __launch_bounds__(WG, 3) void k_ms_push(MsArgs args) {
    MsArgs a = args;
    ms_resolve(a, false);
    constexpr int VPR = WG / K, TV = TILE_EDGES / K, ROUNDS = TILE_EDGES / WG, EMAX = TV + 2;
    __shared__ uint32_t sS[EMAX];
    __shared__ uint64_t sRow[EMAX];
    __shared__ uint32_t sDeg[EMAX];
    __shared__ uint32_t sV[EMAX];
    __shared__ uint16_t sOwn[TV];
    __shared__ uint32_t sQ[TILE_EDGES];
    __shared__ uint32_t sMisc[8];
    Ctl* ctl = a.ctl;
    if (a.act && blockIdx.x == 0 && threadIdx.x < MAXW) a.act[(a.level % 4) * MAXW + threadIdx.x] = 0;
    if (ctl->done) return;
    const unsigned long long pk = ctl->qpack[(a.launch - 1) % 3];
    const uint32_t nE = (uint32_t)(pk & 0xffffffffull), T = (uint32_t)(pk >> 32);
    const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
    if (nE == 0 || T == 0) {
        if (lead) ctl->done = 1;
        return;
    }
    if (lead) {
        ctl->qpack[(a.launch + 1) % 3] = 0;
        ctl->iters = a.level;
    }
    const uint32_t ntiles = (T + TV - 1) / TV;
    const uint32_t c0 = (uint32_t)((uint64_t)blockIdx.x * ntiles / gridDim.x);
    const uint32_t c1 = (uint32_t)((uint64_t)(blockIdx.x + 1) * ntiles / gridDim.x);
    const int grp = threadIdx.x / K, lj = threadIdx.x % K;
    if (c0 >= c1) return;
    uint32_t e_start = wg_search(a.sin, nE, c0 * TV);
    for (uint32_t tile = c0; tile < c1; ++tile) {
        const uint32_t t0 = tile * TV, t1 = min(t0 + (uint32_t)TV, T);
        for (int k = threadIdx.x; k < EMAX; k += WG) {
            const uint64_t i = (uint64_t)e_start + k;
            sS[k] = (i < nE) ? a.sin[i] : 0xffffffffu;
        }
        if (threadIdx.x == 0) {
            sMisc[0] = 1;
            sMisc[1] = 1;
            sMisc[2] = 0;
        }
        __syncthreads();
        for (int k = threadIdx.x; k < EMAX; k += WG) {
            const uint32_t cur = sS[k], nxt = (k + 1 < EMAX) ? sS[k + 1] : 0xffffffffu;
            if (cur <= t1 - 1 && nxt > t1 - 1) sMisc[0] = k + 1;
            if (cur <= t1 && nxt > t1) sMisc[1] = k + 1;
        }
        __syncthreads();
        const int nent = (int)sMisc[0], nnext = (int)sMisc[1];
        for (int k = threadIdx.x; k < nent; k += WG) {
            const uint32_t v = a.qin[e_start + k];
            sV[k] = v;
            sRow[k] = a.row[v];
            sDeg[k] = (uint32_t)(a.row[v + 1] - a.row[v]);
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
        // edge loads of every round first; then, word by word, every visit and seen load of
        // all rounds before any atomic, so each lane keeps ROUNDS requests in flight
        uint32_t ru[ROUNDS], rv[ROUNDS];
        bool rok[ROUNDS], got[ROUNDS];
#pragma unroll
        for (int r = 0; r < ROUNDS; ++r) {
            const int slot = r * VPR + grp;
            const uint32_t t = t0 + slot;
            rok[r] = false;
            got[r] = false;
            ru[r] = 0;
            rv[r] = 0;
            if (t < t1) {
                const int k = sOwn[slot];
                const uint32_t eoff = (t - sS[k]) * K + lj;
                if (eoff < sDeg[k]) {
                    rok[r] = true;
                    rv[r] = sV[k];
                    ru[r] = __builtin_nontemporal_load(&a.col[sRow[k] + eoff]);
                }
            }
        }
        // With interleaved records (R != W) the destination's `next` words and stamp share the
        // cache line of its `seen` words, so a snapshot of them comes with the seen load, and the
        // atomics are issued only for bits and stamps it does not show yet.  Bits of `next` are
        // only added during a push and a stamp only moves to this level's value, so a stale
        // snapshot costs an extra atomic and never a lost bit.  (The atomics are what this
        // kernel waits on: they are performed beyond the L2.)
        const bool snap = a.R != a.W;
        uint32_t sp[ROUNDS];
#pragma unroll
        for (int r = 0; r < ROUNDS; ++r) sp[r] = 0;
        if ((a.W & 1) == 0) {
            // even W: the masks of a vertex are 16-byte aligned, two words per load
            // the masks of half the rounds at a time: fewer registers (with the launch bound, 3
            // waves per SIMD instead of 2), and the atomics this kernel waits on overlap more
            constexpr int HR = ROUNDS / 2;
#pragma unroll
            for (int h = 0; h < ROUNDS; h += HR)
            for (uint32_t w = 0; w < a.W; w += 2) {
                ulonglong2 mv[HR], ms[HR], mn[HR];
#pragma unroll
                for (int j = 0; j < HR; ++j) {
                    const int r = h + j;
                    mv[j] = rok[r] ? *(const ulonglong2*)&a.visit[(uint64_t)rv[r] * a.R + w] : ulonglong2{0, 0};
                }
#pragma unroll
                for (int j = 0; j < HR; ++j) {
                    const int r = h + j;
                    const bool any = (mv[j].x | mv[j].y) != 0;
                    const uint64_t i = (uint64_t)ru[r] * a.R + w;
                    ms[j] = any ? *(const ulonglong2*)&a.seen[i] : ulonglong2{~0ull, ~0ull};
                    mn[j] = (any && snap) ? *(const ulonglong2*)&a.next[i] : ulonglong2{0, 0};
                    if (w == 0 && any && snap) sp[r] = a.stamp[(uint64_t)ru[r] * a.sst];
                }
#pragma unroll
                for (int j = 0; j < HR; ++j) {
                    const int r = h + j;
                    const uint64_t add0 = mv[j].x & ~ms[j].x & ~mn[j].x, add1 = mv[j].y & ~ms[j].y & ~mn[j].y;
                    unsigned long long* nx = (unsigned long long*)&a.next[(uint64_t)ru[r] * a.R + w];
                    if (add0) atomicOr(nx, (unsigned long long)add0);
                    if (add1) atomicOr(nx + 1, (unsigned long long)add1);
                    got[r] = got[r] || (add0 | add1) != 0;
                }
            }
        } else {
            for (uint32_t w = 0; w < a.W; ++w) {
                uint64_t mv[ROUNDS], ms[ROUNDS], mn[ROUNDS];
#pragma unroll
                for (int r = 0; r < ROUNDS; ++r) mv[r] = rok[r] ? a.visit[(uint64_t)rv[r] * a.R + w] : 0ull;
#pragma unroll
                for (int r = 0; r < ROUNDS; ++r) {
                    const uint64_t i = (uint64_t)ru[r] * a.R + w;
                    ms[r] = mv[r] ? a.seen[i] : ~0ull;
                    mn[r] = (mv[r] && snap) ? a.next[i] : 0ull;
                    if (w == 0 && mv[r] && snap) sp[r] = a.stamp[(uint64_t)ru[r] * a.sst];
                }
#pragma unroll
                for (int r = 0; r < ROUNDS; ++r) {
                    const uint64_t add = mv[r] & ~ms[r] & ~mn[r];
                    if (add) {
                        atomicOr((unsigned long long*)&a.next[(uint64_t)ru[r] * a.R + w], (unsigned long long)add);
                        got[r] = true;
                    }
                }
            }
        }
        // a destination that gained bits is listed once per level (its stamp)
#pragma unroll
        for (int r = 0; r < ROUNDS; ++r) {
            const uint32_t u = ru[r];
            const bool candq = got[r] && sp[r] != a.stampv &&
                               atomicExch(&a.stamp[(uint64_t)u * a.sst], a.stampv) != a.stampv;
            const uint64_t mask = __ballot(candq);
            if (mask) {
                const int leader = __ffsll((unsigned long long)mask) - 1;
                uint32_t base = 0;
                if (lane_id() == leader) base = atomicAdd(&sMisc[2], (uint32_t)__popcll(mask));
                base = __shfl(base, leader, WAVE);
                if (candq) sQ[base + (uint32_t)__popcll(mask & lanemask_lt())] = u;
            }
        }
        __syncthreads();
        const uint32_t cnt = sMisc[2];
        if (cnt > 0) {
            if (threadIdx.x == 0) {
                const unsigned long long old = atomicAdd(&ctl->qpack[a.launch % 3], (unsigned long long)cnt);
                sMisc[3] = (uint32_t)(old & 0xffffffffull);
            }
            __syncthreads();
            const uint32_t bq = sMisc[3];
            if ((uint64_t)bq + cnt <= a.qcap) {
                for (uint32_t i = threadIdx.x; i < cnt; i += WG) a.cand[bq + i] = sQ[i];
            } else if (threadIdx.x == 0) {
                atomicOr(&ctl->err, 4u);
            }
        }
        __syncthreads();
        e_start += (uint32_t)(nnext - 1);
    }
}

// Finalize: new = next & ~seen; seen |= new; visit = new; next = 0; record levels; queue the
// vertices with new bits (and out-edges) as the next frontier.  A thread per candidate updates
// the masks; the levels are then written by whole wavefronts, lane = query, one contiguous row
// store per (candidate, word) with new bits, instead of one scattered store per new bit.
__global__ __launch_bounds__(WG) void k_ms_finalize(MsArgs args) {
    MsArgs a = args;
    ms_resolve(a, true);
    __shared__ uint64_t sScan[WG / WAVE + 1];
    __shared__ unsigned long long sBase;
    __shared__ unsigned long long sAct[MAXW];
    __shared__ uint64_t sNew[MAXW][WG];
    __shared__ uint32_t sU[WG];
    Ctl* ctl = a.ctl;
    if (ctl->done) return;
    if (threadIdx.x < MAXW) sAct[threadIdx.x] = 0;
    __syncthreads();
    const uint32_t count = (uint32_t)(ctl->qpack[(a.launch - 1) % 3] & 0xffffffffull);
    const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
    if (count == 0) {
        if (lead) ctl->done = 1;
        return;
    }
    if (lead) ctl->qpack[(a.launch + 1) % 3] = 0;
    const int wv = threadIdx.x / WAVE, lane = lane_id();
    for (uint64_t base = (uint64_t)blockIdx.x * WG; base < count; base += (uint64_t)gridDim.x * WG) {
        const uint64_t i = base + threadIdx.x;
        uint64_t mine = 0;
        uint32_t u = 0, nv = 0;
        bool any = false;
        if (i < count) {
            u = a.cand[i];
            const bool pushes = a.vnc[u] != 0;
            for (uint32_t w = 0; w < a.W; ++w) {
                const uint64_t idx = (uint64_t)u * a.R + w;
                const uint64_t nw = a.next[idx] & ~a.seen[idx];
                a.next[idx] = 0;
                a.visit[idx] = nw;
                sNew[w][threadIdx.x] = nw;
                if (nw) {
                    a.seen[idx] |= nw;
                    any = true;
                    if (a.act && pushes) atomicOr(&sAct[w], (unsigned long long)nw);
                }
            }
            if (any) {
                nv = a.vnc[u];
                if (nv) mine = 1ull | ((uint64_t)nv << 32);
            }
        }
        sU[threadIdx.x] = any ? u : 0xffffffffu;
        uint64_t tot;
        const uint64_t pre = block_excl_scan<uint64_t>(mine, sScan, &tot);
        if (threadIdx.x == 0) sBase = tot ? atomicAdd(&ctl->qpack[a.launch % 3], (unsigned long long)tot) : 0ull;
        __syncthreads();
        if (mine) {
            const uint32_t pos = (uint32_t)(sBase & 0xffffffffull) + (uint32_t)(pre & 0xffffffffull);
            if (pos < a.qcap) {
                a.qout[pos] = u;
                a.sout[pos] = (uint32_t)(sBase >> 32) + (uint32_t)(pre >> 32);
            } else {
                atomicOr(&ctl->err, 8u);
            }
        }
        // levels of this wavefront's candidates
        for (int c = 0; c < WAVE; ++c) {
            const int slot = wv * WAVE + c;
            const uint32_t cu = sU[slot];
            if (cu == 0xffffffffu) continue;
            for (uint32_t w = 0; w < a.W; ++w) {
                const uint64_t nw = sNew[w][slot];
                const uint32_t q = w * 64 + (uint32_t)lane;
                if (((nw >> lane) & 1ull) && q < a.B) a.dist[(uint64_t)cu * a.B + q] = a.level;
            }
        }
        __syncthreads();
    }
    if (a.act && threadIdx.x < a.W && sAct[threadIdx.x])
        atomicOr((unsigned long long*)&a.act[(a.level % 4) * MAXW + threadIdx.x], sAct[threadIdx.x]);
}

// ---------------------------------------------------------------- batched SSSP / SSWP
// Label widths of the planes (LM): 32-bit labels (LM_U32, lane = query), or narrow labels
// packed into 32-bit lane words (four 8-bit or two 16-bit labels), so one wavefront covers a
// vertex's 256 queries in one or two row passes:
//   LM_W8   SSWP with every weight <= 254: u8 labels, 255 = INF (the source), 0 = unreached.
//           A widest-path label is a weight, INF or 0, so nothing is lost.
//   LM_D16  SSSP: u16 distances, 0xffff = INF.  A finite candidate >= 0xffff sets Ctl::err
//           bit 64 and is dropped; labels below it are exact (a vertex's shortest path only
//           passes smaller distances), and the caller reruns the pass on 32-bit planes.
//   LM_D8   SSSP with small weights: u8 distances, 0xff = INF.  A finite candidate >= 0xff is
//           dropped; when the destination is still unreached in that query the query's bit is
//           set in PlaneArgs::ovq.  A query without the bit is exact: every vertex at distance
//           <= 254 is found as above, and a vertex at a finite distance >= 255 would have a
//           first such vertex y on its shortest path, whose predecessor x (distance <= 254,
//           final and pushed) offers y a candidate >= 255 while y is unreached (a label is
//           never below the true distance).  The caller answers the flagged queries again
//           with wider labels.
// Improvements of a lane word are compare-and-swap loops: labels only improve, so every
// failed swap means another improvement landed and the loop ends.
enum LabelMode : int { LM_U32 = 0, LM_W8 = 1, LM_D16 = 2, LM_D8 = 3 };
constexpr uint32_t PLANE_ERR_NARROW = 64u;

template <int LM>
struct Packed;
// Widest-path labels, four per word (bytes8.hpp).
template <>
struct Packed<LM_W8> : Bytes8Widest {};
// 16-bit distances, two per word (bytes8.hpp).
template <>
struct Packed<LM_D16> : Bytes16Dist {};
// 8-bit distances, four per word (bytes8.hpp).
template <>
struct Packed<LM_D8> : Bytes8Dist {};

// Dense variant: one wavefront per virtual node.  LM_U32: every pending 64-query word handled
// with lane = query, so the label reads and atomics of a wavefront are contiguous 256-byte
// rows.  Narrow labels: lane word x holds queries x*QPL .. x*QPL+QPL-1, so one row read (two
// for 256 16-bit distances) covers all pending queries of the vertex.  Best when many queries are in flight together; the
// two-pass kernel below is better when few are.
template <int K, int LM>
// 32-bit labels: 72 VGPRs at K <= 8 (7 waves per SIMD instead of 6 at 77, no spills): the row
// reads are latency-bound, so the extra resident wave is worth more (-7 to -8% per pass on
// RMAT-24).  Narrow labels: at most 64 VGPRs, 8 waves per SIMD (from 4-5 waves at 85-118
// VGPRs: 25-40% faster passes on RMAT-24 x 256).
__global__ __launch_bounds__(WG, K <= 8 ? (LM == LM_U32 ? 7 : 8) : 1) void k_plane_push_wave(
    PlaneArgs args) {
    PlaneArgs a = args;
    plane_resolve(a);
    constexpr int TV = TILE_EDGES / K, EMAX = TV + 2, NW = WG / WAVE;
    __shared__ uint32_t sS[EMAX];
    __shared__ uint64_t sRow[EMAX];
    __shared__ uint32_t sDeg[EMAX];
    __shared__ uint32_t sV[EMAX];
    __shared__ uint32_t sQ[TILE_EDGES];
    __shared__ uint32_t sMisc[8];
    __shared__ uint64_t sScan[WG / WAVE + 1];
    __shared__ uint64_t sOv[MAXW];  // 8-bit distances: flagged queries, which stop pushing
    Ctl* ctl = a.ctl;
    if (a.act && blockIdx.x == 0 && threadIdx.x < MAXW) a.act[(a.it % 4) * MAXW + threadIdx.x] = 0;
    if (ctl->done) return;
    const unsigned long long pk = ctl->qpack[(a.it - 1) % 3];
    const uint32_t nE = (uint32_t)(pk & 0xffffffffull), T = (uint32_t)(pk >> 32);
    const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
    if (nE == 0 || T == 0) {
        if (lead) ctl->done = 1;
        return;
    }
    if (lead) {
        ctl->qpack[(a.it + 1) % 3] = 0;
        ctl->iters = a.it;
    }
    // Tile width: TV virtual nodes, or fewer when the frontier is small, so its virtual nodes
    // spread over the whole grid (a wavefront takes one virtual node at a time, so a few
    // full tiles would leave most of the GPU idle and the rest walking 64 nodes each).  A
    // tile of tv virtual nodes meets at most tv + 2 entries (every entry has one or more).
    const uint32_t tv = max((uint32_t)NW, min((uint32_t)TV, (T + gridDim.x - 1) / gridDim.x));
    const int emax = min(EMAX, (int)tv + 2);
    const uint32_t ntiles = (T + tv - 1) / tv;
    const uint32_t c0 = (uint32_t)((uint64_t)blockIdx.x * ntiles / gridDim.x);
    const uint32_t c1 = (uint32_t)((uint64_t)(blockIdx.x + 1) * ntiles / gridDim.x);
    if (c0 >= c1) return;
    const int wv = threadIdx.x / WAVE, lane = lane_id();
    uint32_t e_start = wg_search(a.sin, nE, c0 * tv);
    for (uint32_t tile = c0; tile < c1; ++tile) {
        const uint32_t t0 = tile * tv, t1 = min(t0 + tv, T);
        for (int k = threadIdx.x; k < emax; k += WG) {
            const uint64_t i = (uint64_t)e_start + k;
            sS[k] = (i < nE) ? a.sin[i] : 0xffffffffu;
        }
        if (threadIdx.x == 0) {
            sMisc[0] = 1;
            sMisc[1] = 1;
            sMisc[2] = 0;
        }
        if constexpr (LM == LM_D8) {
            // a flagged query is answered again wider, so its further pushes are wasted work
            // (its column only feeds itself); a stale copy only costs some of that work
            if (threadIdx.x < MAXW) sOv[threadIdx.x] = __builtin_nontemporal_load(&a.ovq[threadIdx.x]);
        }
        __syncthreads();
        for (int k = threadIdx.x; k < emax; k += WG) {
            const uint32_t cur = sS[k], nxt = (k + 1 < emax) ? sS[k + 1] : 0xffffffffu;
            if (cur <= t1 - 1 && nxt > t1 - 1) sMisc[0] = k + 1;
            if (cur <= t1 && nxt > t1) sMisc[1] = k + 1;
        }
        __syncthreads();
        const int nent = (int)sMisc[0], nnext = (int)sMisc[1];
        for (int k = threadIdx.x; k < nent; k += WG) {
            const uint32_t v = a.qin[e_start + k];
            sV[k] = v;
            sRow[k] = a.row[v];
            sDeg[k] = (uint32_t)(a.row[v + 1] - a.row[v]);
        }
        __syncthreads();
        // virtual nodes of this tile, one wavefront at a time: lanes first fetch the virtual
        // node's <= K edges (one per lane) and broadcast them, then for every 64-query word
        // with pending changes each lane (one query) issues all K destination-label reads
        // before any atomic, so K requests per lane are in flight
        for (uint32_t t = t0 + wv; t < t1; t += NW) {
            int lo = 0, hi = nent;
            while (hi - lo > 1) {
                const int mid = (lo + hi) >> 1;
                if (sS[mid] <= t) lo = mid; else hi = mid;
            }
            const int k = lo;
            const uint32_t v = sV[k];
            const uint32_t e0 = (t - sS[k]) * K;
            const uint32_t ne = min((uint32_t)K, sDeg[k] - e0);
            uint32_t my_u = 0, my_w = 0;
            if ((uint32_t)lane < ne) {
                my_u = a.col[sRow[k] + e0 + lane];
                my_w = a.wgt[sRow[k] + e0 + lane];
            }
            if constexpr (LM != LM_U32) {
                using P = Packed<LM>;
                constexpr uint32_t QPL = P::QPL;     // queries per lane word
                constexpr uint32_t GL = 64 / QPL;    // lane words per 64-query mask word
                constexpr uint32_t GM = (1u << QPL) - 1;
                uint32_t* plane = a.lab;
                const uint32_t nl = a.B / QPL;       // lane words per row
                for (uint32_t x0 = 0; x0 < nl; x0 += WAVE) {  // passes of 64 lane words
                    const uint32_t x = x0 + (uint32_t)lane;
                    uint32_t nib = 0;
                    if (x < nl) nib = (uint32_t)(a.chg[(uint64_t)v * a.W + x / GL] >> (QPL * (x % GL))) & GM;
                    if constexpr (LM == LM_D8) nib &= ~(uint32_t)(sOv[x / GL] >> (QPL * (x % GL)));
                    if (!__ballot(nib != 0)) continue;  // the same answer in every lane
                    const uint32_t lv = nib ? plane[(uint64_t)v * nl + x] : 0u;
                    uint32_t ovf = 0;  // D16: candidates past 16 bits; D8: flagged queries
                    // destinations and weights are taken from the edge lanes again where
                    // needed (outside divergent code) rather than kept in K registers each
                    uint32_t ecur[K];
#pragma unroll
                    for (int e = 0; e < K; ++e) {  // every destination row read before any update
                        const uint32_t u = __builtin_amdgcn_readlane(my_u, e);
                        ecur[e] = (nib && (uint32_t)e < ne) ? plane[(uint64_t)u * nl + x] : 0u;
                    }
                    uint32_t imp[K];
#pragma unroll
                    for (int e = 0; e < K; ++e) {
                        imp[e] = 0;
                        const uint32_t u = __builtin_amdgcn_readlane(my_u, e);
                        const uint32_t wt = __builtin_amdgcn_readlane(my_w, e);
                        if (nib && (uint32_t)e < ne) {
                            uint32_t ov = 0;
                            const uint32_t cd = P::cand(lv, wt, nib, ov);
                            uint32_t c = ecur[e];
                            if constexpr (LM == LM_D8) {
                                if (ov) ovf |= ov & P::unreached(c);
                            } else {
                                ovf |= ov;
                            }
                            uint32_t nv = P::merge(c, cd);
                            if (nv != c) {
                                uint32_t* p = &plane[(uint64_t)u * nl + x];
                                while (true) {
                                    const uint32_t old = atomicCAS(p, c, nv);
                                    if (old == c) {
                                        imp[e] = P::improved(c, cd);
                                        break;
                                    }
                                    c = old;
                                    nv = P::merge(c, cd);
                                    if (nv == c) break;
                                }
                            }
                        }
                    }
                    if constexpr (LM == LM_D8) {
                        if (ovf)
                            atomicOr((unsigned long long*)&a.ovq[x / GL], (unsigned long long)ovf << (QPL * (x % GL)));
                    } else if (ovf) {
                        atomicOr(&ctl->err, PLANE_ERR_NARROW);
                    }
#pragma unroll
                    for (int e = 0; e < K; ++e) {
                        const uint32_t u = __builtin_amdgcn_readlane(my_u, e);
                        if (!__ballot(imp[e] != 0)) continue;
                        // only vertices with out-edges carry changes and are queued (each
                        // queued entry owns >= 1 virtual node); a sink's label is final
                        if (!a.vnc[u]) continue;
                        // the improved queries of each 64-query word: its GL lane words
                        // are neighbouring lanes
                        uint64_t word = (uint64_t)imp[e] << (QPL * (x % GL));
#pragma unroll
                        for (uint32_t off = GL / 2; off >= 1; off >>= 1) word |= __shfl_xor(word, (int)off, WAVE);
                        if ((x % GL) == 0 && word)
                            atomicOr((unsigned long long*)&a.nchg[(uint64_t)u * a.W + x / GL], (unsigned long long)word);
                        if (lane == 0) {
                            if (a.bitmap) {
                                atomicOr(&a.bitmap[u >> 5], 1u << (u & 31u));
                            } else if (atomicExch(&a.stamp[u], a.stampv) != a.stampv) {
                                const uint32_t pos = atomicAdd(&sMisc[2], 1u);
                                if (pos < TILE_EDGES) sQ[pos] = u;
                                else atomicOr(&ctl->err, 16u);
                            }
                        }
                    }
                }
                continue;
            }
            for (uint32_t w = 0; w < a.W; ++w) {
                const uint64_t m = a.chg[(uint64_t)v * a.W + w];
                if (!m) continue;
                const uint32_t q = w * 64 + (uint32_t)lane;
                const bool mine = (m >> lane) & 1ull;
                const uint32_t lv = mine ? a.lab[(uint64_t)v * a.B + q] : 0u;
                const bool ok = mine && (a.widest ? lv > 0 : lv < INF);
                uint32_t eu[K], ecand[K], ecur[K];
#pragma unroll
                for (int e = 0; e < K; ++e) {
                    eu[e] = __shfl(my_u, e, WAVE);
                    const uint32_t wt = __shfl(my_w, e, WAVE);
                    if (a.widest) ecand[e] = lv < wt ? lv : wt;
                    else {
                        const uint64_t sm = (uint64_t)lv + wt;
                        ecand[e] = sm < INF ? (uint32_t)sm : INF;
                    }
                    ecur[e] = (ok && (uint32_t)e < ne) ? a.lab[(uint64_t)eu[e] * a.B + q] : 0u;
                }
                bool eup[K];
#pragma unroll
                for (int e = 0; e < K; ++e) {
                    eup[e] = false;
                    if (ok && (uint32_t)e < ne && (a.widest ? ecand[e] > ecur[e] : ecand[e] < ecur[e])) {
                        uint32_t* p = &a.lab[(uint64_t)eu[e] * a.B + q];
                        const uint32_t old = a.widest ? atomicMax(p, ecand[e]) : atomicMin(p, ecand[e]);
                        eup[e] = a.widest ? ecand[e] > old : ecand[e] < old;
                    }
                }
#pragma unroll
                for (int e = 0; e < K; ++e) {
                    const uint64_t um = __ballot(eup[e]);
                    if (um && lane == 0 && a.vnc[eu[e]]) {
                        const uint32_t u = eu[e];
                        // only vertices with out-edges carry changes and are queued (each
                        // queued entry owns >= 1 virtual node); a sink's label is final
                        atomicOr((unsigned long long*)&a.nchg[(uint64_t)u * a.W + w], (unsigned long long)um);
                        if (a.bitmap) {
                            atomicOr(&a.bitmap[u >> 5], 1u << (u & 31u));
                        } else if (atomicExch(&a.stamp[u], a.stampv) != a.stampv) {
                            const uint32_t pos = atomicAdd(&sMisc[2], 1u);
                            if (pos < TILE_EDGES) sQ[pos] = u;
                            else atomicOr(&ctl->err, 16u);
                        }
                    }
                }
            }
        }
        __syncthreads();
        // queue the improved vertices with their virtual-node counts
        const uint32_t cnt = min(sMisc[2], (uint32_t)TILE_EDGES);
        if (cnt > 0) {
            const uint32_t per = (cnt + WG - 1) / WG, b0 = threadIdx.x * per;
            uint64_t loc = 0;
            for (uint32_t i = 0; i < per; ++i)
                if (b0 + i < cnt) loc += 1ull | ((uint64_t)a.vnc[sQ[b0 + i]] << 32);
            uint64_t tot;
            uint64_t pre = block_excl_scan<uint64_t>(loc, sScan, &tot);
            if (threadIdx.x == 0) {
                const unsigned long long old = atomicAdd(&ctl->qpack[a.it % 3], (unsigned long long)tot);
                sMisc[3] = (uint32_t)(old & 0xffffffffull);
                sMisc[4] = (uint32_t)(old >> 32);
            }
            __syncthreads();
            uint32_t bq = sMisc[3] + (uint32_t)(pre & 0xffffffffull);
            uint32_t bv = sMisc[4] + (uint32_t)(pre >> 32);
            for (uint32_t i = 0; i < per; ++i) {
                if (b0 + i >= cnt) break;
                const uint32_t u = sQ[b0 + i], nv = a.vnc[u];
                if (bq < a.qcap) {
                    a.qout[bq] = u;
                    a.sout[bq] = bv;
                } else {
                    atomicOr(&ctl->err, 32u);
                }
                ++bq;
                bv += nv;
            }
        }
        __syncthreads();
        e_start += (uint32_t)(nnext - 1);
    }
}

// One tile of the frontier's virtual nodes in two passes.
//  Sparse pass: K lanes per virtual node, one edge per lane, as in the single-query kernel.
//    For every 64-query word with at most SPARSE_MAX pending queries the lane relaxes its edge
//    for each of them (source and destination label reads issued together).  A virtual node
//    with a denser word is listed for the dense pass.
//  Dense pass: one wavefront per listed virtual node.  Lanes fetch its <= K edges and
//    broadcast them; then for every dense word each lane (one query) issues all K destination
//    label reads before any atomic, so the reads of a wavefront are contiguous 256-byte rows.
constexpr uint32_t SPARSE_MAX = 8;
// queries in flight above which the dense variant is launched
constexpr uint32_t PLANE_DENSE_FROM = 16;

template <int K>
// At most 168 VGPRs at K <= 8 (3 waves per SIMD instead of 2 at ~193, ~116 B of spill): -10 to -12%
// per pass for 8 and 16 queries on RMAT-24; 4 waves were no faster than 2.
__global__ __launch_bounds__(WG, K <= 8 ? 3 : 1) void k_plane_push(PlaneArgs args) {
    PlaneArgs a = args;
    plane_resolve(a);
    constexpr int TV = TILE_EDGES / K, EMAX = TV + 2, NW = WG / WAVE, VPR = WG / K;
    constexpr int ROUNDS = TILE_EDGES / WG;
    __shared__ uint32_t sS[EMAX];
    __shared__ uint64_t sRow[EMAX];
    __shared__ uint32_t sDeg[EMAX];
    __shared__ uint32_t sV[EMAX];
    __shared__ uint16_t sOwn[TV];
    __shared__ uint16_t sDense[TV];
    __shared__ uint32_t sQ[TILE_EDGES];
    __shared__ uint32_t sMisc[8];
    __shared__ uint64_t sScan[WG / WAVE + 1];
    Ctl* ctl = a.ctl;
    if (a.act && blockIdx.x == 0 && threadIdx.x < MAXW) a.act[(a.it % 4) * MAXW + threadIdx.x] = 0;
    if (ctl->done) return;
    const unsigned long long pk = ctl->qpack[(a.it - 1) % 3];
    const uint32_t nE = (uint32_t)(pk & 0xffffffffull), T = (uint32_t)(pk >> 32);
    const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
    if (nE == 0 || T == 0) {
        if (lead) ctl->done = 1;
        return;
    }
    if (lead) {
        ctl->qpack[(a.it + 1) % 3] = 0;
        ctl->iters = a.it;
    }
    const uint32_t ntiles = (T + TV - 1) / TV;
    const uint32_t c0 = (uint32_t)((uint64_t)blockIdx.x * ntiles / gridDim.x);
    const uint32_t c1 = (uint32_t)((uint64_t)(blockIdx.x + 1) * ntiles / gridDim.x);
    if (c0 >= c1) return;
    const int wv = threadIdx.x / WAVE, lane = lane_id();
    const int grp = threadIdx.x / K, lj = threadIdx.x % K;
    uint32_t e_start = wg_search(a.sin, nE, c0 * TV);
    // an improvement of query q at u: change bit, and u queued once per iteration
    auto improved = [&](uint32_t u, uint32_t w, uint64_t bits) {
        if (!a.vnc[u]) return;  // a vertex without out-edges has nothing to push
        atomicOr((unsigned long long*)&a.nchg[(uint64_t)u * a.W + w], (unsigned long long)bits);
        if (atomicExch(&a.stamp[u], a.stampv) != a.stampv) {
            const uint32_t pos = atomicAdd(&sMisc[2], 1u);
            if (pos < TILE_EDGES) sQ[pos] = u;
            else atomicOr(&ctl->err, 16u);
        }
    };
    for (uint32_t tile = c0; tile < c1; ++tile) {
        const uint32_t t0 = tile * TV, t1 = min(t0 + (uint32_t)TV, T);
        for (int k = threadIdx.x; k < EMAX; k += WG) {
            const uint64_t i = (uint64_t)e_start + k;
            sS[k] = (i < nE) ? a.sin[i] : 0xffffffffu;
        }
        if (threadIdx.x == 0) {
            sMisc[0] = 1;
            sMisc[1] = 1;
            sMisc[2] = 0;
            sMisc[5] = 0;
        }
        __syncthreads();
        for (int k = threadIdx.x; k < EMAX; k += WG) {
            const uint32_t cur = sS[k], nxt = (k + 1 < EMAX) ? sS[k + 1] : 0xffffffffu;
            if (cur <= t1 - 1 && nxt > t1 - 1) sMisc[0] = k + 1;
            if (cur <= t1 && nxt > t1) sMisc[1] = k + 1;
        }
        __syncthreads();
        const int nent = (int)sMisc[0], nnext = (int)sMisc[1];
        for (int k = threadIdx.x; k < nent; k += WG) {
            const uint32_t v = a.qin[e_start + k];
            sV[k] = v;
            sRow[k] = a.row[v];
            sDeg[k] = (uint32_t)(a.row[v + 1] - a.row[v]);
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

        // ---- sparse pass: edge loads of every round first, then word by word
        uint32_t rv[ROUNDS], ru[ROUNDS], rw[ROUNDS];
        bool rin[ROUNDS], rhas[ROUNDS], rdense[ROUNDS];
#pragma unroll
        for (int r = 0; r < ROUNDS; ++r) {
            const int slot = r * VPR + grp;
            const uint32_t t = t0 + slot;
            rin[r] = t < t1;
            rhas[r] = false;
            rdense[r] = false;
            rv[r] = 0;
            ru[r] = 0;
            rw[r] = 0;
            if (rin[r]) {
                const int k = sOwn[slot];
                rv[r] = sV[k];
                const uint32_t eoff = (t - sS[k]) * K + lj;
                if (eoff < sDeg[k]) {
                    rhas[r] = true;
                    ru[r] = a.col[sRow[k] + eoff];
                    rw[r] = a.wgt[sRow[k] + eoff];
                }
            }
        }
        for (uint32_t w = 0; w < a.W; ++w) {
            uint64_t rm[ROUNDS];
#pragma unroll
            for (int r = 0; r < ROUNDS; ++r) rm[r] = rin[r] ? a.chg[(uint64_t)rv[r] * a.W + w] : 0ull;
#pragma unroll
            for (int r = 0; r < ROUNDS; ++r) {
                uint64_t m = rm[r];
                if (!m) continue;
                if ((uint32_t)__popcll((unsigned long long)m) > SPARSE_MAX) {
                    rdense[r] = true;
                    continue;
                }
                if (!rhas[r]) continue;
                const uint32_t u = ru[r], wt = rw[r];
                while (m) {
                    const uint32_t c = (uint32_t)(__ffsll((unsigned long long)m) - 1);
                    m &= m - 1;
                    const uint32_t q = w * 64 + c;
                    uint32_t* p = &a.lab[(uint64_t)u * a.B + q];
                    const uint32_t lv = a.lab[(uint64_t)rv[r] * a.B + q];
                    const uint32_t cur = *p;
                    if (a.widest ? lv == 0 : lv >= INF) continue;
                    uint32_t cand;
                    if (a.widest) cand = lv < wt ? lv : wt;
                    else {
                        const uint64_t sm = (uint64_t)lv + wt;
                        cand = sm < INF ? (uint32_t)sm : INF;
                    }
                    if (a.widest ? cand > cur : cand < cur) {
                        const uint32_t old = a.widest ? atomicMax(p, cand) : atomicMin(p, cand);
                        if (a.widest ? cand > old : cand < old) improved(u, w, 1ull << c);
                    }
                }
            }
        }
#pragma unroll
        for (int r = 0; r < ROUNDS; ++r)
            if (rdense[r] && lj == 0) sDense[atomicAdd(&sMisc[5], 1u)] = (uint16_t)(r * VPR + grp);
        __syncthreads();

        // ---- dense pass: one wavefront per listed virtual node
        const uint32_t nd = sMisc[5];
        for (uint32_t di = (uint32_t)wv; di < nd; di += NW) {
            const uint32_t slot = sDense[di];
            const uint32_t t = t0 + slot;
            const int k = sOwn[slot];
            const uint32_t v = sV[k];
            const uint32_t e0 = (t - sS[k]) * K;
            const uint32_t ne = min((uint32_t)K, sDeg[k] - e0);
            uint32_t my_u = 0, my_w = 0;
            if ((uint32_t)lane < ne) {
                my_u = a.col[sRow[k] + e0 + lane];
                my_w = a.wgt[sRow[k] + e0 + lane];
            }
            for (uint32_t w = 0; w < a.W; ++w) {
                const uint64_t m = a.chg[(uint64_t)v * a.W + w];
                if ((uint32_t)__popcll((unsigned long long)m) <= SPARSE_MAX) continue;  // sparse pass did it
                const uint32_t q = w * 64 + (uint32_t)lane;
                const bool mine = (m >> lane) & 1ull;
                const uint32_t lv = mine ? a.lab[(uint64_t)v * a.B + q] : 0u;
                const bool ok = mine && (a.widest ? lv > 0 : lv < INF);
                uint32_t eu[K], ecand[K], ecur[K];
#pragma unroll
                for (int e = 0; e < K; ++e) {
                    eu[e] = __shfl(my_u, e, WAVE);
                    const uint32_t wt = __shfl(my_w, e, WAVE);
                    if (a.widest) ecand[e] = lv < wt ? lv : wt;
                    else {
                        const uint64_t sm = (uint64_t)lv + wt;
                        ecand[e] = sm < INF ? (uint32_t)sm : INF;
                    }
                    ecur[e] = (ok && (uint32_t)e < ne) ? a.lab[(uint64_t)eu[e] * a.B + q] : 0u;
                }
                bool eup[K];
#pragma unroll
                for (int e = 0; e < K; ++e) {
                    eup[e] = false;
                    if (ok && (uint32_t)e < ne && (a.widest ? ecand[e] > ecur[e] : ecand[e] < ecur[e])) {
                        uint32_t* p = &a.lab[(uint64_t)eu[e] * a.B + q];
                        const uint32_t old = a.widest ? atomicMax(p, ecand[e]) : atomicMin(p, ecand[e]);
                        eup[e] = a.widest ? ecand[e] > old : ecand[e] < old;
                    }
                }
#pragma unroll
                for (int e = 0; e < K; ++e) {
                    const uint64_t um = __ballot(eup[e]);
                    if (um && lane == 0) improved(eu[e], w, um);
                }
            }
        }
        __syncthreads();
        // queue the improved vertices with their virtual-node counts
        const uint32_t cnt = min(sMisc[2], (uint32_t)TILE_EDGES);
        if (cnt > 0) {
            const uint32_t per = (cnt + WG - 1) / WG, b0 = threadIdx.x * per;
            uint64_t loc = 0;
            for (uint32_t i = 0; i < per; ++i)
                if (b0 + i < cnt) loc += 1ull | ((uint64_t)a.vnc[sQ[b0 + i]] << 32);
            uint64_t tot;
            uint64_t pre = block_excl_scan<uint64_t>(loc, sScan, &tot);
            if (threadIdx.x == 0) {
                const unsigned long long old = atomicAdd(&ctl->qpack[a.it % 3], (unsigned long long)tot);
                sMisc[3] = (uint32_t)(old & 0xffffffffull);
                sMisc[4] = (uint32_t)(old >> 32);
            }
            __syncthreads();
            uint32_t bq = sMisc[3] + (uint32_t)(pre & 0xffffffffull);
            uint32_t bv = sMisc[4] + (uint32_t)(pre >> 32);
            for (uint32_t i = 0; i < per; ++i) {
                if (b0 + i >= cnt) break;
                const uint32_t u = sQ[b0 + i], nv = a.vnc[u];
                if (bq < a.qcap) {
                    a.qout[bq] = u;
                    a.sout[bq] = bv;
                } else {
                    atomicOr(&ctl->err, 32u);
                }
                ++bq;
                bv += nv;
            }
        }
        __syncthreads();
        e_start += (uint32_t)(nnext - 1);
    }
}

// Vertices of the frontier just pushed forget their change masks; vertices improved this
// iteration take theirs from nchg.  Runs over the previous frontier, then the new one.
// With dev_it (replayed graph) the queue, slot and activity slot come from the device iteration:
// move = 0 takes the frontier the push read, move = 1 the one it wrote.
__global__ void k_plane_rotate(const uint32_t* q, const Ctl* ctl, int slot,
                               uint64_t* __restrict__ from, uint64_t* __restrict__ to, uint32_t W,
                               int move, uint64_t* act, const uint32_t* dev_it, const uint32_t* qa,
                               const uint32_t* qb, uint64_t* act_base) {
    __shared__ unsigned long long sAct[MAXW];
    if (ctl->done) return;
    if (dev_it) {
        const uint32_t it = dev_it[0];
        const bool odd = ((it - 1) & 1) != 0;
        q = move ? (odd ? qa : qb) : (odd ? qb : qa);
        slot = move ? (int)(it % 3) : (int)((it - 1) % 3);
        act = (move && act_base) ? act_base + (it % 4) * MAXW : nullptr;
    }
    if (threadIdx.x < MAXW) sAct[threadIdx.x] = 0;
    __syncthreads();
    const uint32_t count = (uint32_t)(ctl->qpack[slot] & 0xffffffffull);
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < (uint64_t)count * W;
         i += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t idx = (uint64_t)q[i / W] * W + (i % W);
        if (move) {
            const uint64_t m = from[idx];
            to[idx] = m;
            from[idx] = 0;
            if (act && m) atomicOr(&sAct[i % W], (unsigned long long)m);
        } else {
            to[idx] = 0;
        }
    }
    __syncthreads();
    if (act && threadIdx.x < W && sAct[threadIdx.x])
        atomicOr((unsigned long long*)&act[threadIdx.x], sAct[threadIdx.x]);
}

// out[q * n + u] = planes[perm[u] * B + q0 + q]: answers of queries q0 .. q0+nq-1 in
// query-major, original vertex order, narrow labels widened (their all-ones INF to INF)
__global__ void k_planes_out(const uint32_t* __restrict__ planes, const uint32_t* __restrict__ perm,
                             uint32_t n, uint32_t B, uint32_t q0, uint32_t nq, int lm, uint32_t* __restrict__ out) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < (uint64_t)n * nq;
         i += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t q = (uint32_t)(i / n), u = (uint32_t)(i % n);
        const uint64_t j = (uint64_t)perm[u] * B + q0 + q;
        uint32_t x;
        if (lm == LM_W8 || lm == LM_D8) {
            x = reinterpret_cast<const uint8_t*>(planes)[j];
            if (x == 0xffu) x = INF;
        } else if (lm == LM_D16) {
            x = reinterpret_cast<const uint16_t*>(planes)[j];
            if (x == 0xffffu) x = INF;
        } else {
            x = planes[j];
        }
        out[i] = x;
    }
}
// Largest edge weight (atomicMax into *out, zeroed by the caller).
__global__ void k_max_u32(const uint32_t* __restrict__ a, uint64_t n, uint32_t* __restrict__ out) {
    uint32_t m = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        m = max(m, a[i]);
#pragma unroll
    for (int off = WAVE / 2; off > 0; off >>= 1) m = max(m, (uint32_t)__shfl_xor(m, off, WAVE));
    if (lane_id() == 0 && m) atomicMax(out, m);
}

// Fills with 16-byte stores where the base allows (the planes are many GB).
__global__ void k_fill_u32b(uint32_t* __restrict__ a, uint64_t n, uint32_t v) {
    const uint64_t tid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x, nt = (uint64_t)gridDim.x * blockDim.x;
    uint64_t head = 0;
    if (((uintptr_t)a & 15) == 0) {
        const uint64_t n4 = n / 4;
        uint4* a4 = reinterpret_cast<uint4*>(a);
        for (uint64_t i = tid; i < n4; i += nt) a4[i] = make_uint4(v, v, v, v);
        head = n4 * 4;
    }
    for (uint64_t i = head + tid; i < n; i += nt) a[i] = v;
}
__global__ void k_fill_u64(uint64_t* __restrict__ a, uint64_t n) {
    const uint64_t tid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x, nt = (uint64_t)gridDim.x * blockDim.x;
    uint64_t head = 0;
    if (((uintptr_t)a & 15) == 0) {
        const uint64_t n2 = n / 2;
        ulonglong2* a2 = reinterpret_cast<ulonglong2*>(a);
        for (uint64_t i = tid; i < n2; i += nt) a2[i] = ulonglong2{0, 0};
        head = n2 * 2;
    }
    for (uint64_t i = head + tid; i < n; i += nt) a[i] = 0;
}
// dst[i] = src[idx[i]]
__global__ void k_gather_u32(const uint32_t* __restrict__ src, const uint32_t* __restrict__ idx, uint32_t m,
                             uint32_t* __restrict__ dst) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < m; i += gridDim.x * blockDim.x) dst[i] = src[idx[i]];
}

// Seeds: entries (vertex, query) applied on the device; the first frontier is the distinct
// source vertices with their virtual-node prefix, computed on the host.
__global__ void k_seed(const uint32_t* __restrict__ sv, const uint32_t* __restrict__ sq, uint32_t ns,
                       uint64_t* __restrict__ m1, uint64_t* __restrict__ m2, uint32_t* __restrict__ lab,
                       uint32_t R, uint32_t B, uint32_t val, int lm) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < ns; i += gridDim.x * blockDim.x) {
        const uint64_t idx = (uint64_t)sv[i] * R + sq[i] / 64;
        const uint64_t bit = 1ull << (sq[i] % 64);
        atomicOr((unsigned long long*)&m1[idx], (unsigned long long)bit);
        if (m2) atomicOr((unsigned long long*)&m2[idx], (unsigned long long)bit);
        const uint64_t j = (uint64_t)sv[i] * B + sq[i];
        if (lm == LM_W8 || lm == LM_D8) reinterpret_cast<uint8_t*>(lab)[j] = (uint8_t)(val == INF ? 0xffu : val);
        else if (lm == LM_D16) reinterpret_cast<uint16_t*>(lab)[j] = (uint16_t)(val == INF ? 0xffffu : val);
        else lab[j] = val;
    }
}

__global__ void k_set_frontier(const uint32_t* __restrict__ fv, const uint32_t* __restrict__ fs, uint32_t nf,
                               uint64_t packed, uint32_t* __restrict__ q, uint32_t* __restrict__ s, Ctl* ctl) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nf; i += gridDim.x * blockDim.x) {
        q[i] = fv[i];
        s[i] = fs[i];
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        for (int k = 0; k < 3; ++k) ctl->qpack[k] = 0;
        ctl->qpack[0] = packed;
        ctl->done = 0;
        ctl->iters = 0;
        ctl->err = 0;
    }
}

// Admission of new queries between iterations (one workgroup); see launch_admit.
__global__ __launch_bounds__(WG) void k_admit(Ctl* ctl, int in_slot, const uint32_t* __restrict__ pv,
                                              const uint32_t* __restrict__ pc, uint32_t npairs,
                                              const uint32_t* __restrict__ uniq, uint32_t nuniq,
                                              const uint32_t* __restrict__ vnc, uint32_t R, uint32_t B,
                                              uint64_t* m_always, uint64_t* m_edges, uint32_t* lab, uint32_t val,
                                              uint32_t* __restrict__ q, uint32_t* __restrict__ s, uint32_t qcap) {
    __shared__ uint64_t sScan[WG / WAVE + 1];
    __shared__ unsigned long long sBase;
    if (threadIdx.x == 0 && ctl->done) {  // drained engine: start from empty slots
        for (int k = 0; k < 3; ++k) ctl->qpack[k] = 0;
        ctl->done = 0;
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < npairs; i += WG) {
        const uint32_t v = pv[i], c = pc[i];
        lab[(uint64_t)v * B + c] = val;
        const uint64_t idx = (uint64_t)v * R + c / 64;
        const unsigned long long bit = 1ull << (c % 64);
        if (m_always) atomicOr((unsigned long long*)&m_always[idx], bit);
        if (m_edges && vnc[v]) atomicOr((unsigned long long*)&m_edges[idx], bit);
    }
    for (uint32_t b = 0; b < nuniq; b += WG) {
        const uint32_t i = b + threadIdx.x;
        uint32_t u = 0, nv = 0;
        if (i < nuniq) {
            u = uniq[i];
            nv = vnc[u];
        }
        const uint64_t mine = nv ? (1ull | ((uint64_t)nv << 32)) : 0ull;
        uint64_t tot;
        const uint64_t pre = block_excl_scan<uint64_t>(mine, sScan, &tot);
        if (threadIdx.x == 0) sBase = tot ? atomicAdd(&ctl->qpack[in_slot], (unsigned long long)tot) : 0ull;
        __syncthreads();
        if (mine) {
            const uint32_t pos = (uint32_t)(sBase & 0xffffffffull) + (uint32_t)(pre & 0xffffffffull);
            if (pos < qcap) {
                q[pos] = u;
                s[pos] = (uint32_t)(sBase >> 32) + (uint32_t)(pre >> 32);
            } else {
                atomicOr(&ctl->err, 64u);
            }
        }
        __syncthreads();
    }
}

// Frees columns: their label entries go back to `fill` and their bits are cleared.
__global__ void k_release(uint32_t n, uint32_t W, uint32_t R, uint32_t B, uint64_t c0, uint64_t c1, uint64_t c2,
                          uint64_t c3, uint32_t* __restrict__ lab, uint32_t fill, uint64_t* m0, uint64_t* m1,
                          uint64_t* m2) {
    const uint64_t clr[MAXW] = {c0, c1, c2, c3};
    for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n;
         v += (uint64_t)gridDim.x * blockDim.x) {
        for (uint32_t w = 0; w < W; ++w) {
            uint64_t bits = clr[w];
            if (!bits) continue;
            const uint64_t idx = v * R + w;
            if (m0) m0[idx] &= ~bits;
            if (m1) m1[idx] &= ~bits;
            if (m2) m2[idx] &= ~bits;
            while (bits) {
                const int b = __ffsll((unsigned long long)bits) - 1;
                bits &= bits - 1;
                lab[v * B + w * 64 + (uint32_t)b] = fill;
            }
        }
    }
}

__global__ void k_extract(const uint32_t* __restrict__ lab, uint32_t B, uint32_t c,
                          const uint32_t* __restrict__ perm, uint32_t n, uint32_t offset,
                          uint32_t* __restrict__ out) {
    for (uint64_t u = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; u < n;
         u += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t x = lab[(uint64_t)perm[u] * B + c];
        out[u] = (x < INF) ? x - offset : x;
    }
}

int cu_count_b(int device) {
    int c = 0;
    if (hipDeviceGetAttribute(&c, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess || c <= 0) c = 256;
    return c;
}

}  // namespace

// Batched state of a graph handle (allocated on first use, grown as needed).
struct BatchWork {
    DevMem m1, m3;          // BFS: per-vertex mask records (m1); planes: chg (m1), nchg (m3)
    DevMem planes;          // n x B labels
    DevMem bitmap;          // dense plane passes: vertices improved this iteration (1 bit each)
    int64_t wmax = -1;      // largest edge weight of the graph (-1: not measured yet)
    DevMem ovq;             // 8-bit SSSP passes: queries whose distances outgrew 8 bits
    DevMem rowbuf;          // one answer (n words) on its way to a host row
    bool d8_off = false;    // a previous 8-bit SSSP pass flagged more than a quarter of its queries
    DevMem q[2], s[2], cand, stamp, seeds, ctl;
    DevMem devit;           // replayed chunks: {iteration, stamp base}
    uint32_t epoch = 1;
    hipGraphExec_t exec = nullptr;  // graph of one chunk and the arguments it was built with
    std::vector<uint8_t> key;
    ~BatchWork() {
        if (exec) (void)hipGraphExecDestroy(exec);
    }
};

namespace {

// Iterations per replayed chunk (a multiple of 2 and 3: queue ping-pong and control slots).
constexpr uint32_t BATCH_CHUNK = 6;

// Compaction of a dense plane pass's bitmap into the queue the push would have written.
CompactArgs compact_args(const PlaneArgs& a) {
    CompactArgs ca{};
    ca.bitmap = a.bitmap;
    ca.nwords = (a.n + 31) / 32;
    ca.n = a.n;
    ca.vnc = a.vnc;
    ca.ctl = a.ctl;
    ca.cap = a.qcap;
    return ca;
}
unsigned compact_grid(uint32_t n) {
    const uint64_t words = ((uint64_t)n + 31) / 32;
    return (unsigned)std::max<uint64_t>(1, std::min<uint64_t>((words + WG - 1) / WG, 1024));
}

__global__ void k_tick_b(uint32_t* it) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *it += 1;
}

eeg_status add_node(hipGraph_t graph, hipGraphNode_t* prev, const void* fn, unsigned grid, unsigned block,
                    void** args) {
    hipKernelNodeParams kp{};
    kp.func = const_cast<void*>(fn);
    kp.gridDim = dim3(grid);
    kp.blockDim = dim3(block);
    kp.sharedMemBytes = 0;
    kp.kernelParams = args;
    kp.extra = nullptr;
    hipGraphNode_t node = nullptr;
    EEG_HIP(hipGraphAddKernelNode(&node, graph, *prev ? prev : nullptr, *prev ? 1 : 0, &kp));
    *prev = node;
    return EEG_OK;
}

// Graph of BATCH_CHUNK iterations (BFS: push, finalize; planes: push, two rotations; each
// followed by a counter tick), built node by node and kept until its arguments change.
// Dense plane push for K and label width lm.
const void* dense_push(uint32_t K, int lm) {
    switch (lm) {
        case LM_W8:
            return (K == 4) ? (const void*)k_plane_push_wave<4, LM_W8> : (K == 8) ? (const void*)k_plane_push_wave<8, LM_W8>
                   : (K == 16) ? (const void*)k_plane_push_wave<16, LM_W8> : (const void*)k_plane_push_wave<32, LM_W8>;
        case LM_D16:
            return (K == 4) ? (const void*)k_plane_push_wave<4, LM_D16> : (K == 8) ? (const void*)k_plane_push_wave<8, LM_D16>
                   : (K == 16) ? (const void*)k_plane_push_wave<16, LM_D16> : (const void*)k_plane_push_wave<32, LM_D16>;
        case LM_D8:
            return (K == 4) ? (const void*)k_plane_push_wave<4, LM_D8> : (K == 8) ? (const void*)k_plane_push_wave<8, LM_D8>
                   : (K == 16) ? (const void*)k_plane_push_wave<16, LM_D8> : (const void*)k_plane_push_wave<32, LM_D8>;
        default:
            return (K == 4) ? (const void*)k_plane_push_wave<4, LM_U32> : (K == 8) ? (const void*)k_plane_push_wave<8, LM_U32>
                   : (K == 16) ? (const void*)k_plane_push_wave<16, LM_U32> : (const void*)k_plane_push_wave<32, LM_U32>;
    }
}

eeg_status batch_graph(BatchWork& bw, bool bfs, uint32_t K, unsigned grid, const MsArgs& ma, const PlaneArgs& pa,
                       uint32_t nsrc, int lm, hipGraphExec_t* out) {
    const bool dense = nsrc > PLANE_DENSE_FROM;
    std::vector<uint8_t> key(sizeof(MsArgs) + sizeof(PlaneArgs) + 20);
    memcpy(key.data(), &ma, sizeof(MsArgs));
    memcpy(key.data() + sizeof(MsArgs), &pa, sizeof(PlaneArgs));
    const uint32_t tail[5] = {bfs ? 1u : 0u, K, grid, dense ? 1u : 0u, (uint32_t)lm};
    memcpy(key.data() + sizeof(MsArgs) + sizeof(PlaneArgs), tail, 20);
    if (bw.exec && bw.key == key) {
        *out = bw.exec;
        return EEG_OK;
    }
    if (bw.exec) {
        (void)hipGraphExecDestroy(bw.exec);
        bw.exec = nullptr;
    }
    hipGraph_t graph = nullptr;
    EEG_HIP(hipGraphCreate(&graph, 0));
    MsArgs m = ma;
    PlaneArgs p = pa;
    uint32_t* itp = const_cast<uint32_t*>(bfs ? ma.dev_it : pa.dev_it);
    void* margs[] = {&m};
    void* pargs[] = {&p};
    void* targs[] = {&itp};
    // rotation arguments (queue, slot and activity slot resolved on the device)
    const uint32_t* rq = nullptr;
    const Ctl* rctl = pa.ctl;
    int rslot = 0;
    uint64_t* rnone = nullptr;
    uint64_t* rfrom1 = pa.nchg;
    uint64_t* rto = pa.chg;
    uint32_t rW = pa.W;
    int move0 = 0, move1 = 1;
    uint64_t* ract = nullptr;
    const uint32_t* rit = pa.dev_it;
    const uint32_t* rqa = pa.qa;
    const uint32_t* rqb = pa.qb;
    uint64_t* ractb = pa.act;
    void* r0args[] = {&rq, &rctl, &rslot, &rnone, &rto, &rW, &move0, &ract, &rit, &rqa, &rqb, &ractb};
    // sorted queues: the iteration comes from the device counter (ticked after each iteration)
    CompactArgs cga{};
    if (pa.bitmap) {
        cga = compact_args(pa);
        cga.qa = pa.qa;
        cga.sa = pa.sa;
        cga.qb = pa.qb;
        cga.sb = pa.sb;
        cga.dev_it = pa.dev_it;
    }
    void* cargs[] = {&cga};
    void* r1args[] = {&rq, &rctl, &rslot, &rfrom1, &rto, &rW, &move1, &ract, &rit, &rqa, &rqb, &ractb};
    const void* fpush =
        bfs ? ((K == 4) ? (const void*)k_ms_push<4> : (K == 8) ? (const void*)k_ms_push<8>
               : (K == 16) ? (const void*)k_ms_push<16> : (const void*)k_ms_push<32>)
        : dense ? dense_push(K, lm)
                : ((K == 4) ? (const void*)k_plane_push<4> : (K == 8) ? (const void*)k_plane_push<8>
                   : (K == 16) ? (const void*)k_plane_push<16> : (const void*)k_plane_push<32>);
    hipGraphNode_t prev = nullptr;
    eeg_status rc = EEG_OK;
    for (uint32_t i = 0; i < BATCH_CHUNK && rc == EEG_OK; ++i) {
        if (bfs) {
            rc = add_node(graph, &prev, fpush, grid, WG, margs);
            if (rc == EEG_OK) rc = add_node(graph, &prev, (const void*)k_ms_finalize, grid, WG, margs);
        } else {
            rc = add_node(graph, &prev, fpush, grid, WG, pargs);
            if (rc == EEG_OK && pa.bitmap)
                rc = add_node(graph, &prev, (const void*)k_compact_frontier, compact_grid(pa.n), WG, cargs);
            if (rc == EEG_OK) rc = add_node(graph, &prev, (const void*)k_plane_rotate, 2048, WG, r0args);
            if (rc == EEG_OK) rc = add_node(graph, &prev, (const void*)k_plane_rotate, 2048, WG, r1args);
        }
        if (rc == EEG_OK) rc = add_node(graph, &prev, (const void*)k_tick_b, 1, 64, targs);
    }
    if (rc != EEG_OK) {
        (void)hipGraphDestroy(graph);
        return rc;
    }
    const hipError_t e = hipGraphInstantiate(&bw.exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) {
        bw.exec = nullptr;
        return fail(EEG_ERR_DEVICE, "hipGraphInstantiate failed: %s", hipGetErrorString(e));
    }
    bw.key = key;
    *out = bw.exec;
    return EEG_OK;
}

}  // namespace

// Batched state of a graph handle, created on first use and freed with the handle.  Callers
// hold the handle's lock.
static BatchWork& batch_work(eeg_graph* g) {
    if (!g->batch_state) g->batch_state = std::make_shared<BatchWork>();
    return *static_cast<BatchWork*>(g->batch_state.get());
}

namespace {

// Largest edge weight of the graph, measured once per handle.
eeg_status graph_wmax(eeg_graph* g, BatchWork& bw, hipStream_t s, uint32_t* wmax) {
    if (bw.wmax < 0) {
        DevMem d;
        EEG_TRY(d.alloc(8));
        uint32_t h = 0;
        EEG_HIP(hipMemsetAsync(d.p, 0, 4, s));
        if (g->m) {
            k_max_u32<<<grid_n(g->m), WG, 0, s>>>(g->out.wp(), g->m, d.as<uint32_t>());
            EEG_HIP(hipGetLastError());
        }
        EEG_HIP(hipMemcpyAsync(&h, d.p, 4, hipMemcpyDeviceToHost, s));
        EEG_HIP(hipStreamSynchronize(s));
        bw.wmax = h;
    }
    *wmax = (uint32_t)bw.wmax;
    return EEG_OK;
}

// Where the answers of a pass go: one query-major block (host or device), or one host row per
// query (null rows are skipped).  Neither: no answers are copied out.
struct BatchOut {
    void* out = nullptr;
    bool device = false;
    uint32_t* const* rows = nullptr;
};

// One multi-source pass with labels of width lm, on the state bw and stream s.  A 16-bit pass
// whose distances outgrow its labels sets *overflow and returns EEG_OK without answers; an
// 8-bit pass answers every query and returns the queries to answer again (bit q of
// ovq[q / 64]) in ovq_out.
eeg_status batch_pass(eeg_graph* g, BatchWork& bw, hipStream_t s, uint32_t algo, const uint32_t* sources,
                      uint32_t nsrc, const eeg_opts& o, const BatchOut& bo, eeg_stats* st, int lm, bool* overflow,
                      uint64_t ovq_out[MAXW]) {
    const auto t_host0 = std::chrono::steady_clock::now();
    if (nsrc == 0 || nsrc > 64 * MAXW) return fail(EEG_ERR_INVALID_ARG, "nsrc must be 1..%u per pass", 64 * MAXW);
    if (algo != EEG_BFS && algo != EEG_SSSP && algo != EEG_SSWP)
        return fail(EEG_ERR_UNSUPPORTED, "batched runs support BFS, SSSP and SSWP");
    if ((algo == EEG_SSSP || algo == EEG_SSWP) && !g->weighted)
        return fail(EEG_ERR_INVALID_ARG, "SSSP/SSWP need a weighted graph");
    const uint32_t K = o.K ? o.K : 8;
    if (k_index(K) < 0) return fail(EEG_ERR_INVALID_ARG, "K must be 4, 8, 16 or 32");
    const uint32_t n = g->n;
    for (uint32_t i = 0; i < nsrc; ++i)
        if (sources[i] >= n) return fail(EEG_ERR_INVALID_ARG, "source %u out of range", sources[i]);
    EEG_TRY(ensure_tigr(g, g->out, g->tout, K));
    const Tigr& t = g->tout[k_index(K)];
    const uint32_t W = (nsrc + 63) / 64, B = W * 64;
    const uint64_t nw = (uint64_t)n * W, nb = (uint64_t)n * B;
    const bool bfs = (algo == EEG_BFS);
    const bool widest = (algo == EEG_SSWP);
    // BFS: one record of 4W words per vertex in m1 (seen, next, visit, stamp).  Planes: change
    // masks chg (m1) and nchg (m3), W words per vertex.
    const uint32_t R = bfs ? 4 * W : W;
    const uint64_t nrec = (uint64_t)n * R;
    EEG_TRY(bw.m1.alloc(nrec * 8 + 8));
    if (!bfs) EEG_TRY(bw.m3.alloc(nw * 8 + 8));
    const uint64_t lbytes = (lm == LM_W8 || lm == LM_D8) ? 1 : lm == LM_D16 ? 2 : 4;
    EEG_TRY(bw.planes.alloc(nb * lbytes + 4));
    for (int i = 0; i < 2; ++i) {
        EEG_TRY(bw.q[i].alloc((size_t)n * 4 + 4));
        EEG_TRY(bw.s[i].alloc((size_t)n * 4 + 4));
    }
    EEG_TRY(bw.cand.alloc((size_t)n * 4 + 4));
    if (bw.stamp.bytes < (size_t)n * 4 + 4) {
        EEG_TRY(bw.stamp.alloc((size_t)n * 4 + 4));
        EEG_HIP(hipMemsetAsync(bw.stamp.p, 0, bw.stamp.bytes, s));
        bw.epoch = 1;
    }
    if (bw.epoch > 0xf0000000u) {
        EEG_HIP(hipMemsetAsync(bw.stamp.p, 0, bw.stamp.bytes, s));
        bw.epoch = 1;
    }
    EEG_TRY(bw.ctl.alloc(sizeof(Ctl)));
    EEG_TRY(bw.seeds.alloc((size_t)4 * nsrc * 4 + 16));
    if (lm == LM_D8) {
        EEG_TRY(bw.ovq.alloc(MAXW * 8));
        k_fill_u64<<<1, 64, 0, s>>>(bw.ovq.as<uint64_t>(), MAXW);
        EEG_HIP(hipGetLastError());
    }
    Ctl* ctl = bw.ctl.as<Ctl>();
    EEG_HIP(hipMemsetAsync(ctl, 0, sizeof(Ctl), s));

    // seeds in internal ids; first frontier = distinct sources with out-edges
    std::vector<uint32_t> sv(nsrc), sq(nsrc);
    for (uint32_t i = 0; i < nsrc; ++i) {
        sv[i] = g->h_perm[sources[i]];
        sq[i] = i;
    }
    std::vector<uint32_t> fv(sv);
    std::sort(fv.begin(), fv.end());
    fv.erase(std::unique(fv.begin(), fv.end()), fv.end());
    uint32_t* dseed = bw.seeds.as<uint32_t>();
    // virtual-node counts of the distinct sources, gathered into the seed buffer's frontier
    // half (overwritten below) and read back in one copy
    std::vector<uint32_t> hvnc(fv.size());
    EEG_HIP(hipMemcpyAsync(dseed + 2 * nsrc, fv.data(), fv.size() * 4, hipMemcpyHostToDevice, s));
    k_gather_u32<<<grid_n(fv.size()), WG, 0, s>>>(t.vnc.as<uint32_t>(), dseed + 2 * nsrc, (uint32_t)fv.size(),
                                                 dseed + 3 * nsrc);
    EEG_HIP(hipGetLastError());
    EEG_HIP(hipMemcpyAsync(hvnc.data(), dseed + 3 * nsrc, fv.size() * 4, hipMemcpyDeviceToHost, s));
    EEG_HIP(hipStreamSynchronize(s));
    std::vector<uint32_t> fvv, fss;
    uint32_t acc = 0;
    for (size_t i = 0; i < fv.size(); ++i)
        if (hvnc[i]) {
            fvv.push_back(fv[i]);
            fss.push_back(acc);
            acc += hvnc[i];
        }
    EEG_HIP(hipMemcpyAsync(dseed, sv.data(), nsrc * 4, hipMemcpyHostToDevice, s));
    EEG_HIP(hipMemcpyAsync(dseed + nsrc, sq.data(), nsrc * 4, hipMemcpyHostToDevice, s));
    if (!fvv.empty()) {
        EEG_HIP(hipMemcpyAsync(dseed + 2 * nsrc, fvv.data(), fvv.size() * 4, hipMemcpyHostToDevice, s));
        EEG_HIP(hipMemcpyAsync(dseed + 3 * nsrc, fss.data(), fss.size() * 4, hipMemcpyHostToDevice, s));
    }
    uint64_t* rec = bw.m1.as<uint64_t>();
    k_fill_u64<<<grid_n(nrec), WG, 0, s>>>(rec, nrec);
    if (!bfs) k_fill_u64<<<grid_n(nw), WG, 0, s>>>(bw.m3.as<uint64_t>(), nw);
    // every label unreached: 0 (widest), INF or all ones (narrow INF) for distances; nb is a
    // multiple of 64, so the narrow planes are whole 32-bit words
    const uint32_t unreached = widest ? 0u : (lm == LM_U32 ? INF : 0xffffffffu);
    k_fill_u32b<<<grid_n(nb * lbytes / 4), WG, 0, s>>>(bw.planes.as<uint32_t>(), nb * lbytes / 4, unreached);
    EEG_HIP(hipGetLastError());
    // BFS: seen and visit get the source bits, level 0.  Planes: chg gets them.
    k_seed<<<grid_n(nsrc), WG, 0, s>>>(dseed, dseed + nsrc, nsrc, rec, bfs ? rec + 2 * W : nullptr,
                                       bw.planes.as<uint32_t>(), R, B, widest ? INF : 0u, lm);
    k_set_frontier<<<grid_n(fvv.size() ? fvv.size() : 1), WG, 0, s>>>(
        dseed + 2 * nsrc, dseed + 3 * nsrc, (uint32_t)fvv.size(),
        ((unsigned long long)acc << 32) | (unsigned long long)fvv.size(), bw.q[0].as<uint32_t>(),
        bw.s[0].as<uint32_t>(), ctl);
    EEG_HIP(hipGetLastError());

    const int cus = cu_count_b(g->device);
    // 48: 4-5% faster 32-bit planes than 24; narrow planes (8 waves per SIMD) 3-4% faster at 96
    // than at 48 (RMAT-24 x 256)
    const unsigned grid = (unsigned)cus * (o.wgs_per_cu ? o.wgs_per_cu : (lm == LM_U32 ? 48 : 96));
    const uint32_t base = bw.epoch;
    hipEvent_t e0, e1;
    EEG_HIP(hipEventCreate(&e0));
    EEG_HIP(hipEventCreate(&e1));
    EEG_HIP(hipEventRecord(e0, s));
    Ctl hc{};
    uint32_t launches = 0, level = 0;
    const uint32_t max_iters = o.max_iters ? o.max_iters : n + 2;
    // BFS: one level = push + finalize; planes: one iteration = push + two mask rotations.
    // Chunks of 8 run ahead of the host's look at the control block.
    const uint32_t qcap = n + 1;
    // argument blocks shared by every iteration; the per-iteration fields are set per launch,
    // or derived on the device when a graph of BATCH_CHUNK iterations is replayed
    MsArgs ma{};
    PlaneArgs pa{};
    if (bfs) {
        ma.row = g->out.rowp();
        ma.col = g->out.colp();
        ma.vnc = t.vnc.as<uint32_t>();
        ma.seen = rec;
        ma.next = rec + W;
        ma.visit = rec + 2 * W;
        ma.stamp = reinterpret_cast<uint32_t*>(rec + 3 * W);  // zeroed with the records
        ma.R = R;
        ma.sst = 2 * R;
        ma.cand = bw.cand.as<uint32_t>();
        ma.dist = bw.planes.as<uint32_t>();
        ma.ctl = ctl;
        ma.W = W;
        ma.B = B;
        ma.n = n;
        ma.qcap = qcap;
    } else {
        pa.row = g->out.rowp();
        pa.col = g->out.colp();
        pa.wgt = g->out.wp();
        pa.vnc = t.vnc.as<uint32_t>();
        pa.lab = bw.planes.as<uint32_t>();
        pa.chg = bw.m1.as<uint64_t>();
        pa.nchg = bw.m3.as<uint64_t>();
        pa.stamp = bw.stamp.as<uint32_t>();
        pa.ctl = ctl;
        pa.W = W;
        pa.B = B;
        pa.n = n;
        pa.qcap = qcap;
        pa.widest = widest ? 1 : 0;
        pa.ovq = lm == LM_D8 ? bw.ovq.as<uint64_t>() : nullptr;
        if (nsrc > PLANE_DENSE_FROM) {  // dense passes: sorted queues rebuilt from a bitmap
            const uint64_t nwords = ((uint64_t)n + 31) / 32;
            EEG_TRY(bw.bitmap.alloc(nwords * 4 + 4));
            k_fill_u32b<<<grid_n(nwords), WG, 0, s>>>(bw.bitmap.as<uint32_t>(), nwords, 0u);
            EEG_HIP(hipGetLastError());
            pa.bitmap = bw.bitmap.as<uint32_t>();
        }
    }
    const bool replay = !o.timing;
    hipGraphExec_t exec = nullptr;
    if (replay) {
        EEG_TRY(bw.devit.alloc(8));
        uint32_t* dit = bw.devit.as<uint32_t>();
        k_fill_u32b<<<1, 64, 0, s>>>(dit, 1, 1u);
        k_fill_u32b<<<1, 64, 0, s>>>(dit + 1, 1, base);
        EEG_HIP(hipGetLastError());
        ma.dev_it = pa.dev_it = dit;
        ma.qa = pa.qa = bw.q[0].as<uint32_t>();
        ma.sa = pa.sa = bw.s[0].as<uint32_t>();
        ma.qb = pa.qb = bw.q[1].as<uint32_t>();
        ma.sb = pa.sb = bw.s[1].as<uint32_t>();
        EEG_TRY(batch_graph(bw, bfs, K, grid, ma, pa, nsrc, lm, &exec));
    }
    uint32_t j = 1;
    while (true) {
        if (replay) {
            EEG_HIP(hipGraphLaunch(exec, s));
            j += BATCH_CHUNK;
            launches += BATCH_CHUNK;
        }
        const uint32_t chunk = replay ? 0 : 8;
        for (uint32_t c = 0; c < chunk; ++c, ++j, ++launches) {
            const int cur = (int)((j - 1) & 1);
            if (bfs) {
                MsArgs a = ma;
                a.level = j;
                a.stampv = base + j;
                a.qin = bw.q[cur].as<uint32_t>();
                a.sin = bw.s[cur].as<uint32_t>();
                a.qout = bw.q[cur ^ 1].as<uint32_t>();
                a.sout = bw.s[cur ^ 1].as<uint32_t>();
                EEG_TRY(launch_ms_level(a, K, grid, s));
            } else {
                PlaneArgs a = pa;
                a.it = j;
                a.stampv = base + j;
                a.qin = bw.q[cur].as<uint32_t>();
                a.sin = bw.s[cur].as<uint32_t>();
                a.qout = bw.q[cur ^ 1].as<uint32_t>();
                a.sout = bw.s[cur ^ 1].as<uint32_t>();
                EEG_TRY(launch_plane_iteration(a, K, grid, s, nsrc, lm));
            }
        }
        EEG_HIP(hipMemcpyAsync(&hc, ctl, sizeof(Ctl), hipMemcpyDeviceToHost, s));
        EEG_HIP(hipStreamSynchronize(s));
        if (hc.err == PLANE_ERR_NARROW && lm != LM_U32) {  // distances outgrew the labels
            bw.epoch = base + j + 2;
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            *overflow = true;
            return EEG_OK;
        }
        if (hc.err) return fail(EEG_ERR_DEVICE, "batched frontier overflow (err=%u)", hc.err);
        if (hc.done) break;
        level = hc.iters;
        if (j > max_iters + 16) return fail(EEG_ERR_NOT_CONVERGED, "batched traversal did not converge");
    }
    EEG_HIP(hipEventRecord(e1, s));
    bw.epoch = base + j + 2;
    if (ovq_out) {
        for (uint32_t i = 0; i < MAXW; ++i) ovq_out[i] = 0;
        if (lm == LM_D8) EEG_HIP(hipMemcpyAsync(ovq_out, bw.ovq.p, MAXW * 8, hipMemcpyDeviceToHost, s));
    }
    // answers: query-major, original vertex order
    const uint64_t outn = (uint64_t)nsrc * n;
    if (bo.out) {
        uint32_t* dst = nullptr;
        DevMem tmp;
        if (bo.device) {
            dst = (uint32_t*)bo.out;
        } else {
            EEG_TRY(tmp.alloc(outn * 4 + 4));
            dst = tmp.as<uint32_t>();
        }
        k_planes_out<<<grid_n(outn, 65536), WG, 0, s>>>(bw.planes.as<uint32_t>(), g->perm.as<uint32_t>(), n, B, 0u,
                                                        nsrc, lm, dst);
        EEG_HIP(hipGetLastError());
        if (!bo.device) EEG_HIP(hipMemcpyAsync(bo.out, dst, outn * 4, hipMemcpyDeviceToHost, s));
        EEG_HIP(hipStreamSynchronize(s));
    } else if (bo.rows) {  // one row at a time through one device row
        EEG_TRY(bw.rowbuf.alloc((size_t)n * 4 + 4));
        for (uint32_t q = 0; q < nsrc; ++q) {
            if (!bo.rows[q]) continue;
            k_planes_out<<<grid_n(n, 65536), WG, 0, s>>>(bw.planes.as<uint32_t>(), g->perm.as<uint32_t>(), n, B, q, 1u,
                                                         lm, bw.rowbuf.as<uint32_t>());
            EEG_HIP(hipGetLastError());
            EEG_HIP(hipMemcpyAsync(bo.rows[q], bw.rowbuf.p, (size_t)n * 4, hipMemcpyDeviceToHost, s));
            EEG_HIP(hipStreamSynchronize(s));
        }
    }
    EEG_HIP(hipStreamSynchronize(s));
    float ms = 0;
    EEG_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)level;
    if (st) {
        st->iterations = hc.iters;
        st->launches = launches;
        st->ms_traversal = ms;
        st->ms_total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_host0).count();
        st->label_bits = bfs ? 0u : (uint32_t)(8 * lbytes);
    }
    return EEG_OK;
}

}  // namespace

namespace {

// A pass on lm labels; a 16-bit pass that overflows is answered again on 32-bit labels.
eeg_status pass_or_wider(eeg_graph* g, BatchWork& bw, hipStream_t s, uint32_t algo, const uint32_t* sources,
                         uint32_t nsrc, const eeg_opts& o, const BatchOut& bo, eeg_stats* st, int lm,
                         uint64_t ovq[MAXW]) {
    bool overflow = false;
    EEG_TRY(batch_pass(g, bw, s, algo, sources, nsrc, o, bo, st, lm, &overflow, ovq));
    if (!overflow) return EEG_OK;
    overflow = false;
    return batch_pass(g, bw, s, algo, sources, nsrc, o, bo, st, LM_U32, &overflow, ovq);
}

// Multi-source queries on the state bw, stream s and (for single-query reruns) traversal state
// wk (null: the graph's own state, whose stream must then be s).
eeg_status run_batch_core(eeg_graph* g, BatchWork& bw, Work* wk, hipStream_t s, uint32_t algo,
                          const uint32_t* sources, uint32_t nsrc, const eeg_opts& o, const BatchOut& bo,
                          eeg_stats* st) {
    const auto t0 = std::chrono::steady_clock::now();
    int lm = LM_U32;
    if ((algo == EEG_SSSP || algo == EEG_SSWP) && g->weighted && nsrc > PLANE_DENSE_FROM && nsrc <= 64 * MAXW) {
        if (o.label_bits != 0 && o.label_bits != 32)
            return fail(EEG_ERR_INVALID_ARG, "label_bits must be 0 or 32 (got %u)", o.label_bits);
        if (o.label_bits == 0) {
            uint32_t wmax = 0;
            EEG_TRY(graph_wmax(g, bw, s, &wmax));
            if (algo == EEG_SSSP) {
                // 8-bit distances while the weights leave room for a few hops under 255; a
                // graph whose 8-bit passes flagged many queries goes straight to 16 bits
                lm = (wmax <= 127 && !bw.d8_off) ? LM_D8 : LM_D16;
            } else if (wmax <= 254) {
                lm = LM_W8;
            }
        }
    }
    uint64_t ovq[MAXW];
    EEG_TRY(pass_or_wider(g, bw, s, algo, sources, nsrc, o, bo, st, lm, ovq));
    if (lm == LM_D8) {
        // queries whose distances outgrew 8 bits: answered again, their rows written over the
        // 8-bit answers
        std::vector<uint32_t> qs, srcs;
        for (uint32_t q = 0; q < nsrc; ++q)
            if ((ovq[q / 64] >> (q % 64)) & 1ull) {
                qs.push_back(q);
                srcs.push_back(sources[q]);
            }
        if (st) st->narrow_reruns = (uint32_t)qs.size();
        if (4 * qs.size() > nsrc) bw.d8_off = true;
        if (!qs.empty()) {
            const uint32_t nr = (uint32_t)qs.size(), n = g->n;
            const uint64_t row = (uint64_t)n * 4;
            eeg_stats s2;
            memset(&s2, 0, sizeof(s2));
            if (nr <= 8) {
                // a few: single-query traversals (sync push on sorted queues), each about a
                // third of a one-source pass
                eeg_opts o1;
                memset(&o1, 0, sizeof(o1));
                o1.variant = EEG_SYNC_PUSH_DD;
                o1.frontier = EEG_FRONTIER_SORTED;
                o1.K = o.K;
                Work& w = wk ? *wk : g->work;
                const bool host = bo.rows || (bo.out && !bo.device);
                if (host) EEG_TRY(bw.rowbuf.alloc(row + 4));
                for (uint32_t i = 0; i < nr; ++i) {
                    eeg_stats s1;
                    memset(&s1, 0, sizeof(s1));
                    EEG_TRY(run_traversal(g, algo, g->h_perm[srcs[i]], o1, &s1, nullptr, 0, wk, wk ? s : nullptr));
                    s2.ms_traversal += s1.ms_traversal;
                    s2.launches += s1.launches;
                    void* hdst = bo.rows ? (void*)bo.rows[qs[i]]
                                         : bo.out ? (void*)((uint8_t*)bo.out + (uint64_t)qs[i] * row) : nullptr;
                    if (!hdst) continue;
                    EEG_TRY(unpermute(g, w.lab.p, host ? bw.rowbuf.p : hdst, s));
                    if (host) EEG_HIP(hipMemcpyAsync(hdst, bw.rowbuf.p, row, hipMemcpyDeviceToHost, s));
                    EEG_HIP(hipStreamSynchronize(s));
                }
            } else {
                // more: one pass on 16-bit planes (32-bit for 16 or fewer), answers straight to
                // the flagged queries' rows
                const int lm2 = nr > PLANE_DENSE_FROM ? LM_D16 : LM_U32;
                std::vector<uint32_t*> sub(nr, nullptr);
                DevMem dtmp;
                BatchOut b2;
                if (bo.rows) {
                    for (uint32_t i = 0; i < nr; ++i) sub[i] = bo.rows[qs[i]];
                    b2.rows = sub.data();
                } else if (bo.out && bo.device) {
                    EEG_TRY(dtmp.alloc(nr * row + 4));
                    b2.out = dtmp.p;
                    b2.device = true;
                } else if (bo.out) {
                    for (uint32_t i = 0; i < nr; ++i) sub[i] = (uint32_t*)((uint8_t*)bo.out + (uint64_t)qs[i] * row);
                    b2.rows = sub.data();
                }
                EEG_TRY(pass_or_wider(g, bw, s, algo, srcs.data(), nr, o, b2, &s2, lm2, nullptr));
                if (b2.device) {
                    for (uint32_t i = 0; i < nr; ++i)
                        EEG_HIP(hipMemcpyAsync((uint8_t*)bo.out + (uint64_t)qs[i] * row, (uint8_t*)dtmp.p + (uint64_t)i * row,
                                               row, hipMemcpyDeviceToDevice, s));
                    EEG_HIP(hipStreamSynchronize(s));
                }
            }
            if (st) {
                st->ms_traversal += s2.ms_traversal;
                st->launches += s2.launches;
            }
        }
    }
    if (st) st->ms_total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return EEG_OK;
}

}  // namespace

eeg_status run_batch(eeg_graph* g, uint32_t algo, const uint32_t* sources, uint32_t nsrc, const eeg_opts& o,
                     void* out, bool out_device, eeg_stats* st) {
    BatchOut bo;
    bo.out = out;
    bo.device = out_device;
    return run_batch_core(g, batch_work(g), nullptr, g->stream, algo, sources, nsrc, o, bo, st);
}

std::shared_ptr<BatchWork> make_batch_work() { return std::make_shared<BatchWork>(); }

eeg_status run_batch_rows(eeg_graph* g, BatchWork& bw, Work& wk, hipStream_t s, uint32_t algo,
                          const uint32_t* sources, uint32_t nsrc, const eeg_opts& o, uint32_t* const* rows,
                          eeg_stats* st) {
    BatchOut bo;
    bo.rows = rows;
    return run_batch_core(g, bw, &wk, s, algo, sources, nsrc, o, bo, st);
}

eeg_status launch_ms_level(MsArgs a, uint32_t K, unsigned grid, hipStream_t s) {
    a.launch = 2 * a.level - 1;
    void* args[] = {&a};
    auto fn = (K == 4) ? k_ms_push<4> : (K == 8) ? k_ms_push<8> : (K == 16) ? k_ms_push<16> : k_ms_push<32>;
    EEG_HIP(hipLaunchKernel((const void*)fn, dim3(grid), dim3(WG), args, 0, s));
    a.launch = 2 * a.level;
    k_ms_finalize<<<grid, WG, 0, s>>>(a);
    EEG_HIP(hipGetLastError());
    return EEG_OK;
}

eeg_status launch_plane_iteration(PlaneArgs a, uint32_t K, unsigned grid, hipStream_t s, uint32_t in_flight, int lm) {
    void* args[] = {&a};
    const void* fn = (K == 4) ? (const void*)k_plane_push<4> : (K == 8) ? (const void*)k_plane_push<8>
                     : (K == 16) ? (const void*)k_plane_push<16> : (const void*)k_plane_push<32>;
    if (in_flight > PLANE_DENSE_FROM) fn = dense_push(K, lm);
    EEG_HIP(hipLaunchKernel(fn, dim3(grid), dim3(WG), args, 0, s));
    if (a.bitmap) {  // sorted queue: the push only set bits
        CompactArgs ca = compact_args(a);
        ca.qa = ca.qb = a.qout;
        ca.sa = ca.sb = a.sout;
        ca.it = a.it;
        void* cargs[] = {&ca};
        EEG_HIP(hipLaunchKernel((const void*)k_compact_frontier, dim3(compact_grid(a.n)), dim3(WG), cargs, 0, s));
    }
    // forget the pushed frontier's masks, then move this iteration's improvements in
    k_plane_rotate<<<2048, WG, 0, s>>>(a.qin, a.ctl, (int)((a.it - 1) % 3), nullptr, a.chg, a.W, 0, nullptr,
                                        nullptr, nullptr, nullptr, nullptr);
    k_plane_rotate<<<2048, WG, 0, s>>>(a.qout, a.ctl, (int)(a.it % 3), a.nchg, a.chg, a.W, 1,
                                        a.act ? a.act + (a.it % 4) * MAXW : nullptr, nullptr, nullptr, nullptr,
                                        nullptr);
    EEG_HIP(hipGetLastError());
    return EEG_OK;
}

eeg_status launch_admit(Ctl* ctl, int in_slot, const uint32_t* pv, const uint32_t* pc, uint32_t npairs,
                        const uint32_t* uniq, uint32_t nuniq, const uint32_t* vnc, uint32_t R, uint32_t B,
                        uint64_t* m_always, uint64_t* m_edges, uint32_t* lab, uint32_t val, uint32_t* q,
                        uint32_t* s, uint32_t qcap, hipStream_t st) {
    k_admit<<<1, WG, 0, st>>>(ctl, in_slot, pv, pc, npairs, uniq, nuniq, vnc, R, B, m_always, m_edges, lab, val,
                              q, s, qcap);
    EEG_HIP(hipGetLastError());
    return EEG_OK;
}

eeg_status launch_release(uint32_t n, uint32_t W, uint32_t R, uint32_t B, const uint64_t clr[BATCH_MAXW],
                          uint32_t* lab, uint32_t fill, uint64_t* m0, uint64_t* m1, uint64_t* m2, hipStream_t st) {
    k_release<<<grid_n(n, 65536), WG, 0, st>>>(n, W, R, B, clr[0], W > 1 ? clr[1] : 0, W > 2 ? clr[2] : 0,
                                                W > 3 ? clr[3] : 0, lab, fill, m0, m1, m2);
    EEG_HIP(hipGetLastError());
    return EEG_OK;
}

eeg_status launch_extract(const uint32_t* lab, uint32_t B, uint32_t c, const uint32_t* perm, uint32_t n,
                          uint32_t offset, uint32_t* out, hipStream_t st) {
    k_extract<<<grid_n(n, 65536), WG, 0, st>>>(lab, B, c, perm, n, offset, out);
    EEG_HIP(hipGetLastError());
    return EEG_OK;
}

int batch_cu_count(int device) { return cu_count_b(device); }

}  // namespace eeg
