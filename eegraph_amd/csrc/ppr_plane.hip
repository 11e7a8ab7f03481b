// Personalised PageRank for B restart vertices per pass (B = 16 or 32 columns).
//
// One query is the Jacobi iteration x' = (1-d) e_s + d * sum_{u->v} x_u/outdeg(u), x0 = e_s,
// stopped when ||x' - x||_1 < epsilon (FP64), the same iteration as the single-query pull of
// pagerank.hip.  Here the contributions of B queries are kept vertex-major (c[v*B + q]), so one
// pass over the in-edge CSR reads each source index once for all B queries and gathers one
// B-float row per in-edge instead of B separate 4-byte words.
//
// Work decomposition: the in-edges of every vertex are cut into segments of at most 64; a
// group of B lanes (one per column) sums one segment in edge order with compensated FP32
// (TwoSum), WAVE/B segments per wavefront.  A vertex with one segment is finished at once;
// the partial sums of longer vertices go to spill slots and a second pass adds them in segment
// order; a hub's (more than PPR_HUB_SEGS segments) are first added in chunks of PPR_HUB_SEGS
// segments by separate groups, then the chunk sums in order.  The order of every sum is fixed,
// so results do not depend on the schedule.
// Each column's L1 change is accumulated in FP64 per lane and reduced in a fixed order.
//
// Columns are independent: a column may start (fresh: x0 = e_s, nothing read from the planes)
// at any iteration while others continue, which the serving loop uses to admit queries between
// iterations.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

#include "compensated.hpp"
#include "ppr_plane.hpp"
#include "scan.hpp"

namespace eeg {
namespace {

struct SegLoad {  // (segments << 32) | 1 for a vertex with in- or out-edges
    const uint64_t* row;
    const uint32_t* outdeg;
    __device__ uint64_t operator()(uint64_t v) const {
        const uint64_t d = row[v + 1] - row[v];
        if (d == 0 && outdeg[v] == 0) return 0;
        const uint64_t ns = d ? (d + PPR_SEG - 1) / PPR_SEG : 1;
        return (ns << 32) | 1ull;
    }
};
struct SegStore {
    const uint64_t* row;
    const uint32_t* outdeg;
    uint32_t* ent;
    uint32_t* sscan;
    __device__ void operator()(uint64_t v, uint64_t pre) const {
        if (row[v + 1] > row[v] || outdeg[v] > 0) {
            const uint32_t pos = (uint32_t)(pre & 0xffffffffull);
            ent[pos] = (uint32_t)v;
            sscan[pos] = (uint32_t)(pre >> 32);
        }
    }
};
struct MultiLoad {  // (spill slots << 32) | 1 for an entry with more than one segment
    const uint32_t* sscan;
    __device__ uint64_t operator()(uint64_t i) const {
        const uint64_t ns = sscan[i + 1] - sscan[i];
        return ns > 1 ? ((ns << 32) | 1ull) : 0ull;
    }
};
struct MultiStore {
    const uint32_t* sscan;
    uint32_t* mlist;
    uint32_t* mbase;
    __device__ void operator()(uint64_t i, uint64_t pre) const {
        if (sscan[i + 1] - sscan[i] > 1) {
            mlist[(uint32_t)(pre & 0xffffffffull)] = (uint32_t)i;
            mbase[i] = (uint32_t)(pre >> 32);
        }
    }
};

// Hubs among the multi-segment entries: {entry, first spill slot, segments} appended to out
// (at most cap; *count may pass cap, which the host reads as "too many").
__global__ void k_find_hubs(const uint32_t* __restrict__ mlist, const uint32_t* __restrict__ sscan,
                            const uint32_t* __restrict__ mbase, uint32_t nmulti, uint32_t* __restrict__ count,
                            uint32_t cap, uint4* __restrict__ out) {
    for (uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; j < nmulti; j += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t i = mlist[j], ns = sscan[i + 1] - sscan[i];
        if (ns <= PPR_HUB_SEGS) continue;
        const uint32_t pos = atomicAdd(count, 1u);
        if (pos < cap) out[pos] = make_uint4(i, mbase[i], ns, 0u);
    }
}
__global__ void k_put(uint32_t* p, uint32_t v) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *p = v;
}
__global__ void k_out_degrees(const uint64_t* __restrict__ row, uint32_t n, uint32_t* __restrict__ deg) {
    for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += (uint64_t)gridDim.x * blockDim.x)
        deg[v] = (uint32_t)(row[v + 1] - row[v]);
}
// Descriptor of every segment: first edge, edge count (bit 31: the entry has this segment
// only), vertex, spill slot (entries with several segments).  Edge offsets fit 32 bits: graphs
// hold fewer than 2^32 edges.
__global__ void k_seg_desc(const uint64_t* __restrict__ row, const uint32_t* __restrict__ ent,
                           const uint32_t* __restrict__ sscan, const uint32_t* __restrict__ mbase, uint32_t nent,
                           uint4* __restrict__ desc) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nent; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t v = ent[i], s0 = sscan[i], ns = sscan[i + 1] - s0;
        const uint64_t r0 = row[v], r1 = row[v + 1];
        for (uint32_t k = 0; k < ns; ++k) {
            const uint64_t e0 = r0 + (uint64_t)k * PPR_SEG;
            const uint32_t len = (uint32_t)min((uint64_t)PPR_SEG, r1 - e0);
            desc[s0 + k] = make_uint4((uint32_t)e0, len | (ns == 1 ? 0x80000000u : 0u), v, ns > 1 ? mbase[i] + k : 0u);
        }
    }
}

struct PlaneArgsP {
    const uint64_t* row;     // in-edge CSR
    const uint32_t* col;
    const uint32_t* ent;
    const uint32_t* sscan;
    const uint4* sdesc;
    const uint32_t* mlist;
    const uint32_t* mbase;
    uint32_t nseg, nmulti;
    const uint32_t* outdeg;
    const float* c_old;
    const float2* x_old;     // ranks as compensated pairs (hi, lo)
    float2* x_new;
    float* c_new;
    float2* spill;
    double* res_part;
    float damping;
    PprCols cols;
    uint32_t hub_chunks;     // 1: hubs' spill slots hold chunk sums every PPR_HUB_SEGS slots
    uint32_t S;              // row stride of the planes and spill slots (columns allocated);
                             // a pass of B <= S columns reads and writes the first B of a row
};

// Old rank of entry vertex v in column q (x0 = e_src for a fresh column, nothing read).
template <int B>
__device__ __forceinline__ float2 old_rank(const PlaneArgsP& a, uint32_t v, uint32_t q, uint32_t src, bool fresh) {
    return fresh ? make_float2(v == src ? 1.0f : 0.0f, 0.0f) : a.x_old[(uint64_t)v * a.S + q];
}

// New rank of entry vertex v in column q from its gathered sum; adds the L1 change to res.
template <int B>
__device__ __forceinline__ void finish_col(const PlaneArgsP& a, uint32_t v, uint32_t q, uint32_t src, float2 xo,
                                           uint32_t od, Comp sum, double& res) {
    float hi, lo;
    rank_pair(v == src ? 1.0f - a.damping : 0.0f, a.damping, sum, &hi, &lo);
    const uint64_t o = (uint64_t)v * a.S + q;
    a.x_new[o] = make_float2(hi, lo);
    a.c_new[o] = od ? (hi + lo) / (float)od : 0.0f;
    res += fabs(((double)hi + (double)lo) - ((double)xo.x + (double)xo.y));
}

// Per-column FP64 residual of a workgroup: lanes q, q+B, ... of every wave, in a fixed order.
template <int B>
__device__ __forceinline__ void store_residual(double res, double* __restrict__ out) {
    __shared__ double sRed[WG / WAVE][B];
#pragma unroll
    for (int off = B; off < WAVE; off <<= 1) res += __shfl_xor(res, off, WAVE);
    const int lane = (int)lane_id();
    if (lane < B) sRed[threadIdx.x / WAVE][lane] = res;
    __syncthreads();
    if (threadIdx.x < B) {
        double t = 0;
        for (int w = 0; w < WG / WAVE; ++w) t += sRed[w][threadIdx.x];
        out[(uint64_t)blockIdx.x * B + threadIdx.x] = t;
    }
}

template <int B>
__global__ __launch_bounds__(WG) void k_pps_pull(PlaneArgsP a) {
    constexpr int G = WAVE / B;  // segments per wavefront
    const uint32_t lane = lane_id(), q = lane % B, grp = lane / B;
    const uint32_t src = a.cols.src[q];
    const bool fresh = (a.cols.fresh >> q) & 1u;
    // a fresh column's contributions are e_src / outdeg(src): nothing is read for it
    const uint32_t sod = (src != PPR_IDLE) ? a.outdeg[src] : 0u;
    const float c0 = sod ? 1.0f / (float)sod : 0.0f;
    double res = 0.0;
    const uint64_t nw = (uint64_t)gridDim.x * (WG / WAVE);
    for (uint64_t w = (uint64_t)blockIdx.x * (WG / WAVE) + threadIdx.x / WAVE; w * G < a.nseg; w += nw) {
        const uint64_t t = w * G + grp;
        if (t >= a.nseg) continue;
        const uint4 dsc = a.sdesc[t];
        const uint32_t e0 = dsc.x, len = dsc.y & 0x7fffffffu, v = dsc.z;
        const bool single = (dsc.y >> 31) != 0;
        // what finishing the vertex reads is loaded before the gathers, so its latency overlaps
        float2 xo = make_float2(0.0f, 0.0f);
        uint32_t od = 0;
        if (single) {
            xo = old_rank<B>(a, v, q, src, fresh);
            od = a.outdeg[v];
        }
        Comp acc{0.0f, 0.0f};
        // source indices of the next 8 edges are loaded while the current 8 rows are gathered
        uint32_t u[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) u[k] = ((uint32_t)k < len) ? a.col[e0 + k] : PPR_IDLE;
        for (uint32_t b = 0; b < len; b += 8) {
            float cv[8];
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                if (fresh) cv[k] = (u[k] == src) ? c0 : 0.0f;
                else cv[k] = (u[k] != PPR_IDLE) ? a.c_old[(uint64_t)u[k] * a.S + q] : 0.0f;
            }
#pragma unroll
            for (int k = 0; k < 8; ++k) u[k] = (b + 8 + k < len) ? a.col[e0 + b + 8 + k] : PPR_IDLE;
#pragma unroll
            for (int k = 0; k < 8; ++k)
                if (b + k < len) acc = comp_add(acc, Comp{cv[k], 0.0f});
        }
        if (single) finish_col<B>(a, v, q, src, xo, od, acc, res);
        else a.spill[(uint64_t)dsc.w * a.S + q] = make_float2(acc.s, acc.e);
    }
    store_residual<B>(res, a.res_part);
}

// cnt compensated partials at p, p + stride, ... added in order.  A hub has thousands of
// them (RMAT-24: 5794 segments), so the loads of the next 8 are issued before the current 8
// are added: the chain waits on the adds rather than on a load per partial.
template <int B>
__device__ __forceinline__ Comp sum_slots(const float2* __restrict__ p, uint32_t cnt, uint64_t stride) {
    Comp acc{0.0f, 0.0f};
    float2 cur[8], nxt[8];
#pragma unroll
    for (int r = 0; r < 8; ++r) cur[r] = ((uint32_t)r < cnt) ? p[(uint64_t)r * stride] : make_float2(0.0f, 0.0f);
    for (uint32_t k = 0; k < cnt; k += 8) {
#pragma unroll
        for (int r = 0; r < 8; ++r)
            nxt[r] = (k + 8 + r < cnt) ? p[(uint64_t)(k + 8 + r) * stride] : make_float2(0.0f, 0.0f);
#pragma unroll
        for (int r = 0; r < 8; ++r)
            if (k + r < cnt) acc = comp_add(acc, Comp{cur[r].x, cur[r].y});
#pragma unroll
        for (int r = 0; r < 8; ++r) cur[r] = nxt[r];
    }
    return acc;
}

// Hub chunks: chunk c of hub h (PPR_HUB_SEGS segments from its spill slot base + c *
// PPR_HUB_SEGS) is added in segment order and its sum replaces the chunk's first slot (read
// before the write by the same lanes).  hubs[h] = {entry, first slot, segments, first chunk}.
template <int B>
__global__ __launch_bounds__(WG) void k_pps_hub_chunks(float2* __restrict__ spill, const uint4* __restrict__ hubs,
                                                       uint32_t nhub, uint32_t nch, uint32_t S) {
    constexpr int G = WAVE / B;
    const uint32_t lane = lane_id(), q = lane % B, grp = lane / B;
    const uint64_t nw = (uint64_t)gridDim.x * (WG / WAVE);
    for (uint64_t w = (uint64_t)blockIdx.x * (WG / WAVE) + threadIdx.x / WAVE; w * G < nch; w += nw) {
        const uint64_t t = w * G + grp;
        if (t >= nch) continue;
        uint32_t lo = 0, hi = nhub;  // last hub whose first chunk is <= t
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if (hubs[mid].w <= t) lo = mid; else hi = mid;
        }
        const uint4 h = hubs[lo];
        const uint32_t c = (uint32_t)t - h.w;
        const uint32_t first = c * PPR_HUB_SEGS, cnt = min(PPR_HUB_SEGS, h.z - first);
        float2* p = spill + (uint64_t)(h.y + first) * S + q;
        const Comp acc = sum_slots<B>(p, cnt, (uint64_t)S);
        *p = make_float2(acc.s, acc.e);
    }
}

// Entries with several segments: their partial sums added in segment order.
template <int B>
__global__ __launch_bounds__(WG) void k_pps_fix(PlaneArgsP a, double* res_part) {
    constexpr int G = WAVE / B;
    const uint32_t lane = lane_id(), q = lane % B, grp = lane / B;
    const uint32_t src = a.cols.src[q];
    const bool fresh = (a.cols.fresh >> q) & 1u;
    double res = 0.0;
    const uint64_t nw = (uint64_t)gridDim.x * (WG / WAVE);
    for (uint64_t w = (uint64_t)blockIdx.x * (WG / WAVE) + threadIdx.x / WAVE; w * G < a.nmulti; w += nw) {
        const uint64_t j = w * G + grp;
        if (j >= a.nmulti) continue;
        const uint32_t i = a.mlist[j];
        const uint32_t ns = a.sscan[i + 1] - a.sscan[i], mb = a.mbase[i];
        const uint32_t v = a.ent[i];
        const float2 xo = old_rank<B>(a, v, q, src, fresh);
        const uint32_t od = a.outdeg[v];
        // a hub's chunk sums sit every PPR_HUB_SEGS slots (k_pps_hub_chunks)
        const bool hub = a.hub_chunks && ns > PPR_HUB_SEGS;
        const uint32_t cnt = hub ? (ns + PPR_HUB_SEGS - 1) / PPR_HUB_SEGS : ns;
        const uint64_t stride = hub ? (uint64_t)PPR_HUB_SEGS * a.S : (uint64_t)a.S;
        const Comp acc = sum_slots<B>(a.spill + (uint64_t)mb * a.S + q, cnt, stride);
        finish_col<B>(a, v, q, src, xo, od, acc, res);
    }
    store_residual<B>(res, res_part);
}

// Per-column residual: nparts partials summed in a fixed order (strided per thread, then a
// fixed tree over the threads of the column).
template <int B>
__global__ __launch_bounds__(WG) void k_pps_reduce(const double* __restrict__ part, uint32_t nparts,
                                                   double* __restrict__ resid) {
    constexpr int R = WG / B;
    __shared__ double s[R][B];
    const uint32_t q = threadIdx.x % B, r = threadIdx.x / B;
    double t = 0;
    for (uint32_t p = r; p < nparts; p += R) t += part[(uint64_t)p * B + q];
    s[r][q] = t;
    __syncthreads();
    for (int off = R / 2; off > 0; off >>= 1) {
        if ((int)r < off) s[r][q] += s[r + off][q];
        __syncthreads();
    }
    if (r == 0) resid[q] = s[0][q];
}

// out[u] = rank of original vertex u in column q: the plane value for entries, the restart
// term alone for vertices without edges.
template <int B>
__global__ void k_pps_extract(const float2* __restrict__ x, const uint64_t* __restrict__ in_row,
                              const uint32_t* __restrict__ outdeg, const uint32_t* __restrict__ perm, uint32_t n,
                              uint32_t q, uint32_t src, float restart, float* __restrict__ out) {
    for (uint64_t u = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; u < n; u += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t v = perm[u];
        const bool entry = in_row[v + 1] > in_row[v] || outdeg[v] > 0;
        const float2 r = x[(uint64_t)v * B + q];
        out[u] = entry ? r.x + r.y : (v == src ? restart : 0.0f);
    }
}

unsigned grid_of(uint64_t work, unsigned cap) {
    const uint64_t g = (work + WG - 1) / WG;
    return (unsigned)std::max<uint64_t>(1, std::min<uint64_t>(g, cap));
}

}  // namespace

eeg_status ensure_ppr_seg(eeg_graph* g) {
    PprSeg& t = g->pprseg;
    if (t.built) return EEG_OK;
    hipStream_t st = g->stream;
    const uint32_t n = g->n;
    if (!g->outdeg.p) {
        EEG_TRY(g->outdeg.alloc((size_t)n * 4 + 4));
        if (n) k_out_degrees<<<grid_of(n, 8192), WG, 0, st>>>(g->out.rowp(), n, g->outdeg.as<uint32_t>());
        EEG_HIP(hipGetLastError());
    }
    EEG_TRY(t.ent.alloc((size_t)n * 4 + 4));
    EEG_TRY(t.sscan.alloc((size_t)n * 4 + 8));
    EEG_TRY(t.mbase.alloc((size_t)n * 4 + 4));
    DevMem total;
    EEG_TRY(total.alloc(16));
    EEG_HIP(hipMemsetAsync(total.p, 0, 16, st));
    const uint64_t* row = g->in.rowp();
    const uint32_t* od = g->outdeg.as<uint32_t>();
    EEG_TRY(scan_exclusive(SegLoad{row, od}, SegStore{row, od, t.ent.as<uint32_t>(), t.sscan.as<uint32_t>()}, n, st,
                           total.as<uint64_t>()));
    uint64_t tot[2] = {0, 0};
    EEG_HIP(hipMemcpyAsync(tot, total.p, 8, hipMemcpyDeviceToHost, st));
    EEG_HIP(hipStreamSynchronize(st));
    if ((tot[0] >> 32) >= 0xffffffffull) return fail(EEG_ERR_UNSUPPORTED, "too many PageRank segments");
    t.nent = (uint32_t)(tot[0] & 0xffffffffull);
    t.nseg = (uint32_t)(tot[0] >> 32);
    k_put<<<1, 1, 0, st>>>(t.sscan.as<uint32_t>() + t.nent, t.nseg);
    EEG_TRY(t.mlist.alloc((size_t)t.nent * 4 + 4));
    EEG_HIP(hipMemsetAsync(total.as<uint64_t>() + 1, 0, 8, st));
    if (t.nent)
        EEG_TRY(scan_exclusive(MultiLoad{t.sscan.as<uint32_t>()},
                               MultiStore{t.sscan.as<uint32_t>(), t.mlist.as<uint32_t>(), t.mbase.as<uint32_t>()}, t.nent,
                               st, total.as<uint64_t>() + 1));
    EEG_HIP(hipMemcpyAsync(tot + 1, total.as<uint64_t>() + 1, 8, hipMemcpyDeviceToHost, st));
    EEG_HIP(hipStreamSynchronize(st));
    t.nmulti = (uint32_t)(tot[1] & 0xffffffffull);
    t.nspill = (uint32_t)(tot[1] >> 32);
    // hubs, for the chunked addition of their partial sums
    t.nhub = t.nhubch = 0;
    if (t.nmulti) {
        const uint32_t cap = 1u << 16;
        EEG_TRY(t.hubs.alloc((size_t)cap * 16));
        EEG_HIP(hipMemsetAsync(total.p, 0, 4, st));
        k_find_hubs<<<grid_of(t.nmulti, 8192), WG, 0, st>>>(t.mlist.as<uint32_t>(), t.sscan.as<uint32_t>(),
                                                             t.mbase.as<uint32_t>(), t.nmulti, total.as<uint32_t>(),
                                                             cap, t.hubs.as<uint4>());
        EEG_HIP(hipGetLastError());
        uint32_t nh = 0;
        EEG_HIP(hipMemcpyAsync(&nh, total.p, 4, hipMemcpyDeviceToHost, st));
        EEG_HIP(hipStreamSynchronize(st));
        if (nh > 0 && nh <= cap) {
            std::vector<uint4> h(nh);
            EEG_HIP(hipMemcpyAsync(h.data(), t.hubs.p, (size_t)nh * 16, hipMemcpyDeviceToHost, st));
            EEG_HIP(hipStreamSynchronize(st));
            std::sort(h.begin(), h.end(), [](const uint4& a, const uint4& b) { return a.x < b.x; });
            uint64_t ch = 0;
            for (auto& x : h) {
                x.w = (uint32_t)ch;
                ch += (x.z + PPR_HUB_SEGS - 1) / PPR_HUB_SEGS;
            }
            if (ch < 0xffffffffull) {
                EEG_HIP(hipMemcpyAsync(t.hubs.p, h.data(), (size_t)nh * 16, hipMemcpyHostToDevice, st));
                EEG_HIP(hipStreamSynchronize(st));
                t.nhub = nh;
                t.nhubch = (uint32_t)ch;
            }
        }
    }
    EEG_TRY(t.sdesc.alloc((size_t)t.nseg * 16 + 16));
    if (t.nent)
        k_seg_desc<<<grid_of(t.nent, 8192), WG, 0, st>>>(row, t.ent.as<uint32_t>(), t.sscan.as<uint32_t>(),
                                                         t.mbase.as<uint32_t>(), t.nent, t.sdesc.as<uint4>());
    EEG_HIP(hipGetLastError());
    EEG_HIP(hipStreamSynchronize(st));
    t.built = true;
    return EEG_OK;
}

size_t PprPlane::bytes_for(uint32_t n, uint32_t B, uint32_t nspill) {
    return (size_t)n * B * 4 * 6 + (size_t)nspill * B * 8 + (size_t)B * 8 * 4096;
}

eeg_status PprPlane::init(eeg_graph* g, uint32_t Bc, float d, hipStream_t s) {
    if (Bc != 8 && Bc != 16 && Bc != 32) return fail(EEG_ERR_INVALID_ARG, "PPR planes hold 8, 16 or 32 columns");
    B = Bc;
    n = g->n;
    damping = d;
    const PprSeg& t = g->pprseg;
    const size_t plane = (size_t)n * B * 4 + 16;
    for (int k = 0; k < 2; ++k) {
        EEG_TRY(c[k].alloc(plane));
        EEG_TRY(x[k].alloc(2 * plane));
        for (DevMem* m : {&c[k], &x[k]}) EEG_HIP(hipMemsetAsync(m->p, 0, m->bytes, s));
    }
    EEG_TRY(spill.alloc((size_t)t.nspill * B * 8 + 16));
    int cus = 256;
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, g->device);
    grid_pull = (unsigned)std::max<uint64_t>(1, std::min<uint64_t>((uint64_t)cus * 8, ((uint64_t)t.nseg + 15) / 16));
    grid_fix = (unsigned)std::max<uint64_t>(1, std::min<uint64_t>((uint64_t)cus * 4, ((uint64_t)t.nmulti + 15) / 16));
    EEG_TRY(res_part.alloc((size_t)(grid_pull + grid_fix) * B * 8));
    EEG_TRY(resid.alloc((size_t)B * 8));
    EEG_HIP(hipStreamSynchronize(s));
    return EEG_OK;
}

eeg_status PprPlane::iterate(eeg_graph* g, uint32_t j, const PprCols& cols, hipStream_t s, uint32_t W) {
    if (W == 0) W = B;
    if ((W != 8 && W != 16 && W != 32) || W > B) return fail(EEG_ERR_INVALID_ARG, "pass width %u of %u columns", W, B);
    for (uint32_t q = W; q < B; ++q)
        if (cols.src[q] != PPR_IDLE) return fail(EEG_ERR_INVALID_ARG, "column %u is beyond a pass of %u", q, W);
    const PprSeg& t = g->pprseg;
    const int rb = (int)((j & 1u) ^ 1u), wb = (int)(j & 1u);
    PlaneArgsP a{};
    a.row = g->in.rowp();
    a.col = g->in.colp();
    a.ent = t.ent.as<uint32_t>();
    a.sscan = t.sscan.as<uint32_t>();
    a.sdesc = t.sdesc.as<uint4>();
    a.mlist = t.mlist.as<uint32_t>();
    a.mbase = t.mbase.as<uint32_t>();
    a.nseg = t.nseg;
    a.nmulti = t.nmulti;
    a.outdeg = g->outdeg.as<uint32_t>();
    a.c_old = c[rb].as<float>();
    a.x_old = x[rb].as<float2>();
    a.c_new = c[wb].as<float>();
    a.x_new = x[wb].as<float2>();
    a.spill = spill.as<float2>();
    a.res_part = res_part.as<double>();
    a.damping = damping;
    a.cols = cols;
    a.hub_chunks = t.nhub ? 1u : 0u;
    a.S = B;
    // residual partials and the reduction are W columns wide
    double* fix_part = res_part.as<double>() + (size_t)grid_pull * W;
    const unsigned grid_hub = grid_of((uint64_t)t.nhubch * W, 4096);
    if (W == 8) {
        k_pps_pull<8><<<grid_pull, WG, 0, s>>>(a);
        if (t.nhub) k_pps_hub_chunks<8><<<grid_hub, WG, 0, s>>>(a.spill, t.hubs.as<uint4>(), t.nhub, t.nhubch, B);
        k_pps_fix<8><<<grid_fix, WG, 0, s>>>(a, fix_part);
        k_pps_reduce<8><<<1, WG, 0, s>>>(res_part.as<double>(), grid_pull + grid_fix, resid.as<double>());
    } else if (W == 16) {
        k_pps_pull<16><<<grid_pull, WG, 0, s>>>(a);
        if (t.nhub) k_pps_hub_chunks<16><<<grid_hub, WG, 0, s>>>(a.spill, t.hubs.as<uint4>(), t.nhub, t.nhubch, B);
        k_pps_fix<16><<<grid_fix, WG, 0, s>>>(a, fix_part);
        k_pps_reduce<16><<<1, WG, 0, s>>>(res_part.as<double>(), grid_pull + grid_fix, resid.as<double>());
    } else {
        k_pps_pull<32><<<grid_pull, WG, 0, s>>>(a);
        if (t.nhub) k_pps_hub_chunks<32><<<grid_hub, WG, 0, s>>>(a.spill, t.hubs.as<uint4>(), t.nhub, t.nhubch, B);
        k_pps_fix<32><<<grid_fix, WG, 0, s>>>(a, fix_part);
        k_pps_reduce<32><<<1, WG, 0, s>>>(res_part.as<double>(), grid_pull + grid_fix, resid.as<double>());
    }
    EEG_HIP(hipGetLastError());
    return EEG_OK;
}

uint32_t PprPlane::width_for(const PprCols& cols) const {
    uint32_t top = 0;  // one past the highest busy column
    for (uint32_t q = 0; q < B; ++q)
        if (cols.src[q] != PPR_IDLE) top = q + 1;
    const uint32_t w = top <= 8 ? 8u : top <= 16 ? 16u : 32u;
    return std::min(w, B);
}

eeg_status PprPlane::extract(eeg_graph* g, uint32_t j, uint32_t q, uint32_t src, float* out, hipStream_t s) {
    const int wb = (int)(j & 1u);
    const unsigned gr = grid_of(n, 8192);
    if (B == 8)
        k_pps_extract<8><<<gr, WG, 0, s>>>(x[wb].as<float2>(), g->in.rowp(), g->outdeg.as<uint32_t>(),
                                           g->perm.as<uint32_t>(), n, q, src, 1.0f - damping, out);
    else if (B == 16)
        k_pps_extract<16><<<gr, WG, 0, s>>>(x[wb].as<float2>(), g->in.rowp(), g->outdeg.as<uint32_t>(),
                                            g->perm.as<uint32_t>(), n, q, src, 1.0f - damping, out);
    else
        k_pps_extract<32><<<gr, WG, 0, s>>>(x[wb].as<float2>(), g->in.rowp(), g->outdeg.as<uint32_t>(),
                                            g->perm.as<uint32_t>(), n, q, src, 1.0f - damping, out);
    EEG_HIP(hipGetLastError());
    return EEG_OK;
}

// Batched PPR (eeg_run_batch): up to B columns in flight; every column iterates from e_src
// until its own L1 change drops below epsilon, is read out and takes the next source, so each
// answer is the single-query iteration's, summed in a fixed order.
eeg_status run_ppr_batch(eeg_graph* g, const uint32_t* sources, uint32_t nsrc, const eeg_opts& o, float* out_host,
                         eeg_stats* st) {
    const float d = o.damping > 0 ? o.damping : 0.85f;
    if (!(d > 0 && d < 1)) return fail(EEG_ERR_INVALID_ARG, "damping must be in (0,1)");
    const double eps = o.epsilon > 0 ? o.epsilon : 1e-7;
    const uint32_t max_iters = o.max_iters ? o.max_iters : 10000;
    const uint32_t n = g->n;
    EEG_TRY(ensure_in(g));
    EEG_TRY(ensure_ppr_seg(g));
    PprPlane pl;
    EEG_TRY(pl.init(g, nsrc > 16 ? 32 : nsrc > 8 ? 16 : 8, d, g->stream));
    const uint32_t B = pl.B;
    hipStream_t s = g->stream;
    DevMem dout;
    if (out_host) EEG_TRY(dout.alloc((size_t)n * 4 + 4));
    std::vector<double> hres(B);
    hipEvent_t e0, e1;
    EEG_HIP(hipEventCreate(&e0));
    EEG_HIP(hipEventCreate(&e1));
    EEG_HIP(hipEventRecord(e0, s));
    uint32_t total_iters = 0, max_col_iters = 0;
    eeg_status rc = EEG_OK;
    // columns are refilled as soon as their query converges (the next source starts fresh in
    // the following iteration), so a pass never waits for its slowest column
    PprCols cols;
    for (uint32_t q = 0; q < 32; ++q) cols.src[q] = PPR_IDLE;
    cols.fresh = 0;
    std::vector<uint32_t> colq(B, 0), start(B, 0);
    uint32_t next = 0, left = nsrc;
    for (uint32_t j = 1; left > 0 && rc == EEG_OK; ++j) {
        for (uint32_t q = 0; q < B && next < nsrc; ++q) {
            if (cols.src[q] != PPR_IDLE) continue;
            cols.src[q] = g->h_perm[sources[next]];
            cols.fresh |= 1u << q;
            colq[q] = next++;
            start[q] = j;
        }
        // columns are refilled lowest first, so a pass needs only the width of the busy ones
        // (a pass of 8 or 16 columns costs about 0.3 or 0.6 of one of 32)
        EEG_TRY(pl.iterate(g, j, cols, s, pl.width_for(cols)));
        EEG_HIP(hipMemcpyAsync(hres.data(), pl.resid.p, (size_t)B * 8, hipMemcpyDeviceToHost, s));
        EEG_HIP(hipStreamSynchronize(s));
        ++total_iters;
        cols.fresh = 0;
        for (uint32_t q = 0; q < B; ++q) {
            if (cols.src[q] == PPR_IDLE) continue;
            const uint32_t its = j - start[q] + 1;
            if (!(hres[q] < eps)) {
                if (its >= max_iters) {
                    rc = fail(EEG_ERR_NOT_CONVERGED, "PPR residual above %g after %u iterations", eps, max_iters);
                    break;
                }
                continue;
            }
            max_col_iters = std::max(max_col_iters, its);
            if (out_host) {
                EEG_TRY(pl.extract(g, j, q, cols.src[q], dout.as<float>(), s));
                EEG_HIP(hipMemcpyAsync(out_host + (uint64_t)colq[q] * n, dout.p, (size_t)n * 4, hipMemcpyDeviceToHost, s));
            }
            cols.src[q] = PPR_IDLE;
            --left;
        }
    }
    EEG_HIP(hipEventRecord(e1, s));
    EEG_HIP(hipStreamSynchronize(s));
    float ms = 0;
    (void)hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (st) {
        st->iterations = max_col_iters;
        st->launches = total_iters;
        st->ms_traversal = ms;
        st->edges_relaxed = (uint64_t)total_iters * g->m;
    }
    return rc;
}

}  // namespace eeg
