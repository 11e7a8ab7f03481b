// Device-side graph construction: out-edge CSR, Tigr virtual-node tables, symmetrised and
// in-edge CSRs, the reference Tigr layout for export, and the device RMAT generator.
//
// The reference builds its layout on the host, serially, outside the timers
// (src/virtual_graph.cpp:36-112); here every step runs in HBM:
//   degree histogram (atomics) -> exclusive scan (row offsets) -> stable placement by source
//   -> per-vertex ceil(deg/K) + nonzero flag, packed and scanned -> run-length Tigr tables.
#include <cstring>  // rocprim headers use memset

#include <rocprim/device/device_radix_sort.hpp>

#include <vector>

#include "engine.hpp"
#include "scan.hpp"

namespace eeg {
namespace {

// ---------------------------------------------------------------- CSR build
// Counts edges with an endpoint >= n (one atomic per wavefront that saw one, so the valid case
// has no contention).
__global__ void k_validate(const uint32_t* __restrict__ src, const uint32_t* __restrict__ dst,
                           uint64_t m, uint32_t n, uint32_t* __restrict__ bad) {
    for (uint64_t e0 = (uint64_t)blockIdx.x * blockDim.x; e0 < m; e0 += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t e = e0 + threadIdx.x;
        const bool b = e < m && (src[e] >= n || dst[e] >= n);
        const uint64_t mask = __ballot(b);
        if (mask && lane_id() == 0) atomicAdd(bad, (uint32_t)__popcll(mask));
    }
}

// row[v] = first position of key v in the sorted keys (lower bound), v in [0, n].
__global__ void k_rows_from_sorted(const uint32_t* __restrict__ keys, uint64_t m, uint32_t n,
                                   uint64_t* __restrict__ row) {
    for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v <= n;
         v += (uint64_t)gridDim.x * blockDim.x) {
        uint64_t lo = 0, hi = m;
        while (lo < hi) {
            const uint64_t mid = lo + (hi - lo) / 2;
            if (keys[mid] < v) lo = mid + 1; else hi = mid;
        }
        row[v] = lo;
    }
}

struct RowStore {
    uint64_t* row;
    __device__ void operator()(uint64_t i, uint64_t v) const { row[i] = v; }
};

__global__ void k_iota(uint32_t* __restrict__ a, uint64_t m) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m;
         i += (uint64_t)gridDim.x * blockDim.x)
        a[i] = (uint32_t)i;
}

__global__ void k_gather(const uint32_t* __restrict__ perm, const uint32_t* __restrict__ a,
                         const uint32_t* __restrict__ b, uint32_t* __restrict__ ao,
                         uint32_t* __restrict__ bo, uint64_t m) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m;
         i += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t p = perm[i];
        ao[i] = a[p];
        if (b) bo[i] = b[p];
    }
}

// ---------------------------------------------------------------- degree-ordered relabeling
// Vertices are renumbered by decreasing total degree (ties by id), so the labels and
// contributions that gathers hit most often sit together and stay in each XCD's L2.
// perm[old] = new, inv[new] = old.
__global__ void k_degree_key(const uint64_t* __restrict__ orow, const uint64_t* __restrict__ irow,
                             uint32_t n, uint32_t* __restrict__ key) {
    for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n;
         v += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t d = (orow[v + 1] - orow[v]) + (irow[v + 1] - irow[v]);
        key[v] = 0xffffffffu - (uint32_t)(d > 0xffffffffull ? 0xffffffffull : d);
    }
}
__global__ void k_invert(const uint32_t* __restrict__ inv, uint32_t n, uint32_t* __restrict__ perm) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        perm[inv[i]] = (uint32_t)i;
}
__global__ void k_relabel(const uint32_t* __restrict__ perm, const uint32_t* __restrict__ a,
                          const uint32_t* __restrict__ b, uint32_t* __restrict__ ao,
                          uint32_t* __restrict__ bo, uint64_t m) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m;
         i += (uint64_t)gridDim.x * blockDim.x) {
        ao[i] = perm[a[i]];
        bo[i] = perm[b[i]];
    }
}
// out[u] = in[perm[u]]: results of the relabeled graph in original vertex order
__global__ void k_unpermute(const uint32_t* __restrict__ perm, uint32_t n, const uint32_t* __restrict__ in,
                            uint32_t* __restrict__ out) {
    for (uint64_t u = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; u < n;
         u += (uint64_t)gridDim.x * blockDim.x)
        out[u] = in[perm[u]];
}
struct OrigDegLoad {  // degree of original vertex u (element n: 0)
    const uint64_t* row;
    const uint32_t* perm;
    uint32_t n;
    __device__ uint64_t operator()(uint64_t u) const {
        if (u >= n) return 0;
        const uint32_t v = perm[u];
        return row[v + 1] - row[v];
    }
};
// Rows of the relabeled CSR copied back in original ids and order.
__global__ void k_orig_rows(const uint64_t* __restrict__ row, const uint32_t* __restrict__ col,
                            const uint32_t* __restrict__ w, const uint32_t* __restrict__ perm,
                            const uint32_t* __restrict__ inv, uint32_t n,
                            const uint64_t* __restrict__ orow, uint32_t* __restrict__ ocol,
                            uint32_t* __restrict__ ow) {
    for (uint64_t u = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; u < n;
         u += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t v = perm[u];
        const uint64_t r0 = row[v], d = row[v + 1] - r0, o = orow[u];
        for (uint64_t j = 0; j < d; ++j) {
            ocol[o + j] = inv[col[r0 + j]];
            if (w) ow[o + j] = w[r0 + j];
        }
    }
}

// max out-degree, lowest id among the maxima, and the count of vertices with edges
__global__ void k_degree_stats(const uint64_t* __restrict__ row, uint32_t n,
                               unsigned long long* __restrict__ best, uint32_t* __restrict__ nnz) {
    unsigned long long b = 0;
    uint32_t c = 0;
    for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n;
         v += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t d = row[v + 1] - row[v];
        const uint32_t dd = d > 0xffffffffull ? 0xffffffffu : (uint32_t)d;
        const unsigned long long key = ((unsigned long long)dd << 32) | (uint32_t)(~(uint32_t)v);
        b = key > b ? key : b;
        c += d > 0;
    }
    for (int off = WAVE / 2; off > 0; off >>= 1) {
        unsigned long long o = __shfl_xor(b, off, WAVE);
        b = o > b ? o : b;
    }
    c = wave_sum(c);
    if (lane_id() == 0) {
        atomicMax(best, b);
        atomicAdd(nnz, c);
    }
}

unsigned grid_for(uint64_t work, unsigned cap = 65536) {
    uint64_t g = (work + WG - 1) / WG;
    if (g == 0) g = 1;
    return (unsigned)(g > cap ? cap : g);
}

uint32_t key_bits(uint32_t n) {
    uint32_t b = 1;
    while (b < 32 && (1ull << b) < (uint64_t)n) ++b;
    return b;
}

// Builds rows of `c` from (keys, vals, w) with keys < n: a stable sort of (key, edge index)
// pairs gives the order (input order within a key); row offsets are lower bounds of each
// vertex id in the sorted keys.
eeg_status build_rows(hipStream_t st, uint32_t n, const uint32_t* keys, const uint32_t* vals,
                      const uint32_t* w, uint64_t m, Csr& c) {
    c.n = n;
    c.m = m;
    EEG_TRY(c.row.alloc((size_t)(n + 1ull) * 8));
    EEG_TRY(c.col.alloc(m * 4));
    if (w) EEG_TRY(c.w.alloc(m * 4));
    DevMem bad;
    EEG_TRY(bad.alloc(8));
    EEG_HIP(hipMemsetAsync(bad.p, 0, 8, st));
    k_validate<<<grid_for(m), WG, 0, st>>>(keys, vals, m, n, bad.as<uint32_t>());
    EEG_HIP(hipGetLastError());
    uint32_t nbad = 0;
    EEG_HIP(hipMemcpyAsync(&nbad, bad.p, 4, hipMemcpyDeviceToHost, st));
    EEG_HIP(hipStreamSynchronize(st));
    if (nbad) return fail(EEG_ERR_INVALID_ARG, "%u edges have an endpoint >= n (%u)", nbad, n);
    if (m == 0) {
        EEG_HIP(hipMemsetAsync(c.row.p, 0, (size_t)(n + 1ull) * 8, st));
        EEG_HIP(hipStreamSynchronize(st));
        c.built = true;
        return EEG_OK;
    }
    DevMem kout, idx, perm, tmp;
    EEG_TRY(kout.alloc(m * 4));
    EEG_TRY(idx.alloc(m * 4));
    EEG_TRY(perm.alloc(m * 4));
    k_iota<<<grid_for(m), WG, 0, st>>>(idx.as<uint32_t>(), m);
    EEG_HIP(hipGetLastError());
    size_t tbytes = 0;
    const uint32_t bits = key_bits(n);
    EEG_HIP(rocprim::radix_sort_pairs(nullptr, tbytes, keys, kout.as<uint32_t>(),
                                      idx.as<uint32_t>(), perm.as<uint32_t>(), (size_t)m, 0, bits,
                                      st));
    EEG_TRY(tmp.alloc(tbytes));
    EEG_HIP(rocprim::radix_sort_pairs(tmp.p, tbytes, keys, kout.as<uint32_t>(),
                                      idx.as<uint32_t>(), perm.as<uint32_t>(), (size_t)m, 0, bits,
                                      st));
    k_rows_from_sorted<<<grid_for((uint64_t)n + 1), WG, 0, st>>>(kout.as<uint32_t>(), m, n,
                                                                 c.row.as<uint64_t>());
    EEG_HIP(hipGetLastError());
    k_gather<<<grid_for(m), WG, 0, st>>>(perm.as<uint32_t>(), vals, w, c.col.as<uint32_t>(),
                                         w ? c.w.as<uint32_t>() : nullptr, m);
    EEG_HIP(hipGetLastError());
    EEG_HIP(hipStreamSynchronize(st));
    c.built = true;
    return EEG_OK;
}

// ---------------------------------------------------------------- Tigr tables
struct TigrLoad {  // (ceil(deg/K) << 32) | (deg > 0)
    const uint64_t* row;
    uint32_t K;
    __device__ uint64_t operator()(uint64_t v) const {
        const uint64_t d = row[v + 1] - row[v];
        return ((d + K - 1) / K << 32) | (uint64_t)(d > 0);
    }
};
struct TigrStore {
    const uint64_t* row;
    uint32_t* vlist;
    uint32_t* vscan;
    __device__ void operator()(uint64_t v, uint64_t pre) const {
        if (row[v + 1] > row[v]) {
            const uint32_t pos = (uint32_t)(pre & 0xffffffffull);
            vlist[pos] = (uint32_t)v;
            vscan[pos] = (uint32_t)(pre >> 32);
        }
    }
};

__global__ void k_put_u32(uint32_t* a, uint64_t i, uint32_t v) { a[i] = v; }

__global__ void k_vnode_counts(const uint64_t* __restrict__ row, uint32_t n, uint32_t K,
                               uint32_t* __restrict__ vnc) {
    for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n;
         v += (uint64_t)gridDim.x * blockDim.x)
        vnc[v] = (uint32_t)((row[v + 1] - row[v] + K - 1) / K);
}

// ---------------------------------------------------------------- expand sources
__global__ void k_expand_src(const uint64_t* __restrict__ row, uint32_t n, uint64_t m,
                             uint32_t* __restrict__ src) {
    for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < m;
         e += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t lo = 0, hi = n;  // largest v with row[v] <= e
        while (hi - lo > 1) {
            const uint32_t mid = lo + (hi - lo) / 2;
            if (row[mid] <= e) lo = mid; else hi = mid;
        }
        src[e] = lo;
    }
}

// ---------------------------------------------------------------- reference layout export
// node_pointer[v] = 2*row[v] + v; edge_list[np] = deg; dst/w interleaved after it
// (src/virtual_graph.cpp:50-64, 84-91); parts (v, p) in ascending order (:97-111).
__global__ void k_export_tigr(const uint64_t* __restrict__ row, const uint32_t* __restrict__ col,
                              const uint32_t* __restrict__ w, uint32_t n, uint32_t K,
                              const uint64_t* __restrict__ vnoff, uint64_t* __restrict__ np,
                              uint32_t* __restrict__ el, uint32_t* __restrict__ parts) {
    for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n;
         v += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t r0 = row[v], d = row[v + 1] - r0, base = 2 * r0 + v;
        if (np) np[v] = base;
        if (el) {
            el[base] = (uint32_t)d;
            for (uint64_t j = 0; j < d; ++j) {
                el[base + 1 + 2 * j] = col[r0 + j];
                el[base + 2 + 2 * j] = w ? w[r0 + j] : 1u;
            }
        }
        if (parts) {
            const uint64_t nv = (d + K - 1) / K, o = vnoff[v];
            for (uint64_t p = 0; p < nv; ++p) {
                parts[2 * (o + p)] = (uint32_t)v;
                parts[2 * (o + p) + 1] = (uint32_t)p;
            }
        }
    }
}
struct VnLoad {
    const uint64_t* row;
    uint32_t K;
    __device__ uint64_t operator()(uint64_t v) const {
        return (row[v + 1] - row[v] + K - 1) / K;
    }
};

// ---------------------------------------------------------------- device RMAT
__global__ void k_rmat(uint64_t key, uint64_t wkey, uint32_t scale, uint64_t ta, uint64_t tab,
                       uint64_t tabc, uint64_t m, uint32_t wmin, uint32_t wmax,
                       uint32_t* __restrict__ s, uint32_t* __restrict__ d, uint32_t* __restrict__ w) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m;
         i += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t a, b;
        rmat_edge(key, scale, ta, tab, tabc, i, &a, &b);
        s[i] = a;
        d[i] = b;
        if (w) w[i] = edge_weight(wkey, i, wmin, wmax);
    }
}

}  // namespace

// Row offsets of the list sorted by `keys` (keys only; used for degrees in original ids).
static eeg_status rows_from_keys(hipStream_t st, const uint32_t* keys, uint64_t m, uint32_t n,
                                 DevMem& row) {
    EEG_TRY(row.alloc(((size_t)n + 1) * 8));
    if (m == 0) {
        EEG_HIP(hipMemsetAsync(row.p, 0, ((size_t)n + 1) * 8, st));
        return EEG_OK;
    }
    DevMem kout, tmp;
    EEG_TRY(kout.alloc(m * 4));
    size_t tbytes = 0;
    const uint32_t bits = key_bits(n);
    EEG_HIP(rocprim::radix_sort_keys(nullptr, tbytes, keys, kout.as<uint32_t>(), (size_t)m, 0, bits, st));
    EEG_TRY(tmp.alloc(tbytes));
    EEG_HIP(rocprim::radix_sort_keys(tmp.p, tbytes, keys, kout.as<uint32_t>(), (size_t)m, 0, bits, st));
    k_rows_from_sorted<<<grid_for((uint64_t)n + 1), WG, 0, st>>>(kout.as<uint32_t>(), m, n,
                                                                 row.as<uint64_t>());
    EEG_HIP(hipGetLastError());
    EEG_HIP(hipStreamSynchronize(st));
    return EEG_OK;
}

eeg_status build_csr_from_device_edges(eeg_graph* g, const uint32_t* d_src, const uint32_t* d_dst,
                                       const uint32_t* d_w, uint64_t m) {
    hipStream_t st = g->stream;
    const uint32_t n = g->n;
    {
        DevMem bad;
        EEG_TRY(bad.alloc(8));
        EEG_HIP(hipMemsetAsync(bad.p, 0, 8, st));
        if (m) {
            k_validate<<<grid_for(m), WG, 0, st>>>(d_src, d_dst, m, n, bad.as<uint32_t>());
            EEG_HIP(hipGetLastError());
        }
        uint32_t nbad = 0;
        EEG_HIP(hipMemcpyAsync(&nbad, bad.p, 4, hipMemcpyDeviceToHost, st));
        EEG_HIP(hipStreamSynchronize(st));
        if (nbad) return fail(EEG_ERR_INVALID_ARG, "%u edges have an endpoint >= n (%u)", nbad, n);
    }
    // degrees in original ids, then the degree order
    DevMem orow, irow;
    EEG_TRY(rows_from_keys(st, d_src, m, n, orow));
    EEG_TRY(rows_from_keys(st, d_dst, m, n, irow));
    EEG_TRY(g->perm.alloc((size_t)n * 4 + 4));
    EEG_TRY(g->inv.alloc((size_t)n * 4 + 4));
    if (n) {
        DevMem key, kout, ids, tmp;
        EEG_TRY(key.alloc((size_t)n * 4));
        EEG_TRY(kout.alloc((size_t)n * 4));
        EEG_TRY(ids.alloc((size_t)n * 4));
        k_degree_key<<<grid_for(n), WG, 0, st>>>(orow.as<uint64_t>(), irow.as<uint64_t>(), n,
                                                 key.as<uint32_t>());
        k_iota<<<grid_for(n), WG, 0, st>>>(ids.as<uint32_t>(), n);
        EEG_HIP(hipGetLastError());
        size_t tbytes = 0;
        EEG_HIP(rocprim::radix_sort_pairs(nullptr, tbytes, key.as<uint32_t>(), kout.as<uint32_t>(),
                                          ids.as<uint32_t>(), g->inv.as<uint32_t>(), (size_t)n, 0, 32, st));
        EEG_TRY(tmp.alloc(tbytes));
        EEG_HIP(rocprim::radix_sort_pairs(tmp.p, tbytes, key.as<uint32_t>(), kout.as<uint32_t>(),
                                          ids.as<uint32_t>(), g->inv.as<uint32_t>(), (size_t)n, 0, 32, st));
        k_invert<<<grid_for(n), WG, 0, st>>>(g->inv.as<uint32_t>(), n, g->perm.as<uint32_t>());
        EEG_HIP(hipGetLastError());
        EEG_HIP(hipStreamSynchronize(st));
    }
    irow.release();
    {
        DevMem s2, d2;
        EEG_TRY(s2.alloc(m * 4 + 4));
        EEG_TRY(d2.alloc(m * 4 + 4));
        if (m) {
            k_relabel<<<grid_for(m), WG, 0, st>>>(g->perm.as<uint32_t>(), d_src, d_dst,
                                                  s2.as<uint32_t>(), d2.as<uint32_t>(), m);
            EEG_HIP(hipGetLastError());
        }
        EEG_TRY(build_rows(st, n, s2.as<uint32_t>(), d2.as<uint32_t>(), d_w, m, g->out));
    }
    g->h_perm.resize(n);
    if (n) EEG_HIP(hipMemcpy(g->h_perm.data(), g->perm.p, (size_t)n * 4, hipMemcpyDeviceToHost));
    // statistics in original ids
    DevMem best, nnz;
    EEG_TRY(best.alloc(8));
    EEG_TRY(nnz.alloc(4));
    EEG_HIP(hipMemsetAsync(best.p, 0, 8, st));
    EEG_HIP(hipMemsetAsync(nnz.p, 0, 4, st));
    if (n) {
        k_degree_stats<<<grid_for(n, 4096), WG, 0, st>>>(orow.as<uint64_t>(), n,
                                                         best.as<unsigned long long>(), nnz.as<uint32_t>());
        EEG_HIP(hipGetLastError());
    }
    unsigned long long b = 0;
    uint32_t c = 0;
    EEG_HIP(hipMemcpyAsync(&b, best.p, 8, hipMemcpyDeviceToHost, st));
    EEG_HIP(hipMemcpyAsync(&c, nnz.p, 4, hipMemcpyDeviceToHost, st));
    EEG_HIP(hipStreamSynchronize(st));
    g->max_deg = (uint32_t)(b >> 32);
    g->argmax_deg = n ? ~(uint32_t)(b & 0xffffffffull) : 0;
    g->nonzero_out = c;
    return EEG_OK;
}

// A CSR of the relabeled graph `src` expressed in original ids and original row order.
eeg_status orig_csr(eeg_graph* g, const Csr& src, Csr& out) {
    hipStream_t st = g->stream;
    const uint32_t n = g->n;
    out.n = n;
    out.m = src.m;
    EEG_TRY(out.row.alloc(((size_t)n + 1) * 8));
    EEG_TRY(out.col.alloc(src.m * 4 + 4));
    if (src.w.p) EEG_TRY(out.w.alloc(src.m * 4 + 4));
    EEG_TRY(scan_exclusive(OrigDegLoad{src.rowp(), g->perm.as<uint32_t>(), n},
                           RowStore{out.row.as<uint64_t>()}, (uint64_t)n + 1, st, nullptr));
    if (n) {
        k_orig_rows<<<grid_for(n, 8192), WG, 0, st>>>(src.rowp(), src.colp(), src.wp(),
                                                       g->perm.as<uint32_t>(), g->inv.as<uint32_t>(), n,
                                                       out.rowp(), out.col.as<uint32_t>(),
                                                       src.w.p ? out.w.as<uint32_t>() : nullptr);
        EEG_HIP(hipGetLastError());
    }
    EEG_HIP(hipStreamSynchronize(st));
    out.built = true;
    return EEG_OK;
}

eeg_status unpermute(eeg_graph* g, const void* d_in, void* d_out, hipStream_t st) {
    if (!g->n) return EEG_OK;
    k_unpermute<<<grid_for(g->n, 8192), WG, 0, st>>>(g->perm.as<uint32_t>(), g->n,
                                                      (const uint32_t*)d_in, (uint32_t*)d_out);
    EEG_HIP(hipGetLastError());
    return EEG_OK;
}

eeg_status ensure_tigr(eeg_graph* g, Csr& c, Tigr* table, uint32_t K) {
    const int ki = k_index(K);
    if (ki < 0) return fail(EEG_ERR_INVALID_ARG, "K must be 4, 8, 16 or 32 (got %u)", K);
    Tigr& t = table[ki];
    if (t.built) return EEG_OK;
    hipStream_t st = g->stream;
    t.K = K;
    DevMem total;
    EEG_TRY(total.alloc(8));
    EEG_TRY(t.vlist.alloc((size_t)c.n * 4 + 4));
    EEG_TRY(t.vscan.alloc((size_t)c.n * 4 + 8));
    EEG_HIP(hipMemsetAsync(total.p, 0, 8, st));
    EEG_TRY(scan_exclusive(TigrLoad{c.rowp(), K},
                           TigrStore{c.rowp(), t.vlist.as<uint32_t>(), t.vscan.as<uint32_t>()},
                           c.n, st, total.as<uint64_t>()));
    uint64_t tot = 0;
    EEG_HIP(hipMemcpyAsync(&tot, total.p, 8, hipMemcpyDeviceToHost, st));
    EEG_HIP(hipStreamSynchronize(st));
    t.nz = (uint32_t)(tot & 0xffffffffull);
    t.P = tot >> 32;
    if (t.P >= 0xffffffffull)
        return fail(EEG_ERR_UNSUPPORTED, "%llu virtual nodes exceed 2^32-1; use a larger K",
                    (unsigned long long)t.P);
    k_put_u32<<<1, 1, 0, st>>>(t.vscan.as<uint32_t>(), t.nz, (uint32_t)t.P);
    EEG_HIP(hipGetLastError());
    EEG_TRY(t.vnc.alloc((size_t)c.n * 4 + 4));
    if (c.n) {
        k_vnode_counts<<<grid_for(c.n, 8192), WG, 0, st>>>(c.rowp(), c.n, K, t.vnc.as<uint32_t>());
        EEG_HIP(hipGetLastError());
    }
    EEG_HIP(hipStreamSynchronize(st));
    t.built = true;
    return EEG_OK;
}

static eeg_status expand_sources(eeg_graph* g, DevMem& src) {
    EEG_TRY(src.alloc(g->m * 4 + 4));
    if (g->m) {
        k_expand_src<<<grid_for(g->m), WG, 0, g->stream>>>(g->out.rowp(), g->n, g->m,
                                                           src.as<uint32_t>());
        EEG_HIP(hipGetLastError());
    }
    return EEG_OK;
}

eeg_status ensure_sym(eeg_graph* g) {
    if (g->sym.built) return EEG_OK;
    if (2 * g->m >= 0xffffffffull)
        return fail(EEG_ERR_UNSUPPORTED, "symmetrised edge count exceeds 2^32-1");
    DevMem src, keys, vals;
    EEG_TRY(expand_sources(g, src));
    const uint64_t m = g->m;
    EEG_TRY(keys.alloc(2 * m * 4 + 4));
    EEG_TRY(vals.alloc(2 * m * 4 + 4));
    hipStream_t st = g->stream;
    if (m) {
        EEG_HIP(hipMemcpyAsync(keys.p, src.p, m * 4, hipMemcpyDeviceToDevice, st));
        EEG_HIP(hipMemcpyAsync(keys.as<uint32_t>() + m, g->out.col.p, m * 4,
                               hipMemcpyDeviceToDevice, st));
        EEG_HIP(hipMemcpyAsync(vals.p, g->out.col.p, m * 4, hipMemcpyDeviceToDevice, st));
        EEG_HIP(hipMemcpyAsync(vals.as<uint32_t>() + m, src.p, m * 4, hipMemcpyDeviceToDevice,
                               st));
    }
    src.release();
    return build_rows(st, g->n, keys.as<uint32_t>(), vals.as<uint32_t>(), nullptr, 2 * m, g->sym);
}

eeg_status ensure_in(eeg_graph* g) {
    if (g->in.built) return EEG_OK;
    DevMem src;
    EEG_TRY(expand_sources(g, src));
    return build_rows(g->stream, g->n, g->out.colp(), src.as<uint32_t>(), nullptr, g->m, g->in);
}

eeg_status export_tigr(eeg_graph* g, uint32_t K, uint64_t* node_pointer, uint32_t* edge_list,
                       uint32_t* parts, uint64_t* num_parts) {
    if (k_index(K) < 0) return fail(EEG_ERR_INVALID_ARG, "K must be 4, 8, 16 or 32");
    EEG_TRY(ensure_tigr(g, g->out, g->tout, K));
    const Tigr& t = g->tout[k_index(K)];
    if (num_parts) *num_parts = t.P;
    if (!node_pointer && !edge_list && !parts) return EEG_OK;
    hipStream_t st = g->stream;
    const uint32_t n = g->n;
    const uint64_t m = g->m;
    Csr oc;  // the layout is defined on original ids in input order
    EEG_TRY(orig_csr(g, g->out, oc));
    DevMem np, el, pp, vnoff;
    if (node_pointer) EEG_TRY(np.alloc((size_t)n * 8 + 8));
    if (edge_list) EEG_TRY(el.alloc((2 * m + n) * 4 + 4));
    if (parts) {
        EEG_TRY(pp.alloc(2 * t.P * 4 + 8));
        EEG_TRY(vnoff.alloc((size_t)n * 8 + 8));
        EEG_TRY(scan_exclusive(VnLoad{oc.rowp(), K}, RowStore{vnoff.as<uint64_t>()}, n, st,
                               nullptr));
    }
    if (n) {
        k_export_tigr<<<grid_for(n, 8192), WG, 0, st>>>(
            oc.rowp(), oc.colp(), oc.wp(), n, K,
            parts ? vnoff.as<uint64_t>() : nullptr, node_pointer ? np.as<uint64_t>() : nullptr,
            edge_list ? el.as<uint32_t>() : nullptr, parts ? pp.as<uint32_t>() : nullptr);
        EEG_HIP(hipGetLastError());
    }
    if (node_pointer) EEG_HIP(hipMemcpyAsync(node_pointer, np.p, (size_t)n * 8, hipMemcpyDeviceToHost, st));
    if (edge_list) EEG_HIP(hipMemcpyAsync(edge_list, el.p, (2 * m + n) * 4, hipMemcpyDeviceToHost, st));
    if (parts) EEG_HIP(hipMemcpyAsync(parts, pp.p, 2 * t.P * 4, hipMemcpyDeviceToHost, st));
    EEG_HIP(hipStreamSynchronize(st));
    return EEG_OK;
}

eeg_status generate_rmat_device(uint32_t scale, uint32_t ef, double a, double b, double c,
                                uint64_t seed, uint32_t wmin, uint32_t wmax, uint32_t* d_src,
                                uint32_t* d_dst, uint32_t* d_w, hipStream_t st) {
    const uint64_t m = (uint64_t)ef << scale;
    const auto th = rmat_thresholds(a, b, c);
    k_rmat<<<grid_for(m, 16384), WG, 0, st>>>(stream_key(seed, STREAM_RMAT),
                                              stream_key(seed, STREAM_WEIGHT), scale, th.ta,
                                              th.tab, th.tabc, m, wmin, wmax, d_src, d_dst, d_w);
    EEG_HIP(hipGetLastError());
    return EEG_OK;
}

}  // namespace eeg
