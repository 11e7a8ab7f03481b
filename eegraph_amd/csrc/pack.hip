// Packed graphs: many independent graphs laid side by side in one HBM-resident CSR (vertex ids
// offset by the vertex counts of the graphs before them, edge rows by their edge counts), so a
// single traversal seeded with one source per graph answers every graph in one launch
// sequence.  The graphs are disjoint, so graph i's labels depend on graph i's source alone and
// equal its single-query answer bit for bit; min-label CC labels come back as union ids and
// are shifted back by the graph's vertex offset.  No counterpart exists in the reference,
// whose drivers run one graph per process (src/main.cu:101-121).
//
// The union is cached for the most recent list of graphs (by handle serial), together with its
// Tigr and symmetrised tables, so repeated calls on the same set pay the packing once.
#include <algorithm>
#include <chrono>
#include <cstring>
#include <mutex>
#include <vector>

#include "engine.hpp"

namespace eeg {
namespace {

unsigned grid_pk(uint64_t work, unsigned cap = 16384) {
    uint64_t g = (work + WG - 1) / WG;
    if (g == 0) g = 1;
    return (unsigned)(g > cap ? cap : g);
}

// out[v] = row[v] + eoff for v <= n (entry n is rewritten with the same value by the next graph)
__global__ void k_pack_rows(const uint64_t* __restrict__ row, uint32_t n, uint64_t eoff,
                            uint64_t* __restrict__ out) {
    for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v <= n;
         v += (uint64_t)gridDim.x * blockDim.x)
        out[v] = row[v] + eoff;
}

__global__ void k_add_u32(const uint32_t* __restrict__ in, uint64_t cnt, uint32_t off,
                          uint32_t* __restrict__ out) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < cnt;
         i += (uint64_t)gridDim.x * blockDim.x)
        out[i] = in[i] + off;
}

struct Packed {
    std::vector<uint64_t> serials;
    std::vector<uint32_t> voff;  // ngraphs + 1 vertex offsets
    eeg_graph* g = nullptr;      // the union
    DevMem seeds;
    ~Packed() { delete g; }
};

std::mutex g_pack_mu;
Packed* g_pack = nullptr;  // most recent union (one at a time bounds the HBM held)

eeg_status build_union(eeg_graph* const* gs, uint32_t k, Packed* pk) {
    uint64_t N = 0, M = 0;
    const bool weighted = gs[0]->weighted;
    pk->voff.assign(k + 1, 0);
    for (uint32_t i = 0; i < k; ++i) {
        if (gs[i]->device != gs[0]->device) return fail(EEG_ERR_INVALID_ARG, "graphs live on different devices");
        if (gs[i]->weighted != weighted)
            return fail(EEG_ERR_INVALID_ARG, "weighted and unweighted graphs cannot be packed together");
        pk->voff[i] = (uint32_t)N;
        N += gs[i]->n;
        M += gs[i]->m;
        if (N > 0xfffffff0ull) return fail(EEG_ERR_UNSUPPORTED, "packed vertex count exceeds 2^32 - 16");
    }
    pk->voff[k] = (uint32_t)N;
    eeg_graph* u = nullptr;
    EEG_TRY(new_graph((uint32_t)N, M, weighted, gs[0]->device, &u));
    pk->g = u;
    hipStream_t s = u->stream;
    Csr& c = u->out;
    c.n = (uint32_t)N;
    c.m = M;
    EEG_TRY(c.row.alloc((N + 1) * 8));
    EEG_TRY(c.col.alloc(M * 4 + 4));
    if (weighted) EEG_TRY(c.w.alloc(M * 4 + 4));
    EEG_TRY(u->perm.alloc(N * 4 + 4));
    EEG_TRY(u->inv.alloc(N * 4 + 4));
    u->h_perm.resize(N);
    // the members' streams may still be writing their tables: wait for them first
    for (uint32_t i = 0; i < k; ++i) EEG_HIP(hipStreamSynchronize(gs[i]->stream));
    uint64_t eoff = 0;
    for (uint32_t i = 0; i < k; ++i) {
        const eeg_graph* gi = gs[i];
        const uint32_t off = pk->voff[i];
        if (gi->n) {
            k_pack_rows<<<grid_pk((uint64_t)gi->n + 1), WG, 0, s>>>(gi->out.rowp(), gi->n, eoff, c.row.as<uint64_t>() + off);
            k_add_u32<<<grid_pk(gi->n), WG, 0, s>>>(gi->perm.as<uint32_t>(), gi->n, off, u->perm.as<uint32_t>() + off);
            k_add_u32<<<grid_pk(gi->n), WG, 0, s>>>(gi->inv.as<uint32_t>(), gi->n, off, u->inv.as<uint32_t>() + off);
            EEG_HIP(hipGetLastError());
            for (uint32_t v = 0; v < gi->n; ++v) u->h_perm[off + v] = gi->h_perm[v] + off;
        }
        if (gi->m) {
            k_add_u32<<<grid_pk(gi->m), WG, 0, s>>>(gi->out.colp(), gi->m, off, c.col.as<uint32_t>() + eoff);
            EEG_HIP(hipGetLastError());
            if (weighted)
                EEG_HIP(hipMemcpyAsync(c.w.as<uint32_t>() + eoff, gi->out.wp(), gi->m * 4, hipMemcpyDeviceToDevice, s));
        }
        eoff += gi->m;
        if (i == 0 || gi->max_deg > u->max_deg) {  // lowest id among the maxima
            u->max_deg = gi->max_deg;
            u->argmax_deg = gi->argmax_deg + off;
        }
        u->nonzero_out += gi->nonzero_out;
    }
    if (N == 0) EEG_HIP(hipMemsetAsync(c.row.p, 0, 8, s));
    c.built = true;
    EEG_TRY(pk->seeds.alloc((size_t)k * 4 + 4));
    EEG_HIP(hipStreamSynchronize(s));
    return EEG_OK;
}

__global__ void k_unoffset(uint32_t* __restrict__ lab, uint64_t cnt, uint32_t off) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < cnt;
         i += (uint64_t)gridDim.x * blockDim.x)
        lab[i] -= off;
}

}  // namespace
}  // namespace eeg

using namespace eeg;

extern "C" eeg_status eeg_run_packed(eeg_graph* const* graphs, uint32_t ngraphs, uint32_t algo,
                                     const uint32_t* sources, const eeg_opts* opts,
                                     void* const* labels_out, eeg_stats* stats) {
    clear_error();
    if (!graphs || ngraphs == 0) return fail(EEG_ERR_INVALID_ARG, "no graphs");
    if (algo > EEG_PPR) return fail(EEG_ERR_INVALID_ARG, "unknown algorithm %u", algo);
    if (algo == EEG_PR || algo == EEG_PPR)
        return fail(EEG_ERR_UNSUPPORTED, "packed runs support BFS, SSSP, SSWP and CC");
    const bool sourced = algo != EEG_CC;
    if (sourced && !sources) return fail(EEG_ERR_INVALID_ARG, "null sources");
    for (uint32_t i = 0; i < ngraphs; ++i) {
        if (!graphs[i]) return fail(EEG_ERR_INVALID_ARG, "graph %u is null", i);
        if (sourced && sources[i] >= graphs[i]->n)
            return fail(EEG_ERR_INVALID_ARG, "source %u of graph %u out of range (n = %u)", sources[i], i,
                        graphs[i]->n);
    }
    eeg_opts o;
    memset(&o, 0, sizeof(o));
    if (opts) o = *opts;
    eeg_stats local;
    eeg_stats* st = stats ? stats : &local;
    memset(st, 0, sizeof(*st));
    const auto t0 = std::chrono::steady_clock::now();

    std::lock_guard<std::mutex> lk(g_pack_mu);
    DeviceGuard dg;
    EEG_TRY(select_device(graphs[0]->device));
    std::vector<uint64_t> key(ngraphs);
    for (uint32_t i = 0; i < ngraphs; ++i) key[i] = graphs[i]->serial;
    if (!g_pack || g_pack->serials != key) {
        delete g_pack;
        g_pack = new (std::nothrow) Packed();
        if (!g_pack) return fail(EEG_ERR_OOM, "host allocation failed");
        eeg_status s = build_union(graphs, ngraphs, g_pack);
        if (s != EEG_OK) {
            delete g_pack;
            g_pack = nullptr;
            return s;
        }
        g_pack->serials = key;
    }
    Packed& pk = *g_pack;
    eeg_graph* u = pk.g;
    std::lock_guard<std::mutex> lu(u->mu);
    hipStream_t s = u->stream;

    double e0 = 0, e1 = 0;
    const bool energy = o.energy && eeg_energy_read(u->device, &e0, nullptr) == EEG_OK;
    eeg_status rc;
    if (sourced) {
        std::vector<uint32_t> seeds(ngraphs);
        for (uint32_t i = 0; i < ngraphs; ++i) seeds[i] = u->h_perm[pk.voff[i] + sources[i]];
        EEG_HIP(hipMemcpyAsync(pk.seeds.p, seeds.data(), (size_t)ngraphs * 4, hipMemcpyHostToDevice, s));
        rc = run_traversal(u, algo, 0, o, st, pk.seeds.as<uint32_t>(), ngraphs);
    } else {
        rc = run_traversal(u, algo, 0, o, st);
    }
    if (energy && eeg_energy_read(u->device, &e1, nullptr) == EEG_OK) st->energy_j = (e1 - e0) * 1e-6;
    if (rc != EEG_OK && rc != EEG_ERR_NOT_CONVERGED) return rc;
    const eeg_status keep = rc;
    if (labels_out && u->n) {
        EEG_TRY(u->work.out.alloc((size_t)u->n * 4 + 4));
        EEG_TRY(unpermute(u, u->work.lab.p, u->work.out.p, s));
        uint32_t* outd = u->work.out.as<uint32_t>();
        for (uint32_t i = 0; i < ngraphs; ++i) {
            const uint32_t ni = pk.voff[i + 1] - pk.voff[i];
            if (!ni || !labels_out[i]) continue;
            if (algo == EEG_CC && pk.voff[i]) {
                k_unoffset<<<grid_pk(ni), WG, 0, s>>>(outd + pk.voff[i], ni, pk.voff[i]);
                EEG_HIP(hipGetLastError());
            }
            EEG_HIP(hipMemcpyAsync(labels_out[i], outd + pk.voff[i], (size_t)ni * 4, hipMemcpyDeviceToHost, s));
        }
        EEG_HIP(hipStreamSynchronize(s));
    }
    st->ms_total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (keep != EEG_OK) return fail(keep, "no fixpoint within max_iters");
    return EEG_OK;
}
