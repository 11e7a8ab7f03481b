// C-ABI entry points for device graphs and single queries.  The reference allocates, copies
// and frees every device array inside each eegraph_<alg> call (src/eegraph.cu:61-74,
// 262-268); here the graph handle owns its HBM tables across queries.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstring>
#include <string>

#include "engine.hpp"
#include "ppr_plane.hpp"

namespace eeg {

eeg_status Work::event(size_t i, hipEvent_t* out) {
    while (events.size() <= i) {
        hipEvent_t e;
        EEG_HIP(hipEventCreate(&e));
        events.push_back(e);
    }
    *out = events[i];
    return EEG_OK;
}

eeg_status Work::poll_events() {
    for (int i = 0; i < 2; ++i)
        if (!poll[i]) EEG_HIP(hipEventCreateWithFlags(&poll[i], hipEventDisableTiming));
    return EEG_OK;
}

// Runs end with a stream synchronisation, so no record of these events is pending here.
Work::~Work() {
    for (auto& ge : graphs) (void)hipGraphExecDestroy(ge.exec);
    for (auto e : events) (void)hipEventDestroy(e);
    for (auto e : poll)
        if (e) (void)hipEventDestroy(e);
    if (host_ctl) (void)hipHostFree(host_ctl);
}

eeg_status select_device(int device) {
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count == 0) {
        (void)hipGetLastError();
        return fail(EEG_ERR_NO_DEVICE, "no GPU available (%s)", hipGetErrorString(e));
    }
    if (device < 0 || device >= count)
        return fail(EEG_ERR_INVALID_ARG, "device %d out of range (%d devices)", device, count);
    EEG_HIP(hipSetDevice(device));
    return EEG_OK;
}

eeg_status new_graph(uint32_t n, uint64_t m, bool weighted, int device, eeg_graph** out) {
    if (m >= 0xffffffffull)
        return fail(EEG_ERR_UNSUPPORTED, "graphs are limited to 2^32-1 edges (got %llu)",
                    (unsigned long long)m);
    EEG_TRY(select_device(device));
    auto* g = new (std::nothrow) eeg_graph();
    if (!g) return fail(EEG_ERR_OOM, "host allocation failed");
    g->device = device;
    g->n = n;
    g->m = m;
    g->weighted = weighted;
    hipError_t e = hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete g;
        return fail(EEG_ERR_DEVICE, "hipStreamCreate failed: %s", hipGetErrorString(e));
    }
    static std::atomic<uint64_t> next_serial{1};
    g->serial = next_serial.fetch_add(1);
    *out = g;
    return EEG_OK;
}

}  // namespace eeg

eeg_graph::~eeg_graph() {
    if (stream) {
        (void)hipStreamSynchronize(stream);
        (void)hipStreamDestroy(stream);
    }
}

using namespace eeg;

extern "C" eeg_status eeg_device_count(int* count) {
    clear_error();
    if (!count) return fail(EEG_ERR_INVALID_ARG, "null count");
    int c = 0;
    hipError_t e = hipGetDeviceCount(&c);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        c = 0;
    }
    *count = c;
    return EEG_OK;
}

extern "C" eeg_status eeg_graph_create_device(const uint32_t* d_src, const uint32_t* d_dst,
                                              const uint32_t* d_w, uint64_t m, uint32_t n,
                                              int device, eeg_graph** out) {
    clear_error();
    if (!out || (m && (!d_src || !d_dst))) return fail(EEG_ERR_INVALID_ARG, "null argument");
    *out = nullptr;
    DeviceGuard dg;
    eeg_graph* g = nullptr;
    EEG_TRY(new_graph(n, m, d_w != nullptr, device, &g));
    eeg_status s = build_csr_from_device_edges(g, d_src, d_dst, d_w, m);
    if (s != EEG_OK) {
        delete g;
        return s;
    }
    *out = g;
    return EEG_OK;
}

extern "C" eeg_status eeg_graph_create(const uint32_t* src, const uint32_t* dst,
                                       const uint32_t* w, uint64_t m, uint32_t n, int device,
                                       eeg_graph** out) {
    clear_error();
    if (!out || (m && (!src || !dst))) return fail(EEG_ERR_INVALID_ARG, "null argument");
    *out = nullptr;
    DeviceGuard dg;
    eeg_graph* g = nullptr;
    EEG_TRY(new_graph(n, m, w != nullptr, device, &g));
    eeg_status s;
    {
        DevMem ds, dd, dw;
        s = ds.alloc(m * 4 + 4);
        if (s == EEG_OK) s = dd.alloc(m * 4 + 4);
        if (s == EEG_OK && w) s = dw.alloc(m * 4 + 4);
        if (s == EEG_OK && m) {
            hipError_t e = hipMemcpyAsync(ds.p, src, m * 4, hipMemcpyHostToDevice, g->stream);
            if (e == hipSuccess) e = hipMemcpyAsync(dd.p, dst, m * 4, hipMemcpyHostToDevice, g->stream);
            if (e == hipSuccess && w) e = hipMemcpyAsync(dw.p, w, m * 4, hipMemcpyHostToDevice, g->stream);
            if (e == hipSuccess) e = hipStreamSynchronize(g->stream);
            if (e != hipSuccess) s = fail(EEG_ERR_DEVICE, "edge upload failed: %s", hipGetErrorString(e));
        }
        if (s == EEG_OK)
            s = build_csr_from_device_edges(g, ds.as<uint32_t>(), dd.as<uint32_t>(),
                                            w ? dw.as<uint32_t>() : nullptr, m);
    }
    if (s != EEG_OK) {
        delete g;
        return s;
    }
    *out = g;
    return EEG_OK;
}

extern "C" eeg_status eeg_graph_create_rmat(uint32_t scale, uint32_t edge_factor, double a,
                                            double b, double c, uint64_t seed, uint32_t wmin,
                                            uint32_t wmax, int device, eeg_graph** out) {
    clear_error();
    if (!out || scale == 0 || scale > 31 || edge_factor == 0 || wmin > wmax)
        return fail(EEG_ERR_INVALID_ARG, "bad RMAT parameters");
    if (a < 0 || b < 0 || c < 0 || a + b + c > 1.0 + 1e-12)
        return fail(EEG_ERR_INVALID_ARG, "rmat: probabilities must be >= 0 with a+b+c <= 1");
    *out = nullptr;
    DeviceGuard dg;
    const uint64_t m = (uint64_t)edge_factor << scale;
    eeg_graph* g = nullptr;
    EEG_TRY(new_graph((uint32_t)(1ull << scale), m, true, device, &g));
    eeg_status s;
    {
        DevMem ds, dd, dw;
        s = ds.alloc(m * 4);
        if (s == EEG_OK) s = dd.alloc(m * 4);
        if (s == EEG_OK) s = dw.alloc(m * 4);
        if (s == EEG_OK)
            s = generate_rmat_device(scale, edge_factor, a, b, c, seed, wmin, wmax,
                                     ds.as<uint32_t>(), dd.as<uint32_t>(), dw.as<uint32_t>(),
                                     g->stream);
        if (s == EEG_OK)
            s = build_csr_from_device_edges(g, ds.as<uint32_t>(), dd.as<uint32_t>(),
                                            dw.as<uint32_t>(), m);
    }
    if (s != EEG_OK) {
        delete g;
        return s;
    }
    *out = g;
    return EEG_OK;
}

extern "C" eeg_status eeg_graph_destroy(eeg_graph* g) {
    clear_error();
    if (!g) return EEG_OK;
    DeviceGuard dg;
    (void)hipSetDevice(g->device);
    delete g;
    return EEG_OK;
}

extern "C" eeg_status eeg_graph_get_info(const eeg_graph* g, eeg_graph_info* info) {
    clear_error();
    if (!g || !info) return fail(EEG_ERR_INVALID_ARG, "null argument");
    memset(info, 0, sizeof(*info));
    info->n = g->n;
    info->m = g->m;
    info->weighted = g->weighted ? 1 : 0;
    info->max_out_degree = g->max_deg;
    info->argmax_out_degree = g->argmax_deg;
    info->nonzero_out = g->nonzero_out;
    info->device = (uint32_t)g->device;
    uint64_t b = 0;
    for (const Csr* c : {&g->out, &g->sym, &g->in}) b += c->row.bytes + c->col.bytes + c->w.bytes;
    b += g->perm.bytes + g->inv.bytes;
    for (int k = 0; k < 4; ++k)
        for (const Tigr* t : {&g->tout[k], &g->tsym[k], &g->tin[k]}) b += t->vlist.bytes + t->vscan.bytes;
    info->device_bytes = b;
    return EEG_OK;
}

extern "C" eeg_status eeg_graph_prepare(eeg_graph* g, uint32_t K, uint32_t what) {
    clear_error();
    if (!g) return fail(EEG_ERR_INVALID_ARG, "null graph");
    std::lock_guard<std::mutex> lk(g->mu);
    DeviceGuard dg;
    EEG_HIP(hipSetDevice(g->device));
    if (what & EEG_PREP_SYM) {
        EEG_TRY(ensure_sym(g));
        if (what & EEG_PREP_TIGR) EEG_TRY(ensure_tigr(g, g->sym, g->tsym, K));
    }
    if (what & EEG_PREP_TRANSPOSE) {
        EEG_TRY(ensure_in(g));
        if (what & EEG_PREP_TIGR) EEG_TRY(ensure_tigr(g, g->in, g->tin, K));
    }
    if (what & EEG_PREP_TIGR) EEG_TRY(ensure_tigr(g, g->out, g->tout, K));
    return EEG_OK;
}

extern "C" eeg_status eeg_tigr_export(eeg_graph* g, uint32_t K, uint64_t* node_pointer,
                                      uint32_t* edge_list, uint32_t* parts, uint64_t* num_parts) {
    clear_error();
    if (!g) return fail(EEG_ERR_INVALID_ARG, "null graph");
    std::lock_guard<std::mutex> lk(g->mu);
    DeviceGuard dg;
    EEG_HIP(hipSetDevice(g->device));
    return export_tigr(g, K, node_pointer, edge_list, parts, num_parts);
}

extern "C" eeg_status eeg_csr_export(eeg_graph* g, uint64_t* row, uint32_t* col, uint32_t* w) {
    clear_error();
    if (!g) return fail(EEG_ERR_INVALID_ARG, "null graph");
    std::lock_guard<std::mutex> lk(g->mu);
    DeviceGuard dg;
    EEG_HIP(hipSetDevice(g->device));
    Csr oc;  // HBM holds the degree-ordered graph; export in original ids and order
    EEG_TRY(orig_csr(g, g->out, oc));
    if (row) EEG_HIP(hipMemcpy(row, oc.row.p, ((size_t)g->n + 1) * 8, hipMemcpyDeviceToHost));
    if (col && g->m) EEG_HIP(hipMemcpy(col, oc.col.p, g->m * 4, hipMemcpyDeviceToHost));
    if (w && g->m && oc.w.p) EEG_HIP(hipMemcpy(w, oc.w.p, g->m * 4, hipMemcpyDeviceToHost));
    return EEG_OK;
}

extern "C" eeg_status eeg_csr_export_derived(eeg_graph* g, uint32_t which, uint64_t* row,
                                             uint32_t* col, uint64_t* m_out) {
    clear_error();
    if (!g || (which != EEG_PREP_SYM && which != EEG_PREP_TRANSPOSE))
        return fail(EEG_ERR_INVALID_ARG, "which must be EEG_PREP_SYM or EEG_PREP_TRANSPOSE");
    std::lock_guard<std::mutex> lk(g->mu);
    DeviceGuard dg;
    EEG_HIP(hipSetDevice(g->device));
    Csr* c;
    if (which == EEG_PREP_SYM) {
        EEG_TRY(ensure_sym(g));
        c = &g->sym;
    } else {
        EEG_TRY(ensure_in(g));
        c = &g->in;
    }
    if (m_out) *m_out = c->m;
    if (!row && !col) return EEG_OK;
    Csr oc;
    EEG_TRY(orig_csr(g, *c, oc));
    if (row) EEG_HIP(hipMemcpy(row, oc.row.p, ((size_t)g->n + 1) * 8, hipMemcpyDeviceToHost));
    if (col && c->m) EEG_HIP(hipMemcpy(col, oc.col.p, c->m * 4, hipMemcpyDeviceToHost));
    return EEG_OK;
}

static eeg_status run_common(eeg_graph* g, uint32_t algo, uint32_t source, const eeg_opts* opts,
                             void* out, bool out_is_device, hipStream_t ustream, eeg_stats* stats) {
    clear_error();
    if (!g) return fail(EEG_ERR_INVALID_ARG, "null graph");
    if (algo > EEG_PPR) return fail(EEG_ERR_INVALID_ARG, "unknown algorithm %u", algo);
    const bool sourced = (algo == EEG_BFS || algo == EEG_SSSP || algo == EEG_SSWP || algo == EEG_PPR);
    if (sourced && source >= g->n)
        return fail(EEG_ERR_INVALID_ARG, "source %u out of range (n = %u)", source, g->n);
    eeg_opts o;
    memset(&o, 0, sizeof(o));
    if (opts) o = *opts;
    eeg_stats local;
    eeg_stats* st = stats ? stats : &local;
    uint32_t* act = st->activity;
    const uint32_t actcap = st->activity_cap;
    memset(st, 0, sizeof(*st));
    st->activity = act;
    st->activity_cap = actcap;
    std::lock_guard<std::mutex> lk(g->mu);
    DeviceGuard dg;
    EEG_HIP(hipSetDevice(g->device));
    const auto t0 = std::chrono::steady_clock::now();
    double e0 = 0, e1 = 0;
    const bool energy = o.energy && eeg_energy_read(g->device, &e0, nullptr) == EEG_OK;
    const uint32_t isrc = (sourced && g->n) ? g->h_perm[source] : 0;  // internal id
    eeg_status s = (algo == EEG_PR || algo == EEG_PPR) ? run_pagerank(g, algo, isrc, o, st)
                                                        : run_traversal(g, algo, isrc, o, st);
    // the run's detail message outlives the energy read, which may set its own
    std::string detail = (s != EEG_OK) ? eeg_last_error() : "";
    if (energy && eeg_energy_read(g->device, &e1, nullptr) == EEG_OK) st->energy_j = (e1 - e0) * 1e-6;
    if (s != EEG_OK) fail(s, "%s", detail.c_str());
    else clear_error();
    if (s != EEG_OK && s != EEG_ERR_NOT_CONVERGED) return s;
    if (out && g->n) {  // back to original vertex order
        const size_t bytes = (size_t)g->n * 4;
        if (out_is_device) {
            hipStream_t us = ustream ? ustream : g->stream;
            if (us != g->stream) EEG_HIP(hipStreamSynchronize(g->stream));
            EEG_TRY(unpermute(g, g->work.lab.p, out, us));
            EEG_HIP(hipStreamSynchronize(us));
        } else {
            EEG_TRY(g->work.out.alloc(bytes + 4));
            EEG_TRY(unpermute(g, g->work.lab.p, g->work.out.p, g->stream));
            EEG_HIP(hipMemcpyAsync(out, g->work.out.p, bytes, hipMemcpyDeviceToHost, g->stream));
            EEG_HIP(hipStreamSynchronize(g->stream));
        }
    }
    st->ms_total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return s;
}

extern "C" eeg_status eeg_run(eeg_graph* g, uint32_t algo, uint32_t source, const eeg_opts* opts,
                              void* labels_out, eeg_stats* stats) {
    return run_common(g, algo, source, opts, labels_out, false, nullptr, stats);
}

extern "C" eeg_status eeg_run_device(eeg_graph* g, uint32_t algo, uint32_t source,
                                     const eeg_opts* opts, void* d_labels_out, void* stream,
                                     eeg_stats* stats) {
    return run_common(g, algo, source, opts, d_labels_out, true, (hipStream_t)stream, stats);
}

// ---------------------------------------------------------------- batched queries
namespace eeg {
eeg_status run_batch(eeg_graph* g, uint32_t algo, const uint32_t* sources, uint32_t nsrc, const eeg_opts& o,
                     void* out, bool out_device, eeg_stats* st);
}

extern "C" eeg_status eeg_run_batch(eeg_graph* g, uint32_t algo, const uint32_t* sources, uint32_t nsrc,
                                    const eeg_opts* opts, void* labels_out, eeg_stats* stats) {
    clear_error();
    if (!g || !sources || nsrc == 0) return fail(EEG_ERR_INVALID_ARG, "null graph/sources or nsrc = 0");
    eeg_opts o;
    memset(&o, 0, sizeof(o));
    if (opts) o = *opts;
    eeg_stats local;
    eeg_stats* st = stats ? stats : &local;
    memset(st, 0, sizeof(*st));
    std::lock_guard<std::mutex> lk(g->mu);
    DeviceGuard dg;
    EEG_HIP(hipSetDevice(g->device));
    const auto t0 = std::chrono::steady_clock::now();
    double e0 = 0, e1 = 0;
    const bool energy = o.energy && eeg_energy_read(g->device, &e0, nullptr) == EEG_OK;
    if (algo == EEG_PPR) {
        // planes of up to 32 restart vertices: one gather pass per iteration for all of them
        for (uint32_t i = 0; i < nsrc; ++i)
            if (sources[i] >= g->n) return fail(EEG_ERR_INVALID_ARG, "source %u out of range", sources[i]);
        EEG_TRY(run_ppr_batch(g, sources, nsrc, o, (float*)labels_out, st));
        if (energy && eeg_energy_read(g->device, &e1, nullptr) == EEG_OK) st->energy_j = (e1 - e0) * 1e-6;
        st->ms_total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        return EEG_OK;
    }
    // passes of at most 256 sources (4 mask words per vertex)
    const uint32_t pass = 256;
    for (uint32_t b = 0; b < nsrc; b += pass) {
        const uint32_t cnt = std::min(pass, nsrc - b);
        eeg_stats ps;
        memset(&ps, 0, sizeof(ps));
        void* dst = labels_out ? (void*)((uint32_t*)labels_out + (uint64_t)b * g->n) : nullptr;
        EEG_TRY(run_batch(g, algo, sources + b, cnt, o, dst, false, &ps));
        st->iterations = std::max(st->iterations, ps.iterations);
        st->launches += ps.launches;
        st->ms_traversal += ps.ms_traversal;
        st->label_bits = std::max(st->label_bits, ps.label_bits);
        st->narrow_reruns += ps.narrow_reruns;
    }
    if (energy && eeg_energy_read(g->device, &e1, nullptr) == EEG_OK) st->energy_j = (e1 - e0) * 1e-6;
    st->ms_total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return EEG_OK;
}
