// Host-side view of a device-resident graph and the launch helpers shared by the HIP
// translation units.  Everything allocated here lives in HBM for the life of the handle.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <memory>
#include <mutex>
#include <vector>

#include "device.hpp"
#include "eeg_internal.hpp"

#define EEG_HIP(call)                                                                     \
    do {                                                                                  \
        hipError_t e_ = (call);                                                           \
        if (e_ != hipSuccess)                                                             \
            return ::eeg::fail(e_ == hipErrorOutOfMemory ? EEG_ERR_OOM : EEG_ERR_DEVICE, \
                               "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_),     \
                               __FILE__, __LINE__);                                       \
    } while (0)

namespace eeg {

// Owning device allocation.
struct DevMem {
    void* p = nullptr;
    size_t bytes = 0;
    DevMem() = default;
    DevMem(const DevMem&) = delete;
    DevMem& operator=(const DevMem&) = delete;
    ~DevMem() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    eeg_status alloc(size_t n) {
        if (p && bytes >= n && n > 0) return EEG_OK;
        release();
        if (n == 0) n = 8;
        hipError_t e = hipMalloc(&p, n);
        if (e != hipSuccess) {
            p = nullptr;
            (void)hipGetLastError();
            return fail(EEG_ERR_OOM, "hipMalloc(%zu) failed: %s", n, hipGetErrorString(e));
        }
        bytes = n;
        return EEG_OK;
    }
    template <class T>
    T* as() const { return static_cast<T*>(p); }
};

// Compressed sparse rows of one edge direction.
struct Csr {
    uint32_t n = 0;
    uint64_t m = 0;
    DevMem row;  // u64[n+1]
    DevMem col;  // u32[m]
    DevMem w;    // u32[m] (empty if unweighted)
    bool built = false;
    const uint64_t* rowp() const { return row.as<const uint64_t>(); }
    const uint32_t* colp() const { return col.as<const uint32_t>(); }
    const uint32_t* wp() const { return w.p ? w.as<const uint32_t>() : nullptr; }
};

// Tigr virtual-node tables for one K, run-length encoded: the vertices with at least one
// edge (vlist, ascending) and the exclusive prefix of ceil(deg/K) over them (vscan, nz+1
// entries).  Virtual node t belongs to vlist[i] with vscan[i] <= t < vscan[i+1] and covers
// that vertex's edges [row + (t - vscan[i]) * K, +K).
struct Tigr {
    uint32_t K = 0;
    uint32_t nz = 0;
    uint64_t P = 0;
    DevMem vlist;  // u32[nz]
    DevMem vscan;  // u32[nz+1]
    DevMem vnc;    // u32[n]: ceil(deg/K) per vertex (virtual nodes a frontier entry brings)
    bool built = false;
    DevMem tstart;      // first entry (vlist position) of every traversal tile of the static sweep
    DevMem tblk;        // per tile: first and last 64-vertex block of its entries (lo | hi << 32)
    bool tiles_built = false;
    DevMem span;        // in-edge tables only: positions of vertices crossing a PageRank tile:
                        // nspan crossing few tiles from the front, nspan_long crossing many
                        // from the back (span_cap entries)
    DevMem prstart;     // in-edge tables only: first entry (vlist position) of each PageRank tile
    uint32_t nspan = 0, nspan_long = 0, span_cap = 0;
    bool span_built = false;
};

// Segments of at most 64 in-edges for the many-restart PageRank planes (ppr_plane.hip).
// Entries are the vertices with in- or out-edges (ascending); entry i owns segments
// [sscan[i], sscan[i+1]), max(1, ceil(indeg/64)) of them; sdesc describes every segment (first
// edge, edge count, vertex, spill slot); entries with more than one segment are listed in mlist,
// and their segments' partial sums go to spill slots from mbase[i].
struct PprSeg {
    uint32_t nent = 0, nseg = 0, nmulti = 0, nspill = 0;
    DevMem ent, sscan, sdesc, mlist, mbase;
    // hubs (more than PPR_HUB_SEGS segments): {first spill slot, segments, first chunk, 0} each,
    // ascending by entry; nhub = 0 when there are none or too many to list
    DevMem hubs;
    uint32_t nhub = 0, nhubch = 0;
    bool built = false;
};

// Reusable per-graph traversal state (one query at a time per handle).
struct Work {
    DevMem lab;         // u32[n] (or f32 ranks), internal vertex order
    DevMem out;         // u32[n]: result in original vertex order before a host copy
    DevMem q[2], s[2];  // frontier queues: entry vertex ids and vnode prefix
    DevMem stamp;       // u32[n] iteration stamps for queue dedup
    DevMem flag;        // u32[n] iteration stamps for flag-scan frontiers
    DevMem bitmap;      // u32[(n+31)/32]: vertices improved this iteration (sorted frontiers)
    DevMem vnpush;      // u32[P] last label pushed by each virtual node (async dd)
    DevMem vdirty;      // u32 per 64-vertex block: stamp of the last improvement in the block
    DevMem ctl;         // Ctl
    DevMem fx[2], fl[2], fc[2];  // PR: ranks as (hi, lo) pairs and contributions x/outdeg
    DevMem spill;       // PR: per-tile partial sums
    DevMem red;         // PR: per-workgroup residual partials
    uint32_t epoch = 1; // iteration stamps already used on stamp/flag
    DevMem devit;       // replayed iteration chunks: {iteration, stamp base}
    // executable graphs of replayed iteration chunks, by the argument bytes they were built
    // with (most recently used last, at most kGraphCache): queries that alternate algorithms
    // on one handle reuse their graphs
    struct GraphEntry {
        std::vector<uint8_t> key;
        hipGraphExec_t exec;
    };
    static constexpr size_t kGraphCache = 8;
    std::vector<GraphEntry> graphs;
    Ctl* host_ctl = nullptr;  // pinned mirrors of Ctl (two, for overlapped polling)
    std::vector<hipEvent_t> events;  // pool: one or two per launch
    eeg_status event(size_t i, hipEvent_t* out);
    // Completion markers of the two in-flight control-block copies.  Created once and kept
    // for the life of the work state: a run ends with records still pending on its stream,
    // and an event must not be destroyed while a record of it is pending.
    hipEvent_t poll[2] = {nullptr, nullptr};
    eeg_status poll_events();
    ~Work();
};

}  // namespace eeg

struct eeg_graph {
    int device = 0;
    hipStream_t stream = nullptr;
    bool weighted = false;
    uint32_t n = 0;
    uint64_t m = 0;
    eeg::Csr out, sym, in;
    eeg::Tigr tout[4], tsym[4], tin[4];  // index log2(K) - 2
    eeg::Work work;
    eeg::DevMem perm, inv;           // degree order: perm[original] = internal, inv = perm^-1
    eeg::DevMem outdeg;              // u32 out-degrees (built with the PageRank tables)
    eeg::PprSeg pprseg;              // in-edge segments of the many-restart PageRank planes
    std::vector<uint32_t> h_perm;    // host copy of perm (sources map without a device read)
    uint32_t max_deg = 0, argmax_deg = 0, nonzero_out = 0;
    uint64_t serial = 0;             // unique per handle in this process
    std::shared_ptr<void> batch_state;  // batched-traversal buffers (batch.hip), freed with the handle
    std::mutex mu;
    ~eeg_graph();
};

namespace eeg {

inline int k_index(uint32_t K) {
    switch (K) {
        case 4: return 0;
        case 8: return 1;
        case 16: return 2;
        case 32: return 3;
    }
    return -1;
}

// Handle creation (capi.hip): checks the device and makes it current, creates the stream.
eeg_status select_device(int device);
eeg_status new_graph(uint32_t n, uint64_t m, bool weighted, int device, eeg_graph** out);

struct DeviceGuard {  // restores the caller's current device
    int prev = -1;
    DeviceGuard() { (void)hipGetDevice(&prev); }
    ~DeviceGuard() {
        if (prev >= 0) (void)hipSetDevice(prev);
    }
};

// Graph construction (build.hip).
eeg_status build_csr_from_device_edges(eeg_graph* g, const uint32_t* d_src, const uint32_t* d_dst,
                                       const uint32_t* d_w, uint64_t m);
eeg_status ensure_tigr(eeg_graph* g, Csr& c, Tigr* table, uint32_t K);
eeg_status ensure_sym(eeg_graph* g);
eeg_status ensure_in(eeg_graph* g);
eeg_status ensure_pr_span(eeg_graph* g, uint32_t K);  // pagerank.hip, after ensure_tigr(in)
eeg_status export_tigr(eeg_graph* g, uint32_t K, uint64_t* node_pointer, uint32_t* edge_list,
                       uint32_t* parts, uint64_t* num_parts);
eeg_status orig_csr(eeg_graph* g, const Csr& src, Csr& out);
eeg_status unpermute(eeg_graph* g, const void* d_in, void* d_out, hipStream_t st);
eeg_status generate_rmat_device(uint32_t scale, uint32_t ef, double a, double b, double c,
                                uint64_t seed, uint32_t wmin, uint32_t wmax, uint32_t* d_src,
                                uint32_t* d_dst, uint32_t* d_w, hipStream_t stream);

// Traversals (relax.hip, pagerank.hip).
// `source` is an internal id; d_seeds (device, nseeds distinct internal ids) replaces it when
// given.  The query runs with the graph's own state and stream, or with `wk` on `sk` (several
// queries at once on one graph; the tables they read must have been built beforehand).
eeg_status run_traversal(eeg_graph* g, uint32_t algo, uint32_t source, const eeg_opts& o,
                         eeg_stats* st, const uint32_t* d_seeds = nullptr, uint32_t nseeds = 0,
                         Work* wk = nullptr, hipStream_t sk = nullptr);
eeg_status run_pagerank(eeg_graph* g, uint32_t algo, uint32_t source, const eeg_opts& o,
                        eeg_stats* st, Work* wk = nullptr, hipStream_t sk = nullptr);

// Device buffers of a traversal that allocates nothing (n + 1 words each, Ctl aligned).
struct FixedState {
    uint32_t* lab;
    uint32_t* q[2];
    uint32_t* s[2];
    uint32_t* stamp;
    Ctl* ctl;
};
// BFS/SSSP/SSWP/CC in sync_push_dd (compact queues) or the topology-driven sweep, on `fs`,
// on stream `s`; the graph's tables must exist (eeg_graph_prepare).  Labels stay in fs.lab
// in internal vertex order.
eeg_status run_traversal_fixed(eeg_graph* g, uint32_t algo, uint32_t source, const eeg_opts& o,
                               const FixedState& fs, hipStream_t s, eeg_stats* st);

}  // namespace eeg
