// Batched traversal kernels shared by the one-shot batched entry point (batch.hip) and the
// serving engines (serve.hip): kernel argument blocks and host-side launch wrappers.
#pragma once

#include <memory>

#include "engine.hpp"

namespace eeg {

constexpr uint32_t BATCH_MAXW = 4;  // up to 256 queries per traversal (4 x 64-bit words)

// State of multi-source passes (label planes, masks, queues, the replayed graph of a chunk);
// one per graph handle for eeg_run_batch, one per serving lane that batches its queries.
struct BatchWork;
std::shared_ptr<BatchWork> make_batch_work();
// BFS/SSSP/SSWP from up to 256 sources in one pass (narrow labels as eeg_run_batch chooses
// them) on the state bw, stream s and, for the single-query reruns of 8-bit distances, the
// traversal state wk (also on s).  Row q of the answers (n words, original vertex order) goes
// to the host memory rows[q]; null rows are not copied out.  The graph's tables for K must
// exist.
eeg_status run_batch_rows(eeg_graph* g, BatchWork& bw, Work& wk, hipStream_t s, uint32_t algo,
                          const uint32_t* sources, uint32_t nsrc, const eeg_opts& o, uint32_t* const* rows,
                          eeg_stats* st);

// Multi-source BFS with 64-bit visit masks: a push launch (odd launch index) ORs new bits into
// `next`, a finalize launch (even) turns next & ~seen into the next frontier and levels.
struct MsArgs {
    const uint64_t* row;
    const uint32_t* col;
    const uint32_t* vnc;
    const uint32_t* qin;   // frontier vertices
    const uint32_t* sin;   // their virtual-node prefix
    // masks of vertex u: seen[u*R + w], visit[u*R + w], next[u*R + w] for w < W; stamp of u:
    // stamp[u*sst].  Separate arrays have R = W, sst = 1; the one-shot batched pass interleaves
    // them into one record per vertex (R = 4W: seen, next, visit, stamp), so an edge's mask
    // read, atomic and stamp exchange touch one cache line of the destination
    uint64_t* seen;
    uint64_t* visit;       // bits a frontier vertex pushes this level
    uint64_t* next;        // bits received this level
    uint32_t* stamp;
    uint32_t* cand;        // vertices that received bits (finalize list)
    uint32_t* qout;        // next frontier
    uint32_t* sout;
    uint32_t* dist;        // n x B levels, vertex-major
    Ctl* ctl;
    uint64_t* act;         // optional: 4 slots x W words; slot level%4 gets the columns that
                           // still have frontier work after this level
    uint32_t W, B, n;
    uint32_t R, sst;       // mask record stride (64-bit words), stamp stride (32-bit words)
    uint32_t qcap;         // capacity of qout / cand
    uint32_t launch;       // launch index (slot rotation)
    uint32_t level;        // level being discovered by this push (1-based)
    uint32_t stampv;
    // replayed graph: level = dev_it[0], stamp base = dev_it[1]; launch, stampv and the queue
    // ping-pong over (qa, sa) / (qb, sb) are derived from them
    const uint32_t* dev_it;
    uint32_t *qa, *sa, *qb, *sb;
};

// Batched SSSP / SSWP over vertex-major label planes.
struct PlaneArgs {
    const uint64_t* row;
    const uint32_t* col;
    const uint32_t* wgt;
    const uint32_t* vnc;
    const uint32_t* qin;
    const uint32_t* sin;
    uint32_t* lab;        // n x B, vertex-major
    uint64_t* chg;        // n x W: queries whose label improved and were not pushed yet
    uint64_t* nchg;       // n x W: improvements of this iteration
    uint32_t* stamp;
    uint32_t* qout;
    uint32_t* sout;
    Ctl* ctl;
    uint64_t* act;        // optional: 4 slots x W words; slot it%4 gets the columns with
                          // pending changes after this iteration
    uint32_t W, B, n;
    uint32_t qcap;
    uint32_t it;
    uint32_t stampv;
    int widest;           // 0: SSSP (min, +w), 1: SSWP (max, min(.,w))
    const uint32_t* dev_it;  // replayed graph, as in MsArgs
    uint32_t *qa, *sa, *qb, *sb;
    uint32_t* bitmap;     // dense passes of eeg_run_batch: improved vertices set their bit and
                          // k_compact_frontier builds the next queue in vertex order (null: the
                          // push appends to the queue itself)
    uint64_t* ovq;        // 8-bit distance passes: W words, queries to answer again wider
};

// One BFS level: push then finalize (launch indices 2*level-1 and 2*level).
eeg_status launch_ms_level(MsArgs a, uint32_t K, unsigned grid, hipStream_t s);
// One plane iteration: push, then the two change-mask rotations.  `in_flight` (queries being
// traversed) picks the push variant: one wavefront per virtual node with lane = query for
// many, a two-pass kernel with K lanes per virtual node for few.
// `lm` is the label width of the dense push (0: 32-bit labels, lane = query; narrow widths
// only with the planes of eeg_run_batch).
eeg_status launch_plane_iteration(PlaneArgs a, uint32_t K, unsigned grid, hipStream_t s, uint32_t in_flight,
                                  int lm = 0);

// Admission between iterations (one workgroup): pairs (pv[i], pc[i]) = (internal source,
// column) get label `val`; their column bit is ORed into m_always, and into m_edges when the
// source has out-edges (mask word w of vertex v at index v*R + w).  The distinct sources with out-edges (uniq) are appended to the
// frontier the next push reads (Ctl slot in_slot, queue q/s of capacity qcap).  A drained
// engine (Ctl done) restarts from empty slots.
eeg_status launch_admit(Ctl* ctl, int in_slot, const uint32_t* pv, const uint32_t* pc, uint32_t npairs,
                        const uint32_t* uniq, uint32_t nuniq, const uint32_t* vnc, uint32_t R, uint32_t B,
                        uint64_t* m_always, uint64_t* m_edges, uint32_t* lab, uint32_t val, uint32_t* q,
                        uint32_t* s, uint32_t qcap, hipStream_t st);
// Frees the columns whose bits are set in clr[0..W): label column back to `fill`, bits cleared
// in each non-null mask array (word w of vertex v at index v*R + w).
eeg_status launch_release(uint32_t n, uint32_t W, uint32_t R, uint32_t B, const uint64_t clr[BATCH_MAXW],
                          uint32_t* lab, uint32_t fill, uint64_t* m0, uint64_t* m1, uint64_t* m2, hipStream_t st);
// out[u] = lab[perm[u] * B + c] (minus `offset` for finite values), original vertex order.
eeg_status launch_extract(const uint32_t* lab, uint32_t B, uint32_t c, const uint32_t* perm, uint32_t n,
                          uint32_t offset, uint32_t* out, hipStream_t st);

int batch_cu_count(int device);

}  // namespace eeg
