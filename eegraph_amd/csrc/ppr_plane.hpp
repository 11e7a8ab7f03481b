// Personalised PageRank for many restart vertices at once: B columns of ranks per vertex
// (vertex-major planes), one gather pass over the in-edge CSR per iteration for all of them.
// Used by eeg_run_batch (PPR) and the serving loop's PPR engine.
#pragma once

#include "engine.hpp"

namespace eeg {

constexpr uint32_t PPR_SEG = 64;      // in-edges per segment of the gather
constexpr uint32_t PPR_HUB_SEGS = 64;  // entries with more segments are hubs: their partial sums
                                      // are added in chunks of PPR_HUB_SEGS segments first
constexpr uint32_t PPR_IDLE = 0xffffffffu;

// Columns of one iteration: restart vertex (internal id, PPR_IDLE for an unused column) and the
// columns in their first iteration (x0 = e_src, nothing read from the planes for them).
struct PprCols {
    uint32_t src[32];
    uint32_t fresh;  // bit q: column q starts with this iteration
};

// Segment tables of the in-edge CSR (PprSeg, engine.hpp), built once per graph; call with the
// graph lock held, after ensure_in.
eeg_status ensure_ppr_seg(eeg_graph* g);

// The planes of B columns and the per-iteration launches.  Iteration j (1-based) reads buffer
// (j & 1) ^ 1 and writes buffer j & 1.
struct PprPlane {
    uint32_t B = 0, n = 0;
    float damping = 0.85f;
    DevMem c[2];                // n x B contributions x/outdeg
    DevMem x[2];                // n x B ranks as compensated pairs (hi, lo)
    DevMem spill;               // nspill x B float2
    DevMem res_part;            // (grid_pull + grid_fix) x B doubles
    DevMem resid;               // B doubles: L1 change of each column in the last iteration
    unsigned grid_pull = 0, grid_fix = 0;
    static size_t bytes_for(uint32_t n, uint32_t B, uint32_t nspill);
    eeg_status init(eeg_graph* g, uint32_t B, float damping, hipStream_t s);
    // one Jacobi iteration for every column; resid receives each column's L1 change.  W (8,
    // 16 or 32, at most B; 0 = B) columns per pass: columns W.. must be idle, the pass reads
    // and writes the first W of every row and resid[0..W-1]
    eeg_status iterate(eeg_graph* g, uint32_t j, const PprCols& cols, hipStream_t s, uint32_t W = 0);
    // narrowest pass width that covers every busy column of cols
    uint32_t width_for(const PprCols& cols) const;
    // column q's ranks after iteration j into out (n floats, original vertex order)
    eeg_status extract(eeg_graph* g, uint32_t j, uint32_t q, uint32_t src, float* out, hipStream_t s);
};

// eeg_run_batch for PPR: planes of 8, 16 or 32 columns (the narrowest that holds nsrc, 32 for
// more), columns refilled as queries converge; row q of out_host (n floats, may
// be null) receives the ranks of sources[q].
eeg_status run_ppr_batch(eeg_graph* g, const uint32_t* sources, uint32_t nsrc, const eeg_opts& o, float* out_host,
                         eeg_stats* st);

}  // namespace eeg
