// PageRank / personalised PageRank (Jacobi with compensated sums, residual push).
#include "engine.hpp"

namespace eeg {
eeg_status run_pagerank(eeg_graph*, uint32_t, uint32_t, const eeg_opts&, eeg_stats*) {
    return fail(EEG_ERR_UNSUPPORTED, "PageRank not built yet");
}
}  // namespace eeg
