// Entry points whose implementation has not landed yet.
#include "eeg_internal.hpp"

extern "C" eeg_status eeg_server_create(eeg_graph*, const eeg_server_cfg*, eeg_server**) {
    return eeg::fail(EEG_ERR_UNSUPPORTED, "server not built yet");
}
extern "C" eeg_status eeg_server_submit(eeg_server*, uint32_t, uint32_t, uint64_t*) { return EEG_ERR_UNSUPPORTED; }
extern "C" eeg_status eeg_server_wait(eeg_server*, uint64_t, void*, double, eeg_query_result*) { return EEG_ERR_UNSUPPORTED; }
extern "C" eeg_status eeg_server_cancel(eeg_server*, uint64_t) { return EEG_ERR_UNSUPPORTED; }
extern "C" eeg_status eeg_server_destroy(eeg_server*) { return EEG_ERR_UNSUPPORTED; }
