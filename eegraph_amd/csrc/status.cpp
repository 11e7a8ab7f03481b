// Status strings and the thread-local detail message behind eeg_last_error().
#include <cstdarg>
#include <cstdio>

#include "eeg_internal.hpp"

namespace {
thread_local char g_last_error[512] = "";
}

namespace eeg {

eeg_status fail(eeg_status s, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_last_error, sizeof(g_last_error), fmt, ap);
    va_end(ap);
    return s;
}

void clear_error() { g_last_error[0] = '\0'; }

}  // namespace eeg

extern "C" const char* eeg_last_error(void) { return g_last_error; }

extern "C" const char* eeg_strerror(eeg_status s) {
    switch (s) {
        case EEG_OK: return "ok";
        case EEG_ERR_INVALID_ARG: return "invalid argument";
        case EEG_ERR_OOM: return "out of memory";
        case EEG_ERR_DEVICE: return "device error";
        case EEG_ERR_NOT_CONVERGED: return "not converged";
        case EEG_ERR_REJECTED: return "rejected by admission control";
        case EEG_ERR_NO_DEVICE: return "no GPU device";
        case EEG_ERR_IO: return "i/o error";
        case EEG_ERR_UNSUPPORTED: return "unsupported";
        case EEG_ERR_CANCELLED: return "cancelled";
        case EEG_ERR_TIMEOUT: return "timed out";
    }
    return "unknown status";
}
