// Internal helpers shared by the host-side translation units of libeegraph.
#pragma once

#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/eeg_capi.h"

// Returns the status of `expr` from the enclosing function unless it is EEG_OK.
#define EEG_TRY(expr)                \
    do {                             \
        eeg_status s_ = (expr);      \
        if (s_ != EEG_OK) return s_; \
    } while (0)

namespace eeg {

// Records a detail message for eeg_last_error() and returns the status unchanged.
eeg_status fail(eeg_status s, const char* fmt, ...);
void clear_error();

// ---- counter-based random numbers -------------------------------------------------------
// SplitMix64 output function (Steele, Lea & Flood, "Fast splittable pseudorandom number
// generators", OOPSLA 2014).  A stream (seed, stream id) is a SplitMix64 sequence whose
// starting state is derived from both; value i of the stream is the (i+1)-th output.
// Every draw depends only on (seed, stream, i), so generation is order- and
// thread-count-independent and the device generator reproduces it exactly.
#if defined(__HIP__)
#define EEG_HD __host__ __device__ __forceinline__
#else
#define EEG_HD inline
#endif

EEG_HD uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}
EEG_HD uint64_t stream_key(uint64_t seed, uint64_t stream) {
    return mix64(seed * 0x9e3779b97f4a7c15ULL + stream * 0xd6e8feb86659fd93ULL + 1ULL);
}
EEG_HD uint64_t stream_draw(uint64_t key, uint64_t i) {
    return mix64(key + (i + 1ULL) * 0x9e3779b97f4a7c15ULL);
}
// Maps the high 32 bits of a draw onto [0, range) by multiply-shift (range <= 2^32).
EEG_HD uint32_t draw_below(uint64_t r, uint64_t range) {
    return (uint32_t)(((r >> 32) * range) >> 32);
}

enum : uint64_t {
    STREAM_RMAT = 1,      // quadrant choices, two 32-bit levels per draw
    STREAM_WEIGHT = 2,    // one draw per directed edge
    STREAM_SHORTCUT = 3,  // two draws per shortcut pair (endpoints)
};

// RMAT quadrant thresholds on a 32-bit uniform: u < ta -> (0,0), < tab -> (0,1),
// < tabc -> (1,0), else (1,1).  Computed as floor(p * 2^32) in double precision.
struct RmatThresholds {
    uint64_t ta, tab, tabc;
};
inline RmatThresholds rmat_thresholds(double a, double b, double c) {
    auto t = [](double p) -> uint64_t {
        double v = p * 4294967296.0;
        if (v < 0) v = 0;
        if (v > 4294967296.0) v = 4294967296.0;
        return (uint64_t)v;
    };
    return RmatThresholds{t(a), t(a + b), t(a + b + c)};
}

// Edge i of RMAT(scale): levels are consumed most-significant bit first; level l uses the
// high half (even l) or low half (odd l) of draw i*ceil(scale/2) + l/2 of STREAM_RMAT.
EEG_HD void rmat_edge(uint64_t key, uint32_t scale, uint64_t ta, uint64_t tab, uint64_t tabc,
                      uint64_t i, uint32_t* s_out, uint32_t* d_out) {
    const uint64_t per = (scale + 1) / 2;
    uint32_t s = 0, d = 0;
    uint64_t r = 0;
    for (uint32_t l = 0; l < scale; ++l) {
        if ((l & 1u) == 0) r = stream_draw(key, i * per + (l >> 1));
        const uint64_t u = (l & 1u) ? (r & 0xffffffffULL) : (r >> 32);
        uint32_t bs, bd;
        if (u < ta) { bs = 0; bd = 0; }
        else if (u < tab) { bs = 0; bd = 1; }
        else if (u < tabc) { bs = 1; bd = 0; }
        else { bs = 1; bd = 1; }
        s = (s << 1) | bs;
        d = (d << 1) | bd;
    }
    *s_out = s;
    *d_out = d;
}
EEG_HD uint32_t edge_weight(uint64_t wkey, uint64_t i, uint32_t wmin, uint32_t wmax) {
    return wmin + draw_below(stream_draw(wkey, i), (uint64_t)wmax - wmin + 1ULL);
}

}  // namespace eeg
