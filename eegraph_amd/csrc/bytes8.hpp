// Four 8-bit labels per 32-bit word, handled bytewise on the whole word: pairs of bytes in
// 16-bit halves with packed min/max (v_pk_min_u16 / v_pk_max_u16 on gfx950), no branch per
// label.  Used by the narrow label planes of the batched passes (batch.hip); host-callable so a
// CPU test can check them against per-label loops.
#pragma once
#include <cstdint>

#ifndef EEG_HD
#define EEG_HD __host__ __device__ inline
#endif

namespace eeg {

typedef unsigned short eeg_us2 __attribute__((ext_vector_type(2)));
EEG_HD uint32_t pk_min_u16(uint32_t a, uint32_t b) {
    return __builtin_bit_cast(uint32_t, __builtin_elementwise_min(__builtin_bit_cast(eeg_us2, a),
                                                                  __builtin_bit_cast(eeg_us2, b)));
}
EEG_HD uint32_t pk_max_u16(uint32_t a, uint32_t b) {
    return __builtin_bit_cast(uint32_t, __builtin_elementwise_max(__builtin_bit_cast(eeg_us2, a),
                                                                  __builtin_bit_cast(eeg_us2, b)));
}
// bit 7 of each byte -> bits 0..3
EEG_HD uint32_t pack4(uint32_t hb) { return (((hb >> 7) & 0x01010101u) * 0x01020408u) >> 24; }
// bit 7 set in exactly the bytes of t that are zero
EEG_HD uint32_t zero_bytes(uint32_t t) { return ~(((t & 0x7f7f7f7fu) + 0x7f7f7f7fu) | t | 0x7f7f7f7fu); }
// 0xff in the bytes whose bit of nib is set
EEG_HD uint32_t byte_sel(uint32_t nib) {
    return ((nib & 1u) | ((nib & 2u) << 7) | ((nib & 4u) << 14) | ((nib & 8u) << 21)) * 0xffu;
}

// Widest-path labels (0 = unreached, 255 = INF at the source), weights <= 254.
struct Bytes8Widest {
    static constexpr uint32_t QPL = 4, BITS = 8, FM = 0xffu;  // labels per 32-bit lane word
    // candidates of the labels in `nib` over an edge of weight w: min(label, w); 0 (no
    // candidate) for the others
    EEG_HD static uint32_t cand(uint32_t lv, uint32_t w, uint32_t nib, uint32_t&) {
        const uint32_t l = lv & byte_sel(nib), ww = w * 0x00010001u;
        return pk_min_u16(l & 0x00ff00ffu, ww) | (pk_min_u16((l >> 8) & 0x00ff00ffu, ww) << 8);
    }
    EEG_HD static uint32_t merge(uint32_t cur, uint32_t c) {  // per-label max
        return pk_max_u16(cur & 0x00ff00ffu, c & 0x00ff00ffu) |
               (pk_max_u16((cur >> 8) & 0x00ff00ffu, (c >> 8) & 0x00ff00ffu) << 8);
    }
    EEG_HD static uint32_t improved(uint32_t cur, uint32_t c) {  // labels where c beats cur
        return pack4(~zero_bytes(merge(cur, c) ^ cur) & 0x80808080u);
    }
};

// Distances (255 = INF), weights <= 254.  Pairs of bytes are added in 16-bit halves, so a sum
// up to 255 + 254 keeps its carry in bit 8, which saturates the byte to 255.  cand() sets bit i
// of `ov` for a pending finite label whose sum does not fit.
struct Bytes8Dist {
    static constexpr uint32_t QPL = 4, BITS = 8, FM = 0xffu;
    EEG_HD static uint32_t cand(uint32_t lv, uint32_t w, uint32_t nib, uint32_t& ov) {
        const uint32_t l = lv | ~byte_sel(nib);  // labels not asked for count as INF
        const uint32_t ww = w * 0x00010001u;
        uint32_t se = (l & 0x00ff00ffu) + ww, so = ((l >> 8) & 0x00ff00ffu) + ww;
        se = (se | (((se >> 8) & 0x00010001u) * 0xffu)) & 0x00ff00ffu;
        so = (so | (((so >> 8) & 0x00010001u) * 0xffu)) & 0x00ff00ffu;
        const uint32_t c = se | (so << 8);
        ov |= pack4(zero_bytes(~c) & ~zero_bytes(~l));
        return c;
    }
    EEG_HD static uint32_t merge(uint32_t cur, uint32_t c) {  // per-label min
        return pk_min_u16(cur & 0x00ff00ffu, c & 0x00ff00ffu) |
               (pk_min_u16((cur >> 8) & 0x00ff00ffu, (c >> 8) & 0x00ff00ffu) << 8);
    }
    EEG_HD static uint32_t improved(uint32_t cur, uint32_t c) {  // labels where c beats cur
        return pack4(~zero_bytes(merge(cur, c) ^ cur) & 0x80808080u);
    }
    EEG_HD static uint32_t unreached(uint32_t cur) { return pack4(zero_bytes(~cur)); }  // labels at INF
};

// Distances of 16 bits (0xffff = INF), two per word, any weight: the same results as a loop
// over the two halves (a pending finite half whose sum reaches 0xffff sets its bit of `ov`;
// weights are clamped at 0xffff, which keeps such sums at or past 0xffff), without branches.
struct Bytes16Dist {
    static constexpr uint32_t QPL = 2, BITS = 16, FM = 0xffffu;
    EEG_HD static uint32_t cand(uint32_t lv, uint32_t w, uint32_t nib, uint32_t& ov) {
        const uint32_t wc = w < 0xffffu ? w : 0xffffu;
        const uint32_t lo = (nib & 1u) ? (lv & 0xffffu) : 0xffffu, hi = (nib & 2u) ? (lv >> 16) : 0xffffu;
        const uint32_t slo = lo + wc, shi = hi + wc;  // at most 2 x 0xffff: no carry out
        const bool flo = lo != 0xffffu, fhi = hi != 0xffffu;
        ov |= ((flo && slo >= 0xffffu) ? 1u : 0u) | ((fhi && shi >= 0xffffu) ? 2u : 0u);
        const uint32_t clo = slo < 0xffffu ? slo : 0xffffu, chi = shi < 0xffffu ? shi : 0xffffu;
        return clo | (chi << 16);
    }
    EEG_HD static uint32_t merge(uint32_t cur, uint32_t c) { return pk_min_u16(cur, c); }  // per-label min
    EEG_HD static uint32_t improved(uint32_t cur, uint32_t c) {  // labels where c beats cur
        return ((c & 0xffffu) < (cur & 0xffffu) ? 1u : 0u) | ((c >> 16) < (cur >> 16) ? 2u : 0u);
    }
    EEG_HD static uint32_t unreached(uint32_t cur) {  // labels at INF
        return ((cur & 0xffffu) == 0xffffu ? 1u : 0u) | ((cur >> 16) == 0xffffu ? 2u : 0u);
    }
};

}  // namespace eeg
