"""The branch-free label arithmetic of the narrow label planes (eegraph_amd/csrc/bytes8.hpp)
equals per-label loops: candidates, overflow bits, min/max merges, improved and unreached
masks, on random words (8-bit labels biased to 0, 1, 127, 128, 253, 254 and 255, weights
0..254; 16-bit labels biased to the ends of their range, weights up to 2^32 - 1).  The header
is compiled for the host with the ROCm clang."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLANG = "/opt/rocm/llvm/bin/clang++"

HARNESS = r"""
#include <cstdio>
#include <random>
#include <algorithm>
#define EEG_HD inline
#include "bytes8.hpp"
using namespace eeg;
static uint32_t f(uint32_t x, int i) { return (x >> (8 * i)) & 0xffu; }
// widest path: per-label loops
static uint32_t w_cand(uint32_t lv, uint32_t w, uint32_t nib) {
    uint32_t c = 0;
    for (int i = 0; i < 4; ++i) if ((nib >> i) & 1u) c |= std::min(f(lv, i), w) << (8 * i);
    return c;
}
static uint32_t w_merge(uint32_t a, uint32_t b) { uint32_t r = 0; for (int i = 0; i < 4; ++i) r |= std::max(f(a, i), f(b, i)) << (8 * i); return r; }
static uint32_t w_imp(uint32_t cur, uint32_t c) { uint32_t m = 0; for (int i = 0; i < 4; ++i) m |= (f(c, i) > f(cur, i) ? 1u : 0u) << i; return m; }
// distances: 255 = INF, sums >= 255 of finite pending labels are dropped and flagged
static uint32_t d_cand(uint32_t lv, uint32_t w, uint32_t nib, uint32_t& ov) {
    uint32_t c = 0xffffffffu;
    for (int i = 0; i < 4; ++i) {
        const uint32_t l = f(lv, i);
        if (!((nib >> i) & 1u) || l == 0xffu) continue;
        if (l + w < 0xffu) c = (c & ~(0xffu << (8 * i))) | ((l + w) << (8 * i));
        else ov |= 1u << i;
    }
    return c;
}
static uint32_t d_merge(uint32_t a, uint32_t b) { uint32_t r = 0; for (int i = 0; i < 4; ++i) r |= std::min(f(a, i), f(b, i)) << (8 * i); return r; }
static uint32_t d_imp(uint32_t cur, uint32_t c) { uint32_t m = 0; for (int i = 0; i < 4; ++i) m |= (f(c, i) < f(cur, i) ? 1u : 0u) << i; return m; }
static uint32_t d_unr(uint32_t cur) { uint32_t m = 0; for (int i = 0; i < 4; ++i) m |= (f(cur, i) == 0xffu ? 1u : 0u) << i; return m; }
// 16-bit distances: per-half loop (65535 = INF; a pending finite half whose sum reaches 65535
// is dropped and flagged)
static uint32_t h(uint32_t x, int i) { return (x >> (16 * i)) & 0xffffu; }
static uint32_t e_cand(uint32_t lv, uint32_t w, uint32_t nib, uint32_t& ov) {
    uint32_t c = 0xffffffffu;
    for (int i = 0; i < 2; ++i) {
        const uint32_t l = h(lv, i);
        if (!((nib >> i) & 1u) || l == 0xffffu) continue;
        const uint64_t sm = (uint64_t)l + w;
        if (sm < 0xffffu) c = (c & ~(0xffffu << (16 * i))) | ((uint32_t)sm << (16 * i));
        else ov |= 1u << i;
    }
    return c;
}
static uint32_t e_merge(uint32_t a, uint32_t b) { uint32_t r = 0; for (int i = 0; i < 2; ++i) r |= std::min(h(a, i), h(b, i)) << (16 * i); return r; }
static uint32_t e_imp(uint32_t cur, uint32_t c) { uint32_t m = 0; for (int i = 0; i < 2; ++i) m |= (h(c, i) < h(cur, i) ? 1u : 0u) << i; return m; }
static uint32_t e_unr(uint32_t cur) { uint32_t m = 0; for (int i = 0; i < 2; ++i) m |= (h(cur, i) == 0xffffu ? 1u : 0u) << i; return m; }
static long check16(std::mt19937_64& rng) {
    const uint32_t sp[] = {0, 1, 0x7fff, 0x8000, 0xfffd, 0xfffe, 0xffff};
    const uint32_t wsp[] = {0, 1, 127, 128, 1000, 0xfffe, 0xffff, 0x10000, 0xffffffffu};
    auto half = [&]() -> uint32_t { return (rng() & 1) ? sp[rng() % 7] : (uint32_t)(rng() & 0xffff); };
    long bad = 0;
    for (long it = 0; it < 2000000; ++it) {
        const uint32_t lv = half() | (half() << 16), cur = half() | (half() << 16), nib = rng() & 3;
        const uint32_t w = (rng() & 1) ? (uint32_t)(rng() & 0xffff) : wsp[rng() % 9];
        uint32_t o1 = 0, o2 = 0;
        const uint32_t c1 = Bytes16Dist::cand(lv, w, nib, o1), c2 = e_cand(lv, w, nib, o2);
        const bool ok = c1 == c2 && o1 == o2 && Bytes16Dist::merge(cur, c1) == e_merge(cur, c2) &&
                        Bytes16Dist::improved(cur, c1) == e_imp(cur, c2) && Bytes16Dist::unreached(cur) == e_unr(cur);
        if (!ok && bad++ < 5) printf("16-bit mismatch lv=%08x cur=%08x nib=%x w=%u\n", lv, cur, nib, w);
    }
    return bad;
}
int main() {
    std::mt19937_64 rng(7);
    const uint32_t special[] = {0, 1, 127, 128, 253, 254, 255};
    auto byte = [&]() -> uint32_t { return (rng() & 1) ? special[rng() % 7] : (uint32_t)(rng() & 0xff); };
    auto word = [&]() -> uint32_t { uint32_t x = 0; for (int i = 0; i < 4; ++i) x |= byte() << (8 * i); return x; };
    long bad = 0;
    for (long it = 0; it < 4000000; ++it) {
        const uint32_t lv = word(), cur = word(), nib = rng() & 15;
        const uint32_t w = (rng() & 1) ? (uint32_t)(rng() % 255) : special[rng() % 6];
        uint32_t o1 = 0, o2 = 0, u1 = 0;
        const uint32_t c1 = Bytes8Dist::cand(lv, w, nib, o1), c2 = d_cand(lv, w, nib, o2);
        const uint32_t a1 = Bytes8Widest::cand(lv, w, nib, u1), a2 = w_cand(lv, w, nib);
        const bool ok = c1 == c2 && o1 == o2 && Bytes8Dist::merge(cur, c1) == d_merge(cur, c2) &&
                        Bytes8Dist::improved(cur, c1) == d_imp(cur, c2) && Bytes8Dist::unreached(cur) == d_unr(cur) &&
                        a1 == a2 && u1 == 0 && Bytes8Widest::merge(cur, a1) == w_merge(cur, a2) &&
                        Bytes8Widest::improved(cur, a1) == w_imp(cur, a2);
        if (!ok && bad++ < 5) printf("mismatch lv=%08x cur=%08x nib=%x w=%u\n", lv, cur, nib, w);
    }
    bad += check16(rng);
    printf("mismatches %ld\n", bad);
    return bad != 0;
}
"""


@pytest.mark.skipif(not os.path.exists(CLANG), reason="needs the ROCm clang")
def test_bytes8_equals_per_label_loops(tmp_path):
    src = tmp_path / "bytes8_check.cpp"
    src.write_text(HARNESS)
    exe = tmp_path / "bytes8_check"
    subprocess.run([CLANG, "-O2", "-std=c++17", "-I", os.path.join(ROOT, "eegraph_amd", "csrc"), str(src),
                    "-o", str(exe)], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "mismatches 0" in r.stdout
