// Compensated FP32 arithmetic of the PageRank kernels.
#pragma once

#include "device.hpp"

namespace eeg {

struct Comp {  // value s with the rounding error e that s + e represents exactly
    float s, e;
};
__device__ __forceinline__ Comp two_sum(float a, float b) {
    const float s = a + b;
    const float bb = s - a;
    const float err = (a - (s - bb)) + (b - bb);
    return Comp{s, err};
}
__device__ __forceinline__ Comp comp_add(Comp a, Comp b) {
    const Comp t = two_sum(a.s, b.s);
    return Comp{t.s, t.e + (a.e + b.e)};
}

// Butterfly over a group of K lanes: every lane ends with the same (s, e), because each step
// combines the same two operands on both partner lanes (addition is commutative and the
// TwoSum error is exact).
template <int K>
__device__ __forceinline__ Comp group_reduce(Comp c) {
#pragma unroll
    for (int off = 1; off < K; off <<= 1) {
        Comp o;
        o.s = __shfl_xor(c.s, off, WAVE);
        o.e = __shfl_xor(c.e, off, WAVE);
        c = comp_add(c, o);
    }
    return c;
}

// New rank base + d * sum kept as an FP32 pair (hi, lo): the product d*s is split exactly with
// an FMA, base + d*s with TwoSum, and the small terms folded into lo.  The pair carries ~46
// bits, so the iteration settles on an exact fixed point instead of cycling by one FP32 ulp.
__device__ __forceinline__ void rank_pair(float base, float d, Comp sum, float* hi, float* lo) {
    const float p = d * sum.s;
    const float pe = fmaf(d, sum.s, -p);
    const Comp q = two_sum(base, p);
    const float small = q.e + (pe + d * sum.e);
    const float h = q.s + small;            // fast two-sum: |q.s| >= |small|
    *hi = h;
    *lo = small - (h - q.s);
}

}  // namespace eeg
