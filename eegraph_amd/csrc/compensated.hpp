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

// Value of lane (this lane + OFF) of the same 16-lane row: a DPP row shift (one VALU move, no
// LDS permute) for OFF < 16, a shuffle beyond.  Lanes whose partner lies outside the row get 0.
template <int OFF>
__device__ __forceinline__ float lane_down(float v) {
    if constexpr (OFF < 16) {
        return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x100 + OFF, 0xF, 0xF, false));
    } else {
        return __shfl_down(v, OFF, WAVE);
    }
}

// Compensated sum of the x of a group of K aligned lanes, valid in the group's first lane:
// a pairwise tree, ((x0 + x1) + (x2 + x3)) + ((x4 + x5) + (x6 + x7)) for K = 8, each step a
// TwoSum with the error terms carried.  The tree and the operand order are fixed, so the bits
// do not depend on the schedule.
template <int K, int OFF = 1>
__device__ __forceinline__ Comp group_reduce_step(Comp c) {
    if constexpr (OFF >= K) {
        return c;
    } else {
        Comp o;
        o.s = lane_down<OFF>(c.s);
        o.e = lane_down<OFF>(c.e);
        return group_reduce_step<K, 2 * OFF>(comp_add(c, o));
    }
}
template <int K>
__device__ __forceinline__ Comp group_sum_first_lane(float x) {
    // first level: the inputs carry no error term
    const Comp c = two_sum(x, lane_down<1>(x));
    return group_reduce_step<K, 2>(c);
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
