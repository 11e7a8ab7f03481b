// Device-side building blocks for gfx950 (CDNA4, wave64): wavefront primitives, label rules
// and the control block shared by the traversal kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace eeg {

constexpr int WAVE = 64;
constexpr int WG = 256;            // threads per workgroup (4 waves)
constexpr int TILE_EDGES = 2048;   // edge slots per tile: 8 rounds of one edge per lane
constexpr uint32_t INF = 4294967294u;

__device__ __forceinline__ int lane_id() {
    return (int)__builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
}
__device__ __forceinline__ uint64_t lanemask_lt() {
    const int l = lane_id();
    return (l == 0) ? 0ull : (~0ull >> (64 - l));
}

// Inclusive prefix sum across the 64 lanes of a wavefront.
template <class T>
__device__ __forceinline__ T wave_incl_scan(T v) {
    const int l = lane_id();
#pragma unroll
    for (int off = 1; off < WAVE; off <<= 1) {
        T o = __shfl_up(v, off, WAVE);
        if (l >= off) v += o;
    }
    return v;
}
template <class T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
    for (int off = WAVE / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, WAVE);
    return v;
}

// Exclusive prefix sum over a workgroup (BS threads); returns the total through *total.
// Uses `scratch` (BS/WAVE + 1 entries of LDS).  Every thread must call it.
template <class T, int BS = WG>
__device__ __forceinline__ T block_excl_scan(T v, T* scratch, T* total) {
    const int tid = threadIdx.x;
    const int w = tid / WAVE;
    T incl = wave_incl_scan(v);
    if ((tid % WAVE) == WAVE - 1) scratch[w] = incl;
    __syncthreads();
    if (tid == 0) {
        T run = 0;
        for (int i = 0; i < BS / WAVE; ++i) {
            T t = scratch[i];
            scratch[i] = run;
            run += t;
        }
        scratch[BS / WAVE] = run;
    }
    __syncthreads();
    T r = incl - v + scratch[w];
    *total = scratch[BS / WAVE];
    __syncthreads();
    return r;
}

// Largest i in [0, n) with S[i] <= t, for ascending S with S[0] <= t, found by the whole
// workgroup of BS threads: each round samples BS evenly spaced positions and narrows the
// interval BS-fold.
template <int BS = WG>
__device__ __forceinline__ uint32_t wg_search(const uint32_t* __restrict__ S, uint32_t n, uint32_t t) {
    uint32_t lo = 0, hi = n;
    while (hi - lo > 1) {
        const uint32_t span = hi - lo;
        const uint32_t step = (span + BS - 1) / BS;
        const uint64_t i = (uint64_t)lo + (uint64_t)threadIdx.x * step;
        const int pred = (i < hi) && (S[i] <= t);
        const int cnt = __syncthreads_count(pred);
        const uint32_t nlo = lo + (uint32_t)(cnt - 1) * step;
        const uint32_t nhi = (uint32_t)min((uint64_t)nlo + step, (uint64_t)hi);
        lo = nlo;
        hi = nhi;
    }
    return lo;
}

// Relaxation rules.  cand() returns the label a neighbour would get over an edge of weight w
// from a source labelled `lab`; ok is false when the source cannot improve anything (BFS/SSSP
// relax only from finite labels, which removes the INF+w wrap of src/sssp.cu:62).
enum Algo : int { A_BFS = 0, A_SSSP = 1, A_SSWP = 2, A_CC = 3 };

template <int ALG>
struct Rule;
template <>
struct Rule<A_BFS> {  // src/bfs.cu:60-66: dist[u] + 1, atomicMin
    static constexpr bool weighted = false;
    __device__ static bool usable(uint32_t lab) { return lab < INF; }
    __device__ static uint32_t cand(uint32_t lab, uint32_t) { return lab + 1u; }
    __device__ static bool better(uint32_t a, uint32_t b) { return a < b; }
    __device__ static uint32_t apply(uint32_t* p, uint32_t v) { return atomicMin(p, v); }
};
template <>
struct Rule<A_SSSP> {  // src/sssp.cu:60-66: dist[u] + w, atomicMin (saturating at INF here)
    static constexpr bool weighted = true;
    __device__ static bool usable(uint32_t lab) { return lab < INF; }
    __device__ static uint32_t cand(uint32_t lab, uint32_t w) {
        const uint64_t s = (uint64_t)lab + w;
        return s < INF ? (uint32_t)s : INF;
    }
    __device__ static bool better(uint32_t a, uint32_t b) { return a < b; }
    __device__ static uint32_t apply(uint32_t* p, uint32_t v) { return atomicMin(p, v); }
};
template <>
struct Rule<A_SSWP> {  // src/sswp.cu:60-66: min(dist[u], w), atomicMax
    static constexpr bool weighted = true;
    __device__ static bool usable(uint32_t lab) { return lab > 0; }
    __device__ static uint32_t cand(uint32_t lab, uint32_t w) { return lab < w ? lab : w; }
    __device__ static bool better(uint32_t a, uint32_t b) { return a > b; }
    __device__ static uint32_t apply(uint32_t* p, uint32_t v) { return atomicMax(p, v); }
};
template <>
struct Rule<A_CC> {  // min-label propagation (Subway/shared/gpu_kernels.cu:96-98 rule)
    static constexpr bool weighted = false;
    __device__ static bool usable(uint32_t) { return true; }
    __device__ static uint32_t cand(uint32_t lab, uint32_t) { return lab; }
    __device__ static bool better(uint32_t a, uint32_t b) { return a < b; }
    __device__ static uint32_t apply(uint32_t* p, uint32_t v) { return atomicMin(p, v); }
};

// Control block of one traversal, resident in HBM.  Iteration i (1-based, offset by an
// epoch so per-vertex stamps never need clearing) reads frontier slot (i-1)%3, writes slot
// i%3 and clears slot (i+1)%3, so no host memset is needed between launches.
struct Ctl {
    unsigned long long qpack[3];   // (vnodes << 32) | entries of the frontier queues
    uint32_t changed[4];           // per-iteration "some label improved" flags
    uint32_t done;                 // set once an iteration finds no work
    uint32_t iters;                // iterations that did work
    uint32_t err;                  // nonzero: a bound was hit (never written out of bounds)
    uint32_t rounds;               // residual push: take passes that found work
    unsigned long long stat[4];    // entries staged, vnodes pushed, edges relaxed, updates
    double residual[4];            // PR: L1 residual of the last iterations (slot i&3)
    double mass;                   // PR delta push / PPR: residual mass still unpushed
    double pad2[3];
};

}  // namespace eeg
