// Device-wide exclusive scan over u64 values given by a functor, shared by the graph builders.
#pragma once

#include "engine.hpp"

namespace eeg {

constexpr int SCAN_ITEMS = 16;
constexpr int SCAN_TILE = WG * SCAN_ITEMS;

// ---------------------------------------------------------------- exclusive scan (u64)
// Three phases: per-tile reduce -> scan of tile sums (recursive) -> per-tile rescan with the
// tile's offset.  F(i) yields element i, Wr(i, prefix) consumes its exclusive prefix.
template <class F>
__global__ __launch_bounds__(WG) void k_scan_reduce(F f, uint64_t n, uint64_t* __restrict__ sums) {
    const uint64_t base = (uint64_t)blockIdx.x * SCAN_TILE + (uint64_t)threadIdx.x * SCAN_ITEMS;
    uint64_t s = 0;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; ++k)
        if (base + k < n) s += f(base + k);
    __shared__ uint64_t part[WG / WAVE];
    s = wave_sum(s);
    if (lane_id() == 0) part[threadIdx.x / WAVE] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint64_t t = 0;
        for (int i = 0; i < WG / WAVE; ++i) t += part[i];
        sums[blockIdx.x] = t;
    }
}

template <class F, class Wr>
__global__ __launch_bounds__(WG) void k_scan_apply(F f, Wr wr, uint64_t n,
                                                   const uint64_t* __restrict__ offs,
                                                   uint64_t* __restrict__ total) {
    const uint64_t base = (uint64_t)blockIdx.x * SCAN_TILE + (uint64_t)threadIdx.x * SCAN_ITEMS;
    uint64_t v[SCAN_ITEMS];
    uint64_t s = 0;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; ++k) {
        v[k] = (base + k < n) ? f(base + k) : 0;
        s += v[k];
    }
    __shared__ uint64_t scratch[WG / WAVE + 1];
    uint64_t tot;
    uint64_t run = block_excl_scan<uint64_t>(s, scratch, &tot) + (offs ? offs[blockIdx.x] : 0);
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; ++k) {
        if (base + k < n) wr(base + k, run);
        run += v[k];
    }
    if (total && blockIdx.x == gridDim.x - 1 && threadIdx.x == WG - 1) *total = run;
}

struct LoadArr {
    const uint64_t* a;
    __device__ uint64_t operator()(uint64_t i) const { return a[i]; }
};
struct StoreArr {
    uint64_t* a;
    __device__ void operator()(uint64_t i, uint64_t v) const { a[i] = v; }
};

template <class F, class Wr>
eeg_status scan_exclusive(F f, Wr wr, uint64_t n, hipStream_t st, uint64_t* d_total) {
    if (n == 0) return EEG_OK;
    const uint64_t ntiles = (n + SCAN_TILE - 1) / SCAN_TILE;
    if (ntiles == 1) {
        k_scan_apply<<<1, WG, 0, st>>>(f, wr, n, nullptr, d_total);
        EEG_HIP(hipGetLastError());
        return EEG_OK;
    }
    DevMem sums;
    EEG_TRY(sums.alloc(ntiles * sizeof(uint64_t)));
    k_scan_reduce<<<(unsigned)ntiles, WG, 0, st>>>(f, n, sums.as<uint64_t>());
    EEG_HIP(hipGetLastError());
    EEG_TRY(scan_exclusive(LoadArr{sums.as<uint64_t>()}, StoreArr{sums.as<uint64_t>()}, ntiles,
                           st, nullptr));
    k_scan_apply<<<(unsigned)ntiles, WG, 0, st>>>(f, wr, n, sums.as<uint64_t>(), d_total);
    EEG_HIP(hipGetLastError());
    // sums must outlive the kernels that read it
    EEG_HIP(hipStreamSynchronize(st));
    return EEG_OK;
}

}  // namespace eeg
