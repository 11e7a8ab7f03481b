// Sorted frontier queues, shared by the single-query traversals (relax.hip) and the batched
// plane passes (batch.hip): a kernel that improved a vertex sets its bit in a 1-bit-per-vertex
// bitmap, and k_compact_frontier turns the bits into the next queue in vertex order.
#pragma once

#include "device.hpp"

namespace eeg {
namespace {

// Sorted frontier: the vertices whose bit iteration `it` set, with out-edges, as the next
// queue (buffer it & 1, Ctl slot it % 3), in ascending id order within each block's 8192-vertex
// chunk; chunks take their queue ranges in the order their atomic lands.  Bits are cleared as
// they are read.  One thread per 32-bit word; replayed graphs resolve `it` from dev_it.
struct CompactArgs {
    uint32_t* bitmap;
    uint32_t nwords;
    uint32_t n;
    const uint32_t* vnc;
    Ctl* ctl;
    uint32_t *qa, *sa, *qb, *sb;  // queue buffers 0 and 1
    uint32_t cap;
    uint32_t it;
    const uint32_t* dev_it;
    uint32_t it_off;
};

__global__ __launch_bounds__(WG) void k_compact_frontier(CompactArgs a) {
    const uint32_t it = a.dev_it ? *a.dev_it + a.it_off : a.it;
    Ctl* ctl = a.ctl;
    if (ctl->done) return;
    uint32_t* qout = (it & 1u) ? a.qb : a.qa;
    uint32_t* sout = (it & 1u) ? a.sb : a.sa;
    const int slot = (int)(it % 3u);
    __shared__ uint64_t sScan[WG / WAVE + 1];
    __shared__ uint32_t sBase[2];
    for (uint32_t c0 = blockIdx.x * WG; c0 < a.nwords; c0 += gridDim.x * WG) {
        const uint32_t wi = c0 + threadIdx.x;
        const uint32_t word = wi < a.nwords ? a.bitmap[wi] : 0u;
        if (!__syncthreads_or(word != 0u)) continue;  // the same answer in every thread
        // the word's 32 virtual-node counts, read once: one 128-byte line as 8 vector loads
        // (a last word past n reads only the counts that exist)
        uint32_t cv[32];
        if (word) {
            a.bitmap[wi] = 0u;
            const uint32_t v0 = wi * 32u;
            if (v0 + 32u <= a.n) {
                const uint4* p = reinterpret_cast<const uint4*>(a.vnc + v0);
#pragma unroll
                for (int k = 0; k < 8; ++k) {
                    const uint4 x = p[k];
                    cv[4 * k] = x.x;
                    cv[4 * k + 1] = x.y;
                    cv[4 * k + 2] = x.z;
                    cv[4 * k + 3] = x.w;
                }
            } else {
#pragma unroll
                for (int b = 0; b < 32; ++b) cv[b] = (v0 + (uint32_t)b < a.n) ? a.vnc[v0 + b] : 0u;
            }
        } else {
#pragma unroll
            for (int b = 0; b < 32; ++b) cv[b] = 0u;
        }
        uint32_t cnt = 0, vs = 0;
#pragma unroll
        for (int b = 0; b < 32; ++b) {
            const bool on = ((word >> b) & 1u) && cv[b];
            cnt += on ? 1u : 0u;
            vs += on ? cv[b] : 0u;
        }
        uint64_t tot;
        const uint64_t pre = block_excl_scan<uint64_t>(((uint64_t)vs << 32) | cnt, sScan, &tot);
        if (threadIdx.x == 0) {
            const unsigned long long old = tot ? atomicAdd(&ctl->qpack[slot], (unsigned long long)tot) : 0ull;
            sBase[0] = (uint32_t)(old & 0xffffffffull);
            sBase[1] = (uint32_t)(old >> 32);
        }
        __syncthreads();
        uint32_t e = sBase[0] + (uint32_t)(pre & 0xffffffffull);
        uint32_t vv = sBase[1] + (uint32_t)(pre >> 32);
#pragma unroll
        for (int b = 0; b < 32; ++b) {
            // vertices without out-edges are left out: their labels are final, nothing to push
            if (!((word >> b) & 1u) || !cv[b]) continue;
            if (e < a.cap) {
                qout[e] = wi * 32u + (uint32_t)b;
                sout[e] = vv;
            } else {
                atomicOr(&ctl->err, 2u);
            }
            ++e;
            vv += cv[b];
        }
        __syncthreads();  // sBase is rewritten by the next chunk
    }
}

}  // namespace
}  // namespace eeg
