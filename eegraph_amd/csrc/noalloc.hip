// eeg_run_device_noalloc: the single-query traversal of eeg_run_device without any allocation,
// for callers that cannot allocate (code under HIP graph capture, tight serving loops,
// harnesses that forbid hipMalloc).  No counterpart exists in the reference, whose drivers
// allocate and free every device array inside each call (src/eegraph.cu:61-74, 262-268).
//
// The traversal state (labels, two frontier queues with their virtual-node prefixes, dedup
// stamps, control block) lives in `scratch`, a fixed device array of kScratchBytes that every
// GPU the library runs on holds until the process ends.  Each GPU has its own copy at its own
// address, found once per device with hipGetSymbolAddress; host code uses only that address.
// All calls on one GPU share the array: calls that overlap on one GPU race and must be
// serialised by the caller (the graph lock serialises calls on the same graph only).
#include <cstring>
#include <mutex>
#include <string>

#include "engine.hpp"

namespace eeg {
namespace {

// Arrays of (kMaxVertices + 64) words, each a multiple of 256 bytes, then the control block.
constexpr size_t kWords = (size_t)EEG_NOALLOC_MAX_VERTICES + 64;
constexpr size_t kArrayBytes = kWords * 4;
constexpr size_t kArrays = 6;  // labels, q0, q1, s0, s1, stamps
constexpr size_t kScratchBytes = kArrays * kArrayBytes + 256;
static_assert(kArrayBytes % 256 == 0, "scratch arrays stay 256-byte aligned");
static_assert(sizeof(Ctl) <= 256, "control block fits its slot");

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// Device address of this GPU's `scratch`, looked up on the first call with the device
// current (the array's name in host code is only a stand-in for the symbol).
eeg_status scratch_for(int device, unsigned char** out) {
    if (device < 0 || device >= kMaxDevices) return fail(EEG_ERR_INVALID_ARG, "device %d out of range", device);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            (void)hipGetLastError();
            return fail(EEG_ERR_DEVICE, "hipGetSymbolAddress(scratch) failed on device %d: %s", device,
                        hipGetErrorString(e));
        }
        g_scratch_ptr[device] = p;
    }
    *out = static_cast<unsigned char*>(g_scratch_ptr[device]);
    return EEG_OK;
}

}  // namespace
}  // namespace eeg

using namespace eeg;

extern "C" eeg_status eeg_run_device_noalloc(eeg_graph* g, uint32_t algo, uint32_t source,
                                             const eeg_opts* opts, void* d_labels_out, void* stream,
                                             eeg_stats* stats) {
    clear_error();
    if (!g || !d_labels_out) return fail(EEG_ERR_INVALID_ARG, "null graph or output");
    if (algo > EEG_CC) return fail(EEG_ERR_UNSUPPORTED, "only BFS, SSSP, SSWP and CC run without allocation");
    if (g->n > EEG_NOALLOC_MAX_VERTICES)
        return fail(EEG_ERR_INVALID_ARG, "graph has %u vertices; the allocation-free entry point holds at most %u",
                    g->n, EEG_NOALLOC_MAX_VERTICES);
    const bool sourced = algo != EEG_CC;
    if (sourced && source >= g->n)
        return fail(EEG_ERR_INVALID_ARG, "source %u out of range (n = %u)", source, g->n);
    eeg_opts o;
    memset(&o, 0, sizeof(o));
    if (opts) o = *opts;
    eeg_stats local;
    eeg_stats* st = stats ? stats : &local;
    memset(st, 0, sizeof(*st));
    std::lock_guard<std::mutex> lk(g->mu);
    DeviceGuard dg;
    EEG_HIP(hipSetDevice(g->device));
    unsigned char* base = nullptr;
    EEG_TRY(scratch_for(g->device, &base));
    FixedState fs;
    fs.lab = reinterpret_cast<uint32_t*>(base + 0 * kArrayBytes);
    fs.q[0] = reinterpret_cast<uint32_t*>(base + 1 * kArrayBytes);
    fs.q[1] = reinterpret_cast<uint32_t*>(base + 2 * kArrayBytes);
    fs.s[0] = reinterpret_cast<uint32_t*>(base + 3 * kArrayBytes);
    fs.s[1] = reinterpret_cast<uint32_t*>(base + 4 * kArrayBytes);
    fs.stamp = reinterpret_cast<uint32_t*>(base + 5 * kArrayBytes);
    fs.ctl = reinterpret_cast<Ctl*>(base + kArrays * kArrayBytes);
    hipStream_t s = stream ? (hipStream_t)stream : g->stream;
    const uint32_t isrc = (sourced && g->n) ? g->h_perm[source] : 0u;
    eeg_status rc = run_traversal_fixed(g, algo, isrc, o, fs, s, st);
    if (rc != EEG_OK && rc != EEG_ERR_NOT_CONVERGED) return rc;
    const std::string detail = (rc != EEG_OK) ? eeg_last_error() : "";
    EEG_TRY(unpermute(g, fs.lab, d_labels_out, s));  // back to original vertex order
    EEG_HIP(hipStreamSynchronize(s));
    if (rc != EEG_OK) fail(rc, "%s", detail.c_str());
    return rc;
}
