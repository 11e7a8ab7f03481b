// Device energy through amd-smi's energy accumulator.  The reference samples NVML power every
// 50 us in a thread and integrates it with the trapezoid rule (include/nvmlClass.cuh:121-162,
// 271-355); an accumulating hardware counter needs no sampler: energy over a region is the
// difference of two reads.  libamd_smi is opened at run time so the library still loads (and
// reports EEG_ERR_UNSUPPORTED) where amd-smi is missing or not permitted.
#include <dlfcn.h>

#include <amd_smi/amdsmi.h>

#include <atomic>
#include <mutex>
#include <vector>

#include "eeg_internal.hpp"

namespace {

struct Smi {
    bool ok = false;
    std::vector<amdsmi_processor_handle> gpus;
    amdsmi_status_t (*energy)(amdsmi_processor_handle, uint64_t*, float*, uint64_t*) = nullptr;
};

Smi& smi() {
    static Smi s;
    static std::once_flag once;
    std::call_once(once, [] {
        void* h = dlopen("libamd_smi.so", RTLD_NOW | RTLD_LOCAL);
        if (!h) h = dlopen("/opt/rocm/lib/libamd_smi.so", RTLD_NOW | RTLD_LOCAL);
        if (!h) return;
        auto init = (amdsmi_status_t(*)(uint64_t))dlsym(h, "amdsmi_init");
        auto sockets = (amdsmi_status_t(*)(uint32_t*, amdsmi_socket_handle*))dlsym(h, "amdsmi_get_socket_handles");
        auto procs = (amdsmi_status_t(*)(amdsmi_socket_handle, uint32_t*, amdsmi_processor_handle*))dlsym(
            h, "amdsmi_get_processor_handles");
        auto ptype = (amdsmi_status_t(*)(amdsmi_processor_handle, processor_type_t*))dlsym(
            h, "amdsmi_get_processor_type");
        s.energy = (amdsmi_status_t(*)(amdsmi_processor_handle, uint64_t*, float*, uint64_t*))dlsym(
            h, "amdsmi_get_energy_count");
        if (!init || !sockets || !procs || !ptype || !s.energy) return;
        if (init(AMDSMI_INIT_AMD_GPUS) != AMDSMI_STATUS_SUCCESS) return;
        uint32_t ns = 0;
        if (sockets(&ns, nullptr) != AMDSMI_STATUS_SUCCESS || ns == 0) return;
        std::vector<amdsmi_socket_handle> sh(ns);
        if (sockets(&ns, sh.data()) != AMDSMI_STATUS_SUCCESS) return;
        for (uint32_t i = 0; i < ns; ++i) {
            uint32_t np = 0;
            if (procs(sh[i], &np, nullptr) != AMDSMI_STATUS_SUCCESS || np == 0) continue;
            std::vector<amdsmi_processor_handle> ph(np);
            if (procs(sh[i], &np, ph.data()) != AMDSMI_STATUS_SUCCESS) continue;
            for (uint32_t j = 0; j < np; ++j) {
                processor_type_t t;
                if (ptype(ph[j], &t) == AMDSMI_STATUS_SUCCESS && t == AMDSMI_PROCESSOR_TYPE_AMD_GPU)
                    s.gpus.push_back(ph[j]);
            }
        }
        s.ok = !s.gpus.empty();
    });
    return s;
}

}  // namespace

// Device index i maps to the i-th GPU amd-smi enumerates (one GPU per box here).
extern "C" eeg_status eeg_energy_read(int device, double* microjoules, uint64_t* timestamp_ns) {
    Smi& s = smi();
    if (!s.ok) return eeg::fail(EEG_ERR_UNSUPPORTED, "amd-smi energy counter unavailable");
    if (device < 0 || (size_t)device >= s.gpus.size())
        return eeg::fail(EEG_ERR_INVALID_ARG, "device %d has no amd-smi handle", device);
    uint64_t acc = 0, ts = 0;
    float res = 0;
    if (s.energy(s.gpus[device], &acc, &res, &ts) != AMDSMI_STATUS_SUCCESS)
        return eeg::fail(EEG_ERR_UNSUPPORTED, "amdsmi_get_energy_count failed");
    if (microjoules) *microjoules = (double)acc * (double)res;
    if (timestamp_ns) *timestamp_ns = ts;
    return EEG_OK;
}
