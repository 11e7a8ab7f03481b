// Device energy through amd-smi's energy accumulator.  The reference samples NVML power every
// 50 us in a thread and integrates it with the trapezoid rule (include/nvmlClass.cuh:121-162,
// 271-355); an accumulating hardware counter needs no sampler: energy over a region is the
// difference of two reads.  libamd_smi is opened at run time so the library still loads (and
// reports EEG_ERR_UNSUPPORTED) where amd-smi is missing or not permitted.
#include <dlfcn.h>

#define __HIP_PLATFORM_AMD__ 1
#include <hip/hip_runtime_api.h>

#include <amd_smi/amdsmi.h>

#include <atomic>
#include <cstdio>
#include <mutex>
#include <vector>

#include "eeg_internal.hpp"

namespace {

struct Smi {
    bool ok = false;
    std::vector<amdsmi_processor_handle> gpus;
    std::vector<uint64_t> bdf;  // PCI address of each handle: domain << 16 | bus << 8 | dev << 3 | fn
    amdsmi_status_t (*energy)(amdsmi_processor_handle, uint64_t*, float*, uint64_t*) = nullptr;
    std::mutex mu;
    std::vector<int> handle_of;  // HIP device -> index into gpus (-2 = not looked up, -1 = none)
};

uint64_t pack_bdf(uint64_t domain, uint64_t bus, uint64_t dev, uint64_t fn) {
    return (domain << 16) | ((bus & 0xff) << 8) | ((dev & 0x1f) << 3) | (fn & 0x7);
}

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
        auto getbdf = (amdsmi_status_t(*)(amdsmi_processor_handle, amdsmi_bdf_t*))dlsym(
            h, "amdsmi_get_gpu_device_bdf");
        s.energy = (amdsmi_status_t(*)(amdsmi_processor_handle, uint64_t*, float*, uint64_t*))dlsym(
            h, "amdsmi_get_energy_count");
        if (!init || !sockets || !procs || !ptype || !getbdf || !s.energy) return;
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
                amdsmi_bdf_t b;
                if (ptype(ph[j], &t) == AMDSMI_STATUS_SUCCESS && t == AMDSMI_PROCESSOR_TYPE_AMD_GPU &&
                    getbdf(ph[j], &b) == AMDSMI_STATUS_SUCCESS) {
                    s.gpus.push_back(ph[j]);
                    s.bdf.push_back(pack_bdf(b.domain_number, b.bus_number, b.device_number, b.function_number));
                }
            }
        }
        s.ok = !s.gpus.empty();
    });
    return s;
}

// The amd-smi handle of a HIP device, matched by PCI address: HIP's device numbering follows
// HIP_VISIBLE_DEVICES / ROCR_VISIBLE_DEVICES, amd-smi enumerates every GPU of the machine.
int handle_index(Smi& s, int device) {
    std::lock_guard<std::mutex> lk(s.mu);
    if (device >= (int)s.handle_of.size()) s.handle_of.resize((size_t)device + 1, -2);
    int& h = s.handle_of[(size_t)device];
    if (h != -2) return h;
    h = -1;
    char id[64] = {0};
    if (hipDeviceGetPCIBusId(id, (int)sizeof(id), device) != hipSuccess) {
        (void)hipGetLastError();
        return h;
    }
    unsigned domain = 0, bus = 0, dev = 0, fn = 0;
    if (sscanf(id, "%x:%x:%x.%x", &domain, &bus, &dev, &fn) != 4) return h;
    const uint64_t want = pack_bdf(domain, bus, dev, fn);
    for (size_t i = 0; i < s.bdf.size(); ++i)
        if (s.bdf[i] == want) h = (int)i;
    return h;
}

}  // namespace

extern "C" eeg_status eeg_energy_read(int device, double* microjoules, uint64_t* timestamp_ns) {
    Smi& s = smi();
    if (!s.ok) return eeg::fail(EEG_ERR_UNSUPPORTED, "amd-smi energy counter unavailable");
    if (device < 0) return eeg::fail(EEG_ERR_INVALID_ARG, "device %d out of range", device);
    const int hidx = handle_index(s, device);
    if (hidx < 0)
        return eeg::fail(EEG_ERR_UNSUPPORTED, "no amd-smi GPU has the PCI address of HIP device %d", device);
    uint64_t acc = 0, ts = 0;
    float res = 0;
    if (s.energy(s.gpus[(size_t)hidx], &acc, &res, &ts) != AMDSMI_STATUS_SUCCESS)
        return eeg::fail(EEG_ERR_UNSUPPORTED, "amdsmi_get_energy_count failed");
    if (microjoules) *microjoules = (double)acc * (double)res;
    if (timestamp_ns) *timestamp_ns = ts;
    return EEG_OK;
}
