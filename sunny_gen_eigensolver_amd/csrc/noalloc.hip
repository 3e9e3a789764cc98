// sunny_geneig_launch_noalloc: sunny_geneig_launch for callers that cannot
// allocate device memory (code under hipGraph capture, tight serving loops,
// harnesses that forbid hipMalloc). The split, mixed and staged launches
// need a per-matrix workspace; here it comes from one file-scope device
// array, `scratch`, of kScratchBytes, which each GPU the library runs on
// holds until the process ends. No counterpart in the reference ABI.
//
// The array's name is a host-side stand-in in host code: kernels and HIP
// calls get the device address that hipGetSymbolAddress returns for the
// current device, resolved once per device when a handle is created there
// (noalloc_setup) and kept in g_dev[device].ptr.
//
// Every call on a GPU shares that GPU's array and leaves its contents
// behind; the kernels write every workspace element they read, so nothing
// needs clearing between calls. Calls on one GPU are serialised on the
// device: each waits (hipStreamWaitEvent) for the event the previous call
// recorded, whatever the stream or handle. A caller that bypasses that
// ordering (two processes, or its own copies into the array) would race.
#include "handles.hpp"

#include <hip/hip_runtime.h>

#include <mutex>
#include <stdexcept>
#include <string>

namespace {

constexpr size_t kScratchBytes = SUNNY_GENEIG_NOALLOC_SCRATCH_BYTES;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

struct DeviceScratch {
    bool tried = false;
    void* ptr = nullptr;          // device address of `scratch` on this device
    hipEvent_t last = nullptr;    // recorded after each call on this device
    std::string error;            // why set-up failed, if it did
};

std::mutex g_mu;
DeviceScratch g_dev[kMaxDevices];

}  // namespace

namespace sgeig {

void noalloc_setup(int device) {
    if (device < 0 || device >= kMaxDevices) return;
    std::lock_guard<std::mutex> lk(g_mu);
    DeviceScratch& d = g_dev[device];
    if (d.tried) return;
    d.tried = true;
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
        d.error = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
        return;
    }
    hipEvent_t ev = nullptr;
    e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
    if (e != hipSuccess) {
        d.error = std::string("hipEventCreateWithFlags: ") + hipGetErrorString(e);
        return;
    }
    d.ptr = p;
    d.last = ev;
}

}  // namespace sgeig

extern "C" sunny_geneig_status_t sunny_geneig_launch_noalloc(sunny_geneig_solver_t solver, uintptr_t d_H,
                                                             uintptr_t d_S, uintptr_t d_W, uintptr_t d_U,
                                                             uintptr_t d_info, int batch_size,
                                                             uintptr_t stream) {
    sgeig::api_clear_error();
    if (!solver) {
        sgeig::api_set_error("sunny_geneig_launch_noalloc: solver is NULL");
        return SUNNY_GENEIG_INVALID_ARG;
    }
    if (batch_size < 0) {
        sgeig::api_set_error("sunny_geneig_launch_noalloc: batch_size < 0");
        return SUNNY_GENEIG_INVALID_ARG;
    }
    if (batch_size > 0 && (!d_H || !d_S || !d_W || !d_U || !d_info)) {
        sgeig::api_set_error("sunny_geneig_launch_noalloc: NULL device buffer");
        return SUNNY_GENEIG_INVALID_ARG;
    }
    if (batch_size == 0) return SUNNY_GENEIG_OK;
    sgeig::Solver& s = *solver->impl;
    const int dev = s.device();
    if (dev < 0 || dev >= kMaxDevices) {
        sgeig::api_set_error("sunny_geneig_launch_noalloc: device id out of the supported range");
        return SUNNY_GENEIG_INVALID_ARG;
    }
    try {
        std::lock_guard<std::mutex> lk(g_mu);
        DeviceScratch& d = g_dev[dev];
        if (!d.ptr) {
            sgeig::api_set_error("sunny_geneig_launch_noalloc: no scratch on device " + std::to_string(dev) +
                                 (d.error.empty() ? std::string(" (not set up)") : ": " + d.error));
            return SUNNY_GENEIG_DEVICE_ERROR;
        }
        hipStream_t st = reinterpret_cast<hipStream_t>(stream);
        SG_HIP(hipStreamWaitEvent(st, d.last, 0));
        s.launch_in_scratch(reinterpret_cast<const void*>(d_H), reinterpret_cast<const void*>(d_S),
                            reinterpret_cast<double*>(d_W), reinterpret_cast<void*>(d_U),
                            reinterpret_cast<int*>(d_info), batch_size, st, d.ptr, kScratchBytes);
        SG_HIP(hipEventRecord(d.last, st));
        return SUNNY_GENEIG_OK;
    } catch (const std::invalid_argument& e) {
        sgeig::api_set_error(std::string("sunny_geneig_launch_noalloc: ") + e.what());
        return SUNNY_GENEIG_INVALID_ARG;
    } catch (const std::exception& e) {
        sgeig::api_set_error(std::string("sunny_geneig_launch_noalloc: ") + e.what());
    } catch (...) {
        sgeig::api_set_error("unknown C++ exception in sunny_geneig_launch_noalloc");
    }
    return SUNNY_GENEIG_LAUNCH_ERROR;
}
