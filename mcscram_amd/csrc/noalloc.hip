// Allocation-free probability analysis: the engine's device memory (words
// matrix, descriptors, argument rows, one-bit counters) is carved from one
// fixed, file-scope device array instead of hipMalloc, and its launches go
// to the caller's stream. For callers that must not allocate device memory
// (serving loops that forbid hipMalloc, harnesses with a fixed footprint).
//
// The array, `scratch`, is a per-device global of the code object: every GPU
// the library runs this on holds MCS_NOALLOC_SCRATCH_BYTES until the process
// ends. In host code its name is a host-side stand-in, so its device address
// is looked up once per device with hipGetSymbolAddress, with that device
// current, and only that address is ever handed to the engine.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "mcscram_amd.h"
#include "status.h"

static __device__ __align__(256) unsigned long long scratch[MCS_NOALLOC_SCRATCH_BYTES / 8];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_mutex;
void *g_scratch_ptr[kMaxDevices] = {};

// Device address of this device's copy of `scratch` (looked up on the first
// call with the device current, then kept).
mcs_status scratch_address(int device, void **out) {
    std::lock_guard<std::mutex> lock(g_mutex);
    if (!g_scratch_ptr[device]) {
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return mcs::fail(MCS_ERR_DEVICE, std::string("hipGetSymbolAddress(scratch) failed: ") +
                                                 hipGetErrorString(e));
        g_scratch_ptr[device] = ptr;
    }
    *out = g_scratch_ptr[device];
    return MCS_OK;
}

}  // namespace

extern "C" mcs_status mcs_probability_analysis_noalloc(
    const mcs_pdag *pdag, const mcs_settings *settings, const mcs_sample_shape *shape,
    const int32_t *observe, int32_t n_observe, mcs_tally *out, int32_t n_out, double *p_total,
    uint32_t *passes_run, void *stream) {
    int device = 0;
    MCS_HIP(hipGetDevice(&device));
    if (device < 0 || device >= kMaxDevices)
        return mcs::fail(MCS_ERR_DEVICE, "device ordinal " + std::to_string(device) +
                                             " beyond the scratch table's " +
                                             std::to_string(kMaxDevices));
    void *base = nullptr;
    const mcs_status st = scratch_address(device, &base);
    if (st != MCS_OK) return st;
    const mcs::ExternalMemory ext{base, uint64_t(MCS_NOALLOC_SCRATCH_BYTES),
                                  static_cast<hipStream_t>(stream),
                                  "the fixed no-allocation scratch"};
    return mcs::probability_analysis(pdag, settings, shape, observe, n_observe, out, n_out,
                                     p_total, passes_run, &ext);
}
