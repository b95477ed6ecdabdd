// Error reporting shared by the engine's translation units: every C ABI
// call returns a status and leaves a thread-local message.
#pragma once

#include <hip/hip_runtime_api.h>

#include <string>

#include "mcscram_amd.h"

namespace mcs {

mcs_status fail(mcs_status code, std::string msg);  // engine.cpp
void set_error(const char *msg);

// Device memory and stream an engine uses instead of allocating its own:
// its buffers are carved from [base, base + bytes) and it launches on
// stream. The engine never frees either.
struct ExternalMemory {
    void *base;
    uint64_t bytes;
    hipStream_t stream;
    const char *name;  // how error messages call the region
};

// mcs_engine_create_ex with the engine's device memory taken from ext when
// ext is not null (engine.cpp).
mcs_status create_engine(const mcs_pdag *pdag, const mcs_sample_shape *shape,
                         const int32_t *tally_index, int32_t n_tally, uint32_t key0,
                         uint32_t key1, int32_t device, const ExternalMemory *ext,
                         mcs_engine **out);

// mcs_probability_analysis with the engine built over ext when it is not
// null (analysis.cpp).
mcs_status probability_analysis(const mcs_pdag *pdag, const mcs_settings *settings,
                                const mcs_sample_shape *shape, const int32_t *observe,
                                int32_t n_observe, mcs_tally *out, int32_t n_out,
                                double *p_total, uint32_t *passes_run,
                                const ExternalMemory *ext);

}  // namespace mcs

#define MCS_HIP(expr)                                                                        \
    do {                                                                                     \
        const hipError_t e_ = (expr);                                                        \
        if (e_ != hipSuccess)                                                                \
            return ::mcs::fail(e_ == hipErrorOutOfMemory ? MCS_ERR_OUT_OF_MEMORY : MCS_ERR_DEVICE, \
                               std::string(#expr " failed: ") + hipGetErrorString(e_));      \
    } while (0)
