// Error reporting shared by the engine's translation units: every C ABI
// call returns a status and leaves a thread-local message.
#pragma once

#include <hip/hip_runtime_api.h>

#include <string>

#include "mcscram_amd.h"

namespace mcs {

mcs_status fail(mcs_status code, std::string msg);  // engine.cpp
void set_error(const char *msg);

}  // namespace mcs

#define MCS_HIP(expr)                                                                        \
    do {                                                                                     \
        const hipError_t e_ = (expr);                                                        \
        if (e_ != hipSuccess)                                                                \
            return ::mcs::fail(e_ == hipErrorOutOfMemory ? MCS_ERR_OUT_OF_MEMORY : MCS_ERR_DEVICE, \
                               std::string(#expr " failed: ") + hipGetErrorString(e_));      \
    } while (0)
