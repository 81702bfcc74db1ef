// Error reporting shared by the C-ABI translation units: entry points return
// FX_* codes and leave a message for fx_last_error() (thread-local).
#pragma once

#include <hip/hip_runtime.h>

#include <string>

namespace fx_api {

// sets the message of fx_last_error() and returns code
int fail(int code, const std::string& msg);
// FX_EHIP with the failing call and the HIP error string
int hip_fail(hipError_t e, const char* where);
// the reference's abort message of a device error code
const char* phys_message(int code);

}  // namespace fx_api

#define FX_HIP(call)                                                   \
    do {                                                               \
        hipError_t e_ = (call);                                        \
        if (e_ != hipSuccess) return fx_api::hip_fail(e_, #call);      \
    } while (0)
