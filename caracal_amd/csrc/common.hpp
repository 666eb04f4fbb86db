// Shared host-side helpers of the C ABI: thread-local error text and HIP
// status checking that turns failures into CAM_ERR_* returns.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdarg>
#include <cstdio>
#include <string>
#include "caracal_amd.h"

namespace cam {

void set_error(const char *fmt, ...);
void clear_error();

#define CAM_CHECK_ARG(cond, ...)                 \
    do {                                         \
        if (!(cond)) {                           \
            ::cam::set_error(__VA_ARGS__);       \
            return CAM_ERR_ARG;                  \
        }                                        \
    } while (0)

#define CAM_HIP(call)                                                              \
    do {                                                                           \
        hipError_t e_ = (call);                                                    \
        if (e_ != hipSuccess) {                                                    \
            ::cam::set_error("%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), \
                             __FILE__, __LINE__);                                  \
            return e_ == hipErrorOutOfMemory ? CAM_ERR_NOMEM : CAM_ERR_DEVICE;     \
        }                                                                          \
    } while (0)

// Launch-error check after a kernel launch.
#define CAM_LAUNCH_CHECK() CAM_HIP(hipGetLastError())

}  // namespace cam
