// sm_internal.hpp -- shared internals of libseismod_hip.so (not installed).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <string>

#include "seismod_hip.h"

namespace sm {

// Records a message for sm_last_error() and returns the status.
int fail(int status, const char *fmt, ...);

// Check a HIP call; on error record it and return SM_ERR_HIP from the
// enclosing function.
#define SM_HIP(call)                                                        \
    do {                                                                    \
        hipError_t e_ = (call);                                             \
        if (e_ != hipSuccess)                                               \
            return ::sm::fail(SM_ERR_HIP, "%s: %s (%s:%d)", #call,          \
                              hipGetErrorString(e_), __FILE__, __LINE__);   \
    } while (0)

#define SM_CHECK(cond, status, ...)                                         \
    do {                                                                    \
        if (!(cond)) return ::sm::fail((status), __VA_ARGS__);              \
    } while (0)

}  // namespace sm

struct sm_ctx {
    int device;
    hipStream_t stream;
};

// Device buffer owned by a solver; freed with hipFree in the destructor of
// the solver (explicit release, not RAII, so errors can be reported).
struct sm_dbuf {
    void *ptr = nullptr;
    size_t bytes = 0;
};

namespace sm {
int dalloc(sm_dbuf &b, size_t bytes);   // hipMalloc + zero on the stream
void dfree(sm_dbuf &b);
int set_device(const sm_ctx *ctx);
}  // namespace sm
