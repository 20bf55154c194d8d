// common.hpp - error reporting shared by the library's translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>

#include "../../include/imago_hip.h"

// Records a formatted message for imago_last_error() and returns code.
int imago_fail(int code, const char *fmt, ...);

#define fail imago_fail

#define HIPCHK(x)                                                                       \
    do {                                                                                \
        hipError_t e_ = (x);                                                            \
        if (e_ != hipSuccess)                                                           \
            return fail(e_ == hipErrorOutOfMemory ? IMAGO_E_NOMEM : IMAGO_E_HIP,         \
                        "%s failed: %s (%s:%d)", #x, hipGetErrorString(e_), __FILE__, \
                        __LINE__);                                                      \
    } while (0)

#define BLASCHK(x)                                                                 \
    do {                                                                           \
        rocblas_status s_ = (x);                                                   \
        if (s_ != rocblas_status_success)                                          \
            return fail(IMAGO_E_HIP, "%s failed: rocblas status %d (%s:%d)", #x,  \
                        (int)s_, __FILE__, __LINE__);                              \
    } while (0)

#define RCHK(x)                        \
    do {                               \
        int r_ = (x);                  \
        if (r_ != IMAGO_OK) return r_; \
    } while (0)
