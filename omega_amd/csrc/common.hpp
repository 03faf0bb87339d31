// common.hpp -- error plumbing and small helpers shared by the library's translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>
#include <cstdio>
#include <cstdarg>

#include "../../include/omega_hip.h"

namespace omg {

void set_error(const char* fmt, ...);

#define OMG_CHECK_HIP(call)                                                                          \
    do {                                                                                             \
        hipError_t e_ = (call);                                                                      \
        if (e_ != hipSuccess) {                                                                      \
            ::omg::set_error("%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), __FILE__, __LINE__); \
            return OMG_ERR_HIP;                                                                      \
        }                                                                                            \
    } while (0)

#define OMG_CHECK_LAUNCH()                                                                           \
    do {                                                                                             \
        hipError_t e_ = hipGetLastError();                                                           \
        if (e_ != hipSuccess) {                                                                      \
            ::omg::set_error("kernel launch failed: %s (%s:%d)", hipGetErrorString(e_), __FILE__, __LINE__); \
            return OMG_ERR_HIP;                                                                      \
        }                                                                                            \
    } while (0)

#define OMG_REQUIRE(cond, ...)                                                                       \
    do {                                                                                             \
        if (!(cond)) {                                                                               \
            ::omg::set_error(__VA_ARGS__);                                                           \
            return OMG_ERR_ARG;                                                                      \
        }                                                                                            \
    } while (0)

static inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

static inline unsigned grid_for(int64_t n, int block) { return (unsigned)((n + block - 1) / block); }

}  // namespace omg
