// Internal helpers of the C ABI: per-thread error message, argument checks,
// launch-error capture.  Not part of the public header.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>

namespace pg {

void set_error(const char* fmt, ...);
void clear_error();

// Returns -1 from the calling entry point with a message when `cond` fails.
#define PG_REQUIRE(cond, ...)              \
    do {                                   \
        if (!(cond)) {                     \
            ::pg::set_error(__VA_ARGS__);  \
            return -1;                     \
        }                                  \
    } while (0)

// Checks a HIP runtime status.
#define PG_HIP(call)                                                                             \
    do {                                                                                         \
        hipError_t e_ = (call);                                                                  \
        if (e_ != hipSuccess) {                                                                  \
            ::pg::set_error("%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), __FILE__, __LINE__); \
            return -1;                                                                           \
        }                                                                                        \
    } while (0)

// Checks that the last kernel launch was accepted.
#define PG_LAUNCHED(name)                                                                        \
    do {                                                                                         \
        hipError_t e_ = hipGetLastError();                                                       \
        if (e_ != hipSuccess) {                                                                  \
            ::pg::set_error("launch of %s failed: %s", name, hipGetErrorString(e_));            \
            return -1;                                                                           \
        }                                                                                        \
    } while (0)

inline int blocks_for(int64_t n, int threads) { return (int)((n + threads - 1) / threads); }

}  // namespace pg
