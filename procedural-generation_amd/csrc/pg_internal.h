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

// The serial change-list apply (one lane, droplet order) enqueued behind a
// device gate: it runs only if *gate != 0 when it starts (NULL: always),
// writes each droplet's (clamped, to pending, deficit lost, lost offshore) to
// drop[4 p ...] and, if `count` is given, adds 1 to it when it runs.
// Defined in pg_check.hip.
int enqueue_serial_apply(const int32_t* cl_cell, const double* cl_delta, const float* cl_vol, const float* cl_mom,
                         const int64_t* cl_count, int64_t P, int64_t cap, const double* height, const double* sediment,
                         double* acc, double* pending, float* samp, double* disch_track, double* mom_track,
                         const uint8_t* mask, double iter_erode, double iter_deposit, double fan_slope, int use_route,
                         double dep_floor, double offshore_writeoff, const double* erode_cap, int use_erode_cap,
                         double* lake_load, const int32_t* lake_id, double* lake_room, const uint32_t* gate,
                         double* drop, int32_t* count, hipStream_t stream);

}  // namespace pg
