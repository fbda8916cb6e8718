// Device-side helpers shared by the gfx950 kernels: complex arithmetic on
// interleaved (re, im) pairs, error plumbing for the C ABI.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "rsvdt_hip.h"

namespace rsvdt {

void set_error(const char *where, hipError_t e);
void set_error_msg(const char *msg);

#define RSVDT_HIP_TRY(call)                                      \
    do {                                                         \
        hipError_t _e = (call);                                  \
        if (_e != hipSuccess) {                                  \
            rsvdt::set_error(#call, _e);                         \
            return RSVDT_HIP_ERR_RUNTIME;                        \
        }                                                        \
    } while (0)

#define RSVDT_LAUNCH_CHECK(name)                                 \
    do {                                                         \
        hipError_t _e = hipGetLastError();                       \
        if (_e != hipSuccess) {                                  \
            rsvdt::set_error(name, _e);                          \
            return RSVDT_HIP_ERR_RUNTIME;                        \
        }                                                        \
    } while (0)

template <typename T> struct vec2;
template <> struct vec2<double> { using type = double2; };
template <> struct vec2<float> { using type = float2; };

__device__ __forceinline__ double2 zmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// conj(a) * b
__device__ __forceinline__ double2 zcmul(double2 a, double2 b) {
    return make_double2(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ double2 zadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 zsub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double zabs2(double2 a) { return a.x * a.x + a.y * a.y; }

static inline hipStream_t as_stream(rsvdt_stream_t s) { return (hipStream_t)s; }
static inline unsigned ceil_div(int64_t a, int64_t b) { return (unsigned)((a + b - 1) / b); }

}  // namespace rsvdt
