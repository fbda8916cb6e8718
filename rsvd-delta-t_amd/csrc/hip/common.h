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

// Workgroup -> row-block map.  Dispatch hands workgroup i to XCD i % 8.
// mode 0: identity.  mode 1: XCD x gets one contiguous range of row blocks
// (its L2 sees the stencil neighbours, but the 8 XCDs sweep 8 distant parts
// of the vector at once).  modes 2, 3: XCD x gets chunks x, x+8, x+16, ... of
// 64 (mode 2) or 256 (mode 3) row blocks, so all XCDs sweep the same region
// of the vector together (one reuse window in the shared last-level cache)
// while each chunk's near neighbours stay in one L2; the tail that does not
// fill a whole round of 8 chunks keeps the identity map.
__device__ __forceinline__ int xcd_block(int bx, int nb, int mode)
{
    if (mode == 1) {
        const int q = nb >> 3, r = nb & 7, x = bx & 7, idx = bx >> 3;
        return x * q + (x < r ? x : r) + idx;
    }
    if (mode >= 2) {
        const int ch = mode == 2 ? 64 : 256;
        const int full = nb / (8 * ch);
        if (bx < full * 8 * ch) {
            const int x = bx & 7, idx = bx >> 3;
            return ((idx / ch) * 8 + x) * ch + idx % ch;
        }
    }
    return bx;
}

static inline hipStream_t as_stream(rsvdt_stream_t s) { return (hipStream_t)s; }
static inline unsigned ceil_div(int64_t a, int64_t b) { return (unsigned)((a + b - 1) / b); }

// The Horner stage on the tiled operator form (rsvdt_spmm_desc variant 9),
// in tile.hip.
int horner_tile(const rsvdt_spmm_desc *d, const void *vin, const void *y, const double *ys, const void *g, double c,
                void *out, double *snap, const int *slot, int64_t sstride, hipStream_t s);

}  // namespace rsvdt
