// Runtime part of the thin C ABI: device selection, memory, events, and the
// small elementwise kernels (fills, conversions, strided copies, norms).
#include <stdio.h>
#include <string.h>
#include <mutex>
#include <unordered_map>
#include "common.h"

namespace rsvdt {
static thread_local char g_err[512] = "";
void set_error(const char *where, hipError_t e) {
    snprintf(g_err, sizeof(g_err), "%s: %s (%d)", where, hipGetErrorString(e), (int)e);
}
void set_error_msg(const char *msg) { snprintf(g_err, sizeof(g_err), "%s", msg); }
}  // namespace rsvdt

using namespace rsvdt;

extern "C" const char *rsvdt_hip_last_error(void) { return g_err; }

extern "C" int rsvdt_noalloc_setup(int device);

extern "C" int rsvdt_hip_init(int device) {
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n == 0) {
        set_error_msg("rsvdt_hip_init: no HIP device visible");
        return RSVDT_HIP_ERR_NODEV;
    }
    if (device < 0 || device >= n) {
        set_error_msg("rsvdt_hip_init: device index out of range");
        return RSVDT_HIP_ERR_ARG;
    }
    RSVDT_HIP_TRY(hipSetDevice(device));
    hipDeviceProp_t p;
    RSVDT_HIP_TRY(hipGetDeviceProperties(&p, device));
    if (strncmp(p.gcnArchName, "gfx950", 6) != 0) {
        char msg[256];
        snprintf(msg, sizeof(msg), "rsvdt_hip_init: device is %s, this build targets gfx950 only", p.gcnArchName);
        set_error_msg(msg);
        return RSVDT_HIP_ERR_NODEV;
    }
    return rsvdt_noalloc_setup(device);     // device address of the no-allocation entry point's scratch
}

// Device allocations keep a small cache of freed blocks.  hipFree of a block
// costs 0.3-0.5 ms (the unmapping; it also waits for the device), and an
// action frees a dozen work blocks, which on small operators is as long as
// its time stepping.  A freed block of at most 256 MB is kept instead, after
// the same wait for the device that hipFree performs (so no queued work on any
// stream can still use it), and handed to a later request of its size up to
// 2x.  At most 2 GB (64 blocks) is held; a failing hipMalloc empties the cache
// and tries again; rsvdt_hip_release_cached frees it all (RSVDT_Finalize
// calls it), and rsvdt_hip_mem_info counts it as free.
namespace {
struct cached_blk {
    void *p;
    size_t bytes;
};
constexpr size_t kCacheMaxBlock = (size_t)256 << 20;
constexpr size_t kCacheMaxTotal = (size_t)2 << 30;
constexpr int kCacheSlots = 64;
std::mutex g_alloc_mu;
cached_blk g_cache[kCacheSlots];
int g_ncache = 0;
size_t g_cache_bytes = 0;
std::unordered_map<void *, size_t> g_live;     // this allocator's blocks and their rounded sizes

size_t round_alloc(size_t b) { return (b + 65535) & ~(size_t)65535; }

// with g_alloc_mu held
void cache_release_locked() {
    for (int i = 0; i < g_ncache; ++i) (void)hipFree(g_cache[i].p);
    g_ncache = 0;
    g_cache_bytes = 0;
}
}  // namespace

extern "C" int rsvdt_hip_malloc(void **p, size_t bytes) {
    *p = nullptr;
    const size_t r = round_alloc(bytes ? bytes : 256);
    std::lock_guard<std::mutex> lk(g_alloc_mu);
    if (r <= kCacheMaxBlock) {
        int best = -1;
        for (int i = 0; i < g_ncache; ++i)
            if (g_cache[i].bytes >= r && g_cache[i].bytes <= 2 * r && (best < 0 || g_cache[i].bytes < g_cache[best].bytes))
                best = i;
        if (best >= 0) {
            *p = g_cache[best].p;
            g_live[*p] = g_cache[best].bytes;
            g_cache_bytes -= g_cache[best].bytes;
            g_cache[best] = g_cache[--g_ncache];
            return RSVDT_HIP_OK;
        }
    }
    hipError_t e = hipMalloc(p, r);
    if (e != hipSuccess && g_ncache > 0) {
        (void)hipGetLastError();
        cache_release_locked();
        e = hipMalloc(p, r);
    }
    if (e != hipSuccess) {
        *p = nullptr;
        set_error("hipMalloc", e);
        return RSVDT_HIP_ERR_RUNTIME;
    }
    g_live[*p] = r;
    return RSVDT_HIP_OK;
}
extern "C" int rsvdt_hip_free(void *p) {
    if (!p) return RSVDT_HIP_OK;
    std::lock_guard<std::mutex> lk(g_alloc_mu);
    auto it = g_live.find(p);
    const size_t sz = it != g_live.end() ? it->second : 0;
    if (it != g_live.end()) g_live.erase(it);
    if (sz && sz <= kCacheMaxBlock && g_ncache < kCacheSlots && g_cache_bytes + sz <= kCacheMaxTotal) {
        RSVDT_HIP_TRY(hipDeviceSynchronize());
        g_cache[g_ncache++] = cached_blk{p, sz};
        g_cache_bytes += sz;
        return RSVDT_HIP_OK;
    }
    RSVDT_HIP_TRY(hipFree(p));
    return RSVDT_HIP_OK;
}
extern "C" int rsvdt_hip_release_cached(void) {
    std::lock_guard<std::mutex> lk(g_alloc_mu);
    RSVDT_HIP_TRY(hipDeviceSynchronize());
    cache_release_locked();
    return RSVDT_HIP_OK;
}
extern "C" size_t rsvdt_hip_cached_bytes(void) {
    std::lock_guard<std::mutex> lk(g_alloc_mu);
    return g_cache_bytes;
}
// The iDFT table of CreateDFTiDFTMats: entry (m, j), m < nwe, j < nt2, at
// out[j nwe + m] = conj(exp(alpha m_alias j)) 2/nt2, alpha = -2 pi i/nt2,
// m_alias = m for a real operator, else m below nwe/2 (nwe even; (nwe-1)/2
// inclusive when odd) and m - nwe + nt2 above.  The phase index
// (m_alias j) mod nt2 is exact in integers, so the angle never exceeds 2 pi.
__global__ void k_idft_table(int64_t nwe, int64_t nt2, int real_op, double2 *__restrict__ out)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nwe * nt2) return;
    const int64_t j = e / nwe, m = e - j * nwe;
    int64_t a = m;
    if (!real_op && m > (nwe % 2 == 0 ? nwe / 2 - 1 : (nwe - 1) / 2)) a = m - nwe + nt2;
    a %= nt2;
    const int64_t p = (a * j) % nt2;
    const double ph = (-2.0 * 3.14159265358979323846 / (double)nt2) * (double)p;
    const double sc = 2.0 / (double)nt2;
    out[e] = make_double2(cos(ph) * sc, -sin(ph) * sc);
}
extern "C" int rsvdt_hip_idft_table(int64_t nwe, int64_t nt2, int real_op, double *out, rsvdt_stream_t s)
{
    if (nwe <= 0 || nt2 <= 0 || !out || nt2 > 0x7fffffff) {
        set_error_msg("rsvdt_hip_idft_table: bad arguments");
        return RSVDT_HIP_ERR_ARG;
    }
    hipLaunchKernelGGL(k_idft_table, dim3(ceil_div(nwe * nt2, 256)), dim3(256), 0, as_stream(s), nwe, nt2, real_op,
                       (double2 *)out);
    RSVDT_LAUNCH_CHECK("rsvdt_hip_idft_table");
    return RSVDT_HIP_OK;
}

// Learned sketch columns from the network's FP32 output (interleaved complex,
// point-major, ld floats per point) into the first ncols columns of the
// sketch's frequency blocks (N x k row-major complex FP64 each).
__global__ void k_sketch_cols_f32(int64_t total, int ncols, int64_t k, const float *__restrict__ src, int64_t ld,
                                  double2 *__restrict__ Y)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;      // point * ncols + column
    if (e >= total) return;
    const int64_t pt = e / ncols;
    const int j = (int)(e - pt * ncols);
    Y[pt * k + j] = make_double2((double)src[pt * ld + 2 * j], (double)src[pt * ld + 2 * j + 1]);
}
extern "C" int rsvdt_hip_sketch_cols_f32(int64_t nwe, int64_t N, int64_t ncols, int64_t k, const float *src,
                                         int64_t ld, double *Y, rsvdt_stream_t s)
{
    if (nwe <= 0 || N <= 0 || ncols <= 0) return RSVDT_HIP_OK;
    if (!src || !Y || ncols > k || ld < 2 * ncols || ncols > 0x7fffffff) {
        set_error_msg("rsvdt_hip_sketch_cols_f32: bad arguments");
        return RSVDT_HIP_ERR_ARG;
    }
    const int64_t total = nwe * N * ncols;
    hipLaunchKernelGGL(k_sketch_cols_f32, dim3(ceil_div(total, 256)), dim3(256), 0, as_stream(s), total, (int)ncols,
                       k, src, ld, (double2 *)Y);
    RSVDT_LAUNCH_CHECK("rsvdt_hip_sketch_cols_f32");
    return RSVDT_HIP_OK;
}
// Host<->device copies first wait for the whole device: the engine's work runs
// on non-blocking streams, which the null-stream hipMemcpy does not order with.
extern "C" int rsvdt_hip_memcpy_h2d(void *dst, const void *src, size_t bytes) {
    RSVDT_HIP_TRY(hipDeviceSynchronize());
    if (bytes) RSVDT_HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return RSVDT_HIP_OK;
}
extern "C" int rsvdt_hip_memcpy_d2h(void *dst, const void *src, size_t bytes) {
    RSVDT_HIP_TRY(hipDeviceSynchronize());
    if (bytes) RSVDT_HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return RSVDT_HIP_OK;
}
extern "C" int rsvdt_hip_memcpy_d2d(void *dst, const void *src, size_t bytes, rsvdt_stream_t s) {
    if (bytes) RSVDT_HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, as_stream(s)));
    return RSVDT_HIP_OK;
}
extern "C" int rsvdt_hip_sync(void) {
    RSVDT_HIP_TRY(hipDeviceSynchronize());
    return RSVDT_HIP_OK;
}
extern "C" int rsvdt_hip_mem_info(size_t *f, size_t *t) {
    RSVDT_HIP_TRY(hipMemGetInfo(f, t));
    std::lock_guard<std::mutex> lk(g_alloc_mu);
    *f += g_cache_bytes;                   // cached blocks are given back on demand
    if (*f > *t) *f = *t;
    return RSVDT_HIP_OK;
}
extern "C" int rsvdt_hip_stream_create(rsvdt_stream_t *s) {
    hipStream_t h;
    RSVDT_HIP_TRY(hipStreamCreateWithFlags(&h, hipStreamNonBlocking));   // default priority
    *s = (rsvdt_stream_t)h;
    return RSVDT_HIP_OK;
}
extern "C" int rsvdt_hip_stream_destroy(rsvdt_stream_t s) {
    if (s) RSVDT_HIP_TRY(hipStreamDestroy((hipStream_t)s));
    return RSVDT_HIP_OK;
}
extern "C" int rsvdt_hip_stream_sync(rsvdt_stream_t s) {
    RSVDT_HIP_TRY(hipStreamSynchronize((hipStream_t)s));
    return RSVDT_HIP_OK;
}
// Pinned host memory (hipHostMalloc) for streamed device-to-host transfers.
extern "C" int rsvdt_hip_host_alloc(void **p, size_t bytes) {
    *p = nullptr;
    RSVDT_HIP_TRY(hipHostMalloc(p, bytes ? bytes : 256, hipHostMallocDefault));
    return RSVDT_HIP_OK;
}
extern "C" int rsvdt_hip_host_free(void *p) {
    if (p) RSVDT_HIP_TRY(hipHostFree(p));
    return RSVDT_HIP_OK;
}
// Asynchronous copy into pinned host memory, ordered on stream s.
extern "C" int rsvdt_hip_memcpy_d2h_async(void *dst_host, const void *src, size_t bytes, rsvdt_stream_t s) {
    if (bytes) RSVDT_HIP_TRY(hipMemcpyAsync(dst_host, src, bytes, hipMemcpyDeviceToHost, (hipStream_t)s));
    return RSVDT_HIP_OK;
}
extern "C" int rsvdt_hip_stream_wait_event(rsvdt_stream_t s, void *ev) {
    RSVDT_HIP_TRY(hipStreamWaitEvent((hipStream_t)s, (hipEvent_t)ev, 0));
    return RSVDT_HIP_OK;
}
extern "C" int rsvdt_hip_event_create(void **ev) {
    hipEvent_t e;
    RSVDT_HIP_TRY(hipEventCreate(&e));
    *ev = (void *)e;
    return RSVDT_HIP_OK;
}
extern "C" int rsvdt_hip_event_destroy(void *ev) {
    if (ev) RSVDT_HIP_TRY(hipEventDestroy((hipEvent_t)ev));
    return RSVDT_HIP_OK;
}
extern "C" int rsvdt_hip_event_record(void *ev, rsvdt_stream_t s) {
    RSVDT_HIP_TRY(hipEventRecord((hipEvent_t)ev, as_stream(s)));
    return RSVDT_HIP_OK;
}
extern "C" int rsvdt_hip_event_elapsed_ms(void *a, void *b, float *ms) {
    RSVDT_HIP_TRY(hipEventSynchronize((hipEvent_t)b));
    RSVDT_HIP_TRY(hipEventElapsedTime(ms, (hipEvent_t)a, (hipEvent_t)b));
    return RSVDT_HIP_OK;
}

// ---------------------------------------------------------------- fills
__global__ void k_zero16(uint4 *p, int64_t n16) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (; i < n16; i += stride) p[i] = make_uint4(0, 0, 0, 0);
}
__global__ void k_zero1(unsigned char *p, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}
extern "C" int rsvdt_hip_zero(void *p, size_t bytes, rsvdt_stream_t s) {
    if (!bytes) return RSVDT_HIP_OK;
    uintptr_t a = (uintptr_t)p;
    if ((a & 15) == 0 && (bytes & 15) == 0) {
        int64_t n16 = (int64_t)(bytes / 16);
        unsigned g = (unsigned)((n16 + 255) / 256);
        if (g > 65536u) g = 65536u;
        hipLaunchKernelGGL(k_zero16, dim3(g), dim3(256), 0, as_stream(s), (uint4 *)p, n16);
    } else {
        hipLaunchKernelGGL(k_zero1, dim3(ceil_div((int64_t)bytes, 256)), dim3(256), 0, as_stream(s),
                           (unsigned char *)p, (int64_t)bytes);
    }
    RSVDT_LAUNCH_CHECK("rsvdt_hip_zero");
    return RSVDT_HIP_OK;
}

// ---------------------------------------------------------------- elementwise
__global__ void k_zscale(int64_t n, double a, const double2 *x, double2 *y) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) { double2 v = x[i]; y[i] = make_double2(a * v.x, a * v.y); }
}
extern "C" int rsvdt_hip_zscale(int64_t n, double a, const double *x, double *y, rsvdt_stream_t s) {
    if (n <= 0) return RSVDT_HIP_OK;
    unsigned g = ceil_div(n, 256); if (g > 65536u) g = 65536u;
    hipLaunchKernelGGL(k_zscale, dim3(g), dim3(256), 0, as_stream(s), n, a, (const double2 *)x, (double2 *)y);
    RSVDT_LAUNCH_CHECK("rsvdt_hip_zscale");
    return RSVDT_HIP_OK;
}
// y += a x: one complex entry per thread with nontemporal loads and stores
// (two streams read, one written, nothing reused) -- on gfx950 the fastest
// shape of this 2-read/1-write mix measured (6.56 TB/s over 1 GiB arrays
// against 5.1 TB/s for a grid-stride loop with cached accesses).
typedef double zev2d __attribute__((ext_vector_type(2)));
__global__ __launch_bounds__(256) void k_zaxpy(int64_t n, double a, const zev2d *__restrict__ x, zev2d *y) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const zev2d v = __builtin_nontemporal_load(x + i), w = __builtin_nontemporal_load(y + i);
    const zev2d r = {w.x + a * v.x, w.y + a * v.y};
    __builtin_nontemporal_store(r, y + i);
}
extern "C" int rsvdt_hip_zaxpy(int64_t n, double a, const double *x, double *y, rsvdt_stream_t s) {
    if (n <= 0) return RSVDT_HIP_OK;
    if (n > (int64_t)0x7fffffff * 256) { set_error_msg("rsvdt_hip_zaxpy: n too large"); return RSVDT_HIP_ERR_ARG; }
    hipLaunchKernelGGL(k_zaxpy, dim3(ceil_div(n, 256)), dim3(256), 0, as_stream(s), n, a, (const zev2d *)x,
                       (zev2d *)y);
    RSVDT_LAUNCH_CHECK("rsvdt_hip_zaxpy");
    return RSVDT_HIP_OK;
}
__global__ void k_z2c(int64_t n, const double2 *x, float2 *y) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) { double2 v = x[i]; y[i] = make_float2((float)v.x, (float)v.y); }
}
__global__ void k_c2z(int64_t n, const float2 *x, double2 *y) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) { float2 v = x[i]; y[i] = make_double2((double)v.x, (double)v.y); }
}
extern "C" int rsvdt_hip_z2c(int64_t n, const double *x, float *y, rsvdt_stream_t s) {
    if (n <= 0) return RSVDT_HIP_OK;
    unsigned g = ceil_div(n, 256); if (g > 65536u) g = 65536u;
    hipLaunchKernelGGL(k_z2c, dim3(g), dim3(256), 0, as_stream(s), n, (const double2 *)x, (float2 *)y);
    RSVDT_LAUNCH_CHECK("rsvdt_hip_z2c");
    return RSVDT_HIP_OK;
}
extern "C" int rsvdt_hip_c2z(int64_t n, const float *x, double *y, rsvdt_stream_t s) {
    if (n <= 0) return RSVDT_HIP_OK;
    unsigned g = ceil_div(n, 256); if (g > 65536u) g = 65536u;
    hipLaunchKernelGGL(k_c2z, dim3(g), dim3(256), 0, as_stream(s), n, (const float2 *)x, (double2 *)y);
    RSVDT_LAUNCH_CHECK("rsvdt_hip_c2z");
    return RSVDT_HIP_OK;
}

__global__ void k_zcopy3(int64_t nb, int64_t nr, int64_t nc, const double2 *src, int64_t sb, int64_t sr,
                         int64_t sc, double2 *dst, int64_t db, int64_t dr, int64_t dc) {
    int64_t n = nb * nr * nc;
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) {
        int64_t c = i % nc, t = i / nc, r = t % nr, b = t / nr;
        dst[b * db + r * dr + c * dc] = src[b * sb + r * sr + c * sc];
    }
}
extern "C" int rsvdt_hip_zcopy3(int64_t nb, int64_t nr, int64_t nc, const double *src, int64_t sb, int64_t sr,
                                int64_t sc, double *dst, int64_t db, int64_t dr, int64_t dc, rsvdt_stream_t s) {
    int64_t n = nb * nr * nc;
    if (n <= 0) return RSVDT_HIP_OK;
    unsigned g = ceil_div(n, 256); if (g > 65536u) g = 65536u;
    hipLaunchKernelGGL(k_zcopy3, dim3(g), dim3(256), 0, as_stream(s), nb, nr, nc, (const double2 *)src, sb, sr,
                       sc, (double2 *)dst, db, dr, dc);
    RSVDT_LAUNCH_CHECK("rsvdt_hip_zcopy3");
    return RSVDT_HIP_OK;
}

// Per-row transpose between the block layout [s][r][b] (block s = an
// nrows x kt row-major slab) and the per-batch row-major layout [j][r][s]
// (batch j = nrows x n row-major), for the batch range b0 .. b0+kb-1.  One
// workgroup per row; both sides are read and written in contiguous runs.
__global__ __launch_bounds__(256) void k_transpose_rows(int64_t nrows, int n, int64_t kt, int64_t b0, int kb,
                                                        const double2 *__restrict__ src, double2 *__restrict__ dst,
                                                        int dir)
{
    extern __shared__ double2 tt[];   // [n][kb+1]
    const int64_t r = blockIdx.x;
    const int tot = n * kb;
    if (dir == 0) {
        for (int e = threadIdx.x; e < tot; e += blockDim.x) {
            const int s = e / kb, j = e % kb;
            tt[s * (kb + 1) + j] = src[(int64_t)s * nrows * kt + r * kt + b0 + j];
        }
        __syncthreads();
        for (int e = threadIdx.x; e < tot; e += blockDim.x) {
            const int j = e / n, s = e % n;
            dst[(int64_t)j * nrows * n + r * n + s] = tt[s * (kb + 1) + j];
        }
    } else {
        for (int e = threadIdx.x; e < tot; e += blockDim.x) {
            const int j = e / n, s = e % n;
            tt[s * (kb + 1) + j] = src[(int64_t)j * nrows * n + r * n + s];
        }
        __syncthreads();
        for (int e = threadIdx.x; e < tot; e += blockDim.x) {
            const int s = e / kb, j = e % kb;
            dst[(int64_t)s * nrows * kt + r * kt + b0 + j] = tt[s * (kb + 1) + j];
        }
    }
}
extern "C" int rsvdt_hip_transpose_rows(int64_t nrows, int64_t n, int64_t kt, int64_t b0, int64_t kb,
                                        const double *src, double *dst, int dir, rsvdt_stream_t s)
{
    if (nrows <= 0 || n <= 0 || kb <= 0) return RSVDT_HIP_OK;
    if (b0 < 0 || b0 + kb > kt || nrows > 0x7fffffffLL || n * (kb + 1) > 8192 || (dir != 0 && dir != 1)) {
        set_error_msg("rsvdt_hip_transpose_rows: bad arguments (n*(kb+1) <= 8192, b0+kb <= kt)");
        return RSVDT_HIP_ERR_ARG;
    }
    const size_t lds = (size_t)n * (size_t)(kb + 1) * 16;
    hipLaunchKernelGGL(k_transpose_rows, dim3((unsigned)nrows), dim3(256), lds, as_stream(s), nrows, (int)n, kt, b0,
                       (int)kb, (const double2 *)src, (double2 *)dst, dir);
    RSVDT_LAUNCH_CHECK("rsvdt_hip_transpose_rows");
    return RSVDT_HIP_OK;
}

// Synthetic snapshot sets of the transient-removal estimate
// (SourceCode/TransientRemovalEst.c:84-98), per-batch row-major:
// out[b][r][c] = Q[r*ldq + c + skip] + qss[r] * ph[b*ns + c], c < ns.
__global__ __launch_bounds__(256) void k_trans_synth(int64_t nrows, int64_t ldq, int64_t ns, int64_t skip,
                                                     int64_t total, const double2 *__restrict__ Q,
                                                     const double2 *__restrict__ qss, const double2 *__restrict__ ph,
                                                     double2 *__restrict__ out)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += stride) {
        const int64_t c = e % ns, rb = e / ns, r = rb % nrows, b = rb / nrows;
        const double2 q = Q[r * ldq + c + skip], v = qss[r], p = ph[b * ns + c];
        out[e] = make_double2(q.x + (v.x * p.x - v.y * p.y), q.y + (v.x * p.y + v.y * p.x));
    }
}
extern "C" int rsvdt_hip_trans_synth(int64_t nrows, int64_t ldq, int64_t ns, int64_t skip, int64_t nb,
                                     const double *Q, const double *qss, const double *ph, double *out,
                                     rsvdt_stream_t s)
{
    if (nrows <= 0 || ns <= 0 || nb <= 0) return RSVDT_HIP_OK;
    if (skip < 0 || ns + skip > ldq) {
        set_error_msg("rsvdt_hip_trans_synth: bad arguments (ns + skip <= ldq)");
        return RSVDT_HIP_ERR_ARG;
    }
    const int64_t total = nb * nrows * ns;
    const int64_t g = total / 256 + 1 < 65536 ? total / 256 + 1 : 65536;
    hipLaunchKernelGGL(k_trans_synth, dim3((unsigned)g), dim3(256), 0, as_stream(s), nrows, ldq, ns, skip, total,
                       (const double2 *)Q, (const double2 *)qss, (const double2 *)ph, (double2 *)out);
    RSVDT_LAUNCH_CHECK("rsvdt_hip_trans_synth");
    return RSVDT_HIP_OK;
}

// ---------------------------------------------------------------- graphs
// Snapshot slot of each step of a chunk (SourceCode/SaveSnapshots.c:17-29):
// step i = i0 + t stores a snapshot when i > Nt and (i - Nt) % RR == 0, at
// column (i - Nt)/RR - 1 mod Nstore (reversed for the adjoint); -1 otherwise.
__global__ void k_snap_slots(int *slots, int64_t S, int64_t i0, int64_t Nt, int64_t RR, int64_t Nstore, int dir)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= S) return;
    const int64_t i = i0 + t;
    int v = -1;
    if (i > Nt && (i - Nt) % RR == 0) {
        int64_t pos = ((i - Nt) / RR - 1) % Nstore;
        if (!dir) pos = Nstore - 1 - pos;
        v = (int)pos;
    }
    slots[t] = v;
}
extern "C" int rsvdt_hip_snap_slots(int *slots, int64_t S, int64_t i0, int64_t Nt, int64_t RR, int64_t Nstore,
                                    int dir, rsvdt_stream_t s)
{
    if (S <= 0) return RSVDT_HIP_OK;
    if (RR <= 0 || Nstore <= 0) { set_error_msg("rsvdt_hip_snap_slots: bad arguments"); return RSVDT_HIP_ERR_ARG; }
    hipLaunchKernelGGL(k_snap_slots, dim3(ceil_div(S, 256)), dim3(256), 0, as_stream(s), slots, S, i0, Nt, RR, Nstore,
                       dir);
    RSVDT_LAUNCH_CHECK("rsvdt_hip_snap_slots");
    return RSVDT_HIP_OK;
}

extern "C" int rsvdt_hip_graph_begin(rsvdt_stream_t s)
{
    RSVDT_HIP_TRY(hipStreamBeginCapture(as_stream(s), hipStreamCaptureModeThreadLocal));
    return RSVDT_HIP_OK;
}

extern "C" int rsvdt_hip_graph_end(rsvdt_stream_t s, void **exec)
{
    hipGraph_t g = nullptr;
    *exec = nullptr;
    RSVDT_HIP_TRY(hipStreamEndCapture(as_stream(s), &g));
    hipGraphExec_t e = nullptr;
    hipError_t err = hipGraphInstantiate(&e, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (err != hipSuccess) { set_error("hipGraphInstantiate", err); return RSVDT_HIP_ERR_RUNTIME; }
    *exec = (void *)e;
    return RSVDT_HIP_OK;
}

extern "C" int rsvdt_hip_graph_launch(void *exec, rsvdt_stream_t s)
{
    RSVDT_HIP_TRY(hipGraphLaunch((hipGraphExec_t)exec, as_stream(s)));
    return RSVDT_HIP_OK;
}

extern "C" int rsvdt_hip_graph_destroy(void *exec)
{
    if (exec) RSVDT_HIP_TRY(hipGraphExecDestroy((hipGraphExec_t)exec));
    return RSVDT_HIP_OK;
}

// sum |x|^2: per-block partials into work, then one block sums them (deterministic)
__global__ void k_znorm2_part(int64_t n, const double2 *x, double *part) {
    __shared__ double sh[256];
    double acc = 0;
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) acc += zabs2(x[i]);
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}
__global__ void k_sum_parts(int n, const double *part, double *out) {
    __shared__ double sh[256];
    double acc = 0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) acc += part[i];
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = sh[0];
}
extern "C" int rsvdt_hip_znorm2(int64_t n, const double *x, double *out, void *work, rsvdt_stream_t s) {
    const int nblk = 1024;   // work must hold 1024 doubles
    hipLaunchKernelGGL(k_znorm2_part, dim3(nblk), dim3(256), 0, as_stream(s), n, (const double2 *)x, (double *)work);
    hipLaunchKernelGGL(k_sum_parts, dim3(1), dim3(256), 0, as_stream(s), nblk, (const double *)work, out);
    RSVDT_LAUNCH_CHECK("rsvdt_hip_znorm2");
    return RSVDT_HIP_OK;
}

// ---------------------------------------------------------------- seeded sketch
// Counter-based generator: entry (row r, logical column c) of a rows x cols
// matrix stored in kb-column row-major blocks gets the two uniforms
// u(seed, 2i), u(seed, 2i+1) with i = c*rows + r (its column-major index),
// u(s, j) = (splitmix64(s*0x100000001B3 + j) >> 11 + 0.5) / 2^53, so the
// sketch is independent of the layout and of the launch shape.
__device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
__device__ __forceinline__ double u01(uint64_t seed, uint64_t j) {
    return ((double)(mix64(seed * 0x100000001B3ULL + j) >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}
__global__ void k_fill_random(int64_t rows, int64_t kb, int64_t nblk, int64_t col0, uint64_t seed, int gauss,
                              double2 *out) {
    const int64_t n = rows * kb * nblk;
    int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (; p < n; p += stride) {
        const int64_t b = p / (rows * kb), rem = p % (rows * kb), r = rem / kb, j = rem % kb;
        const uint64_t i = (uint64_t)((b * kb + j + col0) * rows + r);
        const double a = u01(seed, 2 * i), c = u01(seed, 2 * i + 1);
        if (gauss) {
            const double rad = sqrt(-2.0 * log(a)) * 0.70710678118654752440;
            double sn, cs;
            sincos(6.28318530717958647692 * c, &sn, &cs);
            out[p] = make_double2(rad * cs, rad * sn);
        } else {
            out[p] = make_double2(a, c);
        }
    }
}
extern "C" int rsvdt_hip_fill_random(int64_t rows, int64_t kb, int64_t nblk, uint64_t seed, int gauss, double *out,
                                     rsvdt_stream_t s) {
    const int64_t n = rows * kb * nblk;
    if (n <= 0) return RSVDT_HIP_OK;
    unsigned g = ceil_div(n, 256); if (g > 65536u) g = 65536u;
    hipLaunchKernelGGL(k_fill_random, dim3(g), dim3(256), 0, as_stream(s), rows, kb, nblk, (int64_t)0, (uint64_t)seed,
                       gauss, (double2 *)out);
    RSVDT_LAUNCH_CHECK("rsvdt_hip_fill_random");
    return RSVDT_HIP_OK;
}
