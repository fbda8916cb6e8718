// The per-frequency QR (QRAllFreqs) for callers that cannot allocate device
// memory: code under HIP graph capture, tight serving loops, harnesses that
// forbid hipMalloc.  The CholeskyQR work (Gram matrices and their partial
// sums, Cholesky factors and inverses, the pivot counter) lives in one
// file-scope device array of RSVDT_NOALLOC_SCRATCH_BYTES, which every GPU
// the library runs on holds until the process ends.  Its device address is
// looked up once per device (hipGetSymbolAddress) and only that address is
// handed to kernels and fills: the array's name in host code is a host-side
// stand-in, not a device address.
//
// Every call on one GPU shares that GPU's array and leaves the last call's
// contents in it; the pivot counter is cleared at the start of each call.
// Calls that overlap on one GPU (two host threads, or two streams whose
// work is not ordered) race on the array.
#include <stdio.h>
#include "common.h"

extern "C" {
#include "rsvdt_internal.h"
}

static __device__ __align__(256) char scratch[RSVDT_NOALLOC_SCRATCH_BYTES];

#define RSVDT_NOALLOC_MAXDEV 64
static void *g_scratch_ptr[RSVDT_NOALLOC_MAXDEV];   // device address of `scratch`, per device

// Looks up this device's copy of the array (the current device must be
// `device`).  Called by rsvdt_hip_init for its device and by the entry
// point the first time it runs on any other device.
extern "C" int rsvdt_noalloc_setup(int device)
{
    if (device < 0 || device >= RSVDT_NOALLOC_MAXDEV) {
        rsvdt::set_error_msg("rsvdt_noalloc_setup: device index out of range");
        return RSVDT_HIP_ERR_ARG;
    }
    if (g_scratch_ptr[device]) return RSVDT_HIP_OK;
    void *ptr = nullptr;
    hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || ptr == nullptr) {
        rsvdt::set_error("hipGetSymbolAddress(scratch)", e != hipSuccess ? e : hipErrorInvalidSymbol);
        return RSVDT_HIP_ERR_RUNTIME;
    }
    g_scratch_ptr[device] = ptr;
    return RSVDT_HIP_OK;
}

static size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

extern "C" size_t rsvdt_qr_blocks_noalloc_bytes(int64_t nblocks, int64_t N, int64_t k)
{
    if (nblocks <= 0 || N <= 0 || k <= 0) return 0;
    const size_t mat = align256((size_t)nblocks * (size_t)k * (size_t)k * 16);
    return 3 * mat + 256 + align256(rsvdt_hip_bgram_work_bytes(nblocks, N, k, k));
}

extern "C" int rsvdt_qr_blocks_noalloc(double *Y, int64_t N, int64_t k, int64_t nblocks, rsvdt_stream_t s)
{
    if (Y == nullptr || N <= 0 || k <= 0 || nblocks <= 0) {
        rsvdt::set_error_msg("rsvdt_qr_blocks_noalloc: null Y or non-positive size");
        return RSVDT_HIP_ERR_ARG;
    }
    if (k > 64) {
        char msg[200];
        snprintf(msg, sizeof(msg), "rsvdt_qr_blocks_noalloc: k = %lld; at most 64 columns per block (wider blocks "
                 "need an out-of-place apply buffer the size of Y)", (long long)k);
        rsvdt::set_error_msg(msg);
        return RSVDT_HIP_ERR_ARG;
    }
    if (nblocks > 65535) {
        rsvdt::set_error_msg("rsvdt_qr_blocks_noalloc: at most 65535 blocks");
        return RSVDT_HIP_ERR_ARG;
    }
    const size_t need = rsvdt_qr_blocks_noalloc_bytes(nblocks, N, k);
    if (need > (size_t)RSVDT_NOALLOC_SCRATCH_BYTES) {
        char msg[256];
        snprintf(msg, sizeof(msg), "rsvdt_qr_blocks_noalloc: %lld blocks of %lld x %lld need %zu bytes of work, "
                 "more than the %zu-byte static scratch", (long long)nblocks, (long long)N, (long long)k, need,
                 (size_t)RSVDT_NOALLOC_SCRATCH_BYTES);
        rsvdt::set_error_msg(msg);
        return RSVDT_HIP_ERR_ARG;
    }
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) {
        rsvdt::set_error_msg("rsvdt_qr_blocks_noalloc: no current device");
        return RSVDT_HIP_ERR_NODEV;
    }
    int rc = rsvdt_noalloc_setup(dev);
    if (rc) return rc;
    char *base = (char *)g_scratch_ptr[dev];
    const size_t mat = align256((size_t)nblocks * (size_t)k * (size_t)k * 16);
    rsvdt_cholqr_work w;
    w.G = (double *)base;
    w.R = (double *)(base + mat);
    w.Ri = (double *)(base + 2 * mat);
    w.info = (int *)(base + 3 * mat);
    w.work = base + 3 * mat + 256;
    w.Ra = nullptr;
    w.Rt = nullptr;
    // shifted CholeskyQR3, as QRAllFreqs; asynchronous on s, no pivot readback
    const PetscErrorCode ierr = rsvdt_cholqr_bufs(nblocks, N, k, Y, N * k, k, 1, nullptr, 3, 1, 0.0, nullptr,
                                                  (size_t)nblocks * (size_t)N * (size_t)k * 16, &w, s, 0);
    return ierr ? (ierr == RSVDT_ERR_USER ? RSVDT_HIP_ERR_ARG : ierr) : RSVDT_HIP_OK;
}
