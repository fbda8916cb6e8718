// Learned-sketch network for gfx950: a coordinate MLP that maps a grid point
// and frequency of a parameterised operator family, (x, w, mu0, Re U), to 8
// complex forcing columns of the RSVD sketch (the columns CreateRandomMat
// would otherwise draw at random, SourceCode/CreateRandomMat.c:33-39).
//
//   features (64) -> 256 -> 256 -> 256 -> 16 (SiLU between layers)
//
// Training and inference run on BF16 MFMA (v_mfma_f32_16x16x32_bf16, FP32
// accumulation) with FP32 storage of weights and activations: operands are
// rounded to BF16 (round to nearest even) as they are loaded.
#include "common.h"

namespace rsvdt {

typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8;
typedef __attribute__((ext_vector_type(4))) float f4v;

// ----------------------------------------------------------------- features
// [x/xs, w/ws, mu, u, sin(kappa_j x), cos(kappa_j x) (j < 28), 0 pad] with
// kappa_j = kmax (j + 1) / 28 in the units of x.
__global__ void k_mlp_features(int64_t n, const double *__restrict__ q, double xs, double ws, double kmax,
                               float *__restrict__ out)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double x = q[4 * i], w = q[4 * i + 1], mu = q[4 * i + 2], u = q[4 * i + 3];
    float *o = out + i * 64;
    o[0] = (float)(x / xs);
    o[1] = (float)(w / ws);
    o[2] = (float)mu;
    o[3] = (float)u;
    for (int j = 0; j < 28; ++j) {
        const double a = kmax * (double)(j + 1) / 28.0 * x;
        o[4 + 2 * j] = (float)sin(a);
        o[5 + 2 * j] = (float)cos(a);
    }
    for (int j = 60; j < 64; ++j) o[j] = 0.0f;
}

// ----------------------------------------------------------------- BF16 GEMM
// C(m, n) = sum_k A(m, k) B(k, n), A(m, k) = A[m*sam + k*sak], B(k, n) =
// B[k*sbk + n*sbn], FP32 in memory, BF16 into the MFMA, FP32 accumulation.
// A workgroup (4 waves) owns a 64 x 64 tile of C and a wave a 32 x 32
// quarter (2 x 2 MFMA tiles).  Per K-step of 32 the workgroup loads the A and
// B tiles once, rounded to BF16, into LDS with k contiguous -- from either
// memory order of each operand, the loads of a wave running along the
// operand's unit-stride dimension -- so each element of A is read once per
// 64 columns of C and each element of B once per 64 rows.
// MFMA 16x16x32 lane layout (checked on gfx950 with a known-answer product):
//   A(16x32): lane l -> A[l%16][8(l/16)+i];  B(32x16): lane l -> B[8(l/16)+i][l%16];
//   D(16x16): lane l -> D[4(l/16)+j][l%16].
// Epilogues: 0 C = acc; 1 z = acc + bias, Z = z, C = silu(z); 2 C = acc + bias;
// 3 C = acc * silu'(Z) (the backward of epilogue 1).  Long reductions (the
// weight gradients, K = the batch) are split over gridDim.z: each split
// writes its partial tile to `part` and k_gemm_reduce sums the splits in a
// fixed order and applies the epilogue.
enum { EPI_NONE = 0, EPI_BIAS_SILU = 1, EPI_BIAS = 2, EPI_DSILU = 3 };

__device__ __forceinline__ float sigm(float z) { return 1.0f / (1.0f + __expf(-z)); }

template <int EPI>
__device__ __forceinline__ void gemm_epilogue(int64_t m, int64_t n, float v, float *__restrict__ C, int64_t ldc,
                                              const float *__restrict__ bias, float *__restrict__ Z, int64_t ldz)
{
    if (EPI == EPI_BIAS_SILU) {
        const float z = v + bias[n];
        Z[m * ldz + n] = z;
        v = z * sigm(z);
    } else if (EPI == EPI_BIAS) {
        v += bias[n];
    } else if (EPI == EPI_DSILU) {
        const float z = Z[m * ldz + n], s = sigm(z);
        v *= s * (1.0f + z * (1.0f - s));
    }
    C[m * ldc + n] = v;
}

constexpr int GT = 64;          // tile of C per workgroup
constexpr int GK = 32;          // K-step
constexpr int GP = GK + 8;      // LDS row pitch in BF16 (80 B: 16-byte reads of 16 rows spread over the banks)

// One 64 x 32 operand tile, X(r, k) = X[r*sr + k*sk] for rows r0.. (< R) and
// k in [k0, ke): element e = 256 i + tid of the tile (i < 8) runs along the
// operand's unit-stride dimension, so each load instruction of a wave reads
// contiguous memory.  gemm_fetch loads into registers (FP32), gemm_put
// stores them to LDS as [row][k] BF16.
template <bool KF>
__device__ __forceinline__ void gemm_fetch(float (&v)[8], const float *__restrict__ X, int64_t sr, int64_t sk,
                                           int64_t r0, int64_t R, int64_t k0, int64_t ke)
{
    const int tid = threadIdx.x;
    if constexpr (KF) {                // k contiguous: rows 8i + tid/32, k = tid % 32
        const int64_t row = r0 + (tid >> 5), k = k0 + (tid & 31);
        const float *p = X + row * sr + k;
#pragma unroll
        for (int i = 0; i < 8; ++i) v[i] = (row + 8 * i < R && k < ke) ? p[8 * i * sr] : 0.0f;
    } else {                           // rows contiguous: row tid % 64, k = 4i + tid/64
        const int64_t row = r0 + (tid & 63), k = k0 + (tid >> 6);
        const float *p = X + row * sr + k * sk;
#pragma unroll
        for (int i = 0; i < 8; ++i) v[i] = (row < R && k + 4 * i < ke) ? p[4 * i * sk] : 0.0f;
    }
}
template <bool KF>
__device__ __forceinline__ void gemm_put(__bf16 *__restrict__ S, const float (&v)[8])
{
    const int tid = threadIdx.x;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        const int rr = KF ? 8 * i + (tid >> 5) : tid & 63, kk = KF ? tid & 31 : 4 * i + (tid >> 6);
        S[rr * GP + kk] = (__bf16)v[i];
    }
}

template <int EPI, bool AK, bool BK>
__global__ __launch_bounds__(256) void k_gemm_bf16(int M, int N, int K, int kchunk, const float *__restrict__ A,
                                                   int64_t sam, int64_t sak, const float *__restrict__ B, int64_t sbk,
                                                   int64_t sbn, float *__restrict__ C, int64_t ldc,
                                                   const float *__restrict__ bias, float *__restrict__ Z, int64_t ldz,
                                                   float *__restrict__ part)
{
    __shared__ __align__(16) __bf16 As[GT * GP];
    __shared__ __align__(16) __bf16 Bs[GT * GP];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int64_t m0 = (int64_t)blockIdx.y * GT, n0 = (int64_t)blockIdx.x * GT;
    const int wm = (wv >> 1) * 32, wn = (wv & 1) * 32;
    const int64_t kb = (int64_t)blockIdx.z * kchunk;
    const int64_t ke = kb + kchunk < K ? kb + kchunk : K;
    const int r = lane & 15, g = lane >> 4;
    f4v acc[2][2];
#pragma unroll
    for (int t = 0; t < 2; ++t)
#pragma unroll
        for (int u = 0; u < 2; ++u) acc[t][u] = (f4v){0, 0, 0, 0};
    // the next K-step's operands are fetched into registers while this one's
    // MFMAs run
    float va[8], vb[8];
    gemm_fetch<AK>(va, A, sam, sak, m0, M, kb, ke);
    gemm_fetch<BK>(vb, B, sbn, sbk, n0, N, kb, ke);
    for (int64_t k0 = kb; k0 < ke; k0 += GK) {
        gemm_put<AK>(As, va);
        gemm_put<BK>(Bs, vb);
        __syncthreads();
        if (k0 + GK < ke) {
            gemm_fetch<AK>(va, A, sam, sak, m0, M, k0 + GK, ke);
            gemm_fetch<BK>(vb, B, sbn, sbk, n0, N, k0 + GK, ke);
        }
        bf16x8 a[2], b[2];
#pragma unroll
        for (int t = 0; t < 2; ++t) a[t] = *(const bf16x8 *)(As + (wm + 16 * t + r) * GP + 8 * g);
#pragma unroll
        for (int u = 0; u < 2; ++u) b[u] = *(const bf16x8 *)(Bs + (wn + 16 * u + r) * GP + 8 * g);
#pragma unroll
        for (int t = 0; t < 2; ++t)
#pragma unroll
            for (int u = 0; u < 2; ++u) acc[t][u] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[t], b[u], acc[t][u], 0, 0, 0);
        __syncthreads();
    }
#pragma unroll
    for (int t = 0; t < 2; ++t)
#pragma unroll
        for (int u = 0; u < 2; ++u)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int64_t m = m0 + wm + 16 * t + 4 * g + j, n = n0 + wn + 16 * u + r;
                if (m >= M || n >= N) continue;
                if (part != nullptr) part[((int64_t)blockIdx.z * M + m) * N + n] = acc[t][u][j];
                else gemm_epilogue<EPI>(m, n, acc[t][u][j], C, ldc, bias, Z, ldz);
            }
}

// C = epilogue(sum over the nsplit partial products, in split order)
template <int EPI>
__global__ void k_gemm_reduce(int64_t M, int64_t N, int nsplit, const float *__restrict__ part, float *__restrict__ C,
                              int64_t ldc, const float *__restrict__ bias, float *__restrict__ Z, int64_t ldz)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= M * N) return;
    float v = 0.0f;
    for (int sp = 0; sp < nsplit; ++sp) v += part[sp * M * N + e];
    gemm_epilogue<EPI>(e / N, e % N, v, C, ldc, bias, Z, ldz);
}

// out[n] = sum_m A[m*lda + n] (bias gradients); out must be zeroed first.
__global__ void k_colsum(int64_t M, int64_t N, const float *__restrict__ A, int64_t lda, float *__restrict__ out)
{
    const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    const int64_t m0 = (int64_t)blockIdx.y * 256;
    int64_t m1 = m0 + 256;
    if (m1 > M) m1 = M;
    float s = 0.0f;
    for (int64_t m = m0; m < m1; ++m) s += A[m * lda + n];
    atomicAdd(out + n, s);
}

// Phase- and scale-invariant alignment loss of ncol learned columns with
// their target modes, per group of npts points (one operator instance at
// one frequency):  L = sum_{g,j} (1 - |s|^2 / (|f|^2 |v|^2)) / (ngroups ncol),
// s = sum_x conj(v_j(x)) f_j(x).  dF = dL/dRe f + i dL/dIm f.
// F and dF: (ngroups*npts) x (2 ncol) float (re, im interleaved per column);
// V: ngroups x npts x ncol complex FP64.  loss: ngroups floats (per group).
__global__ __launch_bounds__(256) void k_sketch_loss(int npts, int ncol, const float *__restrict__ F,
                                                     const double2 *__restrict__ V, float *__restrict__ dF,
                                                     float *__restrict__ loss, double scale)
{
    __shared__ double tot[4][16];              // per column: Re s, Im s, |f|^2, |v|^2
    const int64_t grp = blockIdx.x;
    const float *Fg = F + grp * (int64_t)npts * 2 * ncol;
    const double2 *Vg = V + grp * (int64_t)npts * ncol;
    float *dFg = dF + grp * (int64_t)npts * 2 * ncol;
    const int j = threadIdx.x % 16, slice = threadIdx.x / 16;      // 16 columns x 16 slices of points
    double sr = 0, si = 0, nf = 0, nv = 0;
    if (j < ncol)
        for (int x = slice; x < npts; x += 16) {
            const double fr = Fg[(int64_t)x * 2 * ncol + 2 * j], fi = Fg[(int64_t)x * 2 * ncol + 2 * j + 1];
            const double2 v = Vg[(int64_t)x * ncol + j];
            sr += v.x * fr + v.y * fi;          // conj(v) f
            si += v.x * fi - v.y * fr;
            nf += fr * fr + fi * fi;
            nv += v.x * v.x + v.y * v.y;
        }
    __shared__ double part[4][256];
    part[0][threadIdx.x] = sr; part[1][threadIdx.x] = si; part[2][threadIdx.x] = nf; part[3][threadIdx.x] = nv;
    __syncthreads();
    if (threadIdx.x < 16) {
        double a[4] = {0, 0, 0, 0};
        for (int sl = 0; sl < 16; ++sl)
            for (int c = 0; c < 4; ++c) a[c] += part[c][sl * 16 + threadIdx.x];
        for (int c = 0; c < 4; ++c) tot[c][threadIdx.x] = a[c];
    }
    __syncthreads();
    // gradient of 1 - |s|^2/(n nv): g = (2/(n nv)) (|s|^2/n f - s v)
    for (int e = threadIdx.x; e < npts * ncol; e += blockDim.x) {
        const int x = e / ncol, jj = e % ncol;
        const double ssr = tot[0][jj], ssi = tot[1][jj], n = tot[2][jj] > 1e-300 ? tot[2][jj] : 1e-300;
        const double nvv = tot[3][jj] > 1e-300 ? tot[3][jj] : 1e-300;
        const double s2 = ssr * ssr + ssi * ssi;
        const double fr = Fg[(int64_t)x * 2 * ncol + 2 * jj], fi = Fg[(int64_t)x * 2 * ncol + 2 * jj + 1];
        const double2 v = Vg[(int64_t)x * ncol + jj];
        const double c = 2.0 / (n * nvv) * scale;
        dFg[(int64_t)x * 2 * ncol + 2 * jj] = (float)(c * (s2 / n * fr - (ssr * v.x - ssi * v.y)));
        dFg[(int64_t)x * 2 * ncol + 2 * jj + 1] = (float)(c * (s2 / n * fi - (ssr * v.y + ssi * v.x)));
    }
    if (threadIdx.x == 0) {
        double L = 0;
        for (int jj = 0; jj < ncol; ++jj) {
            const double s2 = tot[0][jj] * tot[0][jj] + tot[1][jj] * tot[1][jj];
            const double n = tot[2][jj] > 1e-300 ? tot[2][jj] : 1e-300, nvv = tot[3][jj] > 1e-300 ? tot[3][jj] : 1e-300;
            L += 1.0 - s2 / (n * nvv);
        }
        loss[grp] = (float)(L * scale);
    }
}

// Adam (Kingma & Ba), bias-corrected step sizes given as b1t = 1 - beta1^t,
// b2t = 1 - beta2^t.
__global__ void k_adam(int64_t n, float *__restrict__ w, const float *__restrict__ g, float *__restrict__ m,
                       float *__restrict__ v, float lr, float b1, float b2, float eps, float b1t, float b2t)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    w[i] -= lr * (mi / b1t) / (sqrtf(vi / b2t) + eps);
}

}  // namespace rsvdt

using namespace rsvdt;

extern "C" int rsvdt_hip_mlp_features(int64_t n, const double *q, double xs, double ws, double kmax, float *out,
                                      rsvdt_stream_t s)
{
    if (n <= 0) return RSVDT_HIP_OK;
    hipLaunchKernelGGL(k_mlp_features, dim3(ceil_div(n, 256)), dim3(256), 0, as_stream(s), n, q, xs, ws, kmax, out);
    RSVDT_LAUNCH_CHECK("rsvdt_hip_mlp_features");
    return RSVDT_HIP_OK;
}

// Split-K workspace: grown on demand, kept for the process (GEMMs of one
// stream use it one after another).
static float *g_gemm_part = nullptr;
static size_t g_gemm_part_bytes = 0;

extern "C" int rsvdt_hip_gemm_bf16(int64_t M, int64_t N, int64_t K, const float *A, int64_t sam, int64_t sak,
                                   const float *B, int64_t sbk, int64_t sbn, float *C, int64_t ldc, int epilogue,
                                   const float *bias, float *Z, int64_t ldz, rsvdt_stream_t s)
{
    if (M <= 0 || N <= 0) return RSVDT_HIP_OK;
    if (M > 0x7fffffff || N > 0x7fffffff || K > 0x7fffffff || K < 0 || epilogue < 0 || epilogue > 3 ||
        ((epilogue == 1 || epilogue == 2) && !bias) || ((epilogue == 1 || epilogue == 3) && !Z)) {
        set_error_msg("rsvdt_hip_gemm_bf16: bad arguments");
        return RSVDT_HIP_ERR_ARG;
    }
    if (ceil_div(M, GT) > 65535) {
        set_error_msg("rsvdt_hip_gemm_bf16: M above 64 x 65535");
        return RSVDT_HIP_ERR_ARG;
    }
    const int64_t tiles = ceil_div(M, GT) * ceil_div(N, GT);
    // split the reduction when the tiles alone leave most CUs idle and K is long
    int64_t nsplit = 1;
    if (tiles < 256 && K >= 2048) {
        nsplit = 512 / tiles;
        const int64_t kmax = K / 512;                 // at least 512 per split
        if (nsplit > kmax) nsplit = kmax;
        if (nsplit > 64) nsplit = 64;
        if (nsplit < 1) nsplit = 1;
    }
    int64_t kchunk = K;
    if (nsplit > 1) {
        kchunk = ceil_div(ceil_div(K, nsplit), GK) * GK;
        nsplit = ceil_div(K, kchunk);
    }
    if (kchunk < 1) kchunk = 1;
    float *part = nullptr;
    if (nsplit > 1) {
        const size_t need = (size_t)nsplit * (size_t)M * (size_t)N * sizeof(float);
        if (need > g_gemm_part_bytes) {
            if (g_gemm_part) rsvdt_hip_free(g_gemm_part);
            g_gemm_part = nullptr;
            g_gemm_part_bytes = 0;
            if (rsvdt_hip_malloc((void **)&g_gemm_part, need)) return RSVDT_HIP_ERR_RUNTIME;
            g_gemm_part_bytes = need;
        }
        part = g_gemm_part;
    }
    dim3 grid((unsigned)ceil_div(N, GT), (unsigned)ceil_div(M, GT), (unsigned)nsplit);
    hipStream_t st = as_stream(s);
    // AK: A k-contiguous (sak == 1), else read along its rows; BK: likewise for B
    const bool ak = sak == 1, bk = sbk == 1;
#define RSVDT_GEMM_K(E, AKV, BKV)                                                                                \
    hipLaunchKernelGGL((k_gemm_bf16<E, AKV, BKV>), grid, dim3(256), 0, st, (int)M, (int)N, (int)K, (int)kchunk, A, \
                       sam, sak, B, sbk, sbn, C, ldc, bias, Z, ldz, part)
#define RSVDT_GEMM_LAUNCH(E)                                                                                     \
    if (ak && bk) RSVDT_GEMM_K(E, true, true);                                                                  \
    else if (ak) RSVDT_GEMM_K(E, true, false);                                                                  \
    else if (bk) RSVDT_GEMM_K(E, false, true);                                                                  \
    else RSVDT_GEMM_K(E, false, false);                                                                         \
    if (part)                                                                                                    \
        hipLaunchKernelGGL(k_gemm_reduce<E>, dim3((unsigned)ceil_div(M * N, 256)), dim3(256), 0, st, M, N,       \
                           (int)nsplit, part, C, ldc, bias, Z, ldz)
    switch (epilogue) {
    case 0: RSVDT_GEMM_LAUNCH(0); break;
    case 1: RSVDT_GEMM_LAUNCH(1); break;
    case 2: RSVDT_GEMM_LAUNCH(2); break;
    default: RSVDT_GEMM_LAUNCH(3); break;
    }
#undef RSVDT_GEMM_LAUNCH
#undef RSVDT_GEMM_K
    RSVDT_LAUNCH_CHECK("rsvdt_hip_gemm_bf16");
    return RSVDT_HIP_OK;
}

extern "C" int rsvdt_hip_colsum(int64_t M, int64_t N, const float *A, int64_t lda, float *out, rsvdt_stream_t s)
{
    if (N <= 0) return RSVDT_HIP_OK;
    int rc = rsvdt_hip_zero(out, (size_t)N * sizeof(float), s);
    if (rc || M <= 0) return rc;
    if ((M + 255) / 256 > 65535) { set_error_msg("rsvdt_hip_colsum: M too large"); return RSVDT_HIP_ERR_ARG; }
    hipLaunchKernelGGL(k_colsum, dim3(ceil_div(N, 256), (unsigned)((M + 255) / 256)), dim3(256), 0, as_stream(s), M,
                       N, A, lda, out);
    RSVDT_LAUNCH_CHECK("rsvdt_hip_colsum");
    return RSVDT_HIP_OK;
}

extern "C" int rsvdt_hip_sketch_loss(int64_t ngroups, int64_t npts, int64_t ncol, const float *F, const double *V,
                                     float *dF, float *loss, double scale, rsvdt_stream_t s)
{
    if (ngroups <= 0) return RSVDT_HIP_OK;
    if (ncol < 1 || ncol > 16 || npts < 1 || ngroups > 0x7fffffff) {
        set_error_msg("rsvdt_hip_sketch_loss: 1 <= ncol <= 16 required");
        return RSVDT_HIP_ERR_ARG;
    }
    hipLaunchKernelGGL(k_sketch_loss, dim3((unsigned)ngroups), dim3(256), 0, as_stream(s), (int)npts, (int)ncol, F,
                       (const double2 *)V, dF, loss, scale);
    RSVDT_LAUNCH_CHECK("rsvdt_hip_sketch_loss");
    return RSVDT_HIP_OK;
}

extern "C" int rsvdt_hip_adam(int64_t n, float *w, const float *g, float *m, float *v, float lr, float b1, float b2,
                              float eps, float b1t, float b2t, rsvdt_stream_t s)
{
    if (n <= 0) return RSVDT_HIP_OK;
    hipLaunchKernelGGL(k_adam, dim3(ceil_div(n, 256)), dim3(256), 0, as_stream(s), n, w, g, m, v, lr, b1, b2, eps, b1t,
                       b2t);
    RSVDT_LAUNCH_CHECK("rsvdt_hip_adam");
    return RSVDT_HIP_OK;
}
