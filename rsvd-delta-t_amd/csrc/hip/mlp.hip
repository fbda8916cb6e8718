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
// A wave owns a 16 x 16 tile; a workgroup 4 tiles stacked in m.
// MFMA 16x16x32 lane layout (checked on gfx950 with a known-answer product):
//   A(16x32): lane l -> A[l%16][8(l/16)+i];  B(32x16): lane l -> B[8(l/16)+i][l%16];
//   D(16x16): lane l -> D[4(l/16)+j][l%16].
// Epilogues: 0 C = acc; 1 z = acc + bias, Z = z, C = silu(z); 2 C = acc + bias;
// 3 C = acc * silu'(Z) (the backward of epilogue 1).
enum { EPI_NONE = 0, EPI_BIAS_SILU = 1, EPI_BIAS = 2, EPI_DSILU = 3 };

__device__ __forceinline__ float sigm(float z) { return 1.0f / (1.0f + __expf(-z)); }

template <int EPI>
__global__ __launch_bounds__(256) void k_gemm_bf16(int M, int N, int K, const float *__restrict__ A, int64_t sam,
                                                   int64_t sak, const float *__restrict__ B, int64_t sbk,
                                                   int64_t sbn, float *__restrict__ C, int64_t ldc,
                                                   const float *__restrict__ bias, float *__restrict__ Z,
                                                   int64_t ldz)
{
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int64_t m0 = ((int64_t)blockIdx.y * 4 + wv) * 16, n0 = (int64_t)blockIdx.x * 16;
    if (m0 >= M) return;                       // no barrier in this kernel
    const int r = lane & 15, g = lane >> 4;
    const int64_t am = m0 + r, bn = n0 + r;
    f4v acc = {0, 0, 0, 0};
    for (int k0 = 0; k0 < K; k0 += 32) {
        bf16x8 a, b;
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const int64_t k = k0 + 8 * g + i;
            const float av = (am < M && k < K) ? A[am * sam + k * sak] : 0.0f;
            const float bv = (bn < N && k < K) ? B[k * sbk + bn * sbn] : 0.0f;
            a[i] = (__bf16)av;
            b[i] = (__bf16)bv;
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc, 0, 0, 0);
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int64_t m = m0 + 4 * g + j, n = n0 + r;
        if (m >= M || n >= N) continue;
        float v = acc[j];
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
    dim3 grid(ceil_div(N, 16), ceil_div(M, 64));
    hipStream_t st = as_stream(s);
    switch (epilogue) {
    case 0: hipLaunchKernelGGL(k_gemm_bf16<0>, grid, dim3(256), 0, st, (int)M, (int)N, (int)K, A, sam, sak, B, sbk, sbn, C, ldc, bias, Z, ldz); break;
    case 1: hipLaunchKernelGGL(k_gemm_bf16<1>, grid, dim3(256), 0, st, (int)M, (int)N, (int)K, A, sam, sak, B, sbk, sbn, C, ldc, bias, Z, ldz); break;
    case 2: hipLaunchKernelGGL(k_gemm_bf16<2>, grid, dim3(256), 0, st, (int)M, (int)N, (int)K, A, sam, sak, B, sbk, sbn, C, ldc, bias, Z, ldz); break;
    default: hipLaunchKernelGGL(k_gemm_bf16<3>, grid, dim3(256), 0, st, (int)M, (int)N, (int)K, A, sam, sak, B, sbk, sbn, C, ldc, bias, Z, ldz); break;
    }
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
