// Tall complex-FP64 products for gfx950:
//   * coef_gemm: out[p][l] = alpha * sum_q coef(q,p) * in[q][l] -- the forcing
//     synthesis of SourceCode/CreateForcingOnFly.c:17-19 batched over a chunk
//     of half-step samples and over all k test vectors, the snapshot DFT of
//     SourceCode/DFT.c:26-38, and the windowed DFTs of SourceCode/GalerkinProj.c:30-46.
//     L (= N*k complex columns) is huge and K, P <= ~130, so every workgroup
//     streams a slab of `in` once and keeps a P-tile of accumulators per lane;
//     the coefficient tile is wave-uniform and is read through the scalar cache.
//   * bgram: batched conj(X)^T Z over tall row ranges (Gram matrices of
//     CholeskyQR, the projections V^H(.) of the transient removal), reduced
//     deterministically in two passes.
//   * bapply: batched row-local small products Out = alpha*Y + X*C (the
//     triangular applies of CholeskyQR, Q*U of the SVD, y1 - V*a).
#include "common.h"

namespace rsvdt {

// ----------------------------------------------------------------- coef_gemm
// FP64 MFMA (v_mfma_f64_16x16x4f64).  A wave owns a strip of 16 columns l and
// computes all T*16 rows p of the workgroup's row range for it; the complex
// product is four real MFMAs per K-step of 4:
//   Re += Cr*Xr + (-Ci)*Xi,   Im += Cr*Xi + Ci*Xr.
// Operand layouts (16x16x4, wave64): A(16x4): lane i -> A[i%16][i/16],
// B(4x16): lane i -> B[i/16][i%16], D(16x16): lane i holds D[4j+i/16][i%16], j < 4
// (checked on gfx950 with a known-answer product).
// The K x T*16 coefficient block of the workgroup is staged in LDS once and
// reused for every 64-column tile the workgroup visits (grid-stride), so
// `in` is streamed from HBM exactly once per row range.
typedef double dvec4 __attribute__((ext_vector_type(4)));

template <int T, typename OV>
__global__ __launch_bounds__(256) void k_coef_gemm(
    int64_t P, int64_t K, int64_t L, const double2 *__restrict__ coef, int64_t cq, int64_t cp, int64_t p0,
    int sgn, int64_t cmod, double alpha, const double2 *__restrict__ in, int64_t ldi, OV *__restrict__ out,
    int64_t ldo, double beta)
{
    extern __shared__ double2 cs[];                 // [Kpad][T*16 + 4]
    constexpr int PW = T * 16;
    constexpr int PS = PW + 4;                      // row pitch: the 4 rows a K-step reads hit different banks
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int64_t pbase = (int64_t)blockIdx.y * PW;
    const int64_t Kpad = (K + 3) & ~(int64_t)3;
    for (int64_t e = threadIdx.x; e < Kpad * PW; e += 256) {
        const int64_t q = e / PW, pp = e % PW, p = pbase + pp;
        double2 v = make_double2(0, 0);
        if (q < K && p < P) {
            int64_t col = p0 + (int64_t)sgn * p;
            if (col < 0 || col >= cmod) { col %= cmod; if (col < 0) col += cmod; }
            v = coef[q * cq + col * cp];
        }
        cs[q * PS + pp] = v;
    }
    __syncthreads();
    const int kq = lane >> 4, cl = lane & 15;
    const int64_t ntile = (L + 63) / 64;
    // The workgroup's (tile, K-step) sequence is walked as one stream with the
    // streamed operand loaded PF steps ahead, across tile boundaries too
    // (measured on the 64 x 64 x 6.3M forcing chunk: PF = 2, 4, 8, 12, 16 ->
    // 74.6, 69.5, 66.4, 65.0, 65.7 us per sample).
    constexpr int PF = 12;
    const int nks = (int)(Kpad >> 2);
    const int64_t nmine = blockIdx.x < ntile ? (ntile - 1 - blockIdx.x) / gridDim.x + 1 : 0;
    const int64_t nsteps = nmine * nks;
    auto load_step = [&](int64_t gs) -> double2 {
        if (gs >= nsteps) return make_double2(0, 0);
        const int64_t it = gs / nks, q = (gs - it * nks) * 4 + kq;
        const int64_t l = (blockIdx.x + it * gridDim.x) * 64 + wv * 16 + cl;
        return (l < L && q < K) ? in[q * ldi + l] : make_double2(0, 0);
    };
    double2 ring[PF];
#pragma unroll
    for (int p = 0; p < PF; ++p) ring[p] = load_step(p);
    dvec4 ar[T], ai[T];
#pragma unroll
    for (int t = 0; t < T; ++t) { ar[t] = (dvec4){0, 0, 0, 0}; ai[t] = (dvec4){0, 0, 0, 0}; }
    for (int64_t g0 = 0; g0 < nsteps; g0 += PF) {
#pragma unroll
        for (int p = 0; p < PF; ++p) {
            const int64_t gs = g0 + p;
            if (gs < nsteps) {
                const double2 x = ring[p];
                ring[p] = load_step(gs + PF);
                const int64_t it = gs / nks;
                const int ks = (int)(gs - it * nks);
                const double2 *crow = cs + ((int64_t)ks * 4 + kq) * PS + cl;
                double2 c[T];
#pragma unroll
                for (int t = 0; t < T; ++t) c[t] = crow[t * 16];
                // 2T independent accumulators between dependent MFMAs
#pragma unroll
                for (int t = 0; t < T; ++t) {
                    ar[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(c[t].x, x.x, ar[t], 0, 0, 0);
                    ai[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(c[t].x, x.y, ai[t], 0, 0, 0);
                }
#pragma unroll
                for (int t = 0; t < T; ++t) {
                    ar[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(-c[t].y, x.y, ar[t], 0, 0, 0);
                    ai[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(c[t].y, x.x, ai[t], 0, 0, 0);
                }
                if (ks == nks - 1) {                  // tile done: store and restart the sums
                    const int64_t l = (blockIdx.x + it * gridDim.x) * 64 + wv * 16 + cl;
                    if (l < L) {
#pragma unroll
                        for (int t = 0; t < T; ++t)
#pragma unroll
                            for (int j = 0; j < 4; ++j) {
                                const int64_t pp = pbase + t * 16 + 4 * j + kq;
                                if (pp < P) {
                                    OV *o = out + pp * ldo + l;
                                    double orr = alpha * ar[t][j], oi = alpha * ai[t][j];
                                    if (beta != 0.0) { OV v = *o; orr += beta * (double)v.x; oi += beta * (double)v.y; }
                                    OV w;
                                    w.x = orr; w.y = oi;
                                    *o = w;
                                }
                            }
                    }
#pragma unroll
                    for (int t = 0; t < T; ++t) { ar[t] = (dvec4){0, 0, 0, 0}; ai[t] = (dvec4){0, 0, 0, 0}; }
                }
            }
        }
    }
}

template <int T>
static int launch_coef(int64_t P, int64_t K, int64_t L, const double *coef, int64_t cq, int64_t cp, int64_t p0,
                       int sgn, int64_t cmod, double alpha, const double *in, int64_t ldi, void *out, int64_t ldo,
                       int out_prec, double beta, hipStream_t s)
{
    const int64_t Kpad = (K + 3) & ~(int64_t)3;
    const size_t lds = (size_t)Kpad * (T * 16 + 4) * 16;
    const int64_t ntile = (L + 63) / 64;
    // enough workgroups to fill 256 CUs several times; each visits ntile/gx tiles
    const int64_t per_cu = lds > 80 * 1024 ? 1 : (lds > 40 * 1024 ? 2 : 4);
    int64_t gx = 256 * per_cu * 2;
    if (gx > ntile) gx = ntile;
    dim3 grid((unsigned)gx, ceil_div(P, T * 16));
    if (out_prec == 64)
        hipLaunchKernelGGL((k_coef_gemm<T, double2>), grid, dim3(256), lds, s, P, K, L, (const double2 *)coef, cq, cp,
                           p0, sgn, cmod, alpha, (const double2 *)in, ldi, (double2 *)out, ldo, beta);
    else
        hipLaunchKernelGGL((k_coef_gemm<T, float2>), grid, dim3(256), lds, s, P, K, L, (const double2 *)coef, cq, cp,
                           p0, sgn, cmod, alpha, (const double2 *)in, ldi, (float2 *)out, ldo, beta);
    RSVDT_LAUNCH_CHECK("rsvdt_hip_coef_gemm");
    return RSVDT_HIP_OK;
}

// ----------------------------------------------------------------- bgram
#define BG_T 16
__global__ __launch_bounds__(256) void k_bgram_part(
    int64_t nrows, int64_t na, int64_t nc, int64_t chunk, int ntc,
    const double2 *__restrict__ X, int64_t xb, int64_t xr, int64_t xa,
    const double2 *__restrict__ Z, int64_t zb, int64_t zr, int64_t zc, double2 *__restrict__ part)
{
    __shared__ double2 sx[BG_T][BG_T + 1];
    __shared__ double2 sz[BG_T][BG_T + 1];
    const int64_t b = blockIdx.z;
    const int64_t nch = gridDim.x;
    const int tile = blockIdx.y;
    const int64_t a0 = (int64_t)(tile / ntc) * BG_T, c0 = (int64_t)(tile % ntc) * BG_T;
    const int ti = threadIdx.x / BG_T, tj = threadIdx.x % BG_T;
    const int64_t r_begin = (int64_t)blockIdx.x * chunk;
    int64_t r_end = r_begin + chunk;
    if (r_end > nrows) r_end = nrows;
    const double2 *Xb = X + b * xb;
    const double2 *Zb = Z + b * zb;
    double2 acc = make_double2(0, 0);
    for (int64_t r0 = r_begin; r0 < r_end; r0 += BG_T) {
        const int64_t r = r0 + ti;
        const int64_t a = a0 + tj, c = c0 + tj;
        sx[ti][tj] = (r < r_end && a < na) ? Xb[r * xr + a * xa] : make_double2(0, 0);
        sz[ti][tj] = (r < r_end && c < nc) ? Zb[r * zr + c * zc] : make_double2(0, 0);
        __syncthreads();
#pragma unroll
        for (int rr = 0; rr < BG_T; ++rr) {
            const double2 xv = sx[rr][ti], zv = sz[rr][tj];
            acc.x += xv.x * zv.x + xv.y * zv.y;
            acc.y += xv.x * zv.y - xv.y * zv.x;
        }
        __syncthreads();
    }
    const int64_t a = a0 + ti, c = c0 + tj;
    if (a < na && c < nc) part[((b * nch + blockIdx.x) * na + a) * nc + c] = acc;
}

__global__ void k_bgram_reduce(int64_t nbatch, int64_t nch, int64_t na, int64_t nc,
                               const double2 *__restrict__ part, double2 *__restrict__ G)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t per = na * nc;
    if (i >= nbatch * per) return;
    const int64_t b = i / per, e = i % per;
    double2 s = make_double2(0, 0);
    for (int64_t ch = 0; ch < nch; ++ch) {
        const double2 v = part[(b * nch + ch) * per + e];
        s.x += v.x; s.y += v.y;
    }
    G[i] = s;
}

// FP64 MFMA Gram for unit column strides (rows of X and Z contiguous, the
// per-batch row-major layout).  A wave computes a (16 TA) x (16 TC) tile of
// G = X^H Z over its share of a row chunk, complex as four real products:
//   Re G += Xr Zr + Xi Zi,   Im G += Xr Zi - Xi Zr.
// A(16x4): lane i -> X(r0 + i/16, a0 + i%16); B(4x16): lane i -> Z(r0 + i/16,
// c0 + i%16); D: lane i holds G(a0 + 4j + i/16, c0 + i%16), j < 4.
// nsub waves of a workgroup share one tile and split its 4-row steps (small
// Grams); otherwise the 4 waves take 4 tiles.  Each (chunk, sub) writes a
// full partial that k_bgram_reduce sums.  The next step's operands are loaded
// before the current step's MFMAs.
template <int TA, int TC>
__global__ __launch_bounds__(256) void k_bgram_mfma(
    int64_t nrows, int64_t na, int64_t nc, int64_t chunk, int ntc, int ntiles, int nsub,
    const double2 *__restrict__ X, int64_t xb, int64_t xr, const double2 *__restrict__ Z, int64_t zb, int64_t zr,
    double2 *__restrict__ part)
{
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int tile = blockIdx.y * (4 / nsub) + wv / nsub;
    const int sub = wv % nsub;
    if (tile >= ntiles) return;
    const int64_t b = blockIdx.z;
    const int64_t a0 = (int64_t)(tile / ntc) * (16 * TA), c0 = (int64_t)(tile % ntc) * (16 * TC);
    const int64_t r_begin = (int64_t)blockIdx.x * chunk;
    int64_t r_end = r_begin + chunk;
    if (r_end > nrows) r_end = nrows;
    const double2 *Xb = X + b * xb;
    const double2 *Zb = Z + b * zb;
    const int kr = lane >> 4, cl = lane & 15;
    bool aok[TA], cok[TC];
#pragma unroll
    for (int t = 0; t < TA; ++t) aok[t] = a0 + 16 * t + cl < na;
#pragma unroll
    for (int u = 0; u < TC; ++u) cok[u] = c0 + 16 * u + cl < nc;
    dvec4 gr[TA][TC], gi[TA][TC];
#pragma unroll
    for (int t = 0; t < TA; ++t)
#pragma unroll
        for (int u = 0; u < TC; ++u) { gr[t][u] = (dvec4){0, 0, 0, 0}; gi[t][u] = (dvec4){0, 0, 0, 0}; }
    const double2 zero = make_double2(0, 0);
    double2 xv[TA], zv[TC];
    int64_t r0 = r_begin + 4 * sub;
    {
        const int64_t r = r0 + kr;
        const bool rok = r < r_end;
#pragma unroll
        for (int t = 0; t < TA; ++t) xv[t] = (rok && aok[t]) ? Xb[r * xr + a0 + 16 * t + cl] : zero;
#pragma unroll
        for (int u = 0; u < TC; ++u) zv[u] = (rok && cok[u]) ? Zb[r * zr + c0 + 16 * u + cl] : zero;
    }
    for (; r0 < r_end; r0 += 4 * nsub) {
        double2 xn[TA], zn[TC];
        const int64_t r = r0 + 4 * nsub + kr;
        const bool rok = r < r_end;
#pragma unroll
        for (int t = 0; t < TA; ++t) xn[t] = (rok && aok[t]) ? Xb[r * xr + a0 + 16 * t + cl] : zero;
#pragma unroll
        for (int u = 0; u < TC; ++u) zn[u] = (rok && cok[u]) ? Zb[r * zr + c0 + 16 * u + cl] : zero;
#pragma unroll
        for (int t = 0; t < TA; ++t)
#pragma unroll
            for (int u = 0; u < TC; ++u) {
                gr[t][u] = __builtin_amdgcn_mfma_f64_16x16x4f64(xv[t].x, zv[u].x, gr[t][u], 0, 0, 0);
                gr[t][u] = __builtin_amdgcn_mfma_f64_16x16x4f64(xv[t].y, zv[u].y, gr[t][u], 0, 0, 0);
                gi[t][u] = __builtin_amdgcn_mfma_f64_16x16x4f64(xv[t].x, zv[u].y, gi[t][u], 0, 0, 0);
                gi[t][u] = __builtin_amdgcn_mfma_f64_16x16x4f64(-xv[t].y, zv[u].x, gi[t][u], 0, 0, 0);
            }
#pragma unroll
        for (int t = 0; t < TA; ++t) xv[t] = xn[t];
#pragma unroll
        for (int u = 0; u < TC; ++u) zv[u] = zn[u];
    }
    const int64_t nchs = (int64_t)gridDim.x * nsub;
    const int64_t ch = (int64_t)blockIdx.x * nsub + sub;
    double2 *pb = part + (b * nchs + ch) * na * nc;
#pragma unroll
    for (int t = 0; t < TA; ++t)
#pragma unroll
        for (int u = 0; u < TC; ++u)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int64_t a = a0 + 16 * t + 4 * j + kr, c = c0 + 16 * u + cl;
                if (a < na && c < nc) pb[a * nc + c] = make_double2(gr[t][u][j], gi[t][u][j]);
            }
}

// Tile shape of the MFMA Gram: 16/32/48 columns per wave side.
static int bgram_tside(int64_t n) { return n <= 16 ? 1 : (n <= 32 ? 2 : 3); }

struct bgram_mfma_plan {
    int TA, TC, ntc, ntiles, nsub;
    int64_t chunk, nch;
};

static bgram_mfma_plan bgram_mfma_shape(int64_t nbatch, int64_t nrows, int64_t na, int64_t nc)
{
    bgram_mfma_plan p;
    p.TA = bgram_tside(na);
    p.TC = bgram_tside(nc);
    const int64_t nta = (na + 16 * p.TA - 1) / (16 * p.TA);
    p.ntc = (int)((nc + 16 * p.TC - 1) / (16 * p.TC));
    p.ntiles = (int)(nta * p.ntc);
    p.nsub = p.ntiles >= 4 ? 1 : (p.ntiles >= 2 ? 2 : 4);
    const int64_t wg_per_chunk = ((p.ntiles + (4 / p.nsub) - 1) / (4 / p.nsub)) * nbatch;
    int64_t nch = (2048 + wg_per_chunk - 1) / wg_per_chunk;
    if (nch < 1) nch = 1;
    int64_t chunk = (nrows + nch - 1) / nch;
    if (chunk < 64) chunk = 64;
    chunk = (chunk + 3) / 4 * 4;
    p.chunk = chunk;
    p.nch = (nrows + chunk - 1) / chunk;
    return p;
}

template <int TA, int TC>
static void launch_bgram_mfma(const bgram_mfma_plan &p, int64_t nbatch, int64_t nrows, int64_t na, int64_t nc,
                              const double2 *X, int64_t xb, int64_t xr, const double2 *Z, int64_t zb, int64_t zr,
                              double2 *part, hipStream_t s)
{
    const unsigned gy = (unsigned)((p.ntiles + (4 / p.nsub) - 1) / (4 / p.nsub));
    hipLaunchKernelGGL((k_bgram_mfma<TA, TC>), dim3((unsigned)p.nch, gy, (unsigned)nbatch), dim3(256), 0, s, nrows, na,
                       nc, p.chunk, p.ntc, p.ntiles, p.nsub, X, xb, xr, Z, zb, zr, part);
}

static int64_t bgram_chunk(int64_t nbatch, int64_t nrows, int64_t na, int64_t nc)
{
    // aim for ~2048 workgroups in flight in total, chunks of >= 256 rows
    const int64_t tiles = ((na + BG_T - 1) / BG_T) * ((nc + BG_T - 1) / BG_T) * nbatch;
    int64_t nch = (4096 + tiles - 1) / tiles;
    if (nch < 1) nch = 1;
    int64_t chunk = (nrows + nch - 1) / nch;
    if (chunk < 256) chunk = 256;
    chunk = (chunk + BG_T - 1) / BG_T * BG_T;
    return chunk;
}

// ----------------------------------------------------------------- bapply
__global__ __launch_bounds__(256) void k_bapply(
    int64_t nrows, int64_t na, int64_t nc, int ncpad, int rb,
    const double2 *X, int64_t xb, int64_t xr, int64_t xa, const double2 *__restrict__ C, double alpha,
    const double2 *Y, int64_t yb, int64_t yr, int64_t yc, double2 *Out, int64_t ob, int64_t orow, int64_t oc)
{
    extern __shared__ double2 sxr[];          // rb x na
    const int64_t b = blockIdx.y;
    const int64_t r0 = (int64_t)blockIdx.x * rb;
    const double2 *Xb = X + b * xb;
    for (int i = threadIdx.x; i < rb * na; i += blockDim.x) {
        const int rr = i / (int)na, a = i % (int)na;
        const int64_t r = r0 + rr;
        sxr[i] = (r < nrows) ? Xb[r * xr + a * xa] : make_double2(0, 0);
    }
    __syncthreads();
    const int rr = threadIdx.x / ncpad, c = threadIdx.x % ncpad;
    const int64_t r = r0 + rr;
    if (rr >= rb || r >= nrows || c >= nc) return;
    const double2 *Cb = C + b * na * nc;
    double2 acc = make_double2(0, 0);
    if (Y != nullptr) {
        const double2 yv = Y[b * yb + r * yr + c * yc];
        acc.x = alpha * yv.x; acc.y = alpha * yv.y;
    }
    const double2 *xrow = sxr + rr * na;
    for (int64_t a = 0; a < na; ++a) {
        const double2 xv = xrow[a], cv = Cb[a * nc + c];
        acc.x += xv.x * cv.x - xv.y * cv.y;
        acc.y += xv.x * cv.y + xv.y * cv.x;
    }
    Out[b * ob + r * orow + c * oc] = acc;
}

// FP64 MFMA batched apply for unit column strides:
//   Out_b(r, c) = alpha Y_b(r, c) + sum_a X_b(r, a) C_b(a, c).
// grid.y = column tile of 16 TC columns; the workgroup stages C_b(:, tile) in
// LDS once and its waves walk 32-row tiles (grid-stride).  A(16x4): lane i ->
// X(r0 + i%16, a0 + i/16); B(4x16): lane i -> C(a0 + i/16, c0 + i%16);
// D: lane i holds Out(r0 + 4j + i/16, c0 + i%16).  With one column tile a
// wave reads all of its rows before it writes them, so Out may alias X; Out
// may always alias Y (read and written by the same lane).
template <int TC>
__global__ __launch_bounds__(256) void k_bapply_mfma(
    int64_t nrows, int64_t na, int64_t nc, const double2 *X, int64_t xb, int64_t xr,
    const double2 *__restrict__ C, double alpha, const double2 *Y, int64_t yb, int64_t yr,
    double2 *Out, int64_t ob, int64_t orow)
{
    extern __shared__ double2 cs_b[];               // [na4][16 TC]
    constexpr int CW = 16 * TC;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int64_t b = blockIdx.z;
    const int64_t c0 = (int64_t)blockIdx.y * CW;
    const int64_t na4 = (na + 3) & ~(int64_t)3;
    const double2 *Cb = C + b * na * nc;
    for (int64_t e = threadIdx.x; e < na4 * CW; e += blockDim.x) {
        const int64_t a = e / CW, c = c0 + e % CW;
        cs_b[e] = (a < na && c < nc) ? Cb[a * nc + c] : make_double2(0, 0);
    }
    __syncthreads();
    const int kr = lane >> 4, cl = lane & 15;
    const double2 *Xb = X + b * xb;
    const double2 zero = make_double2(0, 0);
    const int64_t ntile = (nrows + 31) / 32;
    for (int64_t tile = (int64_t)blockIdx.x * 4 + wv; tile < ntile; tile += (int64_t)gridDim.x * 4) {
        const int64_t r0 = tile * 32;
        dvec4 dr[2][TC], di[2][TC];
#pragma unroll
        for (int t = 0; t < 2; ++t)
#pragma unroll
            for (int u = 0; u < TC; ++u) { dr[t][u] = (dvec4){0, 0, 0, 0}; di[t][u] = (dvec4){0, 0, 0, 0}; }
        bool rok[2];
#pragma unroll
        for (int t = 0; t < 2; ++t) rok[t] = r0 + 16 * t + cl < nrows;
        double2 xv[2];
#pragma unroll
        for (int t = 0; t < 2; ++t) xv[t] = (rok[t] && kr < na) ? Xb[(r0 + 16 * t + cl) * xr + kr] : zero;
        for (int64_t a0 = 0; a0 < na4; a0 += 4) {
            const int64_t a = a0 + kr, an = a + 4;
            double2 xn[2];
#pragma unroll
            for (int t = 0; t < 2; ++t) xn[t] = (rok[t] && an < na) ? Xb[(r0 + 16 * t + cl) * xr + an] : zero;
            const double2 *crow = cs_b + a * CW + cl;
#pragma unroll
            for (int u = 0; u < TC; ++u) {
                const double2 cv = crow[16 * u];
#pragma unroll
                for (int t = 0; t < 2; ++t) {
                    dr[t][u] = __builtin_amdgcn_mfma_f64_16x16x4f64(xv[t].x, cv.x, dr[t][u], 0, 0, 0);
                    dr[t][u] = __builtin_amdgcn_mfma_f64_16x16x4f64(-xv[t].y, cv.y, dr[t][u], 0, 0, 0);
                    di[t][u] = __builtin_amdgcn_mfma_f64_16x16x4f64(xv[t].x, cv.y, di[t][u], 0, 0, 0);
                    di[t][u] = __builtin_amdgcn_mfma_f64_16x16x4f64(xv[t].y, cv.x, di[t][u], 0, 0, 0);
                }
            }
#pragma unroll
            for (int t = 0; t < 2; ++t) xv[t] = xn[t];
        }
#pragma unroll
        for (int t = 0; t < 2; ++t)
#pragma unroll
            for (int u = 0; u < TC; ++u)
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const int64_t r = r0 + 16 * t + 4 * j + kr, c = c0 + 16 * u + cl;
                    if (r < nrows && c < nc) {
                        double2 o = make_double2(dr[t][u][j], di[t][u][j]);
                        if (Y != nullptr) {
                            const double2 yv = Y[b * yb + r * yr + c];
                            o.x += alpha * yv.x; o.y += alpha * yv.y;
                        }
                        Out[b * ob + r * orow + c] = o;
                    }
                }
    }
}

}  // namespace rsvdt

using namespace rsvdt;

extern "C" int rsvdt_hip_coef_gemm(int64_t P, int64_t K, int64_t L, const double *coef, int64_t cq, int64_t cp,
                                   int64_t p0, int sgn, int64_t cmod, double alpha, const double *in, int64_t ldi,
                                   void *out, int64_t ldo, int out_prec, double beta, rsvdt_stream_t s)
{
    if (P <= 0 || L <= 0) return RSVDT_HIP_OK;
    if (K <= 0 || cmod <= 0 || (out_prec != 64 && out_prec != 32) || P > 65535LL * 16) {
        set_error_msg("rsvdt_hip_coef_gemm: bad arguments");
        return RSVDT_HIP_ERR_ARG;
    }
    hipStream_t st = as_stream(s);
    const int64_t Kpad = (K + 3) & ~(int64_t)3;
    // LDS of the coefficient block for a row tile of 16 T rows (padded pitch)
    auto lds_of = [Kpad](int T) { return Kpad * (16 * T + 4) * 16; };
    if (lds_of(1) > 150 * 1024) {
        set_error_msg("rsvdt_hip_coef_gemm: K too large for the LDS coefficient block (K <= 480)");
        return RSVDT_HIP_ERR_ARG;
    }
    // widest row tile whose coefficient block fits in 144 KB of LDS
    if (P > 64 && lds_of(8) <= 144 * 1024)
        return launch_coef<8>(P, K, L, coef, cq, cp, p0, sgn, cmod, alpha, in, ldi, out, ldo, out_prec, beta, st);
    if (P > 32 && lds_of(4) <= 144 * 1024)
        return launch_coef<4>(P, K, L, coef, cq, cp, p0, sgn, cmod, alpha, in, ldi, out, ldo, out_prec, beta, st);
    if (P > 16 && lds_of(2) <= 144 * 1024)
        return launch_coef<2>(P, K, L, coef, cq, cp, p0, sgn, cmod, alpha, in, ldi, out, ldo, out_prec, beta, st);
    return launch_coef<1>(P, K, L, coef, cq, cp, p0, sgn, cmod, alpha, in, ldi, out, ldo, out_prec, beta, st);
}

extern "C" size_t rsvdt_hip_bgram_work_bytes(int64_t nbatch, int64_t nrows, int64_t na, int64_t nc)
{
    const int64_t chunk = bgram_chunk(nbatch, nrows, na, nc);
    const int64_t nch = (nrows + chunk - 1) / chunk;
    const bgram_mfma_plan p = bgram_mfma_shape(nbatch, nrows, na, nc);
    int64_t parts = p.nch * p.nsub;
    if (nch > parts) parts = nch;
    return (size_t)(nbatch * parts * na * nc) * 16 + 256;
}

extern "C" int rsvdt_hip_bgram(int64_t nbatch, int64_t nrows, int64_t na, int64_t nc, const double *X, int64_t xb,
                               int64_t xr, int64_t xa, const double *Z, int64_t zb, int64_t zr, int64_t zc, double *G,
                               void *work, rsvdt_stream_t s)
{
    if (nbatch <= 0 || na <= 0 || nc <= 0) return RSVDT_HIP_OK;
    if (nrows <= 0) return rsvdt_hip_zero(G, (size_t)(nbatch * na * nc) * 16, s);
    if (nbatch > 65535) { set_error_msg("rsvdt_hip_bgram: nbatch > 65535"); return RSVDT_HIP_ERR_ARG; }
    hipStream_t st = as_stream(s);
    int64_t nparts;
    if (xa == 1 && zc == 1) {
        const bgram_mfma_plan p = bgram_mfma_shape(nbatch, nrows, na, nc);
        const double2 *X2 = (const double2 *)X, *Z2 = (const double2 *)Z;
        double2 *W = (double2 *)work;
#define RSVDT_BG(A, C) \
        if (p.TA == A && p.TC == C) launch_bgram_mfma<A, C>(p, nbatch, nrows, na, nc, X2, xb, xr, Z2, zb, zr, W, st);
        RSVDT_BG(1, 1) RSVDT_BG(1, 2) RSVDT_BG(1, 3) RSVDT_BG(2, 1) RSVDT_BG(2, 2) RSVDT_BG(2, 3)
        RSVDT_BG(3, 1) RSVDT_BG(3, 2) RSVDT_BG(3, 3)
#undef RSVDT_BG
        RSVDT_LAUNCH_CHECK("rsvdt_hip_bgram(mfma)");
        nparts = p.nch * p.nsub;
    } else {
        const int64_t chunk = bgram_chunk(nbatch, nrows, na, nc);
        const int64_t nch = (nrows + chunk - 1) / chunk;
        const int nta = (int)((na + BG_T - 1) / BG_T), ntc = (int)((nc + BG_T - 1) / BG_T);
        hipLaunchKernelGGL(k_bgram_part, dim3((unsigned)nch, (unsigned)(nta * ntc), (unsigned)nbatch), dim3(256), 0,
                           st, nrows, na, nc, chunk, ntc, (const double2 *)X, xb, xr, xa, (const double2 *)Z, zb, zr,
                           zc, (double2 *)work);
        RSVDT_LAUNCH_CHECK("rsvdt_hip_bgram(part)");
        nparts = nch;
    }
    const int64_t tot = nbatch * na * nc;
    hipLaunchKernelGGL(k_bgram_reduce, dim3(ceil_div(tot, 256)), dim3(256), 0, st, nbatch, nparts, na, nc,
                       (const double2 *)work, (double2 *)G);
    RSVDT_LAUNCH_CHECK("rsvdt_hip_bgram(reduce)");
    return RSVDT_HIP_OK;
}

template <int TC>
static int launch_bapply_mfma(int64_t nbatch, int64_t nrows, int64_t na, int64_t nc, const double *X, int64_t xb,
                              int64_t xr, const double *C, double alpha, const double *Y, int64_t yb, int64_t yr,
                              double *Out, int64_t ob, int64_t orow, hipStream_t s)
{
    const int64_t nct = (nc + 16 * TC - 1) / (16 * TC);
    const int64_t na4 = (na + 3) & ~(int64_t)3;
    const size_t lds = (size_t)na4 * 16 * TC * 16;
    const int64_t ntile = (nrows + 31) / 32;
    int64_t gx = (ntile + 3) / 4;
    const int64_t cap = (2048 + nct * nbatch - 1) / (nct * nbatch);
    if (gx > cap) gx = cap < 1 ? 1 : cap;
    hipLaunchKernelGGL((k_bapply_mfma<TC>), dim3((unsigned)gx, (unsigned)nct, (unsigned)nbatch), dim3(256), lds, s,
                       nrows, na, nc, (const double2 *)X, xb, xr, (const double2 *)C, alpha, (const double2 *)Y, yb,
                       yr, (double2 *)Out, ob, orow);
    RSVDT_LAUNCH_CHECK("rsvdt_hip_bapply(mfma)");
    return RSVDT_HIP_OK;
}

extern "C" int rsvdt_hip_bapply(int64_t nbatch, int64_t nrows, int64_t na, int64_t nc, const double *X, int64_t xb,
                                int64_t xr, int64_t xa, const double *C, double alpha, const double *Y, int64_t yb,
                                int64_t yr, int64_t yc, double *Out, int64_t ob, int64_t orow, int64_t oc,
                                rsvdt_stream_t s)
{
    if (nbatch <= 0 || nrows <= 0 || nc <= 0) return RSVDT_HIP_OK;
    if (nbatch > 65535) {
        set_error_msg("rsvdt_hip_bapply: nbatch <= 65535 required");
        return RSVDT_HIP_ERR_ARG;
    }
    hipStream_t st = as_stream(s);
    const bool unit = xa == 1 && oc == 1 && (Y == nullptr || yc == 1);
    const int64_t na4 = (na + 3) & ~(int64_t)3;
    if (unit && na > 0 && na4 * 64 * 16 <= 150 * 1024) {
        const int TC = nc <= 16 ? 1 : (nc <= 32 ? 2 : (nc <= 48 ? 3 : (nc <= 64 ? 4 : 3)));
        const bool multi_tile = nc > 16 * TC;
        if (multi_tile && X == Out) {
            set_error_msg("rsvdt_hip_bapply: in place needs nc <= 64 with unit strides");
            return RSVDT_HIP_ERR_ARG;
        }
        switch (TC) {
        case 1: return launch_bapply_mfma<1>(nbatch, nrows, na, nc, X, xb, xr, C, alpha, Y, yb, yr, Out, ob, orow, st);
        case 2: return launch_bapply_mfma<2>(nbatch, nrows, na, nc, X, xb, xr, C, alpha, Y, yb, yr, Out, ob, orow, st);
        case 3: return launch_bapply_mfma<3>(nbatch, nrows, na, nc, X, xb, xr, C, alpha, Y, yb, yr, Out, ob, orow, st);
        default: return launch_bapply_mfma<4>(nbatch, nrows, na, nc, X, xb, xr, C, alpha, Y, yb, yr, Out, ob, orow, st);
        }
    }
    if (nc > 256 || na > 512) {
        set_error_msg("rsvdt_hip_bapply: nc <= 256, na <= 512 required");
        return RSVDT_HIP_ERR_ARG;
    }
    int ncpad = 1;
    while (ncpad < nc) ncpad <<= 1;
    int rb = 256 / ncpad;
    const int64_t rb_lds = 65536 / ((na > 0 ? na : 1) * 16);
    if (rb > rb_lds) rb = (int)(rb_lds > 0 ? rb_lds : 1);
    const size_t lds = (size_t)rb * (size_t)(na > 0 ? na : 1) * 16;
    hipLaunchKernelGGL(k_bapply, dim3(ceil_div(nrows, rb), (unsigned)nbatch), dim3(256), lds, st, nrows,
                       na, nc, ncpad, rb, (const double2 *)X, xb, xr, xa, (const double2 *)C, alpha,
                       (const double2 *)Y, yb, yr, yc, (double2 *)Out, ob, orow, oc);
    RSVDT_LAUNCH_CHECK("rsvdt_hip_bapply");
    return RSVDT_HIP_OK;
}
