// Batched small dense linear algebra for gfx950, one workgroup per matrix:
//   * Cholesky (with an optional diagonal shift) and triangular inverse: the
//     small factorizations of (shifted) CholeskyQR, which replaces SLEPc's
//     BVOrthogonalize in SourceCode/QRAllFreqs.c:22-32 and QRDecomposition.c:18;
//   * one-sided Jacobi SVD of the k x k triangular factor, replacing SLEPc's
//     SVDSolve in SourceCode/SVDAllFreqsBeforeAdjoint.c:31-49 and SVDAllFreqs.c:34-56;
//   * LU solves of (e^{+-i w m dT} I - M_tilde) a = b for every (test vector,
//     frequency), replacing the PGMRES KSPSolve of SourceCode/GalerkinProj.c:57-62
//     by the exact solve it converges to.
// The matrices are tiny (n <= 130) and there are tens to a few thousand of
// them, so these kernels are latency-bound; one launch covers all of them.
#include "common.h"

namespace rsvdt {

// ------------------------------------------------------------ Cholesky + inverse
__global__ __launch_bounds__(256) void k_bchol_inv(int n, const double2 *__restrict__ G, double shift_rel,
                                                   double drop_tol, double2 *__restrict__ R,
                                                   double2 *__restrict__ Rinv, int *info)
{
    const int64_t b = blockIdx.x;
    const double2 *Gb = G + b * n * n;
    double2 *Rb = R + b * n * n;
    double2 *Xb = Rinv + b * n * n;
    __shared__ double s_shift;
    __shared__ double s_piv;
    __shared__ double s_dmax;
    __shared__ int s_bad;
    if (threadIdx.x == 0) {
        double tr = 0, dm = 0;
        for (int i = 0; i < n; ++i) { tr += Gb[i * n + i].x; dm = fmax(dm, Gb[i * n + i].x); }
        s_shift = shift_rel * tr;
        s_dmax = dm;
        s_bad = 0;
    }
    __syncthreads();
    for (int e = threadIdx.x; e < n * n; e += blockDim.x) {
        const int i = e / n, c = e % n;
        double2 v = (i <= c) ? Gb[e] : make_double2(0, 0);
        if (i == c) { v.x += s_shift; v.y = 0; }
        Rb[e] = v;
    }
    __syncthreads();
    for (int j = 0; j < n; ++j) {
        if (threadIdx.x == 0) {
            double d = Rb[j * n + j].x;
            if (drop_tol > 0 && !(d > drop_tol * s_dmax)) {
                s_piv = 0;                      // column j lies in the span of columns < j: drop it
            } else {
                const double tiny = 1e-300;
                if (!(d > tiny)) { d = tiny; s_bad++; }
                s_piv = sqrt(d);
            }
            Rb[j * n + j] = make_double2(s_piv, 0);
        }
        __syncthreads();
        const double inv = s_piv > 0 ? 1.0 / s_piv : 0.0;
        for (int c = j + 1 + threadIdx.x; c < n; c += blockDim.x) {
            double2 v = Rb[j * n + c];
            Rb[j * n + c] = make_double2(v.x * inv, v.y * inv);
        }
        __syncthreads();
        const int m = n - j - 1;
        for (int e = threadIdx.x; e < m * m; e += blockDim.x) {
            const int i = j + 1 + e / m, c = j + 1 + e % m;
            if (i <= c) {
                const double2 ri = Rb[j * n + i], rc = Rb[j * n + c];
                double2 v = Rb[i * n + c];
                // v -= conj(ri) * rc
                v.x -= ri.x * rc.x + ri.y * rc.y;
                v.y -= ri.x * rc.y - ri.y * rc.x;
                Rb[i * n + c] = v;
            }
        }
        __syncthreads();
    }
    // inverse of the upper-triangular factor, one column per thread
    // (dropped columns have a zero row in R and a zero column in R^-1)
    for (int c = threadIdx.x; c < n; c += blockDim.x) {
        for (int i = c + 1; i < n; ++i) Xb[i * n + c] = make_double2(0, 0);
        const double rcc = Rb[c * n + c].x;
        if (!(rcc > 0)) {
            for (int i = 0; i <= c; ++i) Xb[i * n + c] = make_double2(0, 0);
            continue;
        }
        Xb[c * n + c] = make_double2(1.0 / rcc, 0);
        for (int i = c - 1; i >= 0; --i) {
            const double rii = Rb[i * n + i].x;
            if (!(rii > 0)) { Xb[i * n + c] = make_double2(0, 0); continue; }
            double2 s = make_double2(0, 0);
            for (int l = i + 1; l <= c; ++l) s = zadd(s, zmul(Rb[i * n + l], Xb[l * n + c]));
            Xb[i * n + c] = make_double2(-s.x / rii, -s.y / rii);
        }
    }
    if (threadIdx.x == 0 && info != nullptr && s_bad) atomicAdd(info, s_bad);
}

// The same for n <= 16 NB (NB <= 5) with the matrix in registers: thread
// (tr, tc) = (tid / 16, tid % 16) of the 16 x 16 grid owns the entries
// (tr + 16a, tc + 16b), so that every step's work stays spread over the
// whole workgroup as the trailing block shrinks; the blocks a > b lie below
// the diagonal for every thread and are not held.  Cholesky, right-looking:
// at step j the owners of row j put it in LDS (double-buffered, so one
// barrier per step), every thread takes the pivot from there and updates its
// entries with R(i, c) -= conj(R(j, i)) R(j, c) / d unconditionally, the
// multipliers being zero outside the trailing block, so the step has no
// branches; row j keeps its unscaled values and is scaled once at the end
// from the pivots kept in LDS.  Inverse, solving R X = I from the bottom row
// up the same way: at step i the owners of column i of R and of row i of the
// right-hand side put them in LDS, one barrier, B(r, :) -= R(r, i) X(i, :)
// for r < i with X(i, :) = B(i, :) / R(i, i) from the kept reciprocals.
// 2n barriers, no division inside the steps but the pivot's own.  The steps
// run in blocks of 16 with the block index known at compile time, so the
// finished blocks drop out of the update and the owners of the broadcast
// row or column store it without per-element guards.
__device__ __forceinline__ void zfms(double2 &w, double2 l, double2 u)
{
    // w -= l * u
    w.x = fma(-l.x, u.x, fma(l.y, u.y, w.x));
    w.y = fma(-l.x, u.y, fma(-l.y, u.x, w.y));
}

template <int NB>
__global__ __launch_bounds__(256) void k_bchol_inv_reg(int n, const double2 *__restrict__ G, double shift_rel,
                                                       double drop_tol, double2 *__restrict__ R,
                                                       double2 *__restrict__ Rinv, int *info)
{
    constexpr int NP = 16 * NB;
    __shared__ double2 rowb[2][NP], colb[2][NP];
    __shared__ double s_diag[NP], s_inv[NP], s_piv[NP];
    __shared__ double s_shift, s_dmax;
    const int tid = threadIdx.x, tr = tid / 16, tc = tid % 16;
    const int64_t b = blockIdx.x;
    const double2 *Gb = G + b * n * n;
    double2 Rr[NB][NB];                      // [a][q], a <= q used
#pragma unroll
    for (int a = 0; a < NB; ++a)
#pragma unroll
        for (int q = a; q < NB; ++q) {
            const int i = tr + 16 * a, c = tc + 16 * q;
            Rr[a][q] = (i < n && c < n && i <= c) ? Gb[i * n + c] : make_double2(0, 0);
            if (i == c && i < n) s_diag[i] = Rr[a][q].x;
        }
    __syncthreads();
    if (tid < 64) {                          // trace and largest diagonal entry: one wave
        double tr2 = 0, dm = 0;
        for (int i = tid; i < n; i += 64) { tr2 += s_diag[i]; dm = fmax(dm, s_diag[i]); }
        for (int o = 32; o > 0; o >>= 1) {
            tr2 += __shfl_xor(tr2, o, 64);
            dm = fmax(dm, __shfl_xor(dm, o, 64));
        }
        if (tid == 0) { s_shift = shift_rel * tr2; s_dmax = dm; }
    }
    __syncthreads();
    const double shift = s_shift, dmax = s_dmax;
    if (tr == tc) {
#pragma unroll
        for (int a = 0; a < NB; ++a) { Rr[a][a].x += shift; Rr[a][a].y = 0; }
    }
    int bad = 0;
    // steps in blocks of 16 (a0 at compile time): row j = 16 a0 + t lives in
    // register block a0, and the blocks above a0 are finished
#pragma unroll
    for (int a0 = 0; a0 < NB; ++a0) {
        const int jend = n < 16 * (a0 + 1) ? n : 16 * (a0 + 1);
        for (int j = 16 * a0; j < jend; ++j) {
            const int buf = j & 1;
            if (tr == j - 16 * a0) {         // row j (its columns below j are not read)
#pragma unroll
                for (int q = a0; q < NB; ++q) rowb[buf][tc + 16 * q] = Rr[a0][q];
            }
            __syncthreads();
            double d = rowb[buf][j].x, piv;
            if (drop_tol > 0 && !(d > drop_tol * dmax)) {
                piv = 0;                        // column j lies in the span of columns < j: drop it
            } else {
                if (!(d > 1e-300)) { d = 1e-300; bad++; }
                piv = sqrt(d);
            }
            const double inv = piv > 0 ? 1.0 / piv : 0.0;
            if (tid == 0) { s_piv[j] = piv; s_inv[j] = inv; }
            const double inv2 = inv * inv;
            double2 ri[NB], rc[NB];
#pragma unroll
            for (int a = a0; a < NB; ++a) {
                const int i = tr + 16 * a;
                const double2 v = rowb[buf][i];
                const bool on = i > j && i < n;
                ri[a] = make_double2(on ? v.x * inv2 : 0.0, on ? -v.y * inv2 : 0.0);   // conj(R(j, i)) / d
            }
#pragma unroll
            for (int q = a0; q < NB; ++q) {
                const int c = tc + 16 * q;
                const double2 v = rowb[buf][c];
                const bool on = c > j && c < n;
                rc[q] = make_double2(on ? v.x : 0.0, on ? v.y : 0.0);
            }
#pragma unroll
            for (int a = a0; a < NB; ++a)
#pragma unroll
                for (int q = a; q < NB; ++q) zfms(Rr[a][q], ri[a], rc[q]);
        }
    }
    __syncthreads();                         // pivots and reciprocals visible; row buffers free
    // R(i, i) = piv_i, R(i, c) *= inv_i (c > i)
#pragma unroll
    for (int a = 0; a < NB; ++a) {
        const int i = tr + 16 * a;
        const double inv = i < n ? s_inv[i] : 0.0, piv = i < n ? s_piv[i] : 0.0;
#pragma unroll
        for (int q = a; q < NB; ++q) {
            const int c = tc + 16 * q;
            Rr[a][q] = c == i ? make_double2(piv, 0) : make_double2(Rr[a][q].x * inv, Rr[a][q].y * inv);
        }
    }
    // X = R^-1 from R X = I, bottom row first; B holds X's rows unscaled
    double2 Br[NB][NB];
#pragma unroll
    for (int a = 0; a < NB; ++a)
#pragma unroll
        for (int q = a; q < NB; ++q)
            Br[a][q] = make_double2((tr + 16 * a == tc + 16 * q && tr + 16 * a < n) ? 1.0 : 0.0, 0);
#pragma unroll
    for (int a0 = NB - 1; a0 >= 0; --a0) {
        const int iend = (n < 16 * (a0 + 1) ? n : 16 * (a0 + 1)) - 1;
        for (int i = iend; i >= 16 * a0; --i) {
            const int buf = i & 1;
            if (tc == i - 16 * a0) {           // column i of R (its rows above i are read)
#pragma unroll
                for (int a = 0; a <= a0; ++a) colb[buf][tr + 16 * a] = Rr[a][a0];
            }
            if (tr == i - 16 * a0) {           // row i of B (its columns from i on are read)
#pragma unroll
                for (int q = a0; q < NB; ++q) rowb[buf][tc + 16 * q] = Br[a0][q];
            }
            __syncthreads();
            const double inv = s_inv[i];       // 0 for a dropped column: zero row (and column) of X
            double2 xi[NB], rr[NB];
#pragma unroll
            for (int q = a0; q < NB; ++q) {
                const int c = tc + 16 * q;
                const double2 v = rowb[buf][c];
                const bool on = c >= i && c < n;
                xi[q] = make_double2(on ? v.x * inv : 0.0, on ? v.y * inv : 0.0);
            }
#pragma unroll
            for (int a = 0; a <= a0; ++a) {
                const double2 v = colb[buf][tr + 16 * a];
                const bool on = tr + 16 * a < i;
                rr[a] = make_double2(on ? v.x : 0.0, on ? v.y : 0.0);
            }
#pragma unroll
            for (int a = 0; a <= a0; ++a)
#pragma unroll
                for (int q = a0; q < NB; ++q) zfms(Br[a][q], rr[a], xi[q]);
        }
    }
    double2 *Rb = R + b * n * n, *Xb = Rinv + b * n * n;
#pragma unroll
    for (int a = 0; a < NB; ++a) {
        const int i = tr + 16 * a;
        const double inv = i < n ? s_inv[i] : 0.0;
#pragma unroll
        for (int q = 0; q < NB; ++q) {
            const int c = tc + 16 * q;
            if (i < n && c < n) {
                const bool up = q >= a && i <= c;
                const double2 rv = q >= a ? Rr[a][q] : make_double2(0, 0);
                const double2 bv = q >= a ? Br[a][q] : make_double2(0, 0);
                Rb[i * n + c] = up ? rv : make_double2(0, 0);
                Xb[i * n + c] = up ? make_double2(bv.x * inv, bv.y * inv) : make_double2(0, 0);
            }
        }
    }
    if (tid == 0 && info != nullptr && bad) atomicAdd(info, bad);
}

// D_b = diag(1 where R_b(j,j) is not positive, else 0): the columns that a
// dropping CholeskyQR zeroed (their R has a zero pivot); count += their number.
__global__ void k_bdrop_mask(int64_t tot, int n, const double2 *__restrict__ R, double2 *__restrict__ D, int *count)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= tot) return;
    const int64_t b = e / ((int64_t)n * n);
    const int rem = (int)(e % ((int64_t)n * n));
    const int i = rem / n, c = rem % n;
    double v = 0.0;
    if (i == c && !(R[b * n * n + (int64_t)i * n + i].x > 0)) {
        v = 1.0;
        if (count != nullptr) atomicAdd(count, 1);
    }
    D[e] = make_double2(v, 0.0);
}

__global__ void k_btrmm_upper(int n, const double2 *__restrict__ A, const double2 *__restrict__ B,
                              double2 *__restrict__ C)
{
    const int64_t b = blockIdx.x;
    const double2 *Ab = A + b * n * n, *Bb = B + b * n * n;
    double2 *Cb = C + b * n * n;
    for (int e = threadIdx.x; e < n * n; e += blockDim.x) {
        const int i = e / n, c = e % n;
        double2 s = make_double2(0, 0);
        for (int l = i; l <= c; ++l) s = zadd(s, zmul(Ab[i * n + l], Bb[l * n + c]));
        Cb[e] = s;
    }
}

// ------------------------------------------------------------ Jacobi SVD
// Columns of R live in LDS (column-major); a round-robin tournament pairs the
// columns so that n/2 rotations run concurrently, one thread per pair.
#define JAC_MAXN 64
__global__ __launch_bounds__(256) void k_bjacobi_svd(int n, const double2 *__restrict__ R, double2 *__restrict__ U,
                                                     double *__restrict__ sig)
{
    __shared__ double2 A[JAC_MAXN * JAC_MAXN];
    __shared__ double nrm[JAC_MAXN];
    __shared__ int ord[JAC_MAXN];
    __shared__ int s_rot;
    const int64_t b = blockIdx.x;
    const double2 *Rb = R + b * n * n;
    for (int e = threadIdx.x; e < n * n; e += blockDim.x) {
        const int i = e / n, c = e % n;
        A[c * n + i] = Rb[e];
    }
    __syncthreads();
    const int ne = (n + 1) & ~1;       // players, padded to even
    const int m = ne - 1;
    for (int sweep = 0; sweep < 60; ++sweep) {
        if (threadIdx.x == 0) s_rot = 0;
        __syncthreads();
        for (int rnd = 0; rnd < m; ++rnd) {
            const int p = threadIdx.x;
            if (p < ne / 2) {
                int i, j;
                if (p == 0) { i = rnd; j = m; }
                else { i = (rnd + p) % m; j = (rnd - p + m) % m; }
                if (i > j) { int t = i; i = j; j = t; }
                if (j < n) {
                    double2 *ci = A + i * n, *cj = A + j * n;
                    double al = 0, be = 0;
                    double2 ga = make_double2(0, 0);
                    for (int r = 0; r < n; ++r) {
                        const double2 x = ci[r], y = cj[r];
                        al += zabs2(x);
                        be += zabs2(y);
                        ga.x += x.x * y.x + x.y * y.y;
                        ga.y += x.x * y.y - x.y * y.x;
                    }
                    const double ag = sqrt(zabs2(ga));
                    if (ag > 0 && ag > 1e-15 * sqrt(al * be)) {
                        s_rot = 1;
                        const double2 eph = make_double2(ga.x / ag, -ga.y / ag);   // e^{-i phi}
                        const double zeta = (be - al) / (2.0 * ag);
                        const double t = (zeta >= 0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
                        const double cs = 1.0 / sqrt(1.0 + t * t), sn = cs * t;
                        for (int r = 0; r < n; ++r) {
                            const double2 xp = ci[r];
                            const double2 xq = zmul(cj[r], eph);
                            ci[r] = make_double2(cs * xp.x - sn * xq.x, cs * xp.y - sn * xq.y);
                            cj[r] = make_double2(sn * xp.x + cs * xq.x, sn * xp.y + cs * xq.y);
                        }
                    }
                }
            }
            __syncthreads();
        }
        if (!s_rot) break;
        __syncthreads();
    }
    for (int c = threadIdx.x; c < n; c += blockDim.x) {
        double s = 0;
        for (int r = 0; r < n; ++r) s += zabs2(A[c * n + r]);
        nrm[c] = sqrt(s);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int c = 0; c < n; ++c) ord[c] = c;
        for (int i = 1; i < n; ++i) {          // insertion sort, descending, stable
            int t = ord[i], j = i - 1;
            while (j >= 0 && nrm[ord[j]] < nrm[t]) { ord[j + 1] = ord[j]; --j; }
            ord[j + 1] = t;
        }
    }
    __syncthreads();
    // phase convention: sum_i U(i, j) real and positive (largest-modulus entry
    // when the sum vanishes), so the vectors do not depend on the rotation order
    __shared__ double2 phs[JAC_MAXN];
    for (int c = threadIdx.x; c < n; c += blockDim.x) {
        const double s = nrm[c];
        double2 sum = make_double2(0, 0), bigv = make_double2(1, 0);
        double big = -1;
        for (int r = 0; r < n; ++r) {
            const double2 v = A[c * n + r];
            const double2 u = s > 0 ? make_double2(v.x / s, v.y / s) : make_double2(0, 0);
            sum = zadd(sum, u);
            const double m2 = zabs2(u);
            if (m2 > big) { big = m2; bigv = u; }
        }
        const double as = sqrt(zabs2(sum)), ab = sqrt(zabs2(bigv));
        phs[c] = as > 1e-8 ? make_double2(sum.x / as, -sum.y / as)
                           : (ab > 0 ? make_double2(bigv.x / ab, -bigv.y / ab) : make_double2(1, 0));
    }
    __syncthreads();
    double2 *Ub = U + b * n * n;
    for (int e = threadIdx.x; e < n * n; e += blockDim.x) {
        const int r = e / n, jj = e % n;
        const int c = ord[jj];
        const double s = nrm[c];
        const double2 v = zmul(A[c * n + r], phs[c]);
        Ub[r * n + jj] = s > 0 ? make_double2(v.x / s, v.y / s) : make_double2(0, 0);
    }
    for (int jj = threadIdx.x; jj < n; jj += blockDim.x) sig[b * n + jj] = nrm[ord[jj]];
}

// ------------------------------------------------------------ shifted solves
// W (column-major n x n, in global scratch) = shift*I - M_b; LU with partial
// pivoting applied to the right-hand side as it proceeds, then back substitution.
__global__ __launch_bounds__(256) void k_bshift_solve(int n, int nrhs, const double2 *__restrict__ M,
                                                      const double2 *__restrict__ shift, const double2 *B,
                                                      double2 *X, double2 *__restrict__ work, int *info)
{
    extern __shared__ double2 x[];             // n
    __shared__ double s_val[256];
    __shared__ int s_idx[256];
    const int64_t sys = blockIdx.x;
    const int64_t b = sys / nrhs;
    const int m = (int)(sys % nrhs);
    const double2 *Mb = M + b * n * n;
    double2 *W = work + sys * (int64_t)n * n;
    const double2 sh = shift[b * nrhs + m];
    for (int e = threadIdx.x; e < n * n; e += blockDim.x) {
        const int i = e / n, c = e % n;
        const double2 v = Mb[e];
        double2 w = make_double2(-v.x, -v.y);
        if (i == c) { w.x += sh.x; w.y += sh.y; }
        W[c * n + i] = w;
    }
    for (int i = threadIdx.x; i < n; i += blockDim.x) x[i] = B[(b * n + i) * nrhs + m];
    __syncthreads();
    int bad = 0;
    for (int j = 0; j < n; ++j) {
        // pivot search in column j
        double best = -1;
        int bi = j;
        for (int i = j + threadIdx.x; i < n; i += blockDim.x) {
            const double a = zabs2(W[j * n + i]);
            if (a > best) { best = a; bi = i; }
        }
        s_val[threadIdx.x] = best;
        s_idx[threadIdx.x] = bi;
        __syncthreads();
        for (int o = 128; o > 0; o >>= 1) {
            if ((int)threadIdx.x < o) {
                const double v2 = s_val[threadIdx.x + o];
                const int i2 = s_idx[threadIdx.x + o];
                if (v2 > s_val[threadIdx.x] || (v2 == s_val[threadIdx.x] && i2 < s_idx[threadIdx.x])) {
                    s_val[threadIdx.x] = v2;
                    s_idx[threadIdx.x] = i2;
                }
            }
            __syncthreads();
        }
        const int p = s_idx[0];
        if (!(s_val[0] > 0)) bad = 1;
        if (p != j) {
            for (int c = j + threadIdx.x; c < n; c += blockDim.x) {
                const double2 t = W[c * n + j];
                W[c * n + j] = W[c * n + p];
                W[c * n + p] = t;
            }
            if (threadIdx.x == 0) { const double2 t = x[j]; x[j] = x[p]; x[p] = t; }
        }
        __syncthreads();
        const double2 d = W[j * n + j];
        const double dd = zabs2(d);
        const double2 dinv = dd > 0 ? make_double2(d.x / dd, -d.y / dd) : make_double2(0, 0);
        // multipliers stored in column j below the diagonal
        for (int i = j + 1 + threadIdx.x; i < n; i += blockDim.x) W[j * n + i] = zmul(W[j * n + i], dinv);
        __syncthreads();
        const int mm = n - j - 1;
        for (int e = threadIdx.x; e < mm * mm; e += blockDim.x) {
            const int c = j + 1 + e / mm, i = j + 1 + e % mm;
            W[c * n + i] = zsub(W[c * n + i], zmul(W[j * n + i], W[c * n + j]));
        }
        if (threadIdx.x == 0) {
            const double2 xj = x[j];
            for (int i = j + 1; i < n; ++i) x[i] = zsub(x[i], zmul(W[j * n + i], xj));
        }
        __syncthreads();
    }
    for (int c = n - 1; c >= 0; --c) {
        if (threadIdx.x == 0) {
            const double2 d = W[c * n + c];
            const double dd = zabs2(d);
            x[c] = dd > 0 ? zmul(x[c], make_double2(d.x / dd, -d.y / dd)) : make_double2(0, 0);
        }
        __syncthreads();
        const double2 xc = x[c];
        for (int i = threadIdx.x; i < c; i += blockDim.x) x[i] = zsub(x[i], zmul(W[c * n + i], xc));
        __syncthreads();
    }
    for (int i = threadIdx.x; i < n; i += blockDim.x) X[(b * n + i) * nrhs + m] = x[i];
    if (threadIdx.x == 0 && bad && info != nullptr) atomicAdd(info, 1);
}

// The same for n + 1 <= 16 NB (NB <= 5) with W = shift I - M_b and the
// right-hand side (column n) in registers: thread (tr, tc) = (tid % 16,
// tid / 16) owns the entries (tr + 16a, tc + 16b), so that the 16 owners of
// a column are 16 adjacent lanes of one wave.  Gauss-Jordan elimination with
// partial pivoting and no row exchanges: at step i the owners of column i
// find the pivot row p among the rows not yet used (a 16-lane shuffle
// reduction, lowest index on ties) and put column i in LDS; one barrier; the
// owners of row p put it in LDS; one barrier; every row r != p takes
// W(r, c) -= (W(r, i) / W(p, i)) W(p, c) for c > i.  At the end row p_i holds
// only W(p_i, i) and its right-hand side, so x_i = W(p_i, n) / W(p_i, i).
// 2n barriers, no back substitution; a zero pivot gives a zero component
// and counts in info.  The steps run in blocks of 16 with the block index
// known at compile time, so the finished column blocks drop out of the
// update, and the update is branch-free (multipliers zeroed where unused).
template <int NB>
__global__ __launch_bounds__(256) void k_bshift_solve_reg(int n, int nrhs, const double2 *__restrict__ M,
                                                          const double2 *__restrict__ shift, const double2 *B,
                                                          double2 *X, int *info)
{
    constexpr int NP = 16 * NB;
    __shared__ double2 colb[2][NP], rowb[2][NP], rhs[NP], pivv[NP];
    __shared__ int s_p[2], perm[NP];
    const int tid = threadIdx.x, tr = tid % 16, tc = tid / 16;
    const int64_t sys = blockIdx.x;
    const int64_t b = sys / nrhs;
    const int m = (int)(sys % nrhs);
    const double2 *Mb = M + b * n * n;
    const double2 sh = shift[b * nrhs + m];
    double2 W[NB][NB];
#pragma unroll
    for (int a = 0; a < NB; ++a)
#pragma unroll
        for (int q = 0; q < NB; ++q) {
            const int r = tr + 16 * a, c = tc + 16 * q;
            double2 w = make_double2(0, 0);
            if (r < n && c < n) {
                const double2 v = Mb[r * n + c];
                w = make_double2(-v.x, -v.y);
                if (r == c) { w.x += sh.x; w.y += sh.y; }
            } else if (r < n && c == n) {
                w = B[(b * n + r) * nrhs + m];
            }
            W[a][q] = w;
        }
    unsigned used = 0;                         // bit a: row tr + 16a is a pivot row
    int bad = 0;
    // steps in blocks of 16 (q0 at compile time): column i = 16 q0 + t lives
    // in register block q0, and the blocks left of q0 are finished
#pragma unroll
    for (int q0 = 0; q0 < NB; ++q0) {
        const int iend = n < 16 * (q0 + 1) ? n : 16 * (q0 + 1);
        for (int i = 16 * q0; i < iend; ++i) {
            const int buf = i & 1;
            if (tc == i - 16 * q0) {           // column i: pivot among the rows not yet used
                double best = -1;
                int bi = n;
#pragma unroll
                for (int a = 0; a < NB; ++a) {
                    const int r = tr + 16 * a;
                    colb[buf][r] = W[a][q0];
                    const double x = zabs2(W[a][q0]);
                    const bool cand = r < n && !((used >> a) & 1) && x > best;
                    best = cand ? x : best;
                    bi = cand ? r : bi;
                }
                for (int o = 8; o > 0; o >>= 1) {
                    const double v2 = __shfl_xor(best, o, 64);
                    const int i2 = __shfl_xor(bi, o, 64);
                    if (v2 > best || (v2 == best && i2 < bi)) { best = v2; bi = i2; }
                }
                if (tr == 0) s_p[buf] = bi;
            }
            __syncthreads();
            const int p = s_p[buf], ap = p / 16;
            if (tr == p % 16) {                // row p (its columns up to i are not read)
#pragma unroll
                for (int a = 0; a < NB; ++a)
                    if (a == ap) {
#pragma unroll
                        for (int q = q0; q < NB; ++q) rowb[buf][tc + 16 * q] = W[a][q];
                    }
                used |= 1u << ap;
            }
            __syncthreads();
            const double2 d = colb[buf][p];
            const double dd = zabs2(d);
            if (!(dd > 0)) bad = 1;
            const double2 dinv = dd > 0 ? make_double2(d.x / dd, -d.y / dd) : make_double2(0, 0);
            if (tid == 0) { perm[i] = p; pivv[i] = dinv; }
            double2 l[NB];
#pragma unroll
            for (int a = 0; a < NB; ++a) {
                const int r = tr + 16 * a;
                const double2 lv = zmul(colb[buf][r], dinv);
                const bool on = r < n && r != p;
                l[a] = make_double2(on ? lv.x : 0.0, on ? lv.y : 0.0);
            }
#pragma unroll
            for (int q = q0; q < NB; ++q) {
                const int c = tc + 16 * q;
                const double2 v = rowb[buf][c];
                const bool on = c > i && c <= n;
                const double2 u = make_double2(on ? v.x : 0.0, on ? v.y : 0.0);
#pragma unroll
                for (int a = 0; a < NB; ++a) zfms(W[a][q], l[a], u);
            }
        }
    }
    // x_i = W(p_i, n) / W(p_i, i)
    if (tc == n % 16) {
        const int qn = n / 16;
#pragma unroll
        for (int q = 0; q < NB; ++q)
            if (q == qn) {
#pragma unroll
                for (int a = 0; a < NB; ++a) {
                    const int r = tr + 16 * a;
                    if (r < n) rhs[r] = W[a][q];
                }
            }
    }
    __syncthreads();
    for (int i = tid; i < n; i += 256) X[(b * n + i) * nrhs + m] = zmul(rhs[perm[i]], pivv[i]);
    if (tid == 0 && bad && info != nullptr) atomicAdd(info, 1);
}

// The same with W and the right-hand side in LDS for n <= 96: the pivot
// search is one wave's shuffle reduction, and each elimination step is two
// barriers.  Row-major W(i, c) = shift*I - M_b.
__global__ __launch_bounds__(256) void k_bshift_solve_lds(int n, int nrhs, const double2 *__restrict__ M,
                                                          const double2 *__restrict__ shift, const double2 *B,
                                                          double2 *X, int *info)
{
    extern __shared__ double2 sl[];          // W (n*n), then x (n)
    double2 *W = sl, *x = sl + n * n;
    __shared__ int s_p;
    __shared__ int s_bad;
    const int64_t sys = blockIdx.x;
    const int64_t b = sys / nrhs;
    const int m = (int)(sys % nrhs);
    const double2 *Mb = M + b * n * n;
    const double2 sh = shift[b * nrhs + m];
    for (int e = threadIdx.x; e < n * n; e += blockDim.x) {
        const int i = e / n, c = e % n;
        const double2 v = Mb[e];
        double2 w = make_double2(-v.x, -v.y);
        if (i == c) { w.x += sh.x; w.y += sh.y; }
        W[e] = w;
    }
    for (int i = threadIdx.x; i < n; i += blockDim.x) x[i] = B[(b * n + i) * nrhs + m];
    if (threadIdx.x == 0) s_bad = 0;
    __syncthreads();
    for (int j = 0; j < n; ++j) {
        if (threadIdx.x < 64) {              // pivot: largest |W(i, j)|, i >= j (lowest index on ties)
            double best = -1;
            int bi = j;
            for (int i = j + (int)threadIdx.x; i < n; i += 64) {
                const double a = zabs2(W[i * n + j]);
                if (a > best) { best = a; bi = i; }
            }
            for (int o = 32; o > 0; o >>= 1) {
                const double v2 = __shfl_xor(best, o, 64);
                const int i2 = __shfl_xor(bi, o, 64);
                if (v2 > best || (v2 == best && i2 < bi)) { best = v2; bi = i2; }
            }
            if (threadIdx.x == 0) {
                s_p = bi;
                if (!(best > 0)) s_bad = 1;
            }
        }
        __syncthreads();
        const int p = s_p;
        if (p != j) {                        // swap rows j and p (columns j.. and the rhs)
            for (int c = j + threadIdx.x; c <= n; c += blockDim.x) {
                double2 *a = c < n ? &W[j * n + c] : &x[j];
                double2 *q = c < n ? &W[p * n + c] : &x[p];
                const double2 t = *a;
                *a = *q;
                *q = t;
            }
            __syncthreads();
        }
        const double2 d = W[j * n + j];
        const double dd = zabs2(d);
        const double2 dinv = dd > 0 ? make_double2(d.x / dd, -d.y / dd) : make_double2(0, 0);
        // rows i > j: l = W(i,j)/d, W(i, c) -= l W(j, c) for c > j, x(i) -= l x(j);
        // one thread per (row, column) pair, the multiplier recomputed per pair
        const int mm = n - j - 1;
        for (int e = threadIdx.x; e < mm * (mm + 1); e += blockDim.x) {
            const int i = j + 1 + e / (mm + 1), c = j + 1 + e % (mm + 1);
            const double2 l = zmul(W[i * n + j], dinv);
            if (c < n) W[i * n + c] = zsub(W[i * n + c], zmul(l, W[j * n + c]));
            else x[i] = zsub(x[i], zmul(l, x[j]));
        }
        __syncthreads();
    }
    for (int c = n - 1; c >= 0; --c) {
        if (threadIdx.x == 0) {
            const double2 d = W[c * n + c];
            const double dd = zabs2(d);
            x[c] = dd > 0 ? zmul(x[c], make_double2(d.x / dd, -d.y / dd)) : make_double2(0, 0);
        }
        __syncthreads();
        const double2 xc = x[c];
        for (int i = threadIdx.x; i < c; i += blockDim.x) x[i] = zsub(x[i], zmul(W[i * n + c], xc));
        __syncthreads();
    }
    for (int i = threadIdx.x; i < n; i += blockDim.x) X[(b * n + i) * nrhs + m] = x[i];
    if (threadIdx.x == 0 && s_bad && info != nullptr) atomicAdd(info, 1);
}

}  // namespace rsvdt

using namespace rsvdt;

extern "C" int rsvdt_hip_bchol_inv(int64_t nbatch, int64_t n, const double *G, double shift_rel, double drop_tol,
                                   double *R, double *Rinv, int *info, rsvdt_stream_t s)
{
    if (nbatch <= 0 || n <= 0) return RSVDT_HIP_OK;
    if (n > 1024 || nbatch > 0x7fffffff) { set_error_msg("rsvdt_hip_bchol_inv: n <= 1024"); return RSVDT_HIP_ERR_ARG; }
    const int nb = (int)((n + 15) / 16);
#define RSVDT_CHOL_REG(NB)                                                                                   \
    if (nb == NB) {                                                                                          \
        hipLaunchKernelGGL(k_bchol_inv_reg<NB>, dim3((unsigned)nbatch), dim3(256), 0, as_stream(s), (int)n,  \
                           (const double2 *)G, shift_rel, drop_tol, (double2 *)R, (double2 *)Rinv, info);   \
    } else
    RSVDT_CHOL_REG(1) RSVDT_CHOL_REG(2) RSVDT_CHOL_REG(3) RSVDT_CHOL_REG(4) RSVDT_CHOL_REG(5)
#undef RSVDT_CHOL_REG
        hipLaunchKernelGGL(k_bchol_inv, dim3((unsigned)nbatch), dim3(256), 0, as_stream(s), (int)n,
                           (const double2 *)G, shift_rel, drop_tol, (double2 *)R, (double2 *)Rinv, info);
    RSVDT_LAUNCH_CHECK("rsvdt_hip_bchol_inv");
    return RSVDT_HIP_OK;
}

extern "C" int rsvdt_hip_bdrop_mask(int64_t nbatch, int64_t n, const double *R, double *D, int *count,
                                    rsvdt_stream_t s)
{
    if (nbatch <= 0 || n <= 0) return RSVDT_HIP_OK;
    const int64_t tot = nbatch * n * n;
    hipLaunchKernelGGL(k_bdrop_mask, dim3(ceil_div(tot, 256)), dim3(256), 0, as_stream(s), tot, (int)n,
                       (const double2 *)R, (double2 *)D, count);
    RSVDT_LAUNCH_CHECK("rsvdt_hip_bdrop_mask");
    return RSVDT_HIP_OK;
}

extern "C" int rsvdt_hip_btrmm_upper(int64_t nbatch, int64_t n, const double *Ra, const double *Rb, double *Rout,
                                     rsvdt_stream_t s)
{
    if (nbatch <= 0 || n <= 0) return RSVDT_HIP_OK;
    hipLaunchKernelGGL(k_btrmm_upper, dim3((unsigned)nbatch), dim3(256), 0, as_stream(s), (int)n,
                       (const double2 *)Ra, (const double2 *)Rb, (double2 *)Rout);
    RSVDT_LAUNCH_CHECK("rsvdt_hip_btrmm_upper");
    return RSVDT_HIP_OK;
}

extern "C" int rsvdt_hip_bjacobi_svd(int64_t nbatch, int64_t n, const double *R, double *U, double *sig,
                                     rsvdt_stream_t s)
{
    if (nbatch <= 0 || n <= 0) return RSVDT_HIP_OK;
    if (n > JAC_MAXN) { set_error_msg("rsvdt_hip_bjacobi_svd: n <= 64 required"); return RSVDT_HIP_ERR_ARG; }
    hipLaunchKernelGGL(k_bjacobi_svd, dim3((unsigned)nbatch), dim3(256), 0, as_stream(s), (int)n,
                       (const double2 *)R, (double2 *)U, sig);
    RSVDT_LAUNCH_CHECK("rsvdt_hip_bjacobi_svd");
    return RSVDT_HIP_OK;
}

extern "C" size_t rsvdt_hip_bshift_solve_work_bytes(int64_t nbatch, int64_t n, int64_t nrhs)
{
    return (size_t)(nbatch * nrhs) * (size_t)(n * n) * 16 + 256;
}

extern "C" int rsvdt_hip_bshift_solve(int64_t nbatch, int64_t n, int64_t nrhs, const double *M, const double *shift,
                                      const double *B, double *X, void *work, int *info, rsvdt_stream_t s)
{
    if (nbatch <= 0 || n <= 0 || nrhs <= 0) return RSVDT_HIP_OK;
    if (n > 4096 || nbatch * nrhs > 0x7fffffffLL) { set_error_msg("rsvdt_hip_bshift_solve: bad size"); return RSVDT_HIP_ERR_ARG; }
    const int nb = (int)((n + 16) / 16);      // n + 1 columns with the right-hand side
#define RSVDT_SOLVE_REG(NB)                                                                                  \
    if (nb == NB) {                                                                                          \
        hipLaunchKernelGGL(k_bshift_solve_reg<NB>, dim3((unsigned)(nbatch * nrhs)), dim3(256), 0, as_stream(s), \
                           (int)n, (int)nrhs, (const double2 *)M, (const double2 *)shift, (const double2 *)B,   \
                           (double2 *)X, info);                                                              \
    } else
    RSVDT_SOLVE_REG(1) RSVDT_SOLVE_REG(2) RSVDT_SOLVE_REG(3) RSVDT_SOLVE_REG(4) RSVDT_SOLVE_REG(5)
#undef RSVDT_SOLVE_REG
    if (n <= 96)      // W and the right-hand side in LDS ((n^2 + n) x 16 B <= 149 KB)
        hipLaunchKernelGGL(k_bshift_solve_lds, dim3((unsigned)(nbatch * nrhs)), dim3(256), (size_t)(n * n + n) * 16,
                           as_stream(s), (int)n, (int)nrhs, (const double2 *)M, (const double2 *)shift,
                           (const double2 *)B, (double2 *)X, info);
    else
        hipLaunchKernelGGL(k_bshift_solve, dim3((unsigned)(nbatch * nrhs)), dim3(256), (size_t)n * 16, as_stream(s),
                           (int)n, (int)nrhs, (const double2 *)M, (const double2 *)shift, (const double2 *)B,
                           (double2 *)X, (double2 *)work, info);
    RSVDT_LAUNCH_CHECK("rsvdt_hip_bshift_solve");
    return RSVDT_HIP_OK;
}
