/*
 * rsvdt_oracle.c -- CPU FP64 restatement of the RSVD-Delta-t hot path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  See rsvdt_oracle.h for the
 * pinning status of each function.  The structure deliberately follows the
 * reference (the SourceCode directory): a per-test-vector loop, four SpMVs, two forcing
 * GEMVs and seven AXPYs per RK4 step, a snapshot store and a snapshot-by-DFT
 * GEMM, so that its timing is a faithful CPU baseline of that algorithm.
 * Row loops are OpenMP-parallel, mirroring PETSc's MPI row distribution.
 */
#include "rsvdt_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define PI_D 3.14159265358979323846264338327950288

static void *xmalloc(size_t n) { void *p = malloc(n ? n : 1); if (!p) abort(); return p; }
static void *xcalloc(size_t n, size_t s) { void *p = calloc(n ? n : 1, s ? s : 1); if (!p) abort(); return p; }

/* ------------------------------------------------------------------------ */
/* SetupTimeFreqGrid.c:16-21                                                 */
/* ------------------------------------------------------------------------ */
void orc_setup_grid(double w, int64_t Nw, double TL, double dt_in,
                    double *dt, int64_t *Ns, int64_t *Nt, int64_t *RR)
{
    double T_ss = 2.0 * PI_D / w;
    double dt_w = T_ss / (double)Nw;
    double h = dt_w / ceil(dt_w / dt_in);
    *dt = h;
    *Ns = (int64_t)round(T_ss / h);
    *Nt = (int64_t)round(TL / h);
    *RR = *Ns / Nw;
}

/* ------------------------------------------------------------------------ */
/* CreateDFTiDFTMats.c:18-93.  idft is built as a 2Ns x Nw_eff matrix and    */
/* transposed in place (:93); we store the transposed form directly:         */
/* idft[m + Nw_eff*j] = conj(exp(alpha*m_alias*j)) * 2/(2Ns), alpha=-2 pi i/(2Ns). */
/* dft[s + Nw*m] = exp(-2 pi i m s / Nw) (:68-85).                          */
/* ------------------------------------------------------------------------ */
static void build_dft(orc_ctx *c)
{
    int64_t Nt2 = 2 * c->Ns;           /* "Nt" in CreateDFTiDFTMats.c:25 */
    int64_t Nwe = c->Nw_eff;
    orc_cplx alpha = -2.0 * PI_D * I / (double)Nt2;
    c->idft = xmalloc(sizeof(orc_cplx) * (size_t)(Nwe * Nt2));
    for (int64_t iw = 0; iw < Nwe; iw++) {
        orc_cplx v;
        if (c->real_op) {
            v = alpha * (double)iw;
        } else if (Nwe % 2 == 0) {
            v = (iw <= Nwe / 2 - 1) ? alpha * (double)iw : alpha * (double)((iw - Nwe) + Nt2);
        } else {
            v = (iw <= (Nwe - 1) / 2) ? alpha * (double)iw : alpha * (double)((iw - Nwe) + Nt2);
        }
        for (int64_t j = 0; j < Nt2; j++) {
            orc_cplx e = cexp(v * (double)j);
            c->idft[iw + Nwe * j] = conj(e) * (2.0 / (double)Nt2);
        }
    }
    /* dft: Nw x Nw_eff; for a real operator J has 2*Nw_eff = Nw entries and
     * alpha = -2 pi i/(2 Nw_eff) (CreateDFTiDFTMats.c:61-79). */
    int64_t Nw = c->Nw;
    double denom = c->real_op ? (double)(2 * Nwe) : (double)Nwe;
    orc_cplx a2 = -2.0 * PI_D * I / denom;
    c->dft = xmalloc(sizeof(orc_cplx) * (size_t)(Nw * Nwe));
    for (int64_t iw = 0; iw < Nwe; iw++) {
        orc_cplx v = a2 * (double)iw;
        for (int64_t s = 0; s < Nw; s++) c->dft[s + Nw * iw] = cexp(v * (double)s);
    }
}

/* ------------------------------------------------------------------------ */
/* CSR helpers                                                               */
/* ------------------------------------------------------------------------ */
static void csr_conj_transpose(const orc_csr *A, orc_csr *AH)
{
    int64_t n = A->n;
    AH->n = n; AH->nnz = A->nnz;
    AH->rowptr = xcalloc((size_t)(n + 1), sizeof(int64_t));
    AH->col = xmalloc(sizeof(int64_t) * (size_t)A->nnz);
    AH->val = xmalloc(sizeof(orc_cplx) * (size_t)A->nnz);
    for (int64_t e = 0; e < A->nnz; e++) AH->rowptr[A->col[e] + 1]++;
    for (int64_t r = 0; r < n; r++) AH->rowptr[r + 1] += AH->rowptr[r];
    int64_t *fill = xmalloc(sizeof(int64_t) * (size_t)n);
    memcpy(fill, AH->rowptr, sizeof(int64_t) * (size_t)n);
    for (int64_t r = 0; r < n; r++)
        for (int64_t e = A->rowptr[r]; e < A->rowptr[r + 1]; e++) {
            int64_t cc = A->col[e];
            int64_t p = fill[cc]++;
            AH->col[p] = r;
            AH->val[p] = conj(A->val[e]);
        }
    free(fill);
}

/* MatMultAdd / MatMultHermitianTransposeAdd (TSRK4.c:18-42): y = op(A) x + f */
void orc_spmv_add(const orc_ctx *c, int dir, const orc_cplx *x, const orc_cplx *f, orc_cplx *y)
{
    const orc_csr *M = dir ? &c->A : &c->AH;
    int64_t n = M->n;
#pragma omp parallel for schedule(static) if (n > 8192)
    for (int64_t r = 0; r < n; r++) {
        orc_cplx s = 0;
        for (int64_t e = M->rowptr[r]; e < M->rowptr[r + 1]; e++) s += M->val[e] * x[M->col[e]];
        y[r] = f ? s + f[r] : s;
    }
}

/* CreateForcingOnFly.c:17-19: f = F_hat_k * idft(:, jt) */
static void forcing_on_fly(const orc_ctx *c, const orc_cplx *Fk /* N x Nw_eff */, int64_t jt, orc_cplx *f)
{
    int64_t N = c->N, Nwe = c->Nw_eff;
    const orc_cplx *wcol = c->idft + Nwe * jt;
#pragma omp parallel for schedule(static) if (N > 8192)
    for (int64_t r = 0; r < N; r++) {
        orc_cplx s = 0;
        for (int64_t m = 0; m < Nwe; m++) s += Fk[r + N * m] * wcol[m];
        f[r] = s;
    }
}

/* VecWAXPY: w = a*x + y */
static void waxpy(int64_t n, orc_cplx *w, double a, const orc_cplx *x, const orc_cplx *y)
{
#pragma omp parallel for schedule(static) if (n > 8192)
    for (int64_t r = 0; r < n; r++) w[r] = a * x[r] + y[r];
}
/* VecAXPY: y = y + a*x */
static void axpy(int64_t n, orc_cplx *y, double a, const orc_cplx *x)
{
#pragma omp parallel for schedule(static) if (n > 8192)
    for (int64_t r = 0; r < n; r++) y[r] = y[r] + a * x[r];
}

typedef struct { orc_cplx *F1, *F2, *F3, *k1, *k2, *k3, *k4, *yt; } ts_mats;

static void ts_mats_create(ts_mats *t, int64_t N)   /* CreateTSMats.c */
{
    t->F1 = xcalloc((size_t)N, sizeof(orc_cplx)); t->F2 = xcalloc((size_t)N, sizeof(orc_cplx));
    t->F3 = xcalloc((size_t)N, sizeof(orc_cplx)); t->k1 = xcalloc((size_t)N, sizeof(orc_cplx));
    t->k2 = xcalloc((size_t)N, sizeof(orc_cplx)); t->k3 = xcalloc((size_t)N, sizeof(orc_cplx));
    t->k4 = xcalloc((size_t)N, sizeof(orc_cplx)); t->yt = xcalloc((size_t)N, sizeof(orc_cplx));
}
static void ts_mats_destroy(ts_mats *t)              /* DestroyTSMats.c */
{
    free(t->F1); free(t->F2); free(t->F3); free(t->k1); free(t->k2); free(t->k3); free(t->k4); free(t->yt);
}

static int64_t pmod(int64_t a, int64_t m) { int64_t r = a % m; return r < 0 ? r + m : r; }

/* TSRK4.c:18-47: one forced RK4 step (i is 1-based). */
static void ts_rk4(const orc_ctx *c, const orc_cplx *Fk, ts_mats *t, int dir, int64_t i, orc_cplx *y)
{
    int64_t N = c->N, twoNs = 2 * c->Ns;
    double h = c->dt;
    orc_spmv_add(c, dir, y, t->F3, t->k1);
    /* direct: fmod(2(i-1)+1, 2Ns); adjoint: fmod(1e6*Ns-1-2(i-1), 2Ns) == (1-2i) mod 2Ns */
    int64_t jt = dir ? pmod(2 * (i - 1) + 1, twoNs) : pmod(-1 - 2 * (i - 1), twoNs);
    forcing_on_fly(c, Fk, jt, t->F2);
    jt = dir ? pmod(2 * (i - 1) + 2, twoNs) : pmod(-1 - (2 * (i - 1) + 1), twoNs);
    forcing_on_fly(c, Fk, jt, t->F3);
    waxpy(N, t->yt, h / 2, t->k1, y);
    orc_spmv_add(c, dir, t->yt, t->F2, t->k2);
    waxpy(N, t->yt, h / 2, t->k2, y);
    orc_spmv_add(c, dir, t->yt, t->F2, t->k3);
    waxpy(N, t->yt, h, t->k3, y);
    orc_spmv_add(c, dir, t->yt, t->F3, t->k4);
    axpy(N, y, h / 6, t->k1);
    axpy(N, y, h / 6 * 2, t->k2);
    axpy(N, y, h / 6 * 2, t->k3);
    axpy(N, y, h / 6, t->k4);
}

/* TSTransRK4.c:15-36: one homogeneous RK4 step. */
static void ts_trans_rk4(const orc_ctx *c, ts_mats *t, int dir, orc_cplx *y)
{
    int64_t N = c->N;
    double h = c->dt;
    orc_spmv_add(c, dir, y, NULL, t->k1);
    waxpy(N, t->yt, h / 2, t->k1, y);
    orc_spmv_add(c, dir, t->yt, NULL, t->k2);
    waxpy(N, t->yt, h / 2, t->k2, y);
    orc_spmv_add(c, dir, t->yt, NULL, t->k3);
    waxpy(N, t->yt, h, t->k3, y);
    orc_spmv_add(c, dir, t->yt, NULL, t->k4);
    axpy(N, y, h / 6, t->k1);
    axpy(N, y, h / 6 * 2, t->k2);
    axpy(N, y, h / 6 * 2, t->k3);
    axpy(N, y, h / 6, t->k4);
}

/* SaveSnapshots.c:17-29: Yk is N x Nstore (column-major). */
static void save_snapshots(const orc_ctx *c, const orc_cplx *y, int64_t i, int dir, orc_cplx *Yk)
{
    int64_t Nstore = c->transient_removal ? c->Nw + 1 : c->Nw;
    if (i > c->Nt) {
        int64_t pos = i - c->Nt;
        if (pos % c->RR == 0) {
            pos = pos / c->RR - 1;
            pos = pmod(pos, Nstore);
            pos = dir ? pos : Nstore - 1 - pos;
            pos = pmod(pos, Nstore);
            memcpy(Yk + c->N * pos, y, sizeof(orc_cplx) * (size_t)c->N);
        }
    }
}

/* ------------------------------------------------------------------------ */
/* Dense linear algebra (column-major)                                       */
/* ------------------------------------------------------------------------ */

/* Householder QR; Q (m x n, thin) overwrites a, R (n x n) optional.
 * The phases are fixed so that diag(R) is real positive, which makes Q the
 * unique orthonormal factor that Gram-Schmidt (SLEPc BVOrthogonalize) also
 * produces for a full-rank input. */
int orc_qr(int64_t m, int64_t n, orc_cplx *a, orc_cplx *r)
{
    if (n > m) return 1;
    orc_cplx *v = xmalloc(sizeof(orc_cplx) * (size_t)(m * n));   /* Householder vectors */
    double *beta = xmalloc(sizeof(double) * (size_t)n);
    orc_cplx *R = xcalloc((size_t)(n * n), sizeof(orc_cplx));
    for (int64_t j = 0; j < n; j++) {
        orc_cplx *x = a + m * j;
        double nx = 0;
        for (int64_t i = j; i < m; i++) nx += creal(x[i]) * creal(x[i]) + cimag(x[i]) * cimag(x[i]);
        nx = sqrt(nx);
        orc_cplx *vj = v + m * j;
        for (int64_t i = 0; i < m; i++) vj[i] = 0;
        if (nx == 0) { beta[j] = 0; R[j + n * j] = 0; continue; }
        orc_cplx x0 = x[j];
        double ax0 = cabs(x0);
        orc_cplx ph = ax0 > 0 ? x0 / ax0 : 1.0;
        orc_cplx alpha = -ph * nx;                 /* H x = alpha e_j */
        for (int64_t i = j; i < m; i++) vj[i] = x[i];
        vj[j] = x0 - alpha;
        double nv = 0;
        for (int64_t i = j; i < m; i++) nv += creal(vj[i]) * creal(vj[i]) + cimag(vj[i]) * cimag(vj[i]);
        beta[j] = nv > 0 ? 2.0 / nv : 0;
        /* apply H = I - beta v v^H to columns j..n-1 */
        for (int64_t cc = j; cc < n; cc++) {
            orc_cplx *y = a + m * cc;
            orc_cplx d = 0;
            for (int64_t i = j; i < m; i++) d += conj(vj[i]) * y[i];
            d *= beta[j];
            for (int64_t i = j; i < m; i++) y[i] -= d * vj[i];
        }
        for (int64_t cc = j; cc < n; cc++) R[j + n * cc] = a[j + m * cc];
    }
    /* form thin Q = H_0 ... H_{n-1} [I; 0] */
    for (int64_t cc = 0; cc < n; cc++) {
        orc_cplx *q = a + m * cc;
        for (int64_t i = 0; i < m; i++) q[i] = (i == cc) ? 1.0 : 0.0;
        for (int64_t j = n - 1; j >= 0; j--) {
            orc_cplx *vj = v + m * j;
            orc_cplx d = 0;
            for (int64_t i = j; i < m; i++) d += conj(vj[i]) * q[i];
            d *= beta[j];
            for (int64_t i = j; i < m; i++) q[i] -= d * vj[i];
        }
    }
    /* make diag(R) real positive: Q(:,j) *= ph_j, R(j,:) *= conj(ph_j) */
    for (int64_t j = 0; j < n; j++) {
        orc_cplx d = R[j + n * j];
        double ad = cabs(d);
        orc_cplx ph = ad > 0 ? d / ad : 1.0;
        for (int64_t i = 0; i < m; i++) a[i + m * j] *= ph;
        for (int64_t cc = 0; cc < n; cc++) R[j + n * cc] *= conj(ph);
    }
    if (r) memcpy(r, R, sizeof(orc_cplx) * (size_t)(n * n));
    free(v); free(beta); free(R);
    return 0;
}

/* Left singular vectors (descending) of an m x n (m >= n) matrix: Householder
 * QR, then one-sided (Hestenes) Jacobi on R.  a is overwritten by U (m x n). */
int orc_svd_left(int64_t m, int64_t n, orc_cplx *a, double *s)
{
    orc_cplx *R = xmalloc(sizeof(orc_cplx) * (size_t)(n * n));
    if (orc_qr(m, n, a, R)) { free(R); return 1; }
    /* Jacobi sweeps on the columns of R */
    for (int sweep = 0; sweep < 80; sweep++) {
        double off = 0;
        for (int64_t p = 0; p < n - 1; p++)
            for (int64_t q = p + 1; q < n; q++) {
                orc_cplx *rp = R + n * p, *rq = R + n * q;
                double al = 0, be = 0; orc_cplx ga = 0;
                for (int64_t i = 0; i < n; i++) {
                    al += creal(rp[i] * conj(rp[i]));
                    be += creal(rq[i] * conj(rq[i]));
                    ga += conj(rp[i]) * rq[i];
                }
                double ag = cabs(ga);
                if (ag == 0 || ag <= 1e-15 * sqrt(al * be)) continue;
                double rel = ag / sqrt(al * be);
                if (rel > off) off = rel;
                orc_cplx eph = conj(ga) / ag;          /* e^{-i phi} */
                double zeta = (be - al) / (2.0 * ag);
                double t = (zeta >= 0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
                double cs = 1.0 / sqrt(1.0 + t * t), sn = cs * t;
                for (int64_t i = 0; i < n; i++) {
                    orc_cplx xp = rp[i], xq = rq[i] * eph;
                    rp[i] = cs * xp - sn * xq;
                    rq[i] = sn * xp + cs * xq;
                }
            }
        if (off < 1e-15) break;
    }
    /* singular values = column norms; U_R = R / sigma; sort descending */
    double *sig = xmalloc(sizeof(double) * (size_t)n);
    int64_t *ord = xmalloc(sizeof(int64_t) * (size_t)n);
    for (int64_t j = 0; j < n; j++) {
        double nn = 0;
        for (int64_t i = 0; i < n; i++) nn += creal(R[i + n * j] * conj(R[i + n * j]));
        sig[j] = sqrt(nn);
        ord[j] = j;
    }
    for (int64_t i = 1; i < n; i++) {      /* insertion sort, descending */
        int64_t t = ord[i]; int64_t j = i - 1;
        while (j >= 0 && sig[ord[j]] < sig[t]) { ord[j + 1] = ord[j]; j--; }
        ord[j + 1] = t;
    }
    /* Phase convention (the reference's SLEPc phase is unspecified): each
     * column u of U_R = R W / sigma is scaled by the unit phase that makes
     * sum_i u_i real and positive (the largest-modulus entry when that sum
     * vanishes).  The transient removal is not linear across frequencies, so
     * the adjoint action of U -- and the final gains, at the level of the
     * removal's residual -- depends on this choice. */
    for (int64_t j = 0; j < n; j++) {
        double sj = sig[j];
        orc_cplx sum = 0; double big = -1; orc_cplx bigv = 1;
        for (int64_t i = 0; i < n; i++) {
            orc_cplx u = sj > 0 ? R[i + n * j] / sj : 0;
            sum += u;
            if (cabs(u) > big) { big = cabs(u); bigv = u; }
        }
        orc_cplx ph = cabs(sum) > 1e-8 ? conj(sum) / cabs(sum) : (cabs(bigv) > 0 ? conj(bigv) / cabs(bigv) : 1.0);
        for (int64_t i = 0; i < n; i++) R[i + n * j] *= ph;
    }
    orc_cplx *Q = xmalloc(sizeof(orc_cplx) * (size_t)(m * n));
    memcpy(Q, a, sizeof(orc_cplx) * (size_t)(m * n));
    for (int64_t jj = 0; jj < n; jj++) {
        int64_t j = ord[jj];
        double sj = sig[j];
        s[jj] = sj;
        for (int64_t i = 0; i < m; i++) {
            orc_cplx acc = 0;
            for (int64_t l = 0; l < n; l++) acc += Q[i + m * l] * R[l + n * j];
            a[i + m * jj] = sj > 0 ? acc / sj : 0;
        }
    }
    free(Q); free(R); free(sig); free(ord);
    return 0;
}

/* Dense LU with partial pivoting; replaces the PGMRES solve of
 * GalerkinProj.c:57-62 (rtol 1e-15) by the exact solve it converges to. */
int orc_lu_solve(int64_t n, orc_cplx *a, orc_cplx *b)
{
    int64_t *piv = xmalloc(sizeof(int64_t) * (size_t)n);
    for (int64_t j = 0; j < n; j++) {
        int64_t p = j; double mx = cabs(a[j + n * j]);
        for (int64_t i = j + 1; i < n; i++) if (cabs(a[i + n * j]) > mx) { mx = cabs(a[i + n * j]); p = i; }
        piv[j] = p;
        if (mx == 0) { free(piv); return 1; }
        if (p != j) {
            for (int64_t cc = 0; cc < n; cc++) { orc_cplx t = a[j + n * cc]; a[j + n * cc] = a[p + n * cc]; a[p + n * cc] = t; }
            orc_cplx t = b[j]; b[j] = b[p]; b[p] = t;
        }
        orc_cplx d = a[j + n * j];
        for (int64_t i = j + 1; i < n; i++) {
            orc_cplx l = a[i + n * j] / d;
            a[i + n * j] = l;
            for (int64_t cc = j + 1; cc < n; cc++) a[i + n * cc] -= l * a[j + n * cc];
            b[i] -= l * b[j];
        }
    }
    for (int64_t i = n - 1; i >= 0; i--) {
        orc_cplx s = b[i];
        for (int64_t cc = i + 1; cc < n; cc++) s -= a[i + n * cc] * b[cc];
        b[i] = s / a[i + n * i];
    }
    free(piv);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Context                                                                   */
/* ------------------------------------------------------------------------ */
int orc_init(orc_ctx *c, int64_t N, int64_t nnz, const int64_t *rowptr,
             const int64_t *col, const double *val_ri,
             int64_t k, int64_t q, int64_t Nw, double w, double TL, double dt,
             int transient_removal, int nthreads)
{
    memset(c, 0, sizeof(*c));
    c->N = N; c->k = k; c->q = q; c->Nw = Nw; c->w = w; c->TL = TL;
    c->transient_removal = transient_removal;
    c->nthreads = nthreads;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
    c->A.n = N; c->A.nnz = nnz;
    c->A.rowptr = xmalloc(sizeof(int64_t) * (size_t)(N + 1));
    c->A.col = xmalloc(sizeof(int64_t) * (size_t)nnz);
    c->A.val = xmalloc(sizeof(orc_cplx) * (size_t)nnz);
    memcpy(c->A.rowptr, rowptr, sizeof(int64_t) * (size_t)(N + 1));
    memcpy(c->A.col, col, sizeof(int64_t) * (size_t)nnz);
    double im2 = 0;
    for (int64_t e = 0; e < nnz; e++) {
        c->A.val[e] = val_ri[2 * e] + I * val_ri[2 * e + 1];
        im2 += val_ri[2 * e + 1] * val_ri[2 * e + 1];
    }
    /* LNSStructure.c:37-44: real iff ||Im A||_F < 1e-12 */
    c->real_op = sqrt(im2) < 1e-12;
    c->Nw_eff = c->real_op ? Nw / 2 : Nw;
    csr_conj_transpose(&c->A, &c->AH);
    orc_setup_grid(w, Nw, TL, dt, &c->dt, &c->Ns, &c->Nt, &c->RR);
    build_dft(c);
    return 0;
}

void orc_free(orc_ctx *c)
{
    free(c->A.rowptr); free(c->A.col); free(c->A.val);
    free(c->AH.rowptr); free(c->AH.col); free(c->AH.val);
    free(c->idft); free(c->dft);
    memset(c, 0, sizeof(*c));
}

void orc_info(const orc_ctx *c, int64_t *Ns, int64_t *Nt, int64_t *RR,
              int64_t *Nw_eff, int *real_op, double *dt)
{
    *Ns = c->Ns; *Nt = c->Nt; *RR = c->RR; *Nw_eff = c->Nw_eff; *real_op = c->real_op; *dt = c->dt;
}

void orc_get_dft(const orc_ctx *c, double *idft_ri, double *dft_ri)
{
    if (idft_ri) memcpy(idft_ri, c->idft, sizeof(orc_cplx) * (size_t)(c->Nw_eff * 2 * c->Ns));
    if (dft_ri) memcpy(dft_ri, c->dft, sizeof(orc_cplx) * (size_t)(c->Nw * c->Nw_eff));
}

/* PermuteMat.c:24-38: column ik + iw*k  ->  ik*Nw + iw */
static void permute(int64_t N, int64_t k, int64_t Nw, const orc_cplx *in, orc_cplx *out)
{
    for (int64_t ik = 0; ik < k; ik++)
        for (int64_t iw = 0; iw < Nw; iw++)
            memcpy(out + N * (ik * Nw + iw), in + N * (ik + iw * k), sizeof(orc_cplx) * (size_t)N);
}
/* ReversePermuteMat.c:24-38: column ik*Nw + iw  ->  ik + iw*k */
static void reverse_permute(int64_t N, int64_t k, int64_t Nw, const orc_cplx *in, orc_cplx *out)
{
    for (int64_t ik = 0; ik < k; ik++)
        for (int64_t iw = 0; iw < Nw; iw++)
            memcpy(out + N * (ik + iw * k), in + N * (ik * Nw + iw), sizeof(orc_cplx) * (size_t)N);
}

/* TSActionRK4.c:38-130: time stepping of all test vectors; Y_all is
 * N x (Nstore*k) with column ik*Nstore + s. */
static int time_stepping(orc_ctx *c, int dir, const orc_cplx *F_hat, orc_cplx *Y_all)
{
    int64_t N = c->N, k = c->k, Nwe = c->Nw_eff;
    int64_t rend = c->Nt + c->Ns + (c->transient_removal ? c->RR : 0);
    int64_t Nstore = c->transient_removal ? c->Nw + 1 : c->Nw;
    if (c->max_steps > 0 && c->max_steps < rend) rend = c->max_steps;
    orc_cplx *Fp = xmalloc(sizeof(orc_cplx) * (size_t)(N * k * Nwe));
    permute(N, k, Nwe, F_hat, Fp);                       /* :76 */
    orc_cplx *y = xmalloc(sizeof(orc_cplx) * (size_t)N);
    ts_mats t; ts_mats_create(&t, N);
    for (int64_t ik = 0; ik < k; ik++) {                 /* :89 */
        memset(y, 0, sizeof(orc_cplx) * (size_t)N);
        const orc_cplx *Fk = Fp + N * Nwe * ik;          /* :93-97 */
        forcing_on_fly(c, Fk, 0, t.F3);                  /* :103 */
        orc_cplx *Yk = Y_all + N * Nstore * ik;          /* :126-130 (written in place) */
        for (int64_t i = 1; i <= rend; i++) {            /* :116 */
            ts_rk4(c, Fk, &t, dir, i, y);
            save_snapshots(c, y, i, dir, Yk);
        }
    }
    ts_mats_destroy(&t);
    free(y); free(Fp);
    return 0;
}

/* DFT.c:18-39: Y_hat_k = RR * Y_all_k * dft, then reverse-permute. */
static void dft_op(orc_ctx *c, const orc_cplx *Y_all, orc_cplx *Y_hat)
{
    int64_t N = c->N, k = c->k, Nw = c->Nw, Nwe = c->Nw_eff;
    orc_cplx *Yk = xmalloc(sizeof(orc_cplx) * (size_t)(N * k * Nwe));
    for (int64_t ik = 0; ik < k; ik++) {
        const orc_cplx *Y = Y_all + N * Nw * ik;
#pragma omp parallel for schedule(static) if (N > 8192)
        for (int64_t r = 0; r < N; r++)
            for (int64_t m = 0; m < Nwe; m++) {
                orc_cplx s = 0;
                for (int64_t j = 0; j < Nw; j++) s += Y[r + N * j] * c->dft[j + Nw * m];
                Yk[r + N * (ik * Nwe + m)] = s * (double)c->RR;
            }
    }
    reverse_permute(N, k, Nwe, Yk, Y_hat);
    free(Yk);
}

/* TransientRemovalStrategy.c:30-40 with CreateMtilde.c, TSExpADeltaT.c,
 * GalerkinProjAllModes.c and GalerkinProj.c.  The QR of each snapshot block
 * is computed once (the reference recomputes the identical factor). */
static int transient_removal(orc_ctx *c, int dir, const orc_cplx *Y_all, orc_cplx *Y_hat)
{
    int64_t N = c->N, k = c->k, Nw = c->Nw, Nwe = c->Nw_eff;
    int64_t Ns1 = Nw + 1;
    double deltaT = c->dt * (double)c->RR;
    orc_cplx *V = xmalloc(sizeof(orc_cplx) * (size_t)(N * Ns1));
    orc_cplx *MV = xmalloc(sizeof(orc_cplx) * (size_t)(N * Ns1));
    orc_cplx *Mt = xmalloc(sizeof(orc_cplx) * (size_t)(Ns1 * Ns1));
    orc_cplx *Mtmp = xmalloc(sizeof(orc_cplx) * (size_t)(Ns1 * Ns1));
    orc_cplx *y1 = xmalloc(sizeof(orc_cplx) * (size_t)N);
    orc_cplx *y2 = xmalloc(sizeof(orc_cplx) * (size_t)N);
    orc_cplx *yt = xmalloc(sizeof(orc_cplx) * (size_t)N);
    orc_cplx *b = xmalloc(sizeof(orc_cplx) * (size_t)Ns1);
    orc_cplx *Yup = xmalloc(sizeof(orc_cplx) * (size_t)(N * k * Nwe));
    ts_mats t; ts_mats_create(&t, N);
    int rc = 0;
    for (int64_t ik = 0; ik < k && !rc; ik++) {
        const orc_cplx *Yk = Y_all + N * Ns1 * ik;
        memcpy(V, Yk, sizeof(orc_cplx) * (size_t)(N * Ns1));
        if (orc_qr(N, Ns1, V, NULL)) { rc = 1; break; }        /* CreateMtilde.c:46 */
        memcpy(MV, V, sizeof(orc_cplx) * (size_t)(N * Ns1));
        for (int64_t s = 0; s < Ns1; s++)                     /* TSExpADeltaT.c:29-37 */
            for (int64_t it = 0; it < c->RR; it++) ts_trans_rk4(c, &t, dir, MV + N * s);
        for (int64_t a = 0; a < Ns1; a++)                     /* CreateMtilde.c:48-51 */
            for (int64_t bb = 0; bb < Ns1; bb++) {
                orc_cplx s = 0;
                for (int64_t r = 0; r < N; r++) s += conj(V[r + N * a]) * MV[r + N * bb];
                Mt[a + Ns1 * bb] = s;
            }
        const orc_cplx *W1 = dir ? Yk : Yk + N;               /* GalerkinProj.c:30-32 */
        const orc_cplx *W2 = dir ? Yk + N : Yk;               /* GalerkinProj.c:40-41 */
        for (int64_t iw = 0; iw < Nwe && !rc; iw++) {
            const orc_cplx *dcol = c->dft + Nw * iw;
#pragma omp parallel for schedule(static) if (N > 8192)
            for (int64_t r = 0; r < N; r++) {
                orc_cplx s1 = 0, s2 = 0;
                for (int64_t j = 0; j < Nw; j++) { s1 += W1[r + N * j] * dcol[j]; s2 += W2[r + N * j] * dcol[j]; }
                y1[r] = s1 * (double)c->RR;
                y2[r] = s2 * (double)c->RR;
            }
            orc_cplx ph = dir ? cexp(I * c->w * (double)iw * deltaT) : cexp(-I * c->w * (double)iw * deltaT);
            for (int64_t a = 0; a < Ns1 * Ns1; a++) Mtmp[a] = -Mt[a];           /* :48 */
            for (int64_t a = 0; a < Ns1; a++) Mtmp[a + Ns1 * a] += ph;          /* :49-50 */
            for (int64_t r = 0; r < N; r++) yt[r] = ph * y1[r] - y2[r];         /* :51-54 */
            for (int64_t a = 0; a < Ns1; a++) {                                 /* :55 */
                orc_cplx s = 0;
                for (int64_t r = 0; r < N; r++) s += conj(V[r + N * a]) * yt[r];
                b[a] = s;
            }
            if (orc_lu_solve(Ns1, Mtmp, b)) { rc = 2; break; }                 /* :57-62 */
            orc_cplx *out = Yup + N * (ik * Nwe + iw);                          /* :65-71 */
            for (int64_t r = 0; r < N; r++) {
                orc_cplx s = 0;
                for (int64_t a = 0; a < Ns1; a++) s += V[r + N * a] * b[a];
                out[r] = y1[r] - s;
            }
        }
    }
    if (!rc) reverse_permute(N, k, Nwe, Yup, Y_hat);      /* TransientRemovalStrategy.c:38-40 */
    ts_mats_destroy(&t);
    free(V); free(MV); free(Mt); free(Mtmp); free(y1); free(y2); free(yt); free(b); free(Yup);
    return rc;
}

static orc_cplx *to_cplx(const double *ri, int64_t n)
{
    orc_cplx *p = xmalloc(sizeof(orc_cplx) * (size_t)n);
    memcpy(p, ri, sizeof(orc_cplx) * (size_t)n);
    return p;
}

int orc_action_snapshots(orc_ctx *c, int dir, const double *F_hat_ri, double *Y_all_ri)
{
    int64_t Nstore = c->transient_removal ? c->Nw + 1 : c->Nw;
    memset(Y_all_ri, 0, sizeof(orc_cplx) * (size_t)(c->N * Nstore * c->k));
    return time_stepping(c, dir, (const orc_cplx *)F_hat_ri, (orc_cplx *)Y_all_ri);
}

int orc_action(orc_ctx *c, int dir, const double *F_hat_ri, double *Y_hat_ri)
{
    int64_t Nstore = c->transient_removal ? c->Nw + 1 : c->Nw;
    orc_cplx *Y_all = xcalloc((size_t)(c->N * Nstore * c->k), sizeof(orc_cplx));
    int rc = time_stepping(c, dir, (const orc_cplx *)F_hat_ri, Y_all);
    if (!rc) {
        if (c->transient_removal) rc = transient_removal(c, dir, Y_all, (orc_cplx *)Y_hat_ri);
        else dft_op(c, Y_all, (orc_cplx *)Y_hat_ri);
    }
    free(Y_all);
    return rc;
}

int orc_qr_allfreqs(orc_ctx *c, double *Y_hat_ri)
{
    orc_cplx *Y = (orc_cplx *)Y_hat_ri;
    for (int64_t iw = 0; iw < c->Nw_eff; iw++)
        if (orc_qr(c->N, c->k, Y + c->N * c->k * iw, NULL)) return 1;
    return 0;
}

int orc_svd_before_adjoint(orc_ctx *c, double *Y_hat_ri)
{
    orc_cplx *Y = (orc_cplx *)Y_hat_ri;
    double *s = xmalloc(sizeof(double) * (size_t)c->k);
    for (int64_t iw = 0; iw < c->Nw_eff; iw++)
        if (orc_svd_left(c->N, c->k, Y + c->N * c->k * iw, s)) { free(s); return 1; }
    free(s);
    return 0;
}

int orc_svd_allfreqs(orc_ctx *c, const double *Y_hat_ri, double *V_hat_ri, double *S_hat)
{
    int64_t blk = c->N * c->k;
    memcpy(V_hat_ri, Y_hat_ri, sizeof(orc_cplx) * (size_t)(blk * c->Nw_eff));
    orc_cplx *V = (orc_cplx *)V_hat_ri;
    for (int64_t iw = 0; iw < c->Nw_eff; iw++)
        if (orc_svd_left(c->N, c->k, V + blk * iw, S_hat + c->k * iw)) return 1;
    return 0;
}

int orc_rsvdt_run(orc_ctx *c, const double *F_hat_ri, double *U_hat_ri,
                  double *V_hat_ri, double *S_hat)
{
    size_t n = (size_t)(c->N * c->k * c->Nw_eff);
    orc_cplx *Y = to_cplx(F_hat_ri, (int64_t)n);
    orc_cplx *Z = xmalloc(sizeof(orc_cplx) * n);
    int rc = orc_action(c, 1, (double *)Y, (double *)Z);           /* main.c:134 */
    for (int64_t iq = 0; iq < c->q && !rc; iq++) {                   /* PowerIterationRK4.c:20-29 */
        rc = orc_qr_allfreqs(c, (double *)Z);
        if (!rc) rc = orc_action(c, 0, (double *)Z, (double *)Y);
        if (!rc) rc = orc_qr_allfreqs(c, (double *)Y);
        if (!rc) rc = orc_action(c, 1, (double *)Y, (double *)Z);
    }
    if (!rc) rc = orc_svd_before_adjoint(c, (double *)Z);            /* main.c:138 */
    if (!rc) memcpy(U_hat_ri, Z, sizeof(orc_cplx) * n);              /* StoreU.c:15 */
    if (!rc) rc = orc_action(c, 0, (double *)Z, (double *)Y);        /* main.c:142 */
    if (!rc) rc = orc_svd_allfreqs(c, (double *)Y, V_hat_ri, S_hat); /* main.c:144 */
    free(Y); free(Z);
    return rc;
}

/* Timing sample of the reference's hot loop (TSActionRK4.c:89-122 without the
 * snapshot store): nvec repetitions of nsteps forced RK4 steps of one test
 * vector whose sketch block Fk is N x Nw_eff (column-major).  Returns the
 * wall-clock seconds of the loop. */
#include <time.h>
double orc_time_steps(orc_ctx *c, int dir, const double *Fk_ri, int64_t nvec, int64_t nsteps)
{
    int64_t N = c->N;
    const orc_cplx *Fk = (const orc_cplx *)Fk_ri;
    orc_cplx *y = xcalloc((size_t)N, sizeof(orc_cplx));
    ts_mats t; ts_mats_create(&t, N);
    struct timespec a, b;
    clock_gettime(CLOCK_MONOTONIC, &a);
    for (int64_t v = 0; v < nvec; v++) {
        memset(y, 0, sizeof(orc_cplx) * (size_t)N);
        forcing_on_fly(c, Fk, 0, t.F3);
        for (int64_t i = 1; i <= nsteps; i++) ts_rk4(c, Fk, &t, dir, i, y);
    }
    clock_gettime(CLOCK_MONOTONIC, &b);
    ts_mats_destroy(&t);
    free(y);
    return (double)(b.tv_sec - a.tv_sec) + 1e-9 * (double)(b.tv_nsec - a.tv_nsec);
}

/* ------------------------------------------------------------------ transient run */
/* TransientRemovalEst.c:26-118.  Qt: N x Nw snapshots of the last period
 * (column-major), qss: the synthetic steady state.  init_norm[iw] = ||column
 * iw of Qt*dft/Nw||; upd_norm[iw] = ||qss - P qss|| / ||qss|| with P the
 * orthogonal projector onto the span of the synthetic snapshots
 * Qt[:, is + (period_index == 1)] + qss e^{i w iw is deltaT}, is < Nw - (period_index == 1). */
int orc_trans_est(const orc_ctx *c, const orc_cplx *Qt, const orc_cplx *qss, int64_t period_index,
                  double *init_norm, double *upd_norm)
{
    const int64_t N = c->N, Nw = c->Nw, nh = c->Nw / 2;
    const double deltaT = 2 * PI_D / c->w / (double)Nw;
    const int64_t skip = period_index == 1 ? 1 : 0, ns = Nw - skip;
    for (int64_t iw = 0; iw < nh && init_norm; iw++) {   /* :62-74, Q_hat = Qt dft / Nw */
        double s2 = 0;
        for (int64_t r = 0; r < N; r++) {
            orc_cplx acc = 0;
            for (int64_t s = 0; s < Nw; s++) acc += Qt[r + N * s] * c->dft[s + Nw * iw];
            acc *= 1.0 / (double)Nw;
            s2 += creal(acc) * creal(acc) + cimag(acc) * cimag(acc);
        }
        init_norm[iw] = sqrt(s2);
    }
    double nq = 0;
    for (int64_t r = 0; r < N; r++) nq += creal(qss[r]) * creal(qss[r]) + cimag(qss[r]) * cimag(qss[r]);
    nq = sqrt(nq);
    orc_cplx *Q = xmalloc(sizeof(orc_cplx) * (size_t)(N * (ns > 0 ? ns : 1)));
    orc_cplx *b = xmalloc(sizeof(orc_cplx) * (size_t)(ns > 0 ? ns : 1));
    int rc = 0;
    for (int64_t iw = 0; iw < nh && upd_norm && !rc; iw++) {   /* :80-114 */
        const double w = c->w * (double)iw;
        for (int64_t is = 0; is < ns; is++) {
            orc_cplx ph = cexp(I * (w * (double)is * deltaT));
            for (int64_t r = 0; r < N; r++) Q[r + N * is] = Qt[r + N * (is + skip)] + ph * qss[r];
        }
        if (orc_qr(N, ns, Q, NULL)) { rc = 1; break; }
        for (int64_t j = 0; j < ns; j++) {
            orc_cplx d = 0;
            for (int64_t r = 0; r < N; r++) d += conj(Q[r + N * j]) * qss[r];
            b[j] = d;
        }
        double s2 = 0;
        for (int64_t r = 0; r < N; r++) {
            orc_cplx p = 0;
            for (int64_t j = 0; j < ns; j++) p += Q[r + N * j] * b[j];
            orc_cplx d = qss[r] - p;
            s2 += creal(d) * creal(d) + cimag(d) * cimag(d);
        }
        upd_norm[iw] = sqrt(s2) / nq;
    }
    free(Q); free(b);
    return rc;
}

/* TransientRunRK4.c:39-190: homogeneous RK4 from q0 over TransPeriods periods
 * (dt re-adjusted to divide T/Nw, :40-44), with the norm monitors every
 * save_mod steps (NaN / divergence / convergence stop the run before the norm
 * is recorded) and, when est, the period-end transient-removal estimates on
 * the last period's Nw snapshots (qss = -q0).  Outputs (any may be NULL):
 * norms[Nt_saved] (entry 0 = ||q0||, unrecorded entries 0), last[N],
 * snaps[Nt_saved x N] (the saved snapshots, TransSave), init/upd[periods x Nw/2]
 * (row p-1 for the estimate at the end of period p); info[0] = the step at
 * which the run stopped (rend when it ran out), info[1] = estimates made,
 * info[2] = Nt_saved, info[3] = Nt_period, info[4] = stop reason (0 none,
 * 1 NaN, 2 diverged, 3 converged); *dt_out = the step used. */
int orc_trans_run(const orc_ctx *c, int dir, int64_t periods, int64_t save_mod, double div_val, double con_val,
                  int est, const double *q0_ri, double *norms, double *last_ri, double *snaps_ri,
                  double *init_norms, double *upd_norms, int64_t *info, double *dt_out)
{
    orc_ctx cc = *c;
    const int64_t N = c->N, Ns = c->Nw;
    const double TSS = 2 * PI_D / c->w, deltaT = TSS / (double)Ns;
    cc.dt = deltaT / ceil(deltaT / c->dt);
    const int64_t Nt_period = (int64_t)round(TSS / cc.dt);
    const int64_t delta = Nt_period / Ns;
    const int64_t rend = Nt_period * periods;
    const int64_t Nt_saved = (int64_t)floor((double)(rend / save_mod)) + 1;
    if (dt_out) *dt_out = cc.dt;
    orc_cplx *y = to_cplx(q0_ri, N);
    orc_cplx *Qt = est ? xcalloc((size_t)(N * Ns), sizeof(orc_cplx)) : NULL;
    orc_cplx *qss = est ? xmalloc(sizeof(orc_cplx) * (size_t)N) : NULL;
    if (est) for (int64_t r = 0; r < N; r++) qss[r] = -y[r];
    double nrm = 0;
    for (int64_t r = 0; r < N; r++) nrm += creal(y[r]) * creal(y[r]) + cimag(y[r]) * cimag(y[r]);
    if (norms) { memset(norms, 0, sizeof(double) * (size_t)Nt_saved); norms[0] = sqrt(nrm); }
    if (snaps_ri) memcpy(snaps_ri, y, sizeof(orc_cplx) * (size_t)N);
    if (est) memcpy(Qt, y, sizeof(orc_cplx) * (size_t)N);
    ts_mats t;
    ts_mats_create(&t, N);
    int64_t nest = 0, stop = rend, reason = 0;
    int rc = 0;
    for (int64_t i = 1; i <= rend && !rc; i++) {
        ts_trans_rk4(&cc, &t, dir, y);
        if (i % save_mod == 0) {
            int64_t pos = i / save_mod;
            double s2 = 0;
            for (int64_t r = 0; r < N; r++) s2 += creal(y[r]) * creal(y[r]) + cimag(y[r]) * cimag(y[r]);
            double nv = sqrt(s2);
            if (isnan(nv)) { stop = i; reason = 1; break; }
            if (nv > div_val) { stop = i; reason = 2; break; }
            if (nv < con_val) { stop = i; reason = 3; break; }
            if (snaps_ri) memcpy(snaps_ri + 2 * (size_t)pos * (size_t)N, y, sizeof(orc_cplx) * (size_t)N);
            if (norms) norms[pos] = nv;
        }
        if (est && i % delta == 0) {
            int64_t pos = (i / delta) % Ns;
            if (pos == 0) {
                int64_t p = i / Nt_period;
                rc = orc_trans_est(&cc, Qt, qss, p, init_norms ? init_norms + (p - 1) * (Ns / 2) : NULL,
                                   upd_norms ? upd_norms + (p - 1) * (Ns / 2) : NULL);
                nest++;
            }
            memcpy(Qt + (size_t)pos * (size_t)N, y, sizeof(orc_cplx) * (size_t)N);
        }
    }
    if (last_ri) memcpy(last_ri, y, sizeof(orc_cplx) * (size_t)N);
    if (info) { info[0] = stop; info[1] = nest; info[2] = Nt_saved; info[3] = Nt_period; info[4] = reason; }
    ts_mats_destroy(&t);
    free(y); free(Qt); free(qss);
    return rc;
}
