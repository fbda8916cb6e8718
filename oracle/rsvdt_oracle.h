/*
 * rsvdt_oracle.h -- CPU FP64 restatement of the RSVD-Delta-t hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP
 * implementation: only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it.  The product library never links it and
 * never falls back to it.
 *
 * Every routine follows one reference source file (cited per function in
 * rsvdt_oracle.c).  Dense matrices are column-major, complex FP64, exactly
 * the layout of PETSc's MATDENSE: entry (r,c) of an n-row matrix is a[r + n*c].
 * A sketch / response matrix (F_hat, Y_hat, U_hat, V_hat) is N x (k*Nw_eff)
 * in frequency-major column order (column m*k + ik), as in the reference.
 *
 * Pinning status (what independently checks each function; see tests/):
 *   orc_setup_grid          closed form of the time grid (test_oracle_grid)
 *   orc_build_idft/dft      closed-form exponentials (test_oracle_grid)
 *   orc_action (no TR)      exact discrete solution incl. transient, dense
 *                           linear algebra in numpy (test_oracle_action)
 *   orc_action (TR)         exactness when Nw+1 = N (projection onto the full
 *                           space), transient-error reduction, rank-deficient
 *                           snapshot sets (test_oracle_action)
 *   orc_qr                  numpy/LAPACK QR with positive diagonal (test_oracle_linalg)
 *   orc_svd_left            numpy/LAPACK SVD, phase-invariant (test_oracle_linalg)
 *   orc_lu_solve            numpy.linalg.solve (test_oracle_linalg)
 *   orc_rsvdt_run           dense SVD of the RK4 transfer function (test_oracle_e2e)
 *   orc_trans_run           dense matrix powers of the RK4 step polynomial,
 *                           stop rules against the closed-form norms (test_oracle_transrun)
 *   orc_trans_est           FFT of the snapshots and least-squares residuals
 *                           (numpy lstsq) of the synthetic snapshot sets (test_oracle_transrun)
 *   PGMRES iterate of GalerkinProj.c and SLEPc rounding: parity unpinned
 *   (the reference's PETSc/SLEPc stack cannot be built here; the oracle uses
 *   a direct LU solve and Householder QR whose results are basis-independent).
 */
#ifndef RSVDT_ORACLE_H
#define RSVDT_ORACLE_H

#include <stdint.h>
#include <complex.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef double _Complex orc_cplx;

typedef struct {
    int64_t n;          /* rows == cols */
    int64_t nnz;
    int64_t *rowptr;    /* n+1 */
    int64_t *col;       /* nnz */
    orc_cplx *val;      /* nnz */
} orc_csr;

typedef struct {
    /* operator */
    orc_csr A;
    orc_csr AH;          /* conjugate transpose, built once */
    /* RSVD variables (Variables.h RSVD_vars / TS_vars) */
    int64_t N, k, q, Nw, Nw_eff;
    double w;
    int real_op;
    int transient_removal;
    double dt, TL;
    int64_t Ns, Nt, RR;
    /* DFT matrices */
    orc_cplx *idft;      /* Nw_eff x 2Ns column-major (after the in-place transpose) */
    orc_cplx *dft;       /* Nw x Nw_eff column-major */
    int nthreads;        /* OpenMP threads for row loops (0 = default) */
    int64_t max_steps;   /* >0: stop each test vector after this many steps (timing samples) */
} orc_ctx;

/* SetupTimeFreqGrid.c:16-21 */
void orc_setup_grid(double w, int64_t Nw, double TL, double dt_in,
                    double *dt, int64_t *Ns, int64_t *Nt, int64_t *RR);

/* Initialise a context from CSR arrays (copied).  Real-operator detection
 * (LNSStructure.c:37-44) and the DFT/iDFT matrices (CreateDFTiDFTMats.c) are
 * built here.  Returns 0 on success. */
int orc_init(orc_ctx *c, int64_t N, int64_t nnz, const int64_t *rowptr,
             const int64_t *col, const double *val_ri /* 2*nnz interleaved */,
             int64_t k, int64_t q, int64_t Nw, double w, double TL, double dt,
             int transient_removal, int nthreads);
void orc_free(orc_ctx *c);

/* Read back the grid/frequency info. */
void orc_info(const orc_ctx *c, int64_t *Ns, int64_t *Nt, int64_t *RR,
              int64_t *Nw_eff, int *real_op, double *dt);
/* Copy the iDFT (Nw_eff x 2Ns) and DFT (Nw x Nw_eff) matrices out. */
void orc_get_dft(const orc_ctx *c, double *idft_ri, double *dft_ri);

/* One resolvent action (TSActionRK4.c, with DFT.c or TransientRemovalStrategy.c).
 * dir = 1 direct (DirectActionRK4.c), 0 adjoint (AdjointActionRK4.c).
 * F_hat_ri: N x (k*Nw_eff) frequency-major, interleaved re/im, column-major.
 * Y_hat_ri: output, same shape. */
int orc_action(orc_ctx *c, int dir, const double *F_hat_ri, double *Y_hat_ri);

/* Snapshot store only (no DFT / TR): Y_all_ri is N x (Nstore*k), k-major
 * (column ik*Nstore + s) as built by TSActionRK4.c:126-130. */
int orc_action_snapshots(orc_ctx *c, int dir, const double *F_hat_ri, double *Y_all_ri);

/* QRAllFreqs.c: in-place orthonormalisation of each N x k frequency block. */
int orc_qr_allfreqs(orc_ctx *c, double *Y_hat_ri);

/* SVDAllFreqsBeforeAdjoint.c: Y_hat_m <- left singular vectors (descending). */
int orc_svd_before_adjoint(orc_ctx *c, double *Y_hat_ri);

/* SVDAllFreqs.c: V_hat_m <- left singular vectors, S_hat(:,m) <- sigma. */
int orc_svd_allfreqs(orc_ctx *c, const double *Y_hat_ri, double *V_hat_ri, double *S_hat /* k x Nw_eff col-major */);

/* main.c:134-144: Direct -> PowerIteration -> SVD -> StoreU -> Adjoint -> SVD. */
int orc_rsvdt_run(orc_ctx *c, const double *F_hat_ri, double *U_hat_ri,
                  double *V_hat_ri, double *S_hat);

/* CPU-baseline timing sample: nvec x nsteps forced RK4 steps of one test
 * vector with sketch block Fk (N x Nw_eff); returns wall seconds. */
double orc_time_steps(orc_ctx *c, int dir, const double *Fk_ri, int64_t nvec, int64_t nsteps);

/* Dense kernels exposed for pinning against LAPACK. */
int orc_qr(int64_t m, int64_t n, orc_cplx *a /* m x n col-major, overwritten by Q */, orc_cplx *r /* n x n or NULL */);
int orc_svd_left(int64_t m, int64_t n, orc_cplx *a /* overwritten by U */, double *s);
int orc_lu_solve(int64_t n, orc_cplx *a /* n x n col-major, destroyed */, orc_cplx *b /* n, overwritten */);

/* TransientRemovalEst.c:26-118 (Qt: N x Nw column-major snapshots of one
 * period; init/upd: Nw/2 norms each). */
int orc_trans_est(const orc_ctx *c, const orc_cplx *Qt, const orc_cplx *qss, int64_t period_index,
                  double *init_norm, double *upd_norm);
/* TransientRunRK4.c:39-190 (see rsvdt_oracle.c for the outputs). */
int orc_trans_run(const orc_ctx *c, int dir, int64_t periods, int64_t save_mod, double div_val, double con_val,
                  int est, const double *q0_ri, double *norms, double *last_ri, double *snaps_ri,
                  double *init_norms, double *upd_norms, int64_t *info, double *dt_out);

/* Raw operator applications, for unit tests: y = A x + f (dir=1) or A^H x + f. */
void orc_spmv_add(const orc_ctx *c, int dir, const orc_cplx *x, const orc_cplx *f, orc_cplx *y);

#ifdef __cplusplus
}
#endif
#endif
