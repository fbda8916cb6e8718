/*
 * Transient-run mode: the homogeneous RK4 run that chooses the transient
 * length and the decay before a resolvent analysis
 * (SourceCode/TransientRunRK4.c:39-227, SourceCode/TransientRemovalEst.c:26-127,
 * SourceCode/TSTransRK4.c:15-36), on the device.
 *
 * The state is one device vector; the steps between two monitor events (a
 * norm check every TransSaveMod steps, a snapshot every Nt_period/Nw steps
 * for the removal estimate) run as one call of the homogeneous SpMM-RK4
 * stepper without host round trips.  The period-end removal estimate builds
 * the Nw/2 synthetic snapshot sets of all frequencies at once and
 * orthonormalises them in one batched CholeskyQR.
 */
#include "rsvdt_internal.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define PI_D 3.14159265358979323846264338327950288

/* ------------------------------------------------------------------ Vec handles */
PetscErrorCode RSVDT_VecCreate(PetscInt n, const double *host, Vec *out)
{
    RSVDT_CHK(rsvdt_ensure_device());
    if (n < 0) return rsvdt_err(RSVDT_ERR_USER, "VecCreate", "negative size");
    Vec v = calloc(1, sizeof(*v));
    if (!v) return rsvdt_err(RSVDT_ERR_MEM, "VecCreate", "out of memory");
    v->n = n;
    size_t bytes = (size_t)(n > 0 ? n : 1) * 16;
    if (rsvdt_hip_malloc((void **)&v->d, bytes)) { free(v); return rsvdt_err(RSVDT_ERR_MEM, "VecCreate", "%s", rsvdt_hip_last_error()); }
    int rc = host ? rsvdt_hip_memcpy_h2d(v->d, host, (size_t)n * 16) : rsvdt_hip_zero(v->d, bytes, g_rsvdt.stream);
    if (!rc && !host) rc = rsvdt_hip_sync();
    if (rc) { rsvdt_hip_free(v->d); free(v); return rsvdt_err(rc, "VecCreate", "%s", rsvdt_hip_last_error()); }
    *out = v;
    return 0;
}

PetscErrorCode RSVDT_VecGetValues(Vec v, double *host)
{
    if (!v) return rsvdt_err(RSVDT_ERR_USER, "VecGetValues", "null Vec");
    RSVDT_HIPCHK("VecGetValues", rsvdt_hip_sync());
    RSVDT_HIPCHK("VecGetValues", rsvdt_hip_memcpy_d2h(host, v->d, (size_t)v->n * 16));
    return 0;
}

PetscErrorCode RSVDT_VecGetSize(Vec v, PetscInt *n)
{
    if (!v) return rsvdt_err(RSVDT_ERR_USER, "VecGetSize", "null Vec");
    *n = v->n;
    return 0;
}

PetscErrorCode RSVDT_VecDestroy(Vec *pv)
{
    if (!pv || !*pv) return 0;
    rsvdt_hip_free((*pv)->d);
    free(*pv);
    *pv = NULL;
    return 0;
}

#define VEC_FILE_CLASSID 1211214

static uint64_t rd_be(const unsigned char *p, int nb)
{
    uint64_t v = 0;
    for (int i = 0; i < nb; i++) v = (v << 8) | p[i];
    return v;
}
static void wr_be64(unsigned char *p, uint64_t v)
{
    for (int i = 7; i >= 0; i--) { p[i] = (unsigned char)(v & 0xff); v >>= 8; }
}

/* VecLoad of a PETSc binary Vec: [classid][n] (64- or 32-bit, detected) then
 * n big-endian scalars (complex, or real when the file size says so). */
PetscErrorCode RSVDT_VecLoad(const char *path, Vec *out)
{
    FILE *f = fopen(path, "rb");
    if (!f) return rsvdt_err(RSVDT_ERR_FILE, "VecLoad", "cannot open '%s'", path);
    fseek(f, 0, SEEK_END);
    long long fsize = ftell(f);
    fseek(f, 0, SEEK_SET);
    unsigned char hb[16];
    if (fread(hb, 1, 16, f) < 8) { fclose(f); return rsvdt_err(RSVDT_ERR_FILE, "VecLoad", "'%s': truncated header", path); }
    int ib;
    int64_t n;
    if ((int64_t)rd_be(hb, 8) == VEC_FILE_CLASSID) { ib = 8; n = (int64_t)rd_be(hb + 8, 8); }
    else if ((int32_t)rd_be(hb, 4) == VEC_FILE_CLASSID) { ib = 4; n = (int32_t)rd_be(hb + 4, 4); }
    else { fclose(f); return rsvdt_err(RSVDT_ERR_FILE, "VecLoad", "'%s' is not a PETSc binary Vec (bad classid)", path); }
    if (n < 0) { fclose(f); return rsvdt_err(RSVDT_ERR_FILE, "VecLoad", "'%s': bad size in header", path); }
    long long off = 2LL * ib;
    int sb;
    if (fsize >= off + 16LL * n) sb = 16;
    else if (fsize == off + 8LL * n) sb = 8;
    else { fclose(f); return rsvdt_err(RSVDT_ERR_FILE, "VecLoad", "'%s': truncated vector data", path); }
    fseek(f, off, SEEK_SET);
    unsigned char *buf = malloc((size_t)(n > 0 ? n : 1) * (size_t)sb);
    double *v = malloc((size_t)(n > 0 ? n : 1) * 16);
    if (!buf || !v) { free(buf); free(v); fclose(f); return rsvdt_err(RSVDT_ERR_MEM, "VecLoad", "out of memory"); }
    size_t got = fread(buf, (size_t)sb, (size_t)n, f);
    fclose(f);
    if (got != (size_t)n) { free(buf); free(v); return rsvdt_err(RSVDT_ERR_FILE, "VecLoad", "'%s': read error", path); }
    for (int64_t i = 0; i < n; i++) {
        uint64_t a = rd_be(buf + (size_t)sb * i, 8), b = sb == 16 ? rd_be(buf + (size_t)sb * i + 8, 8) : 0;
        memcpy(&v[2 * i], &a, 8);
        memcpy(&v[2 * i + 1], &b, 8);
    }
    free(buf);
    PetscErrorCode ierr = RSVDT_VecCreate(n, v, out);
    free(v);
    return ierr;
}

static PetscErrorCode vec_view_host(const double *v, int64_t n, const char *path)
{
    FILE *f = fopen(path, "wb");
    if (!f) return rsvdt_err(RSVDT_ERR_FILE, "VecView", "cannot create '%s'", path);
    unsigned char *buf = malloc(16 + (size_t)n * 16);
    if (!buf) { fclose(f); return rsvdt_err(RSVDT_ERR_MEM, "VecView", "out of memory"); }
    wr_be64(buf, VEC_FILE_CLASSID);
    wr_be64(buf + 8, (uint64_t)n);
    for (int64_t i = 0; i < 2 * n; i++) {
        uint64_t u;
        memcpy(&u, &v[i], 8);
        wr_be64(buf + 16 + 8 * i, u);
    }
    size_t w = fwrite(buf, 1, 16 + (size_t)n * 16, f);
    free(buf);
    if (fclose(f) != 0 || w != 16 + (size_t)n * 16) return rsvdt_err(RSVDT_ERR_FILE, "VecView", "write error on '%s'", path);
    return 0;
}

/* VecView: PETSc binary Vec, 64-bit header, complex big-endian scalars. */
PetscErrorCode RSVDT_VecView(Vec v, const char *path)
{
    if (!v) return rsvdt_err(RSVDT_ERR_USER, "VecView", "null Vec");
    double *h = malloc((size_t)(v->n > 0 ? v->n : 1) * 16);
    if (!h) return rsvdt_err(RSVDT_ERR_MEM, "VecView", "out of memory");
    PetscErrorCode ierr = RSVDT_VecGetValues(v, h);
    if (!ierr) ierr = vec_view_host(h, v->n, path);
    free(h);
    return ierr;
}

/* a real vector written as a complex Vec (imaginary parts zero) */
static PetscErrorCode view_real(const double *r, int64_t n, const char *path)
{
    double *h = calloc((size_t)(n > 0 ? n : 1) * 2, sizeof(double));
    if (!h) return rsvdt_err(RSVDT_ERR_MEM, "VecView", "out of memory");
    for (int64_t i = 0; i < n; i++) h[2 * i] = r[i];
    PetscErrorCode ierr = vec_view_host(h, n, path);
    free(h);
    return ierr;
}

/* ------------------------------------------------------------------ monitors */
typedef struct {
    double *nrm_dev;   /* 1 double */
    void *work;        /* rsvdt_hip_znorm2 partials (1024 doubles) */
} norm_ws;

static PetscErrorCode vnorm(norm_ws *ws, int64_t n, const double *x, double *out)
{
    double s2 = 0;
    RSVDT_HIPCHK("VecNorm", rsvdt_hip_znorm2(n, x, ws->nrm_dev, ws->work, g_rsvdt.stream));
    RSVDT_HIPCHK("VecNorm", rsvdt_hip_sync());
    RSVDT_HIPCHK("VecNorm", rsvdt_hip_memcpy_d2h(&s2, ws->nrm_dev, sizeof(double)));
    *out = sqrt(s2);
    return 0;
}

/* ------------------------------------------------------------------ TSTransRK4 */
/* SourceCode/TSTransRK4.c:15-36: one homogeneous RK4 step of y in place with
 * A (TS.DirAdj set) or A^H.  TS_mat is not used: the stage vectors live in
 * the library's device work buffers. */
PetscErrorCode TSTransRK4(LNS_vars *LNS, TS_matrices *TS_mat, RSVDt_vars *V, PetscInt i, Vec y)
{
    (void)TS_mat; (void)i;
    if (!LNS || !LNS->A || !V || !y) return rsvdt_err(RSVDT_ERR_USER, "TSTransRK4", "null argument");
    if (y->n != LNS->A->rows) return rsvdt_err(RSVDT_ERR_USER, "TSTransRK4", "vector length %lld != operator size %lld",
                                               (long long)y->n, (long long)LNS->A->rows);
    return rsvdt_homogeneous_steps(LNS->A, V->TS.DirAdj ? 1 : 0, V->TS.dt, 1, y->d, y->n, 1, 1);
}

/* ------------------------------------------------------------------ TransientRemovalEst */
/* SourceCode/TransientRemovalEst.c:26-127.  Q_transient: N x Nw snapshots of
 * one period (column s = snapshot s), qss: the synthetic steady state.
 * Writes Initial_transient_norm_period_<p> (||column iw of Q_transient*dft||/Nw)
 * and Updated_transient_norm_period_<p> (||qss - P qss||/||qss||, P the
 * projector onto the synthetic snapshots Q_transient[:, is + skip] +
 * qss e^{i w iw is deltaT}, skip = 1 in the first period) for iw < Nw/2.
 * init_out / upd_out (host, Nw/2 each) may be NULL. */
static PetscErrorCode trans_est(Mat Qt, Vec qss, PetscInt p, RSVDt_vars *V, DFT_matrices *D, Directories *dirs,
                                double *init_out, double *upd_out)
{
    const int64_t N = Qt->rows, Nw = Qt->cols, nh = V->RSVD.Nw / 2, Nwe = V->RSVD.Nw_eff;
    if (Nw != V->RSVD.Nw || qss->n != N || !D || !D->dft)
        return rsvdt_err(RSVDT_ERR_USER, "TransientRemovalEst", "inconsistent sizes");
    const double deltaT = 2 * PI_D / V->RSVD.w / (double)Nw;
    const int64_t skip = p == 1 ? 1 : 0, ns = Nw - skip;
    rsvdt_printf("**** deltaT = %g ****\n", deltaT);
    rsvdt_printf("**** Subspace dim = %d ****\n", (int)ns);
    rsvdt_printf("**** period index = %d ****\n", (int)p);
    rsvdt_printf("**** Displaying transient norms for (positive) frequencies ****\n");
    rsvdt_stream_t q = g_rsvdt.stream;
    double *init = calloc((size_t)(nh > 0 ? nh : 1), sizeof(double));
    double *upd = calloc((size_t)(nh > 0 ? nh : 1), sizeof(double));
    double *hg = calloc((size_t)(nh > 0 ? nh : 1) * 2, sizeof(double));
    if (!init || !upd || !hg) { free(init); free(upd); free(hg); return rsvdt_err(RSVDT_ERR_MEM, "TransientRemovalEst", "out of memory"); }
    PetscErrorCode ierr = 0;
    double *Qhat = NULL, *G = NULL, *Qrm = NULL, *X = NULL, *Qb = NULL, *E = NULL, *T = NULL, *ph = NULL;
    void *work = NULL;
    int h = 0;
    /* :62-74: Q_hat = Q_transient dft / Nw, column norms */
    if (nh > 0) {
        h = rsvdt_hip_malloc((void **)&Qhat, (size_t)nh * (size_t)N * 16);
        if (!h) h = rsvdt_hip_malloc((void **)&G, (size_t)nh * 16 * (size_t)(ns > 1 ? ns : 1));
        if (!h) {
            size_t wa = rsvdt_hip_bgram_work_bytes(nh, N, 1, 1), wb = rsvdt_hip_bgram_work_bytes(nh, N, ns > 1 ? ns : 1, 1);
            h = rsvdt_hip_malloc(&work, wa > wb ? wa : wb);
        }
        if (!h) h = rsvdt_hip_coef_gemm(nh, Nw, N, D->dft->d, Nwe, 1, 0, 1, Nwe, 1.0 / (double)Nw, Qt->d, N, Qhat, N,
                                        64, 0.0, q);
        if (!h) h = rsvdt_hip_bgram(nh, N, 1, 1, Qhat, N, 1, 1, Qhat, N, 1, 1, G, work, q);
        if (!h) h = rsvdt_hip_sync();
        if (!h) h = rsvdt_hip_memcpy_d2h(hg, G, (size_t)nh * 16);
        if (h) ierr = rsvdt_err(h, "TransientRemovalEst", "device call failed: %s", rsvdt_hip_last_error());
    }
    for (int64_t iw = 0; iw < nh && !ierr; iw++) {
        init[iw] = sqrt(hg[2 * iw]);
        rsvdt_printf("**** Norm of the ratio norm(qt)/norm(qs): %g @ w = %g ****\n", init[iw], V->RSVD.w * (double)iw);
    }
    /* :76-114: synthetic snapshot sets of all frequencies, batched */
    double nq = 0;
    norm_ws ws = {0};
    if (!ierr) {
        h = rsvdt_hip_malloc((void **)&ws.nrm_dev, sizeof(double));
        if (!h) h = rsvdt_hip_malloc(&ws.work, 1024 * sizeof(double));
        if (h) ierr = rsvdt_err(h, "TransientRemovalEst", "%s", rsvdt_hip_last_error());
        else ierr = vnorm(&ws, N, qss->d, &nq);
    }
    if (!ierr && nh > 0)
        rsvdt_printf("**** Displaying updated transient norms for (positive) frequencies ****\n");
    if (!ierr && nh > 0 && ns > 0) {
        /* row-major copy of the snapshots (N x Nw), then groups of frequencies sized to memory */
        h = rsvdt_hip_malloc((void **)&Qrm, (size_t)N * (size_t)Nw * 16);
        if (!h) h = rsvdt_hip_transpose_rows(N, Nw, 1, 0, 1, Qt->d, Qrm, 0, q);
        size_t fr = 0, tot = 0;
        rsvdt_hip_mem_info(&fr, &tot);
        const double per = 4.0 * (double)N * (double)ns * 16.0;
        int64_t g = nh;
        while (g > 1 && (double)g * per > 0.4 * (double)fr) g = (g + 1) / 2;
        const size_t bytes = (size_t)g * (size_t)N * (size_t)ns * 16;
        if (!h) h = rsvdt_hip_malloc((void **)&X, bytes);
        if (!h) h = rsvdt_hip_malloc((void **)&Qb, bytes);
        if (!h) h = rsvdt_hip_malloc((void **)&E, bytes);
        if (!h) h = rsvdt_hip_malloc((void **)&T, bytes);
        if (!h) h = rsvdt_hip_malloc((void **)&ph, (size_t)g * (size_t)ns * 16);
        if (!h) {
            size_t wb = rsvdt_hip_bgram_work_bytes(g, N, ns, 1), wb1 = rsvdt_hip_bgram_work_bytes(g, N, 1, 1);
            rsvdt_hip_free(work);
            work = NULL;
            h = rsvdt_hip_malloc(&work, wb > wb1 ? wb : wb1);
        }
        double *hph = malloc((size_t)g * (size_t)ns * 16);
        if (!hph) h = RSVDT_ERR_MEM;
        if (h) ierr = rsvdt_err(h, "TransientRemovalEst", "cannot allocate work: %s", rsvdt_hip_last_error());
        for (int64_t b0 = 0; b0 < nh && !ierr; b0 += g) {
            const int64_t gb = g < nh - b0 ? g : nh - b0;
            /* phases e^{i w is deltaT}, w = w0 iw, evaluated as in :92 */
            for (int64_t b = 0; b < gb; b++) {
                const double w = V->RSVD.w * (double)(b0 + b);
                for (int64_t is = 0; is < ns; is++) {
                    const double a = w * (double)is * deltaT;
                    hph[2 * (b * ns + is)] = cos(a);
                    hph[2 * (b * ns + is) + 1] = sin(a);
                }
            }
            h = rsvdt_hip_memcpy_h2d(ph, hph, (size_t)gb * (size_t)ns * 16);
            if (!h) h = rsvdt_hip_trans_synth(N, Nw, ns, skip, gb, Qrm, qss->d, ph, X, q);
            if (h) { ierr = rsvdt_err(h, "TransientRemovalEst", "device call failed: %s", rsvdt_hip_last_error()); break; }
            const int64_t xb = N * ns;
            ierr = rsvdt_span_basis(gb, N, ns, X, Qb, E, T, xb, ns, 1, (size_t)gb * (size_t)xb * 16, 1e-14);
            if (ierr) break;
            /* b = Q^H qss; E = Q b - qss (= -(qdiff)); ||E||^2 */
            h = rsvdt_hip_bgram(gb, N, ns, 1, Qb, xb, ns, 1, qss->d, 0, 1, 1, G, work, q);
            if (!h) h = rsvdt_hip_bapply(gb, N, ns, 1, Qb, xb, ns, 1, G, -1.0, qss->d, 0, 1, 1, E, N, 1, 1, q);
            if (!h) h = rsvdt_hip_bgram(gb, N, 1, 1, E, N, 1, 1, E, N, 1, 1, G, work, q);
            if (!h) h = rsvdt_hip_sync();
            if (!h) h = rsvdt_hip_memcpy_d2h(hg, G, (size_t)gb * 16);
            if (h) { ierr = rsvdt_err(h, "TransientRemovalEst", "device call failed: %s", rsvdt_hip_last_error()); break; }
            for (int64_t b = 0; b < gb; b++) {
                upd[b0 + b] = sqrt(hg[2 * b]) / nq;
                rsvdt_printf("**** Norm of the ratio norm(qt_up)/norm(qs): %g @ w = %g ****\n", upd[b0 + b],
                             V->RSVD.w * (double)(b0 + b));
            }
        }
        free(hph);
    }
    rsvdt_hip_free(Qhat); rsvdt_hip_free(G); rsvdt_hip_free(work); rsvdt_hip_free(Qrm); rsvdt_hip_free(X);
    rsvdt_hip_free(Qb); rsvdt_hip_free(E); rsvdt_hip_free(T); rsvdt_hip_free(ph);
    rsvdt_hip_free(ws.nrm_dev); rsvdt_hip_free(ws.work);
    if (!ierr && dirs) {
        snprintf(dirs->IO_dir, PETSC_MAX_PATH_LEN, "%sInitial_transient_norm_period_%d", dirs->FolderDir, (int)p);
        ierr = view_real(init, nh, dirs->IO_dir);
        if (!ierr) {
            snprintf(dirs->IO_dir, PETSC_MAX_PATH_LEN, "%sUpdated_transient_norm_period_%d", dirs->FolderDir, (int)p);
            ierr = view_real(upd, nh, dirs->IO_dir);
        }
    }
    if (!ierr && init_out) memcpy(init_out, init, sizeof(double) * (size_t)nh);
    if (!ierr && upd_out) memcpy(upd_out, upd, sizeof(double) * (size_t)nh);
    free(init); free(upd); free(hg);
    return ierr;
}

PetscErrorCode TransientRemovalEst(Mat Q_transient, Vec qss, PetscInt period_index, RSVDt_vars *V,
                                   DFT_matrices *D, Directories *dirs)
{
    if (!Q_transient || !qss || !V || !D) return rsvdt_err(RSVDT_ERR_USER, "TransientRemovalEst", "null argument");
    if (Q_transient->kind != RSVDT_MAT_DENSE) return rsvdt_err(RSVDT_ERR_USER, "TransientRemovalEst", "Q_transient must be dense");
    RSVDT_CHK(rsvdt_mat_to_layout(Q_transient, 1, 0));
    return trans_est(Q_transient, qss, period_index, V, D, dirs, NULL, NULL);
}

/* ------------------------------------------------------------------ TransientRunRK4 */
static PetscErrorCode save_vec(Directories *dirs, const char *name, Vec v)
{
    snprintf(dirs->IO_dir, PETSC_MAX_PATH_LEN, "%s%s", dirs->FolderDir, name);
    return RSVDT_VecView(v, dirs->IO_dir);
}

/* SourceCode/TransientRunRK4.c:39-227: q' = A q (A^H when TS.DirAdj is 0)
 * from the initial vector TransICDir (TransICFlg) or a seeded uniform random
 * vector normalised to 1, over TransPeriods periods of 2 pi / w with dt
 * re-adjusted to divide T/Nw.  Every TransSaveMod steps the norm is checked
 * (NaN, above TransDivVal or below TransConVal stop the run before it is
 * recorded) and recorded, and the state saved when TransSave; with
 * TransRemovalEst the Nw snapshots of each period feed TransientRemovalEst
 * at the period's end (qss = -q0).  Files in dirs->FolderDir:
 * q_transient_<pos>, q_transient_norms, q_transient_last_snapshot,
 * Initial/Updated_transient_norm_period_<p>, and the inputs copy. */
PetscErrorCode TransientRunRK4(TransRun_vars *TR, RSVDt_vars *V, LNS_vars *LNS, DFT_matrices *D, Directories *dirs)
{
    if (!TR || !V || !LNS || !LNS->A || !D || !dirs) return rsvdt_err(RSVDT_ERR_USER, "TransientRunRK4", "null argument");
    if (TR->TransSaveMod <= 0 || TR->TransPeriods < 0 || V->RSVD.Nw <= 0 || V->RSVD.w <= 0 || V->TS.dt <= 0)
        return rsvdt_err(RSVDT_ERR_USER, "TransientRunRK4", "TransSaveMod, Nw, w and dt must be positive");
    RSVDT_CHK(rsvdt_ensure_device());
    rsvdt_printf("\n********************************************\n************** Transient run ***************\n"
                 "********************************************\n\n");
    const int64_t N = LNS->A->rows, Ns = V->RSVD.Nw;
    const double TSS = 2 * PI_D / V->RSVD.w, deltaT = TSS / (double)Ns;
    V->TS.dt = deltaT / ceil(deltaT / V->TS.dt);
    const int64_t Nt_period = (int64_t)round(TSS / V->TS.dt);
    const int64_t delta = Nt_period / Ns;
    const int64_t rend = Nt_period * TR->TransPeriods;
    const int64_t Nt_saved = (int64_t)floor((double)(rend / TR->TransSaveMod)) + 1;
    const int est = TR->TransRemovalEst ? 1 : 0;
    const int dir = V->TS.DirAdj ? 1 : 0;
    rsvdt_printf("Total integration time = %g = %d x %g with dt = %g\n", V->TS.dt * (double)rend, (int)TR->TransPeriods,
                 TSS, V->TS.dt);
    if (TR->TransSave) rsvdt_printf("Total snapshots to be saved = %d\n", (int)Nt_saved);
    else rsvdt_printf("Saving snapshots is not requested\n");

    Vec q0 = NULL, qss = NULL;
    Mat Qall = NULL;
    double *norms = calloc((size_t)Nt_saved, sizeof(double));
    if (!norms) return rsvdt_err(RSVDT_ERR_MEM, "TransientRunRK4", "out of memory");
    norm_ws ws = {0};
    PetscErrorCode ierr = 0;
    int h = rsvdt_hip_malloc((void **)&ws.nrm_dev, sizeof(double));
    if (!h) h = rsvdt_hip_malloc(&ws.work, 1024 * sizeof(double));
    if (h) ierr = rsvdt_err(h, "TransientRunRK4", "%s", rsvdt_hip_last_error());
    if (!ierr && est) {
        ierr = rsvdt_mat_alloc_dense(N, Ns, 1, 0, &Qall);
        if (!ierr) ierr = RSVDT_VecCreate(N, NULL, &qss);
    }
    /* :79-98: initial condition */
    if (!ierr) {
        if (TR->TransICFlg) {
            rsvdt_printf("Reading the initial vector\n");
            snprintf(dirs->IO_dir, PETSC_MAX_PATH_LEN, "%s%s%s", dirs->RootDir, dirs->ResultsDir, dirs->TransICDir);
            ierr = RSVDT_VecLoad(dirs->IO_dir, &q0);
            if (!ierr && q0->n != N)
                ierr = rsvdt_err(RSVDT_ERR_USER, "TransientRunRK4", "initial vector has %lld entries, the operator %lld",
                                 (long long)q0->n, (long long)N);
        } else {
            rsvdt_printf("Random vector is generated\n\n");
            ierr = RSVDT_VecCreate(N, NULL, &q0);
            double n0 = 0;
            if (!ierr) {
                h = rsvdt_hip_fill_random(N, 1, 1, (uint64_t)V->RSVD.RandSeed, 0, q0->d, g_rsvdt.stream);
                if (h) ierr = rsvdt_err(h, "TransientRunRK4", "%s", rsvdt_hip_last_error());
            }
            if (!ierr) ierr = vnorm(&ws, N, q0->d, &n0);
            if (!ierr && n0 > 0) {
                h = rsvdt_hip_zscale(N, 1.0 / n0, q0->d, q0->d, g_rsvdt.stream);
                if (h) ierr = rsvdt_err(h, "TransientRunRK4", "%s", rsvdt_hip_last_error());
            }
        }
    }
    if (!ierr && est) {   /* :100-103 qss = -q0 */
        h = rsvdt_hip_zscale(N, -1.0, q0->d, qss->d, g_rsvdt.stream);
        if (h) ierr = rsvdt_err(h, "TransientRunRK4", "%s", rsvdt_hip_last_error());
    }
    double nv = 0;
    if (!ierr) ierr = vnorm(&ws, N, q0->d, &nv);
    if (!ierr) {
        rsvdt_printf("Norm of the initial snapshot: %f @ t = %d\n", nv, 0);
        norms[0] = nv;
    }
    if (!ierr && est) {
        h = rsvdt_hip_memcpy_d2d(Qall->d, q0->d, (size_t)N * 16, g_rsvdt.stream);
        if (h) ierr = rsvdt_err(h, "TransientRunRK4", "%s", rsvdt_hip_last_error());
    }
    if (!ierr && TR->TransSave) ierr = save_vec(dirs, "q_transient_0", q0);

    /* :126-184: time stepping; the steps between two monitor events run in one call */
    double t1 = RSVDT_Time();
    int64_t i = 1;
    while (!ierr && i <= rend) {
        int64_t e = rend;
        int64_t ns_ = ((i + TR->TransSaveMod - 1) / TR->TransSaveMod) * TR->TransSaveMod;
        if (ns_ < e) e = ns_;
        if (est && delta > 0) {
            int64_t nd = ((i + delta - 1) / delta) * delta;
            if (nd < e) e = nd;
        }
        ierr = rsvdt_homogeneous_steps(LNS->A, dir, V->TS.dt, e - i + 1, q0->d, N, 1, 1);
        if (ierr) break;
        i = e;
        int stop = 0;
        if (i % TR->TransSaveMod == 0) {
            const int64_t pos = i / TR->TransSaveMod;
            ierr = vnorm(&ws, N, q0->d, &nv);
            if (ierr) break;
            if (isnan(nv)) {
                rsvdt_printf("Stopped! The norm of the last snapshot is NAN!\n");
                stop = 1;
            } else if (nv > TR->TransDivVal) {
                rsvdt_printf("Diverged! The norm of the last snapshot = %g exceeds the divergence threshold = %g\n", nv,
                             TR->TransDivVal);
                stop = 1;
            } else if (nv < TR->TransConVal) {
                rsvdt_printf("Converged! The norm of the last snapshot = %g is smaller than the convergence threshold = %g\n",
                             nv, TR->TransConVal);
                stop = 1;
            }
            if (stop) break;
            if (TR->TransSave) {
                char name[64];
                snprintf(name, sizeof(name), "q_transient_%d", (int)pos);
                ierr = save_vec(dirs, name, q0);
                if (ierr) break;
            }
            rsvdt_printf("Norm of the snapshot: %f @ t = %f\n", nv, (double)i * V->TS.dt);
            norms[pos] = nv;
        }
        if (est && delta > 0 && i % delta == 0) {
            const int64_t pos = (i / delta) % Ns;
            if (pos == 0) ierr = trans_est(Qall, qss, i / Nt_period, V, D, dirs, NULL, NULL);
            if (ierr) break;
            h = rsvdt_hip_memcpy_d2d(Qall->d + 2 * (size_t)pos * (size_t)N, q0->d, (size_t)N * 16, g_rsvdt.stream);
            if (h) { ierr = rsvdt_err(h, "TransientRunRK4", "%s", rsvdt_hip_last_error()); break; }
        }
        i++;
    }
    if (!ierr) {
        double el = RSVDT_Time() - t1;
        int hh = (int)(el / 3600), mm = (int)((el - 3600 * hh) / 60), ss = (int)(el - 3600 * hh - 60 * mm);
        rsvdt_printf("*** Transient simulation elapsed time = %02d:%02d:%02d ***\n", hh, mm, ss);
        rsvdt_printf("\nThe results directory is: %s\n\n", dirs->FolderDir);
        /* :196-205 */
        snprintf(dirs->IO_dir, PETSC_MAX_PATH_LEN, "%sq_transient_norms", dirs->FolderDir);
        ierr = view_real(norms, Nt_saved, dirs->IO_dir);
        if (!ierr) ierr = save_vec(dirs, "q_transient_last_snapshot", q0);
    }
    RSVDT_VecDestroy(&q0);
    RSVDT_VecDestroy(&qss);
    RSVDT_MatDestroy(&Qall);
    rsvdt_hip_free(ws.nrm_dev);
    rsvdt_hip_free(ws.work);
    free(norms);
    if (!ierr) ierr = SaveInputVarsCopy(dirs);
    if (!ierr) rsvdt_printf("DONE :))\n\n");
    return ierr;
}
