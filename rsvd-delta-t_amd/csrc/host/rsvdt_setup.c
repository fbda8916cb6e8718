/*
 * Pre-processing entry points: user input, time/frequency grid, operator
 * load, DFT/iDFT tables, weight matrices, the random (or loaded, or learned)
 * sketch and the results directory.  Each function keeps the reference's
 * name, arguments, defaults, warnings and errors (file:line per function).
 */
#define _POSIX_C_SOURCE 200809L
#include "rsvdt_internal.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>
#include <unistd.h>

#define PI_D 3.14159265358979323846264338327950288

/* ------------------------------------------------------------------ ReadUserInput */
#define REQ_REAL(key, dst, msg)                                                           \
    do {                                                                                  \
        int ok; double _v;                                                                \
        if (!rsvdt_opt_get_real(key, &_v, &ok)) return rsvdt_err(RSVDT_ERR_USER, "ReadUserInput", msg); \
        if (!ok) return rsvdt_err(RSVDT_ERR_USER, "ReadUserInput", "'%s' is not a real number", key); \
        (dst) = _v;                                                                       \
    } while (0)
#define OPT_BOOL(key, dst, defv, warn)                                                    \
    do {                                                                                  \
        int ok; PetscBool _b;                                                             \
        if (!rsvdt_opt_get_bool(key, &_b, &ok)) {                                         \
            (dst) = (defv);                                                               \
            if (warn) rsvdt_printf("Warning: '%s' variable not found. Setting '%s' to default value: %d\n", key, key, (int)(defv)); \
        } else if (!ok) {                                                                 \
            return rsvdt_err(RSVDT_ERR_USER, "ReadUserInput", "'%s' is not a boolean", key); \
        } else (dst) = _b;                                                                \
    } while (0)
#define OPT_INT(key, dst, defv, warn)                                                     \
    do {                                                                                  \
        int ok; PetscInt _i;                                                              \
        if (!rsvdt_opt_get_int(key, &_i, &ok)) {                                          \
            (dst) = (defv);                                                               \
            if (warn) rsvdt_printf("Warning: '%s' variable not found. Setting '%s' to default value: %d\n", key, key, (int)(defv)); \
        } else if (!ok) {                                                                 \
            return rsvdt_err(RSVDT_ERR_USER, "ReadUserInput", "'%s' is not an integer", key); \
        } else (dst) = _i;                                                                \
    } while (0)
#define REQ_STR(key, dst, msg)                                                            \
    do {                                                                                  \
        if (!rsvdt_opt_get_string(key, (dst), PETSC_MAX_PATH_LEN))                        \
            return rsvdt_err(RSVDT_ERR_USER, "ReadUserInput", msg);                       \
    } while (0)

static PetscErrorCode read_flag_dir(const char *flag, const char *dirkey, PetscBool *flg, char *dir)
{
    OPT_BOOL(flag, *flg, PETSC_FALSE, 1);
    if (*flg && !rsvdt_opt_get_string(dirkey, dir, PETSC_MAX_PATH_LEN))
        return rsvdt_err(RSVDT_ERR_USER, "ReadUserInput", "Must specify '%s'", dirkey);
    return 0;
}

/* SourceCode/ReadUserInput.c:5-175.  Extension keys (no counterpart in the
 * reference): Precision (fp64|fp32), ForcingChunk, SpMMLanes, UseGraph,
 * Autotune, XCDRemap, Verbose, RK4Form (horner|stages). */
PetscErrorCode ReadUserInput(RSVDt_vars *RSVDt, Weight_matrices *W, LNS_vars *LNS, Resolvent_matrices *Res,
                             TransRun_vars *TR, Directories *dirs)
{
    (void)Res;
    rsvdt_printf("\n*** Reading user inputs ***\n");
    char fname[PETSC_MAX_PATH_LEN];
    if (!rsvdt_opt_get_string("inputs", fname, sizeof(fname)) || !fname[0])
        return rsvdt_err(RSVDT_ERR_USER, "ReadUserInput", "Must specify inputs variables via -inputs");
    snprintf(g_rsvdt.inputs_file, sizeof(g_rsvdt.inputs_file), "%s", fname);
    RSVDT_CHK(rsvdt_yaml_insert_file(fname));
    memset(dirs, 0, sizeof(*dirs));
    REQ_REAL("TransientLength", RSVDt->TS.TransientLength, "Must specify 'TransientLength'");
    REQ_REAL("dt", RSVDt->TS.dt, "Must specify 'dt'");
    REQ_REAL("w", RSVDt->RSVD.w, "Must specify 'dt'");          /* the reference's message (ReadUserInput.c:28) */
    RSVDT_CHK(read_flag_dir("InputForcingFlg", "InputForcingDir", &RSVDt->RSVD.InputForcingFlg, dirs->InputForcingDir));
    RSVDT_CHK(read_flag_dir("InvInputWeightFlg", "InvInputWeightDir", &W->InvInputWeightFlg, dirs->InvInputWeightDir));
    RSVDT_CHK(read_flag_dir("OutputWeightFlg", "OutputWeightDir", &W->OutputWeightFlg, dirs->OutputWeightDir));
    RSVDT_CHK(read_flag_dir("InvOutputWeightFlg", "InvOutputWeightDir", &W->InvOutputWeightFlg, dirs->InvOutputWeightDir));
    RSVDT_CHK(read_flag_dir("InputMatrixFlg", "InputMatrixDir", &W->InputMatrixFlg, dirs->InputMatrixDir));
    RSVDT_CHK(read_flag_dir("OutputMatrixFlg", "OutputMatrixDir", &W->OutputMatrixFlg, dirs->OutputMatrixDir));
    OPT_BOOL("TwoPI", RSVDt->RSVD.TwoPI, PETSC_FALSE, 1);
    RSVDt->RSVD.w *= RSVDt->RSVD.TwoPI ? 2 * PI_D : 1;
    OPT_BOOL("TransRun", TR->TransRun, PETSC_FALSE, 1);
    OPT_BOOL("TransientRemoval", RSVDt->TS.TransientRemoval, PETSC_FALSE, 1);
    OPT_BOOL("DiscFlg", LNS->RSVDt.Disc.DiscFlg, PETSC_FALSE, 1);
    OPT_INT("SaveResultsOpt", RSVDt->SaveResultsOpt, 1, 1);
    REQ_STR("RootDir", dirs->RootDir, "Must specify 'RootDir'");
    REQ_STR("ResultsDir", dirs->ResultsDir, "Must specify 'ResultsDir'");
    REQ_STR("OperatorDir", dirs->OperatorDir, "Must specify 'OperatorDir'");
    OPT_INT("k", RSVDt->RSVD.k, 3, 1);
    OPT_INT("q", RSVDt->RSVD.q, 0, 1);
    {
        int ok; PetscInt v;
        if (!rsvdt_opt_get_int("Nw", &v, &ok)) return rsvdt_err(RSVDT_ERR_USER, "ReadUserInput", "Must specify 'Nw'");
        if (!ok) return rsvdt_err(RSVDT_ERR_USER, "ReadUserInput", "'Nw' is not an integer");
        RSVDt->RSVD.Nw = v;
    }
    OPT_INT("RandSeed", RSVDt->RSVD.RandSeed, 1373, 1);
    OPT_INT("Display", RSVDt->Display, 2, 1);
    {
        int ok; double b = 0;
        int set = rsvdt_opt_get_real("beta", &b, &ok);
        if (set) LNS->RSVDt.Disc.beta = b;
        if (!set && LNS->RSVDt.Disc.DiscFlg)
            return rsvdt_err(RSVDT_ERR_USER, "ReadUserInput", "Discounting flag is on! Either set Discounting to zero or specify 'beta'");
        if (LNS->RSVDt.Disc.beta < 0 && LNS->RSVDt.Disc.DiscFlg)
            return rsvdt_err(RSVDT_ERR_USER, "ReadUserInput", "'beta' must be positive, current value: %g", LNS->RSVDt.Disc.beta);
    }
    {   /* ReadUserInput.c:134-141: TransICDir is required unless TransRun is set without TransICFlg */
        int ok; PetscBool b = PETSC_FALSE;
        int set = rsvdt_opt_get_bool("TransICFlg", &b, &ok);
        if (!set && TR->TransRun) {
            TR->TransICFlg = PETSC_FALSE;
            rsvdt_printf("Warning: 'TransICFlg' variable not found. Setting 'TransICFlg' to default value: 0\n");
        } else {
            if (set) TR->TransICFlg = b;
            if (!rsvdt_opt_get_string("TransICDir", dirs->TransICDir, PETSC_MAX_PATH_LEN))
                return rsvdt_err(RSVDT_ERR_USER, "ReadUserInput", "Must specify 'TransICDir'");
        }
    }
    {
        int ok; double v;
        if (!rsvdt_opt_get_real("TransDivVal", &v, &ok)) { if (TR->TransRun) TR->TransDivVal = 1e3; } else TR->TransDivVal = v;
        if (!rsvdt_opt_get_real("TransConVal", &v, &ok)) { if (TR->TransRun) TR->TransConVal = 1e-8; } else TR->TransConVal = v;
    }
    OPT_BOOL("TransSave", TR->TransSave, PETSC_FALSE, TR->TransRun);
    OPT_BOOL("TransRemovalEst", TR->TransRemovalEst, PETSC_FALSE, TR->TransRun);
    OPT_INT("TransSaveMod", TR->TransSaveMod, 100, TR->TransRun);
    OPT_INT("TransPeriods", TR->TransPeriods, 1, TR->TransRun);

    /* engine extension keys */
    RSVDT_Options o = g_rsvdt.opt;
    char s[64];
    if (rsvdt_opt_get_string("Precision", s, sizeof(s))) {
        if (!strcmp(s, "fp32") || !strcmp(s, "32")) o.precision = 32;
        else if (!strcmp(s, "fp64") || !strcmp(s, "64")) o.precision = 64;
        else return rsvdt_err(RSVDT_ERR_USER, "ReadUserInput", "'Precision' must be fp64 or fp32");
    }
    { int ok; PetscInt v; if (rsvdt_opt_get_int("ForcingChunk", &v, &ok) && ok) o.chunk_steps = (int)v; }
    { int ok; PetscInt v; if (rsvdt_opt_get_int("SpMMLanes", &v, &ok) && ok) o.spmm_lanes = (int)v; }
    { int ok; PetscBool b; if (rsvdt_opt_get_bool("UseGraph", &b, &ok) && ok) o.use_graph = b; }
    { int ok; PetscBool b; if (rsvdt_opt_get_bool("FuseSmall", &b, &ok) && ok) o.fuse_small = b; }
    { int ok; PetscInt v; if (rsvdt_opt_get_int("SpMMVariant", &v, &ok) && ok) o.spmm_variant = (int)v; }
    { int ok; PetscInt v; if (rsvdt_opt_get_int("TRGroup", &v, &ok) && ok) o.tr_group = (int)(v > 0 ? v : 0); }
    { int ok; PetscBool b; if (rsvdt_opt_get_bool("Autotune", &b, &ok) && ok) o.autotune = b; }
    { int ok; PetscBool b; if (rsvdt_opt_get_bool("XCDRemap", &b, &ok) && ok) o.xcd_remap = b; }
    { int ok; PetscBool b; if (rsvdt_opt_get_bool("Verbose", &b, &ok) && ok) o.verbose = b; }
    if (rsvdt_opt_get_string("RK4Form", s, sizeof(s))) {
        if (!strcmp(s, "horner") || !strcmp(s, "0")) o.rk4_form = 0;
        else if (!strcmp(s, "stages") || !strcmp(s, "1")) o.rk4_form = 1;
        else return rsvdt_err(RSVDT_ERR_USER, "ReadUserInput", "'RK4Form' must be horner or stages");
    }
    if (o.precision == 32 && RSVDt->TS.TransientRemoval) {
        /* kept in FP32, the snapshots would carry the state's rounding into
         * the removal's correction of the transient (on the 2D flow operator
         * the gains then differ from the FP64 form by up to 4e-5) */
        rsvdt_printf("Note: 'Precision: fp32' with transient removal: the steps before the first snapshot run in FP32, "
                     "the snapshot period in FP64\n");
    }
    RSVDT_CHK(RSVDT_SetOptions(&o));
    rsvdt_printf("\n");
    return 0;
}

/* ------------------------------------------------------------------ SetupTimeFreqGrid */
/* SourceCode/SetupTimeFreqGrid.c:16-21 */
PetscErrorCode SetupTimeFreqGrid(RSVDt_vars *RSVDt, TransRun_vars *TR)
{
    if (RSVDt->RSVD.w <= 0 || RSVDt->RSVD.Nw <= 0 || RSVDt->TS.dt <= 0)
        return rsvdt_err(RSVDT_ERR_USER, "SetupTimeFreqGrid", "w, Nw and dt must be positive");
    double T_ss = 2 * PI_D / RSVDt->RSVD.w;
    double dt_w = T_ss / (double)RSVDt->RSVD.Nw;
    RSVDt->TS.dt = dt_w / ceil(dt_w / RSVDt->TS.dt);
    RSVDt->TS.Ns = (PetscInt)round(T_ss / RSVDt->TS.dt);
    RSVDt->TS.Nt = (PetscInt)round(RSVDt->TS.TransientLength / RSVDt->TS.dt);
    RSVDt->TS.ResRatio = RSVDt->TS.Ns / RSVDt->RSVD.Nw;
    if (!TR || !TR->TransRun) {
        rsvdt_printf("RSVD parameters:\n@ k        = %d\n@ q        = %d\n@ w        = %g\n\n", (int)RSVDt->RSVD.k,
                     (int)RSVDt->RSVD.q, RSVDt->RSVD.w);
        rsvdt_printf("Time-stepping parameters:\n@ TransientLength = %g\n@ dt              = %g\n@ Ns              = %d\n"
                     "@ Nt              = %d\n", RSVDt->TS.TransientLength, RSVDt->TS.dt, (int)RSVDt->TS.Ns,
                     (int)RSVDt->TS.Nt);
        rsvdt_printf(RSVDt->TS.TransientRemoval ? "Transient removal strategy is requested\n\n"
                                                : "Transient removal strategy is NOT requested\n\n");
    }
    return 0;
}

/* ------------------------------------------------------------------ LNSStructure */
/* SourceCode/LNSStructure.c:22-56: load A, detect a real operator
 * (||Im A||_F < 1e-12), set Nw_eff, apply discounting A <- A - beta I. */
PetscErrorCode LNSStructure(LNS_vars *LNS, RSVDt_vars *RSVDt, TransRun_vars *TR, Directories *dirs)
{
    double t1 = RSVDT_Time();
    rsvdt_printf("Loading the operator: %s%s \n", dirs->RootDir, dirs->OperatorDir);
    snprintf(dirs->IO_dir, PETSC_MAX_PATH_LEN, "%s%s", dirs->RootDir, dirs->OperatorDir);
    Mat A;
    RSVDT_CHK(RSVDT_MatLoad(dirs->IO_dir, 1, &A));
    if (A->rows != A->cols)
        return rsvdt_err(RSVDT_ERR_USER, "LNSStructure", "operator must be square (%lld x %lld)", (long long)A->rows,
                         (long long)A->cols);
    if (LNS->RSVDt.Disc.DiscFlg) {
        /* MatShift: add -beta on the diagonal, inserting missing diagonal entries */
        PetscInt n = A->rows, extra = 0;
        for (PetscInt r = 0; r < n; r++) {
            int has = 0;
            for (int64_t e = A->h_rowptr[r]; e < A->h_rowptr[r + 1]; e++) if (A->h_col[e] == r) has = 1;
            if (!has) extra++;
        }
        int64_t *rp = malloc(sizeof(int64_t) * (size_t)(n + 1));
        int64_t *ci = malloc(sizeof(int64_t) * (size_t)(A->nnz + extra + 1));
        double *v = malloc(sizeof(double) * 2 * (size_t)(A->nnz + extra + 1));
        if (!rp || !ci || !v) return rsvdt_err(RSVDT_ERR_MEM, "LNSStructure", "out of memory");
        int64_t p = 0;
        rp[0] = 0;
        for (PetscInt r = 0; r < n; r++) {
            int done = 0;
            for (int64_t e = A->h_rowptr[r]; e < A->h_rowptr[r + 1]; e++) {
                if (!done && A->h_col[e] > r) { ci[p] = r; v[2 * p] = -LNS->RSVDt.Disc.beta; v[2 * p + 1] = 0; p++; done = 1; }
                ci[p] = A->h_col[e]; v[2 * p] = A->h_val[2 * e]; v[2 * p + 1] = A->h_val[2 * e + 1];
                if (A->h_col[e] == r) { v[2 * p] -= LNS->RSVDt.Disc.beta; done = 1; }
                p++;
            }
            if (!done) { ci[p] = r; v[2 * p] = -LNS->RSVDt.Disc.beta; v[2 * p + 1] = 0; p++; }
            rp[r + 1] = p;
        }
        Mat A2;
        PetscErrorCode ierr = RSVDT_MatCreateAIJ(n, n, rp, ci, v, &A2);
        free(rp); free(ci); free(v);
        RSVDT_CHK(ierr);
        RSVDT_MatDestroy(&A);
        A = A2;
    }
    LNS->A = A;
    RSVDt->RSVD.N = A->rows;
    rsvdt_printf("The operator is loaded in %f seconds (N = %d) \n", RSVDT_Time() - t1, (int)RSVDt->RSVD.N);
    RSVDt->TS.RealOperator = A->is_real ? PETSC_TRUE : PETSC_FALSE;
    if (!TR || !TR->TransRun)
        rsvdt_printf(RSVDt->TS.RealOperator ? "\nThe LNS operator is real-valued!\n" : "\nThe LNS operator is complex-valued!\n");
    RSVDt->RSVD.Nw_eff = RSVDt->TS.RealOperator ? RSVDt->RSVD.Nw / 2 : RSVDt->RSVD.Nw;
    if (!TR || !TR->TransRun) rsvdt_printf("The effective number of frequencies, Nw = %d\n", (int)RSVDt->RSVD.Nw_eff);
    if (LNS->RSVDt.Disc.DiscFlg) rsvdt_printf("\n---- Discounting with beta = %g ----\n\n", LNS->RSVDt.Disc.beta);
    return 0;
}

/* ------------------------------------------------------------------ CreateDFTiDFTMats */
/* SourceCode/CreateDFTiDFTMats.c:18-93.  idft (after the reference's in-place
 * transpose) is Nw_eff x 2Ns with idft(m, j) = conj(exp(alpha*m_alias*j)) * 2/(2Ns),
 * alpha = -2 pi i/(2Ns); dft is Nw x Nw_eff with dft(s, m) = exp(-2 pi i m s/Nw).
 * Device layouts: idft column j contiguous (block width 1), dft row-major. */
PetscErrorCode CreateDFTiDFTMats(RSVDt_vars *RSVDt, DFT_matrices *D, LNS_vars *LNS)
{
    (void)LNS;
    PetscInt Nt2 = 2 * RSVDt->TS.Ns, Nwe = RSVDt->RSVD.Nw_eff, Nw = RSVDt->RSVD.Nw;
    if (Nwe <= 0 || Nt2 <= 0) return rsvdt_err(RSVDT_ERR_USER, "CreateDFTiDFTMats", "grid not set up");
    double *dft = malloc(sizeof(double) * 2 * (size_t)(Nw * Nwe));
    if (!dft) return rsvdt_err(RSVDT_ERR_MEM, "CreateDFTiDFTMats", "out of memory");
    /* idft: built on the device from the exact integer phase (m_alias j) mod
     * 2Ns of each entry (k_idft_table), Nw_eff x 2Ns column-major == block
     * width 1; ordered before every reader on the engine's stream */
    PetscErrorCode ierr = rsvdt_mat_alloc_dense(Nwe, Nt2, 1, 0, &D->idft);
    if (ierr) { free(dft); return ierr; }
    int hr = rsvdt_hip_idft_table(Nwe, Nt2, RSVDt->TS.RealOperator ? 1 : 0, D->idft->d, g_rsvdt.stream);
    if (hr) { free(dft); RSVDT_MatDestroy(&D->idft); return rsvdt_err(hr, "CreateDFTiDFTMats", "%s", rsvdt_hip_last_error()); }
    const double a2 = -2.0 * PI_D / (RSVDt->TS.RealOperator ? (double)(2 * Nwe) : (double)Nwe);
    for (PetscInt iw = 0; iw < Nwe; iw++)
        for (PetscInt s = 0; s < Nw; s++) {
            const double ph = a2 * (double)iw * (double)s;
            dft[2 * ((size_t)iw * Nw + s)] = cos(ph);
            dft[2 * ((size_t)iw * Nw + s) + 1] = sin(ph);
        }
    /* dft: Nw x Nw_eff, stored row-major (one block of Nw_eff columns) */
    ierr = RSVDT_MatCreateDenseBlocked(Nw, Nwe, Nwe, dft, &D->dft);
    free(dft);
    return ierr;
}

/* ------------------------------------------------------------------ ReadWeightMats */
/* SourceCode/ReadWeightMats.c:26-99 (same checks and messages). */
PetscErrorCode ReadWeightMats(RSVDt_vars *RSVDt, Weight_matrices *W, Directories *dirs)
{
    PetscInt row1, col1, row2, col2;
    PetscInt N = RSVDt->RSVD.N;
    W->W_f_sqrt_inv = W->B = W->W_q_sqrt_inv = W->W_q_sqrt = W->C = NULL;
    if (W->InvInputWeightFlg) {
        snprintf(dirs->IO_dir, PETSC_MAX_PATH_LEN, "%s%s", dirs->RootDir, dirs->InvInputWeightDir);
        rsvdt_printf("Reading the inverse input weight matrix  : %s\n", dirs->IO_dir);
        RSVDT_CHK(RSVDT_MatLoad(dirs->IO_dir, 1, &W->W_f_sqrt_inv));
        row1 = W->W_f_sqrt_inv->rows; col1 = W->W_f_sqrt_inv->cols;
        if (row1 != col1)
            return rsvdt_err(RSVDT_ERR_USER, "ReadWeightMats", "Input weight matrix (W_f_sqrt_inv) must be square, current size = %d x %d", (int)row1, (int)col1);
    } else { row1 = N; col1 = N; }
    if (W->InputMatrixFlg) {
        snprintf(dirs->IO_dir, PETSC_MAX_PATH_LEN, "%s%s", dirs->RootDir, dirs->InputMatrixDir);
        rsvdt_printf("Reading the input matrix                 : %s\n", dirs->IO_dir);
        RSVDT_CHK(RSVDT_MatLoad(dirs->IO_dir, 1, &W->B));
        row2 = W->B->rows; col2 = W->B->cols;
        if (col2 != col1)
            return rsvdt_err(RSVDT_ERR_USER, "ReadWeightMats", "Size mismatch between input matrix (B) and input weight matrix (W_f_sqrt_inv), %d != %d", (int)col2, (int)col1);
        if (row2 != N)
            return rsvdt_err(RSVDT_ERR_USER, "ReadWeightMats", "Input matrix (B) must have %d rows, current size = %d x %d", (int)N, (int)row2, (int)col2);
        RSVDt->RSVD.Nb = col2;
    } else {
        RSVDt->RSVD.Nb = N;
        if (W->InvInputWeightFlg && RSVDt->RSVD.Nb != col1)
            return rsvdt_err(RSVDT_ERR_USER, "ReadWeightMats", "Size mismatch between input matrix (B) and input weight matrix (W_f_sqrt_inv)");
    }
    if (W->InvOutputWeightFlg) {
        snprintf(dirs->IO_dir, PETSC_MAX_PATH_LEN, "%s%s", dirs->RootDir, dirs->InvOutputWeightDir);
        rsvdt_printf("Reading the inverse output weight matrix : %s\n", dirs->IO_dir);
        RSVDT_CHK(RSVDT_MatLoad(dirs->IO_dir, 1, &W->W_q_sqrt_inv));
        row1 = W->W_q_sqrt_inv->rows; col1 = W->W_q_sqrt_inv->cols;
        if (row1 != col1)
            return rsvdt_err(RSVDT_ERR_USER, "ReadWeightMats", "Inverse output weight matrix (W_q_sqrt_inv) must be square, current size = %d x %d", (int)row1, (int)col1);
    }
    if (W->OutputWeightFlg) {
        snprintf(dirs->IO_dir, PETSC_MAX_PATH_LEN, "%s%s", dirs->RootDir, dirs->OutputWeightDir);
        rsvdt_printf("Reading the output weight matrix         : %s\n", dirs->IO_dir);
        RSVDT_CHK(RSVDT_MatLoad(dirs->IO_dir, 1, &W->W_q_sqrt));
        row1 = W->W_q_sqrt->rows; col1 = W->W_q_sqrt->cols;
        if (row1 != col1)
            return rsvdt_err(RSVDT_ERR_USER, "ReadWeightMats", "Output weight matrix (W_q_sqrt) must be square, current size = %d x %d", (int)row1, (int)col1);
    } else { row1 = N; col1 = N; }
    if (W->OutputMatrixFlg) {
        snprintf(dirs->IO_dir, PETSC_MAX_PATH_LEN, "%s%s", dirs->RootDir, dirs->OutputMatrixDir);
        rsvdt_printf("Reading the output matrix                : %s\n", dirs->IO_dir);
        RSVDT_CHK(RSVDT_MatLoad(dirs->IO_dir, 1, &W->C));
        row2 = W->C->rows; col2 = W->C->cols;
        if (col2 != N)
            return rsvdt_err(RSVDT_ERR_USER, "ReadWeightMats", "Output matrix (C) must have %d columns, current size = %d x %d", (int)N, (int)row2, (int)col2);
        RSVDt->RSVD.Nc = row2;
        if (row2 != row1)
            return rsvdt_err(RSVDT_ERR_USER, "ReadWeightMats", "Size mismatch between output matrix (C) and output weight matrix (W_q_sqrt), %d != %d", (int)row2, (int)row1);
        if (W->InvOutputWeightFlg && row2 != W->W_q_sqrt_inv->rows)
            return rsvdt_err(RSVDT_ERR_USER, "ReadWeightMats", "Size mismatch between output matrix (C) and inverse output weight matrix (W_q_sqrt_inv), %d != %d", (int)row2, (int)W->W_q_sqrt_inv->rows);
    } else {
        RSVDt->RSVD.Nc = N;
        if (W->OutputWeightFlg && RSVDt->RSVD.Nc != row1)
            return rsvdt_err(RSVDT_ERR_USER, "ReadWeightMats", "Size mismatch between output matrix (C) and output weight matrix (W_q_sqrt)");
        if (W->InvOutputWeightFlg && RSVDt->RSVD.Nc != W->W_q_sqrt_inv->rows)
            return rsvdt_err(RSVDT_ERR_USER, "ReadWeightMats", "Size mismatch between output matrix (C) and inverse output weight matrix (W_q_sqrt_inv)");
    }
    return 0;
}

/* ------------------------------------------------------------------ CreateRandomMat */
/* SourceCode/CreateRandomMat.c:21-40.  The sketch is read from
 * InputForcingDir (k = cols/Nw_eff) or filled with seeded random numbers.
 * Deliberate difference: the reference ignores RandSeed for the sketch
 * (MatSetRandom(Y_hat, NULL), :39); here RandSeed seeds it, so runs are
 * reproducible.  RandDist (extension key): uniform (default; PETSc's
 * [0,1) + i[0,1)) or gaussian (complex N(0,1)).  Learned columns set by
 * RSVDT_SetSketchColumns replace the first test vectors. */
PetscErrorCode CreateRandomMat(RSVD_matrices *R, RSVDt_vars *RSVDt, Directories *dirs)
{
    PetscInt Nb = RSVDt->RSVD.Nb > 0 ? RSVDt->RSVD.Nb : RSVDt->RSVD.N;
    PetscInt Nwe = RSVDt->RSVD.Nw_eff;
    if (RSVDt->RSVD.InputForcingFlg) {
        snprintf(dirs->IO_dir, PETSC_MAX_PATH_LEN, "%s%s", dirs->RootDir, dirs->InputForcingDir);
        rsvdt_printf("Reading the input forcing matrix         : %s\n", dirs->IO_dir);
        Mat F;
        RSVDT_CHK(RSVDT_MatLoad(dirs->IO_dir, 0, &F));
        if (F->rows != Nb) {
            PetscInt rr = F->rows, cc = F->cols;
            RSVDT_MatDestroy(&F);
            return rsvdt_err(RSVDT_ERR_USER, "CreateRandomMat", "Input forcing must be have %d rows, current size = %d x %d",
                             (int)Nb, (int)rr, (int)cc);
        }
        if (F->cols % Nwe)
            return rsvdt_err(RSVDT_ERR_USER, "CreateRandomMat", "Input forcing has %d columns, not a multiple of Nw_eff = %d",
                             (int)F->cols, (int)Nwe);
        RSVDt->RSVD.k = F->cols / Nwe;
        rsvdt_printf("Based on the input forcing: k = %d\n", (int)RSVDt->RSVD.k);
        RSVDT_CHK(rsvdt_mat_to_layout(F, RSVDt->RSVD.k, 0));
        R->Y_hat = F;
        return 0;
    }
    rsvdt_printf("Generating a random forcing matrix\n");
    PetscInt k = RSVDt->RSVD.k;
    char dist[32] = "uniform";
    rsvdt_opt_get_string("RandDist", dist, sizeof(dist));
    int gauss = !strcmp(dist, "gaussian") || !strcmp(dist, "normal");
    uint64_t seed = (uint64_t)RSVDt->RSVD.RandSeed;
    Mat Y;
    RSVDT_CHK(rsvdt_mat_alloc_dense(Nb, k * Nwe, k, 0, &Y));
    RSVDT_HIPCHK("CreateRandomMat", rsvdt_hip_fill_random(Nb, k, Nwe, seed, gauss, Y->d, g_rsvdt.stream));
    PetscInt kl = 0;
    if ((g_rsvdt.sketch_host || g_rsvdt.sketch_dev) && g_rsvdt.sketch_N == Nb && g_rsvdt.sketch_Nw_eff == Nwe)
        kl = g_rsvdt.sketch_cols;
    if (kl > k) kl = k;
    if (kl > 0 && g_rsvdt.sketch_dev) {
        /* device columns (FP32 interleaved, point-major): converted in place into
         * Y's blocks, after the whole device is idle -- the caller may have
         * produced them on any stream, which the engine's stream does not wait for */
        RSVDT_HIPCHK("CreateRandomMat", rsvdt_hip_sync());
        RSVDT_HIPCHK("CreateRandomMat", rsvdt_hip_sketch_cols_f32(Nwe, Nb, kl, k, g_rsvdt.sketch_dev, g_rsvdt.sketch_ld,
                                                                  Y->d, g_rsvdt.stream));
    } else if (kl > 0) {
        /* learned columns: host N x (kl*Nwe) column-major, column m*kl + j -> block m, column j */
        size_t n = (size_t)Nb * (size_t)kl * (size_t)Nwe;
        double *tmp = malloc(sizeof(double) * 2 * n);
        double *dtmp = NULL;
        if (!tmp) return rsvdt_err(RSVDT_ERR_MEM, "CreateRandomMat", "out of memory");
        for (PetscInt m = 0; m < Nwe; m++)
            for (PetscInt j = 0; j < kl; j++)
                for (PetscInt r = 0; r < Nb; r++) {
                    const double *src = g_rsvdt.sketch_host + 2 * ((size_t)(m * kl + j) * Nb + r);
                    double *dst = tmp + 2 * (((size_t)m * Nb + r) * kl + j);
                    dst[0] = src[0]; dst[1] = src[1];
                }
        int rc = rsvdt_hip_malloc((void **)&dtmp, sizeof(double) * 2 * n);
        if (!rc) rc = rsvdt_hip_memcpy_h2d(dtmp, tmp, sizeof(double) * 2 * n);
        if (!rc) rc = rsvdt_hip_zcopy3(Nwe, Nb, kl, dtmp, Nb * kl, kl, 1, Y->d, Nb * k, k, 1, g_rsvdt.stream);
        if (!rc) rc = rsvdt_hip_sync();
        free(tmp);
        rsvdt_hip_free(dtmp);
        if (rc) return rsvdt_err(rc, "CreateRandomMat", "%s", rsvdt_hip_last_error());
    }
    R->Y_hat = Y;
    return 0;
}

/* ------------------------------------------------------------------ CreateResultsDir */
/* SourceCode/CreateResultsDir.c:20-30: RootDir/ResultsDir/<FileName><i>/ with
 * the first unused i. */
PetscErrorCode CreateResultsDir(Directories *dirs, const char *FileName)
{
    int ind = 0;
    for (;;) {
        snprintf(dirs->FolderDir, PETSC_MAX_PATH_LEN, "%s%s%s%d/", dirs->RootDir, dirs->ResultsDir, FileName, ind);
        if (access(dirs->FolderDir, F_OK) != 0) break;
        ind++;
    }
    if (mkdir(dirs->FolderDir, 0755) != 0)
        return rsvdt_err(RSVDT_ERR_FILE, "CreateResultsDir", "cannot create '%s'", dirs->FolderDir);
    return 0;
}

/* ------------------------------------------------------------------ PreProcessing */
/* SourceCode/PreProcessing.c:12-71 (the ReadUserInput return code is
 * checked here, unlike the reference, :27-29). */
PetscErrorCode PreProcessing(RSVDt_vars *RSVDt, Weight_matrices *W, LNS_vars *LNS, RSVD_matrices *R,
                             Resolvent_matrices *Res, TransRun_vars *TR, DFT_matrices *D, Directories *dirs)
{
    memset(R, 0, sizeof(*R));
    memset(Res, 0, sizeof(*Res));
    memset(D, 0, sizeof(*D));
    RSVDT_CHK(ReadUserInput(RSVDt, W, LNS, Res, TR, dirs));
    if (!TR->TransRun)
        rsvdt_printf("********************************************\n*************** Problem info ***************\n"
                     "********************************************\n\n");
    RSVDT_CHK(SetupTimeFreqGrid(RSVDt, TR));
    RSVDT_CHK(LNSStructure(LNS, RSVDt, TR, dirs));
    LNS->RSVDt.RSVD = RSVDt->RSVD;
    LNS->RSVDt.TS = RSVDt->TS;
    RSVDT_CHK(CreateDFTiDFTMats(RSVDt, D, LNS));
    if (!TR->TransRun) RSVDT_CHK(ReadWeightMats(RSVDt, W, dirs));
    if (!TR->TransRun) RSVDT_CHK(CreateRandomMat(R, RSVDt, dirs));
    RSVDT_CHK(CreateResultsDir(dirs, TR->TransRun ? "TransientSnapshots_" : "ResolventModes_"));
    return 0;
}
