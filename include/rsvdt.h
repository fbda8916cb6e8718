/*
 * rsvdt.h -- drop-in C interface of the MI355X RSVD-Delta-t engine.
 *
 * The structs, field names and entry-point signatures are those of the
 * reference (SourceCode/Variables.h:5-110 and the one-line prototype headers
 * SourceCode/<Name>.h:5), so a driver written against the reference (its
 * main.c sequence, SourceCode/main.c:107-150) compiles and links against
 * librsvdt unchanged, apart from including this header in place of the
 * PETSc/SLEPc ones and calling RSVDT_Initialize/RSVDT_Finalize in place of
 * SlepcInitialize/SlepcFinalize.
 *
 * PETSc types are re-declared with the reference build's configuration
 * (--with-scalar-type=complex --with-64-bit-indices): PetscInt is int64_t,
 * PetscScalar double complex.  Mat and Vec are this library's opaque handles
 * to device-resident (HBM) data; nothing here depends on PETSc.
 *
 * Every entry point returns 0 on success and a nonzero PetscErrorCode on
 * failure (RSVDT_ERR_USER = 83 for invalid input, as PETSC_ERR_USER; the
 * RSVDT_HIP_* codes of rsvdt_hip.h for device failures), after printing a
 * message that names the entry point.  Entry points are not re-entrant
 * (TS.DirAdj is a mode flag, as in the reference); one host thread drives
 * one GPU on one stream.
 */
#ifndef RSVDT_H
#define RSVDT_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#ifndef PETSC_MAX_PATH_LEN
#define PETSC_MAX_PATH_LEN 4096
#endif

typedef int PetscErrorCode;
typedef int64_t PetscInt;
typedef double PetscReal;
typedef double PetscLogDouble;
typedef enum { PETSC_FALSE = 0, PETSC_TRUE = 1 } PetscBool;
typedef struct _p_RSVDT_Mat *Mat;
typedef struct _p_RSVDT_Vec *Vec;

#define RSVDT_ERR_USER 83
#define RSVDT_ERR_FILE 65
#define RSVDT_ERR_MEM 55

/* ---- SourceCode/Variables.h:5-110 (field names and order kept) ----------- */
typedef struct {
    PetscInt N, Nb, Nc, k, q, Nw, Nw_eff;
    PetscReal w;
    PetscBool TwoPI, InputForcingFlg;
    PetscInt RandSeed;
} RSVD_vars;

typedef struct {
    PetscInt Ns, Nt, ResRatio;
    PetscReal TransientLength, dt;
    PetscBool RealOperator, TransientRemoval, DirAdj;
} TS_vars;

typedef struct {
    PetscBool DiscFlg;
    PetscReal beta;
} Discounting;

typedef struct {
    RSVD_vars RSVD;
    TS_vars TS;
    Discounting Disc;
    PetscInt Display;
    PetscInt SaveResultsOpt;
} RSVDt_vars;

typedef struct {
    PetscBool TransRun, TransSave, TransRemovalEst, TransICFlg;
    PetscInt TransSaveMod, TransPeriods;
    PetscReal TransDivVal, TransConVal;
} TransRun_vars;

typedef struct { Mat idft, dft; } DFT_matrices;

typedef struct {
    Mat A;
    RSVDt_vars RSVDt;
} LNS_vars;

typedef struct { Vec F1, F2, F3, k1, k2, k3, k4, y_temp; } TS_matrices;

typedef struct {
    Mat W_q_sqrt, W_q_sqrt_inv, W_f_sqrt_inv, B, C;
    PetscBool InvInputWeightFlg, OutputWeightFlg, InvOutputWeightFlg, InputMatrixFlg, OutputMatrixFlg;
} Weight_matrices;

typedef struct { Mat F_hat, Y_hat; } RSVD_matrices;
typedef struct { Mat M_tilde_all, Y_hat_up; } TS_removal_matrices;
typedef struct { Mat U_hat, V_hat, S_hat; } Resolvent_matrices;

typedef struct {
    char RootDir[PETSC_MAX_PATH_LEN];
    char ResultsDir[PETSC_MAX_PATH_LEN];
    char OperatorDir[PETSC_MAX_PATH_LEN];
    char IO_dir[PETSC_MAX_PATH_LEN];
    char FolderDir[PETSC_MAX_PATH_LEN];
    char InputForcingDir[PETSC_MAX_PATH_LEN];
    char InvInputWeightDir[PETSC_MAX_PATH_LEN];
    char OutputWeightDir[PETSC_MAX_PATH_LEN];
    char InvOutputWeightDir[PETSC_MAX_PATH_LEN];
    char InputMatrixDir[PETSC_MAX_PATH_LEN];
    char OutputMatrixDir[PETSC_MAX_PATH_LEN];
    char TransICDir[PETSC_MAX_PATH_LEN];
} Directories;

/* ---- reference entry points (SourceCode/<Name>.h:5) ---------------------- */
/* replaces SourceCode/PreProcessing.h:5-6 */
PetscErrorCode PreProcessing(RSVDt_vars *, Weight_matrices *, LNS_vars *, RSVD_matrices *,
                             Resolvent_matrices *, TransRun_vars *, DFT_matrices *, Directories *);
/* replaces SourceCode/ReadUserInput.h:5 */
PetscErrorCode ReadUserInput(RSVDt_vars *, Weight_matrices *, LNS_vars *, Resolvent_matrices *,
                             TransRun_vars *, Directories *);
/* replaces SourceCode/SetupTimeFreqGrid.h:5 */
PetscErrorCode SetupTimeFreqGrid(RSVDt_vars *, TransRun_vars *);
/* replaces SourceCode/LNSStructure.h:5 */
PetscErrorCode LNSStructure(LNS_vars *, RSVDt_vars *, TransRun_vars *, Directories *);
/* replaces SourceCode/CreateDFTiDFTMats.h:5 */
PetscErrorCode CreateDFTiDFTMats(RSVDt_vars *, DFT_matrices *, LNS_vars *);
/* replaces SourceCode/ReadWeightMats.h:5 */
PetscErrorCode ReadWeightMats(RSVDt_vars *, Weight_matrices *, Directories *);
/* replaces SourceCode/CreateRandomMat.h:5 (learned-sketch hook, see RSVDT_SetSketchColumns) */
PetscErrorCode CreateRandomMat(RSVD_matrices *, RSVDt_vars *, Directories *);
/* replaces SourceCode/CreateResultsDir.h:5 */
PetscErrorCode CreateResultsDir(Directories *, const char *);

/* replaces SourceCode/DirectActionRK4.h:5 */
PetscErrorCode DirectActionRK4(RSVD_matrices *, RSVDt_vars *, LNS_vars *, DFT_matrices *,
                               Weight_matrices *, Directories *, TS_removal_matrices *);
/* replaces SourceCode/AdjointActionRK4.h:5 */
PetscErrorCode AdjointActionRK4(RSVD_matrices *, RSVDt_vars *, LNS_vars *, DFT_matrices *,
                                Weight_matrices *, Directories *, TS_removal_matrices *);
/* replaces SourceCode/PowerIterationRK4.h:5 */
PetscErrorCode PowerIterationRK4(RSVD_matrices *, RSVDt_vars *, LNS_vars *, DFT_matrices *,
                                 Weight_matrices *, Directories *, TS_removal_matrices *);
/* replaces SourceCode/ApplyWeightMats.h:5 */
PetscErrorCode ApplyWeightMats(RSVD_matrices *, RSVDt_vars *, Weight_matrices *, PetscBool);
/* replaces SourceCode/TSActionRK4.h:5 */
PetscErrorCode TSActionRK4(RSVD_matrices *, DFT_matrices *, LNS_vars *, RSVDt_vars *, Directories *,
                           TS_removal_matrices *);
/* replaces SourceCode/DFT.h:5 */
PetscErrorCode DFT(Mat, DFT_matrices *, RSVD_matrices *, RSVDt_vars *);
/* replaces SourceCode/PermuteMat.h:5 and SourceCode/ReversePermuteMat.h:5 */
PetscErrorCode PermuteMat(Mat, RSVDt_vars *);
PetscErrorCode ReversePermuteMat(Mat, RSVDt_vars *);

/* replaces SourceCode/TransientRemovalStrategy.h:5 */
PetscErrorCode TransientRemovalStrategy(Mat, LNS_vars *, RSVD_matrices *, RSVDt_vars *,
                                        TS_removal_matrices *, DFT_matrices *);
/* replaces SourceCode/CreateMtilde.h:5 */
PetscErrorCode CreateMtilde(Mat, LNS_vars *, RSVDt_vars *, TS_removal_matrices *);
/* replaces SourceCode/GalerkinProjAllModes.h:5 */
PetscErrorCode GalerkinProjAllModes(Mat, LNS_vars *, RSVDt_vars *, TS_removal_matrices *, DFT_matrices *);
/* replaces SourceCode/TSExpADeltaT.h:5 */
PetscErrorCode TSExpADeltaT(LNS_vars *, RSVDt_vars *, Mat);
/* replaces SourceCode/QRDecomposition.h:5 */
PetscErrorCode QRDecomposition(Mat);

/* replaces SourceCode/QRAllFreqs.h:5 */
PetscErrorCode QRAllFreqs(RSVD_matrices *, RSVDt_vars *);
/* replaces SourceCode/SVDAllFreqsBeforeAdjoint.h:5 */
PetscErrorCode SVDAllFreqsBeforeAdjoint(RSVD_matrices *, RSVDt_vars *);
/* replaces SourceCode/StoreU.h:5 */
PetscErrorCode StoreU(RSVD_matrices *, Resolvent_matrices *);
/* replaces SourceCode/SVDAllFreqs.h:5 */
PetscErrorCode SVDAllFreqs(RSVD_matrices *, RSVDt_vars *, Weight_matrices *, Resolvent_matrices *);
/* replaces SourceCode/SaveResults.h:5 */
PetscErrorCode SaveResults(Resolvent_matrices *, RSVDt_vars *, Directories *);
/* replaces SourceCode/SaveInputVarsCopy.h:5 */
PetscErrorCode SaveInputVarsCopy(Directories *);

/* replaces SourceCode/TransientRunRK4.h:5 (TS.DirAdj selects A or A^H, as in
 * TSTransRK4; the reference leaves it unset in this mode, the rsvdt driver
 * sets it to 1) */
PetscErrorCode TransientRunRK4(TransRun_vars *, RSVDt_vars *, LNS_vars *, DFT_matrices *, Directories *);
/* replaces SourceCode/TransientRemovalEst.h:5 */
PetscErrorCode TransientRemovalEst(Mat, Vec, PetscInt, RSVDt_vars *, DFT_matrices *, Directories *);
/* replaces SourceCode/TSTransRK4.h:5 (TS_matrices unused, may be NULL) */
PetscErrorCode TSTransRK4(LNS_vars *, TS_matrices *, RSVDt_vars *, PetscInt, Vec);

/* ---- runtime (in place of SlepcInitialize / SlepcFinalize, SourceCode/main.c:107,148) -- */
/* Parses "-inputs <yaml>" and "-Key value" pairs into the options table,
 * selects the GPU (env RSVDT_DEVICE, default 0) and creates the stream. */
PetscErrorCode RSVDT_Initialize(int *argc, char ***argv);
PetscErrorCode RSVDT_Finalize(void);
/* Options table (the role of PETSc's options database) */
PetscErrorCode RSVDT_OptionsSetValue(const char *key, const char *value);
PetscErrorCode RSVDT_OptionsClear(void);
/* Wall-clock seconds (PetscTime). */
PetscLogDouble RSVDT_Time(void);

/* ---- Mat / Vec handles (device-resident; host transfer for tests and I/O) -- */
/* Dense complex FP64 matrix, column-major host data (re,im interleaved), as
 * PETSc's MATDENSE.  host may be NULL (zero-filled). */
PetscErrorCode RSVDT_MatCreateDense(PetscInt rows, PetscInt cols, const double *host_colmajor, Mat *out);
/* Dense matrix whose columns form nb blocks of kb columns (a frequency-major
 * sketch: kb = k); the same host layout as RSVDT_MatCreateDense. */
PetscErrorCode RSVDT_MatCreateDenseBlocked(PetscInt rows, PetscInt cols, PetscInt kb, const double *host_colmajor,
                                           Mat *out);
PetscErrorCode RSVDT_MatGetDense(Mat m, double *host_colmajor);
/* Sparse AIJ (CSR) matrix from host arrays, complex values (re,im interleaved). */
PetscErrorCode RSVDT_MatCreateAIJ(PetscInt rows, PetscInt cols, const int64_t *rowptr, const int64_t *colidx,
                                  const double *vals, Mat *out);
/* Host copy of an AIJ Mat (any argument may be NULL); is_real reports the
 * ||Im A||_F < 1e-12 test of SourceCode/LNSStructure.c:37-40. */
PetscErrorCode RSVDT_MatGetAIJ(Mat m, PetscInt *nnz, int64_t *rowptr, int64_t *colidx, double *vals, int *is_real);
PetscErrorCode RSVDT_MatGetSize(Mat m, PetscInt *rows, PetscInt *cols);
PetscErrorCode RSVDT_MatDestroy(Mat *m);
/* MatLoad / MatView (PETSc binary, big-endian, 64-bit or 32-bit indices). */
PetscErrorCode RSVDT_MatLoad(const char *path, int as_sparse, Mat *out);
PetscErrorCode RSVDT_MatView(Mat m, const char *path);

/* Vec handles: device-resident complex FP64 vectors (VecCreate/VecLoad/VecView). */
PetscErrorCode RSVDT_VecCreate(PetscInt n, const double *host, Vec *out);   /* host NULL: zeros */
PetscErrorCode RSVDT_VecGetValues(Vec v, double *host);
PetscErrorCode RSVDT_VecGetSize(Vec v, PetscInt *n);
PetscErrorCode RSVDT_VecDestroy(Vec *v);
PetscErrorCode RSVDT_VecLoad(const char *path, Vec *out);
PetscErrorCode RSVDT_VecView(Vec v, const char *path);

/* ---- engine controls (no counterpart in the reference) ------------------- */
typedef struct {
    int precision;        /* 64: complex FP64 state (reference); 32: complex FP32 RK4 state and SpMM */
    int chunk_steps;      /* RK4 steps per forcing chunk (0 = automatic) */
    int spmm_lanes;       /* lanes per row in the SpMM (0 = automatic / autotuned) */
    int xcd_remap;        /* contiguous row ranges per XCD */
    int use_graph;        /* capture each forcing chunk's RK4 steps in a hipGraph */
    int autotune;         /* time SpMM lane widths once per operator and keep the fastest */
    int verbose;          /* print phase timings */
    int overlap_forcing;  /* compute the next forcing chunk on a second stream during the RK4 steps */
    int rk4_form;         /* 0: RK4 as y' = P(hA) y + g (Horner stages, forcing folded into one
                             precomputed term per step); 1: the reference's stage order */
    int fuse_small;       /* operators whose state fits in LDS: whole chunks of steps per launch */
    int spmm_variant;     /* -1: autotuned; else force the stage-kernel variant (rsvdt_hip.h) */
    int tr_group;         /* most test vectors per transient-removal group (0: sized to free memory) */
} RSVDT_Options;
PetscErrorCode RSVDT_GetOptions(RSVDT_Options *o);
PetscErrorCode RSVDT_SetOptions(const RSVDT_Options *o);

/* Per-phase device timings of the last action, milliseconds (HIP events). */
typedef struct {
    double rk4_ms;         /* all RK4 stage kernels */
    double forcing_ms;     /* forcing-chunk GEMMs */
    double dft_ms;         /* DFT or transient-removal windows */
    double tr_ms;          /* transient removal total */
    double total_ms;       /* the whole action */
    int64_t steps;         /* RK4 steps taken */
    int64_t rk4_launches;  /* RK4 stage launches */
    int chunk_steps;       /* RK4 steps per forcing chunk */
    int spmm_lanes, spmm_variant, spmm_xcd_remap;   /* SpMM launch shape used */
    int rk4_form;          /* stepping form used (RSVDT_Options.rk4_form) */
} RSVDT_ActionStats;
PetscErrorCode RSVDT_GetActionStats(RSVDT_ActionStats *s);

/* Time `nsteps` forced RK4 steps of the current operator/sketch without the
 * snapshot DFT: the hot loop of TSActionRK4 on the device-resident F_hat
 * (used by benchmarks).  ms_out: device time (HIP events). */
PetscErrorCode RSVDT_BenchSteps(RSVD_matrices *, RSVDt_vars *, LNS_vars *, DFT_matrices *, PetscInt nsteps,
                                PetscInt warmup, double *ms_out, double *rk4_ms_out, double *forcing_ms_out);

/* Learned-sketch hook of CreateRandomMat: the first ncols_learned of the k
 * test vectors are taken from this host matrix (N x ncols_learned*Nw_eff,
 * frequency-major, column-major complex) and the remaining ones are random. */
PetscErrorCode RSVDT_SetSketchColumns(PetscInt N, PetscInt ncols_learned, PetscInt Nw_eff, const double *host);
/* The same from device memory in the sketch network's output layout, FP32
 * complex interleaved, point-major: column j of frequency block m at row r is
 * dev[(m N + r) ld + 2j] + i dev[(m N + r) ld + 2j + 1].  The array must stay
 * valid until CreateRandomMat has run; no host copy is made, and
 * CreateRandomMat waits for the whole device before reading it (the columns
 * may come from work on any stream).  Setting either
 * form (or RSVDT_SetSketchColumns(0, 0, 0, NULL)) replaces the other. */
PetscErrorCode RSVDT_SetSketchColumnsDevice(PetscInt N, PetscInt ncols_learned, PetscInt Nw_eff, const float *dev,
                                            PetscInt ld);

#ifdef __cplusplus
}
#endif
#endif
