/*
 * Internal declarations of the C host layer: the Mat/Vec handles, the
 * options table, the device-resident operator and the error helpers.
 */
#ifndef RSVDT_INTERNAL_H
#define RSVDT_INTERNAL_H

#include <stdio.h>
#include <stdint.h>
#include "rsvdt.h"
#include "rsvdt_hip.h"

/* ---- Mat --------------------------------------------------------------- */
enum { RSVDT_MAT_DENSE = 1, RSVDT_MAT_AIJ = 2 };

/* Tiled form of an operator for the LDS-staged stage kernel (rsvdt_spmm_desc
 * variant 9): tiles of `rows` consecutive rows, each tile's distinct columns
 * (d_tcol from d_tptr[t]) and each nonzero's position in its tile's list
 * (d_lidx); hmax / emax: the most columns / nonzeros of a tile.  rows = 0:
 * not built; rows < 0: not buildable (a tile with more than 65535 columns). */
typedef struct {
    int rows, hmax, emax;
    int32_t *d_tptr, *d_tcol;
    uint16_t *d_lidx;
} rsvdt_tiled;
/* tile heights the tiled stage kernel launches: (256 / columns per pass) x
 * rows per slot */
#define RSVDT_NTILE 5
extern const int rsvdt_tile_rows[RSVDT_NTILE];

/* Dense storage: the columns form nb = cols/kb blocks of kb columns; block b
 * is a rows x kb row-major slab at d + b*rows*kb (complex FP64).  order 0:
 * logical column c is (block c/kb, column c%kb) -- the frequency-major
 * sketch; order 1: logical column c = j*nb + b is (block b, column j) -- the
 * k-major view the reference's PermuteMat produces (SourceCode/PermuteMat.c:24-38),
 * and the layout of the snapshot store Y_all.  Permutations therefore only
 * relabel the handle. */
struct _p_RSVDT_Mat {
    int kind;
    PetscInt rows, cols;
    /* dense */
    PetscInt kb;
    int order;
    double *d;
    /* AIJ (CSR): host copy (int64, complex) and device copies (int32 indices,
     * FP64 values, real when is_real) of the matrix and of its conjugate
     * transpose. */
    PetscInt nnz;
    int is_real;             /* ||Im||_F < 1e-12: values kept real on the device */
    int64_t *h_rowptr, *h_col;
    double *h_val;           /* 2*nnz */
    int32_t *d_rowptr, *d_col, *d_rowptrH, *d_colH;
    void *d_val64, *d_valH64;
    double *d_valz, *d_valzH;   /* complex expansion of real values (weight products), built on first use */
    /* row-grouped form of A and A^H for the grouped-gather stage kernel
     * (group = 0: not built, the grouping would not save gathers) */
    int group, groupH;
    PetscInt maxrow, maxrowH;   /* longest row of A and of A^H (nonzeros) */
    int32_t *d_gptr, *d_gptrH;
    uint32_t *d_gword, *d_gwordH;
    void *d_gval, *d_gvalH;
    /* autotuned SpMM launch shape per [stepping form][precision 64/32]
     * [direct/adjoint]: lanes per row, kernel variant, XCD remap, tile
     * height (variant 9); tuned_ok marks a tuned entry, valid for blocks of
     * tuned_k columns */
    int tuned_lanes[2][2][2], tuned_variant[2][2][2], tuned_remap[2][2][2], tuned_ok[2][2][2];
    int tuned_tile[2][2][2];
    PetscInt tuned_k[2][2][2];   /* block width the entry was tuned for */
    /* tiled forms [0: A, 1: A^H][tile height index], built on first use,
     * from the host CSR structure of A and of A^H */
    int64_t *h_rowptrH, *h_colH;
    rsvdt_tiled tiled[2][RSVDT_NTILE];
    /* launch shape of the FP64 Horner stage on several blocks at once (the
     * e^{A dT} steps of the removal, the folded forcing), per [A, A^H], tuned
     * for blocks of mt_k columns: lanes, variant, tile height, blocks per
     * workgroup */
    int mt_ok[2], mt_lanes[2], mt_variant[2], mt_tile[2], mt_bpw[2];
    PetscInt mt_k[2];
};

struct _p_RSVDT_Vec {
    PetscInt n;
    double *d;
};

/* ---- global runtime ---------------------------------------------------- */
typedef struct {
    int initialized;
    int device;
    rsvdt_stream_t stream;      /* all work of a run */
    rsvdt_stream_t stream_frc;  /* forcing-chunk GEMMs, overlapped with the RK4 steps */
    RSVDT_Options opt;
    RSVDT_ActionStats stats;
    /* learned-sketch columns for CreateRandomMat */
    double *sketch_host;
    const float *sketch_dev;     /* device columns (RSVDT_SetSketchColumnsDevice), ld floats per point */
    PetscInt sketch_N, sketch_cols, sketch_Nw_eff, sketch_ld;
    char inputs_file[PETSC_MAX_PATH_LEN];
} rsvdt_runtime;

extern rsvdt_runtime g_rsvdt;

PetscErrorCode rsvdt_ensure_device(void);

/* ---- errors ------------------------------------------------------------ */
PetscErrorCode rsvdt_err(PetscErrorCode code, const char *fn, const char *fmt, ...);
#define RSVDT_CHK(x)                                  \
    do {                                              \
        PetscErrorCode _ierr = (x);                   \
        if (_ierr) return _ierr;                      \
    } while (0)
#define RSVDT_HIPCHK(fn, x)                                                                   \
    do {                                                                                      \
        int _h = (x);                                                                         \
        if (_h) return rsvdt_err(_h, fn, "device call failed: %s", rsvdt_hip_last_error()); \
    } while (0)

int rsvdt_printf(const char *fmt, ...);

/* ---- options table ----------------------------------------------------- */
int rsvdt_opt_get_string(const char *key, char *out, size_t n);   /* returns 1 if set */
int rsvdt_opt_get_real(const char *key, double *out, int *ok);
int rsvdt_opt_get_int(const char *key, PetscInt *out, int *ok);
int rsvdt_opt_get_bool(const char *key, PetscBool *out, int *ok);
PetscErrorCode rsvdt_yaml_insert_file(const char *path);

/* ---- Mat helpers ------------------------------------------------------- */
PetscErrorCode rsvdt_mat_alloc_dense(PetscInt rows, PetscInt cols, PetscInt kb, int order, Mat *out);
PetscErrorCode rsvdt_mat_dup(Mat src, Mat *out);
PetscErrorCode rsvdt_mat_to_layout(Mat m, PetscInt kb, int order);   /* physical relayout if needed */
PetscErrorCode rsvdt_aij_upload(Mat m);
/* The tiled form of A (dir 1) or A^H (dir 0) with tiles of tile_rows rows,
 * built and uploaded on first use; *out is NULL when it cannot be built
 * (a tile with more than 65535 distinct columns, an unsupported height). */
PetscErrorCode rsvdt_tiled_get(Mat m, int dir, int tile_rows, const rsvdt_tiled **out);
PetscErrorCode rsvdt_aij_real_check(Mat m);

/* ---- kernels of the action (rsvdt_action.c) ----------------------------- */
/* Stepping of nb homogeneous N x k blocks by nsteps RK4 steps in place. */
PetscErrorCode rsvdt_homogeneous_steps(Mat A, int dir, double h, PetscInt nsteps, double *X, PetscInt N,
                                       PetscInt k, PetscInt nb);
/* Shifted CholeskyQR3 of nbatch tall N x n matrices given by strides, in
 * place; R (nbatch x n x n, device) optional.  drop_tol > 0 (bases that may
 * be rank-deficient): columns whose second/third-pass pivot falls below
 * drop_tol times the largest Gram diagonal are zeroed instead of normalized. */
PetscErrorCode rsvdt_cholqr(PetscInt nbatch, PetscInt nrows, PetscInt n, double *X, int64_t xb, int64_t xr,
                            int64_t xa, double *R_out, double drop_tol);
/* The CholeskyQR passes on caller-provided device buffers (G, R, Ri:
 * nbatch x n x n complex FP64; Ra, Rt only when R_out is wanted; work of
 * rsvdt_hip_bgram_work_bytes bytes; info 256 bytes), enqueued on stream q.
 * report: read the non-positive pivot count back (synchronises) and warn. */
typedef struct { double *G, *R, *Ri, *Ra, *Rt; void *work; int *info; } rsvdt_cholqr_work;
PetscErrorCode rsvdt_cholqr_bufs(PetscInt nb, PetscInt nrows, PetscInt n, double *X, int64_t xb, int64_t xr,
                                 int64_t xa, double *R_out, int npass, int shift_first, double drop_tol, double *T,
                                 size_t bytes, const rsvdt_cholqr_work *w, rsvdt_stream_t q, int report);
/* Orthonormal basis Q of the column span of nbatch tall matrices X (same
 * strides; Q and the work arrays E and T each cover `bytes` bytes like X)
 * that may be numerically rank-deficient: CholeskyQR3 with dropping, then
 * the same on the residuals (I - Q Q^H) X of the dropped columns,
 * reorthogonalised, in the dropped columns' places.  Columns dependent to
 * rounding stay zero. */
PetscErrorCode rsvdt_span_basis(PetscInt nbatch, PetscInt nrows, PetscInt n, const double *X, double *Q, double *E,
                                double *T, int64_t xb, int64_t xr, int64_t xa, size_t bytes, double drop_tol);
/* Left singular vectors (descending) of nbatch tall matrices, in place, and
 * their singular values (device, nbatch x n). */
PetscErrorCode rsvdt_svd_left(PetscInt nbatch, PetscInt nrows, PetscInt n, double *X, int64_t xb, int64_t xr,
                              int64_t xa, double *sig_dev);

/* device scratch that is reused across calls (grown on demand, freed at
 * Finalize), in RSVDT_NSCRATCH numbered slots */
#define RSVDT_NSCRATCH 32
void *rsvdt_scratch(int slot, size_t bytes);
void rsvdt_scratch_free_all(void);
size_t rsvdt_scratch_size(int slot);     /* bytes held by a slot (0 if none) */
void rsvdt_action_release(void);         /* device buffers the transient removal keeps between calls */
void rsvdt_scratch_release(int slot);    /* free one slot's buffer */

#endif
