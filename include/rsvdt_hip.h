/*
 * rsvdt_hip.h -- thin C ABI over the hand-written gfx950 HIP kernels.
 *
 * Plain pointers, sizes and an opaque stream; no torch or PETSc types.
 * Every function returns 0 on success or a nonzero RSVDT_HIP_* code, and
 * never allocates device memory unless its name says so.  All pointers are
 * device pointers unless the parameter name ends in _host.
 *
 * Storage conventions
 *   complex values are interleaved (re, im) pairs of double (z) or float (c);
 *   a "block-vector array" holds nb blocks of an N x k row-major block, block
 *   b at offset b*N*k: entry (b, row, j) at [b*N*k + row*k + j].  The sketch
 *   F_hat / response Y_hat is such an array with one block per frequency, the
 *   snapshot store with one block per snapshot.
 *
 * These functions sit below the reference-compatible entry points declared
 * in rsvdt.h; the reference has no counterpart for them (it calls PETSc).
 */
#ifndef RSVDT_HIP_H
#define RSVDT_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    RSVDT_HIP_OK = 0,
    RSVDT_HIP_ERR_RUNTIME = 101,     /* a HIP runtime call failed */
    RSVDT_HIP_ERR_ARG = 102,         /* unsupported size / bad argument */
    RSVDT_HIP_ERR_NODEV = 103,       /* no GPU visible */
    RSVDT_HIP_ERR_NUMERIC = 104      /* breakdown (non-positive pivot, singular solve) */
};

typedef void *rsvdt_stream_t;        /* hipStream_t, or NULL for the null stream */

/* ---- runtime ---------------------------------------------------------- */
int rsvdt_hip_init(int device);                       /* select device, check gfx950 */
const char *rsvdt_hip_last_error(void);               /* text of the last failure */
/* device memory; freed blocks of at most 256 MB are kept (2 GB in all) for
 * later requests of their size up to 2x, after the device wait hipFree itself
 * makes; release_cached frees them (RSVDT_Finalize calls it); cached_bytes
 * is what the cache holds, which mem_info counts as free */
int rsvdt_hip_malloc(void **p, size_t bytes);
int rsvdt_hip_free(void *p);
int rsvdt_hip_release_cached(void);
size_t rsvdt_hip_cached_bytes(void);
/* Y block m (N x k row-major complex FP64), columns j < ncols, from src (FP32
 * complex interleaved, point-major): Y[(m N + r) k + j] = src[(m N + r) ld + 2j]
 * + i src[(m N + r) ld + 2j + 1] (learned sketch columns, no counterpart) */
int rsvdt_hip_sketch_cols_f32(int64_t nwe, int64_t N, int64_t ncols, int64_t k, const float *src, int64_t ld,
                              double *Y, rsvdt_stream_t s);
/* iDFT table of CreateDFTiDFTMats (SourceCode/CreateDFTiDFTMats.c:42-57,93)
 * built on the device: out[j nwe + m] = conj(exp(-2 pi i m_alias j / nt2)) 2/nt2 */
int rsvdt_hip_idft_table(int64_t nwe, int64_t nt2, int real_op, double *out, rsvdt_stream_t s);
/* synchronous; both wait for all queued device work first */
int rsvdt_hip_memcpy_h2d(void *dst, const void *src_host, size_t bytes);
int rsvdt_hip_memcpy_d2h(void *dst_host, const void *src, size_t bytes);
int rsvdt_hip_memcpy_d2d(void *dst, const void *src, size_t bytes, rsvdt_stream_t s);
int rsvdt_hip_zero(void *p, size_t bytes, rsvdt_stream_t s);  /* fill kernel, graph-safe */
int rsvdt_hip_sync(void);
int rsvdt_hip_mem_info(size_t *free_bytes, size_t *total_bytes);

/* streams (non-blocking, default priority) and cross-stream ordering */
int rsvdt_hip_stream_create(rsvdt_stream_t *s);
int rsvdt_hip_stream_destroy(rsvdt_stream_t s);
int rsvdt_hip_stream_wait_event(rsvdt_stream_t s, void *ev);
int rsvdt_hip_stream_sync(rsvdt_stream_t s);
/* pinned host memory, and asynchronous device-to-host copies into it */
int rsvdt_hip_host_alloc(void **p, size_t bytes);
int rsvdt_hip_host_free(void *p);
int rsvdt_hip_memcpy_d2h_async(void *dst_host, const void *src, size_t bytes, rsvdt_stream_t s);

/* events for timing on a given stream */
int rsvdt_hip_event_create(void **ev);
int rsvdt_hip_event_destroy(void *ev);
int rsvdt_hip_event_record(void *ev, rsvdt_stream_t s);
int rsvdt_hip_event_elapsed_ms(void *start, void *stop, float *ms);

/* ---- fused CSR SpMM + RK4 stage (replaces TSRK4.c:18-47) --------------- *
 * One of the four classical RK4 stages on nb independent N x k blocks,
 * sharing the operator (CSR, int32 indices):
 *   s      = A * uin + F                (F may be NULL: homogeneous)
 *   stage 1: uout = y + c_next*s ; acc  = y   + c_acc*s     (uin == y)
 *   stage 2,3: uout = y + c_next*s; acc = acc + c_acc*s
 *   stage 4: y    = acc + c_acc*s ; snap (if non-NULL) = y (FP64 copy)
 * prec: 64 = complex FP64 state, 32 = complex FP32 state.
 * a_real: values are real (length nnz) rather than complex (2*nnz).
 * Value arrays must be of the state precision.  Block strides are N*k.
 * lanes: lanes cooperating on one row (power of two, 0 = automatic).  */
typedef struct {
    int64_t nrows;
    int64_t k;            /* columns per block */
    int64_t nb;           /* number of blocks stepped together */
    const int32_t *rowptr;
    const int32_t *col;
    const void *val;
    int a_real;
    int prec;
    int lanes;            /* 0 = default choice */
    int xcd_remap;        /* workgroup -> rows map: 0 identity, 1 one contiguous row range per XCD,
                             2 / 3 chunks of 64 / 256 row blocks dealt round-robin to the XCDs */
    int variant;          /* 0: 16-nonzero gather batches; 1: + nontemporal row-local
                             streams; 2: 8-nonzero batches; 3: 8 + nontemporal; 4: 12-nonzero
                             batches; 5: 12 + nontemporal (variants 1-3 apply to the RK4 stage
                             when k <= lanes and to every Horner stage; 4 and 5 only to the
                             Horner stage with lanes >= 16 and k <= 2 lanes); 6: the
                             Horner stage on the row-grouped operator below (k <= lanes,
                             16..64 lanes); 7: the FP32 Horner stage with two adjacent
                             columns per lane (16-byte gathers; k/2 in 4, 8, 16, 32 lanes);
                             8: the same with two such pairs per lane (k/4 in 4, 8, 16 lanes);
                             9: the Horner stage on the tiled form below (lanes = columns per
                             pass: 8, 16 or 32, 0 = automatic); 10: the same with the next
                             pass's first distinct rows loaded into registers while the
                             current pass computes */
    /* row-grouped form of the same operator (variant 6): groups of `group`
     * consecutive rows (2, 3, 4, 5 or 8); group q's distinct columns are
     * gword[gptr[q] .. gptr[q+1]-1], each (column << 8 | mask of the group's
     * rows that have an entry there), ascending by column; gval holds the
     * values in (column, row) order from offset rowptr[q*group] on (real or
     * complex FP64 like val).  Columns must be < 2^24. */
    int group;
    const int32_t *gptr;
    const uint32_t *gword;
    const void *gval;
    /* tiled form of the same operator (variant 9): tiles of tile_rows
     * consecutive rows; tile t's distinct columns are tcol[tptr[t] ..
     * tptr[t+1]-1], ascending, and lidx[e] is the position of nonzero e's
     * column in its tile's list.  tile_hmax / tile_emax: the most columns /
     * nonzeros of any tile.  A workgroup stages its tile's rows of the
     * gathered vector in LDS once and forms every row of the tile from
     * there. */
    int tile_rows;
    int tile_hmax, tile_emax;
    const int32_t *tptr;
    const int32_t *tcol;
    const uint16_t *lidx;
    /* tiled form with nb > 1 blocks: each workgroup forms its tile in
     * tile_bpw consecutive blocks (0 or 1: one block per workgroup), staging
     * the tile's metadata once for all of them */
    int tile_bpw;
} rsvdt_spmm_desc;

int rsvdt_hip_rk4_stage(const rsvdt_spmm_desc *d, int stage,
                        const void *uin, const void *y_or_null_stage4, const void *F,
                        void *acc, void *out, double c_next, double c_acc,
                        double *snap, rsvdt_stream_t s);

/* One Horner stage of the RK4 step of a linear system written as
 *   y_{i+1} = P(hA) y_i + g_i,  P(z) = 1 + z(1 + z/2(1 + z/3(1 + z/4))):
 *   v1 = y + (h/4) A y;  v2 = y + (h/3) A v1;  v3 = y + (h/2) A v2;
 *   y' = y + h A v3 + g_i.
 * out_b = ys_b * y_b + c * (A vin_b) (+ g_b) for every block b < d->nb
 * (blocks N x k row-major at stride N*k; the same N x k block layout for
 * vin, y, g, out and snap).  yscale: nb complex FP64 block scales, NULL = 1.
 * g: NULL = no forcing term.  out may alias y (never vin).  snap (complex
 * FP64, may be NULL) receives a copy of out.  d->variant and d->lanes pick
 * the launch shape as for rsvdt_hip_rk4_stage. */
int rsvdt_hip_horner_stage(const rsvdt_spmm_desc *d, const void *vin, const void *y, const double *yscale,
                           const void *g, double c, void *out, double *snap, rsvdt_stream_t s);
/* As rsvdt_hip_horner_stage, with the snapshot destination read at run time:
 * when snap_slot is not NULL, the kernel reads the int *snap_slot and writes
 * the copy of out to snap + slot*snap_stride (complex entries) when slot >= 0,
 * nowhere when it is negative (graph-captured steps). */
int rsvdt_hip_horner_stage_ex(const rsvdt_spmm_desc *d, const void *vin, const void *y, const double *yscale,
                              const void *g, double c, void *out, double *snap, const int *snap_slot,
                              int64_t snap_stride, rsvdt_stream_t s);

/* Plain block SpMM: out_b = M * in_b for nb blocks of k columns (complex FP64
 * values, int32 CSR).  in blocks are ncols x k (stride in_bstride), out blocks
 * nrows x k (stride out_bstride).  Used for the B, C and weight matrices.   */
/* Whole runs of Horner steps for operators that fit in LDS with the state of
 * one column (rsvdt_hip_horner_fused_fits: the CSR matrix of nnz nonzeros,
 * y and two stage vectors within 160 KB): y (d->nb blocks of N x k, FP64 or
 * FP32 per d->prec) is advanced by nsteps steps in place; coef = {h/4, h/3,
 * h/2, h}; step s adds the forcing sample g + s*gstride (N x k, state
 * precision; g may be NULL) at its last stage and, when step i0+s stores a
 * snapshot (i > Nt, (i-Nt) % RR == 0, column (i-Nt)/RR-1 mod Nstore,
 * reversed when dir == 0), writes y in FP64 to snap + pos*snap_stride
 * (complex entries; snap may be NULL).  One launch, one workgroup per
 * (column, block).  maxrow: the longest row's nonzero count (-1 if unknown);
 * rows of at most 4 nonzeros in operators of at most 1024 rows keep their
 * nonzeros in registers. */
int rsvdt_hip_horner_fused_fits(int64_t nrows, int64_t nnz, int a_real, int prec);
int rsvdt_hip_horner_fused(const rsvdt_spmm_desc *d, int64_t nnz, int64_t maxrow, void *y, const void *g,
                           int64_t gstride, const double *coef, int64_t nsteps, double *snap, int64_t snap_stride,
                           int64_t i0, int64_t Nt, int64_t RR, int64_t Nstore, int dir, rsvdt_stream_t s);
int rsvdt_hip_spmm_blocks(int64_t nrows, int64_t k, int64_t nb, const int32_t *rowptr, const int32_t *col,
                          const double *val_complex, const double *in, int64_t in_bstride,
                          double *out, int64_t out_bstride, rsvdt_stream_t s);

/* ---- small-coefficient complex GEMM (forcing chunks, DFT, windows) ----- *
 * out[p][l] = alpha * sum_{q<K} coef(q,p) * in[q*ldi + l]   (+ beta*out)
 * for p < P, l < L (complex entries), with the coefficient column map
 *   coef(q,p) = coef[q*cq + ((p0 + sgn*p) mod cmod)*cp]
 * (sgn = +1 or -1), which lets the forcing read iDFT samples straight from
 * the 2Ns-periodic table in either time direction.  coef and in are complex
 * FP64; out is FP64 (out_prec 64) or FP32 (out_prec 32).  Used for the
 * forcing chunk (in = F_hat frequency blocks, coef = iDFT table), the DFT
 * (in = snapshots, coef = RR*dft) and the transient-removal windows.      */
int rsvdt_hip_coef_gemm(int64_t P, int64_t K, int64_t L,
                        const double *coef, int64_t cq, int64_t cp,
                        int64_t p0, int sgn, int64_t cmod, double alpha,
                        const double *in, int64_t ldi,
                        void *out, int64_t ldo, int out_prec,
                        double beta, rsvdt_stream_t s);

/* ---- batched tall-skinny inner products (Gram / projections) ---------- *
 * G_b[a][c] = sum_{r<nrows} conj(X_b(r,a)) * Z_b(r,c), a < na, c < nc,
 * X_b(r,a) = X[b*xb + r*xr + a*xa] (complex FP64), same for Z.
 * G is nbatch x na x nc complex FP64 (row-major per batch).  Deterministic
 * two-pass reduction; work must hold rsvdt_hip_bgram_work_bytes(...) bytes. */
size_t rsvdt_hip_bgram_work_bytes(int64_t nbatch, int64_t nrows, int64_t na, int64_t nc);
int rsvdt_hip_bgram(int64_t nbatch, int64_t nrows, int64_t na, int64_t nc,
                    const double *X, int64_t xb, int64_t xr, int64_t xa,
                    const double *Z, int64_t zb, int64_t zr, int64_t zc,
                    double *G, void *work, rsvdt_stream_t s);

/* ---- batched row-local small products --------------------------------- *
 * Out_b(r,c) = alpha * Y_b(r,c) + sum_{a<na} X_b(r,a) * C_b[a][c]
 * (Y may be NULL).  C is nbatch x na x nc complex FP64 row-major.
 * Out may alias X (in-place X <- X*C) when the strides agree.            */
int rsvdt_hip_bapply(int64_t nbatch, int64_t nrows, int64_t na, int64_t nc,
                     const double *X, int64_t xb, int64_t xr, int64_t xa,
                     const double *C,
                     double alpha, const double *Y, int64_t yb, int64_t yr, int64_t yc,
                     double *Out, int64_t ob, int64_t orow, int64_t oc,
                     rsvdt_stream_t s);

/* ---- batched small dense kernels (one workgroup per matrix) ----------- */
/* Cholesky of (G_b + shift_b*I) (upper R, G = R^H R) then Rinv_b = R^-1.
 * G, R, Rinv: nbatch x n x n complex FP64 row-major.  shift_rel >= 0:
 * shift_b = shift_rel * trace(G_b).  drop_tol > 0: a pivot below drop_tol
 * times the largest diagonal entry of G_b marks its column as lying in the
 * span of the previous ones; it gets a zero row in R and a zero column in
 * R^-1 (so X R^-1 has a zero column there).  info (device int, may be NULL)
 * receives the count of other non-positive pivots (clamped).             */
int rsvdt_hip_bchol_inv(int64_t nbatch, int64_t n, const double *G, double shift_rel, double drop_tol,
                        double *R, double *Rinv, int *info, rsvdt_stream_t s);
/* D_b = diag(d_j), d_j = 1 where R_b(j,j) is not positive (a column that a
 * dropping Cholesky zeroed), else 0; *count (device int, may be NULL) is
 * increased by the number of such columns.  R, D: nbatch x n x n.         */
int rsvdt_hip_bdrop_mask(int64_t nbatch, int64_t n, const double *R, double *D, int *count, rsvdt_stream_t s);
/* Rout_b = Ra_b * Rb_b for upper-triangular n x n (row-major). */
int rsvdt_hip_btrmm_upper(int64_t nbatch, int64_t n, const double *Ra, const double *Rb,
                          double *Rout, rsvdt_stream_t s);
/* One-sided Jacobi SVD of R_b (n x n): R = U diag(sig) W^H, sig descending.
 * U: nbatch x n x n row-major; sig: nbatch x n.  Each column of U is scaled
 * so that the sum of its entries is real and positive.  n <= 64.          */
int rsvdt_hip_bjacobi_svd(int64_t nbatch, int64_t n, const double *R, double *U,
                          double *sig, rsvdt_stream_t s);
/* Solve (shift_{b,m} I - M_b) X = B for every (b, m): M is nbatch x n x n,
 * B and X are nbatch x n x nrhs (row-major, X may alias B), shift is
 * nbatch x nrhs complex.  LU with partial pivoting; work must hold
 * rsvdt_hip_bshift_solve_work_bytes(...) bytes.                           */
size_t rsvdt_hip_bshift_solve_work_bytes(int64_t nbatch, int64_t n, int64_t nrhs);
int rsvdt_hip_bshift_solve(int64_t nbatch, int64_t n, int64_t nrhs,
                           const double *M, const double *shift,
                           const double *B, double *X, void *work, int *info,
                           rsvdt_stream_t s);

/* ---- per-frequency QR without device allocation ----------------------- *
 * No counterpart in the reference: the in-place QR of QRAllFreqs
 * (shifted CholeskyQR3 of each frequency block, the same kernels and the
 * same result) for callers that cannot allocate device memory -- code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc.
 * Y holds nblocks N x k row-major blocks (block b at b*N*k complex entries),
 * k <= 64; the call is asynchronous on s and reads nothing back (a block
 * without full column rank is not reported).  Its work lives in one static
 * device array of RSVDT_NOALLOC_SCRATCH_BYTES per GPU, held until the process
 * ends; sizes whose work (rsvdt_qr_blocks_noalloc_bytes) exceeds it are
 * rejected with RSVDT_HIP_ERR_ARG.  Calls share that array: calls that
 * overlap on one GPU (unordered streams, several host threads) race.     */
#define RSVDT_NOALLOC_SCRATCH_BYTES ((size_t)192 << 20)   /* covers k <= 64 at any N, ~300 blocks of k = 64 */
size_t rsvdt_qr_blocks_noalloc_bytes(int64_t nblocks, int64_t N, int64_t k);
int rsvdt_qr_blocks_noalloc(double *Y, int64_t N, int64_t k, int64_t nblocks, rsvdt_stream_t s);

/* ---- elementwise ------------------------------------------------------ */
/* y = a*x (complex entries, FP64), n entries */
int rsvdt_hip_zscale(int64_t n, double a, const double *x, double *y, rsvdt_stream_t s);
/* y += a*x (complex entries, FP64, real a), n entries */
int rsvdt_hip_zaxpy(int64_t n, double a, const double *x, double *y, rsvdt_stream_t s);
/* precision conversions of complex arrays */
int rsvdt_hip_z2c(int64_t n, const double *x, float *y, rsvdt_stream_t s);
int rsvdt_hip_c2z(int64_t n, const float *x, double *y, rsvdt_stream_t s);
/* strided complex copy: dst[b*db + r*dr + c*dc] = src[b*sb + r*sr + c*sc] */
int rsvdt_hip_zcopy3(int64_t nb, int64_t nr, int64_t nc,
                     const double *src, int64_t sb, int64_t sr, int64_t sc,
                     double *dst, int64_t db, int64_t dr, int64_t dc, rsvdt_stream_t s);
/* Seeded sketch: a rows x (kb*nblk) complex matrix in kb-column row-major
 * blocks; entry (r, c) is drawn from the counter-based stream at index
 * c*rows + r (uniform [0,1) + i[0,1) when gauss = 0, complex N(0,1) when 1). */
int rsvdt_hip_fill_random(int64_t rows, int64_t kb, int64_t nblk, uint64_t seed, int gauss, double *out,
                          rsvdt_stream_t s);
/* Per-row transpose between the block layout [s][r][b] (n blocks, block s
 * an nrows x kt row-major slab) and the per-batch row-major layout [j][r][s]
 * (kb batches, batch j an nrows x n row-major matrix), for the batches
 * b0 .. b0+kb-1 of the block layout.  dir 0: block -> per-batch;
 * dir 1: per-batch -> block.  n*(kb+1) <= 8192.                           */
int rsvdt_hip_transpose_rows(int64_t nrows, int64_t n, int64_t kt, int64_t b0, int64_t kb, const double *src,
                             double *dst, int dir, rsvdt_stream_t s);
/* Synthetic snapshot sets of the transient-removal estimate
 * (SourceCode/TransientRemovalEst.c:84-98): nb per-batch row-major
 * nrows x ns matrices out[b][r][c] = Q[r*ldq + c + skip] + qss[r]*ph[b*ns + c]
 * (Q: nrows x ldq row-major snapshots, ph: nb x ns phases), ns + skip <= ldq. */
int rsvdt_hip_trans_synth(int64_t nrows, int64_t ldq, int64_t ns, int64_t skip, int64_t nb, const double *Q,
                          const double *qss, const double *ph, double *out, rsvdt_stream_t s);
/* Snapshot slot of the steps i0 .. i0+S-1 (SourceCode/SaveSnapshots.c:17-29):
 * slots[t] = the Y_all column of step i0+t, or -1 when that step stores none. */
int rsvdt_hip_snap_slots(int *slots, int64_t S, int64_t i0, int64_t Nt, int64_t RR, int64_t Nstore, int dir,
                         rsvdt_stream_t s);
/* HIP graph capture of a stream's launches: begin, end (instantiates into
 * *exec), launch, destroy. */
int rsvdt_hip_graph_begin(rsvdt_stream_t s);
int rsvdt_hip_graph_end(rsvdt_stream_t s, void **exec);
int rsvdt_hip_graph_launch(void *exec, rsvdt_stream_t s);
int rsvdt_hip_graph_destroy(void *exec);
/* sum of |x|^2 over n complex entries into *out (device double) */
int rsvdt_hip_znorm2(int64_t n, const double *x, double *out_dev, void *work, rsvdt_stream_t s);

/* ---- learned sketch (coordinate MLP, BF16 MFMA) ------------------------ *
 * No counterpart in the reference: the network that proposes part of the
 * initial forcing sketch (the random columns of SourceCode/CreateRandomMat.c:39
 * otherwise).  All arrays FP32 unless noted.                              */
/* features (n x 64) of points q (n x 4 FP64: x, w, mu0, Re U):
 * [x/xs, w/ws, mu0, U, sin(kj x), cos(kj x) for kj = kmax (j+1)/28, j < 28, 0 x 4] */
int rsvdt_hip_mlp_features(int64_t n, const double *q, double xs, double ws, double kmax, float *out,
                           rsvdt_stream_t s);
/* C(m,n) = sum_k A(m,k) B(k,n) with A(m,k) = A[m*sam + k*sak], B(k,n) = B[k*sbk + n*sbn],
 * BF16 operands (rounded to nearest even), FP32 accumulation, C row stride ldc.
 * epilogue 0: C = acc; 1: z = acc + bias[n], Z(m,n) = z, C = z sigmoid(z);
 * 2: C = acc + bias[n]; 3: C = acc * silu'(Z(m,n)).  Z row stride ldz.
 * Long reductions (K >= 2048 with few output tiles) are split and summed in a
 * fixed order through one process-wide workspace: such calls must not
 * overlap on different streams. */
int rsvdt_hip_gemm_bf16(int64_t M, int64_t N, int64_t K, const float *A, int64_t sam, int64_t sak, const float *B,
                        int64_t sbk, int64_t sbn, float *C, int64_t ldc, int epilogue, const float *bias, float *Z,
                        int64_t ldz, rsvdt_stream_t s);
/* out[n] = sum_{m<M} A[m*lda + n] */
int rsvdt_hip_colsum(int64_t M, int64_t N, const float *A, int64_t lda, float *out, rsvdt_stream_t s);
/* Alignment loss of ncol learned columns F (ngroups*npts x 2 ncol, re/im
 * interleaved) with target columns V (ngroups x npts x ncol complex FP64):
 * loss[g] = scale sum_j (1 - |<v_j, f_j>|^2 / (|f_j|^2 |v_j|^2)); dF its
 * gradient (d/dRe f + i d/dIm f) in the layout of F.  ncol <= 16. */
int rsvdt_hip_sketch_loss(int64_t ngroups, int64_t npts, int64_t ncol, const float *F, const double *V, float *dF,
                          float *loss, double scale, rsvdt_stream_t s);
/* One Adam step on n parameters; b1t = 1 - beta1^t, b2t = 1 - beta2^t. */
int rsvdt_hip_adam(int64_t n, float *w, const float *g, float *m, float *v, float lr, float b1, float b2, float eps,
                   float b1t, float b2t, rsvdt_stream_t s);

#ifdef __cplusplus
}
#endif
#endif
