/*
 * The hot path: resolvent actions by RK4 time stepping (TSActionRK4 and its
 * callers), the snapshot DFT, the Galerkin transient removal and the power
 * iteration.  The reference steps one test vector at a time with four SpMVs,
 * two forcing GEMVs and seven AXPYs per step (SourceCode/TSActionRK4.c:89-122,
 * TSRK4.c:18-47); here every step is four fused SpMM+stage launches over the
 * whole N x k block, and the forcing of a chunk of steps is synthesized by one
 * small-coefficient GEMM.
 */
#define _POSIX_C_SOURCE 200809L
#include "rsvdt_internal.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define PI_D 3.14159265358979323846264338327950288

/* scratch slots */
enum { SCR_Y = 0, SCR_U1, SCR_U2, SCR_ACC, SCR_FBUF, SCR_FBUF2, SCR_G, SCR_R, SCR_RINV, SCR_RACC, SCR_RTMP,
       SCR_WORK, SCR_INFO, SCR_MISC, SCR_MISC2, SCR_PROJ, SCR_MASK, SCR_RMASK, SCR_CNT, SCR_FSCALE, SCR_W1, SCR_W2, SCR_QRT, SCR_GINFO, SCR_SLOTS, SCR_NSLOTS };
_Static_assert(SCR_NSLOTS <= RSVDT_NSCRATCH, "scratch slot table too small");

/* ------------------------------------------------------------------ stepper */
/* Two forms of the RK4 step of SourceCode/TSRK4.c:18-47 on the whole N x k
 * block of test vectors.
 *
 * form 0 (default).  For the linear system y' = A y + f(t) the step is
 *   y_{i+1} = P(hA) y_i + g_i,   P(z) = 1 + z + z^2/2 + z^3/6 + z^4/24,
 * and, since the stage forcings of step i are the iDFT samples at half-steps
 * 2(i-1), 2i-1, 2i (TSRK4.c:21-27), i.e. (1, z_m, z_m^2) z_m^{2(i-1)} F_hat_m/Ns
 * per frequency m with z_m the half-step phase, the forcing term is
 *   g_i = sum_m idft(m, 2(i-1)) G_m,   G_m = Q(z_m) F_hat_m,
 *   Q(z) = h/6 [(1+4z+z^2) + (1+2z) hA + (1+z)/2 (hA)^2 + (hA)^3/4],
 * the forced RK4 step from y = 0 written out.  G is computed once per action
 * (3 SpMMs per frequency block) in place of F_hat, a forcing chunk holds one
 * sample per step, and P(hA) y is evaluated by four Horner stages that each
 * read y and one gathered vector (12 block reads/writes per step instead of
 * the 20 of the stage form).  Same polynomial, same forcing, different
 * rounding order.
 *
 * form 1: the reference's stage order (k_rk4_stage), forcing chunks of 2S+1
 * half-step samples.
 *
 * Forcing chunks are double-buffered: while the steps of chunk c run on the
 * main stream from buffer c%2, the forcing GEMM of chunk c+2 fills the other
 * buffer on the forcing stream as soon as chunk c's steps are done, ordered by
 * events. */
typedef struct {
    Mat A;
    int dir, prec, form;
    double h;
    PetscInt N, k, Nwe, twoNs, S;
    rsvdt_spmm_desc desc;
    void *y, *u1, *u2, *acc, *fbuf[2];
    void *ev_frc[2], *ev_rk[2];
    size_t esz;                  /* bytes per complex entry of the state */
    const double *Fin;           /* Nwe FP64 blocks: F_hat (form 1) or G (form 0) */
    const double *idft;          /* [2Ns][Nwe] FP64 complex */
} stepper;

static void stepper_free(stepper *st);


/* Columns per pass of the tiled stage kernel: lanes, or by default 8 for
 * blocks of at most 8 columns and 16 otherwise; the tile heights its launch
 * shapes allow: (256 / columns per pass) x 1, 2, 3 or 4 rows per slot (32
 * columns per pass: 1, 2, 4; 8 columns: 1). */
static int tile_kc(const rsvdt_spmm_desc *d) { return d->lanes ? d->lanes : (d->k <= 8 ? 8 : 16); }
static int tile_ok(int kc, int tr)
{
    if (kc == 8) return tr == 32;
    if (kc == 16) return tr == 16 || tr == 32 || tr == 48 || tr == 64;
    if (kc == 32) return tr == 8 || tr == 16 || tr == 32;
    return 0;
}

/* Points the descriptor at the tiled form of A (dir 1) or A^H with tiles of
 * `tile` rows (0: 32 rows, or 16 with 32 columns per pass), building it on
 * first use.  Returns 0 when there is none (the caller falls back). */
static int tile_attach(rsvdt_spmm_desc *d, Mat A, int dir, int tile)
{
    const int kc = tile_kc(d);
    if (!tile) tile = kc == 32 ? 16 : 32;
    if (!tile_ok(kc, tile)) return 0;
    const rsvdt_tiled *t = NULL;
    if (rsvdt_tiled_get(A, dir, tile, &t) || !t) return 0;
    d->tile_rows = t->rows;
    d->tile_hmax = t->hmax;
    d->tile_emax = t->emax;
    d->tptr = t->d_tptr;
    d->tcol = t->d_tcol;
    d->lidx = t->d_lidx;
    return 1;
}

static void desc_for(rsvdt_spmm_desc *d, Mat A, int dir, int prec, PetscInt k, PetscInt nb)
{
    memset(d, 0, sizeof(*d));
    d->nrows = A->rows;
    d->k = k;
    d->nb = nb;
    d->rowptr = dir ? A->d_rowptr : A->d_rowptrH;
    d->col = dir ? A->d_col : A->d_colH;
    d->val = dir ? A->d_val64 : A->d_valH64;      /* FP64 values for both state precisions */
    d->a_real = A->is_real;
    d->prec = prec;
    d->group = dir ? A->group : A->groupH;
    d->gptr = dir ? A->d_gptr : A->d_gptrH;
    d->gword = dir ? A->d_gword : A->d_gwordH;
    d->gval = dir ? A->d_gval : A->d_gvalH;
    const int p = prec == 32, q = dir ? 0 : 1, f = g_rsvdt.opt.rk4_form ? 1 : 0;
    int tile = 0;
    if (A->tuned_ok[f][p][q] && A->tuned_k[f][p][q] == k && nb == 1) {
        d->lanes = A->tuned_lanes[f][p][q];
        d->variant = A->tuned_variant[f][p][q];
        d->xcd_remap = A->tuned_remap[f][p][q];
        tile = A->tuned_tile[f][p][q];
    } else {
        /* measured best on the 2D flow operator: 8-nonzero gather batches with
         * nontemporal row-local streams, lanes = k */
        d->lanes = 0;
        d->variant = 3;
        d->xcd_remap = g_rsvdt.opt.xcd_remap;
    }
    if (g_rsvdt.opt.spmm_lanes) d->lanes = g_rsvdt.opt.spmm_lanes;
    if (g_rsvdt.opt.spmm_variant >= 0) d->variant = g_rsvdt.opt.spmm_variant;
    if (d->variant == 6 && (!d->group || g_rsvdt.opt.rk4_form || k > 64)) d->variant = 3;   /* no grouped form */
    if (d->variant == 7 && (prec != 32 || g_rsvdt.opt.rk4_form || k % 2 || !(k / 2 == 4 || k / 2 == 8 ||
                                                                         k / 2 == 16 || k / 2 == 32)))
        d->variant = 3;                                                         /* no column pairs */
    if (d->variant == 8 && (prec != 32 || g_rsvdt.opt.rk4_form || k % 4 || !(k / 4 == 4 || k / 4 == 8 || k / 4 == 16)))
        d->variant = 3;                                                         /* no double pairs */
    if (d->variant == 9 || d->variant == 10) {
        if (g_rsvdt.opt.rk4_form || !tile_attach(d, A, dir, tile)) d->variant = 3;   /* no tiled form */
    }
}

/* Launch shape of the stages that run on several blocks at once (the folded
 * forcing terms, the e^{A dT} steps of the transient removal), which the
 * autotuner does not see: two columns per lane (half of k rounded up to a
 * power of two, at least 4 lanes) with 8-nonzero batches, the shape it picks
 * for the 16- and 32-column blocks of the flow operators.  An explicit
 * SpMMVariant option keeps its variant with the default lanes (the grouped
 * and column-pair variants fix their own); an explicit SpMMLanes is kept. */
static void multi_block_shape(rsvdt_spmm_desc *d, PetscInt k)
{
    d->xcd_remap = 0;
    if (g_rsvdt.opt.spmm_variant >= 0) {
        d->lanes = 0;
        return;
    }
    d->variant = 3;
    if (g_rsvdt.opt.spmm_lanes) return;
    int l = 4;
    while (l < k && l < 64) l <<= 1;
    d->lanes = l / 2 >= 4 ? l / 2 : 4;
}

/* Shape of the FP64 Horner stage on nb > 1 blocks of k columns (A or A^H):
 * multi_block_shape's gather shape or the tiled kernel with 8 or 16 columns
 * per pass and 1, 4 or 8 blocks per workgroup, whichever three launches on
 * the caller's buffers (x gathered, out written) time fastest; kept on the
 * Mat for blocks of k columns.  An explicit SpMMVariant or SpMMLanes option
 * keeps multi_block_shape, and so do blocks of less than 32 MB in all: a
 * launch there takes about 10 us, and the ~40 timed candidate launches (and
 * building the tiled forms) would cost more than any shape saves. */
static void multi_block_tuned(rsvdt_spmm_desc *d, Mat A, int dir, PetscInt k, const void *x, void *out)
{
    multi_block_shape(d, k);
    if (g_rsvdt.opt.rk4_form || g_rsvdt.opt.spmm_variant >= 0 || g_rsvdt.opt.spmm_lanes || !g_rsvdt.opt.autotune)
        return;
    if ((double)d->nrows * (double)k * (double)d->nb * 16.0 < 32e6) return;
    const int q = dir ? 0 : 1;
    if (!(A->mt_ok[q] && A->mt_k[q] == k)) {
        static const int cand[][2] = {{0, 0}, {8, 32}, {16, 16}, {16, 32}};
        void *e0 = NULL, *e1 = NULL;
        if (rsvdt_hip_event_create(&e0) || rsvdt_hip_event_create(&e1)) {
            rsvdt_hip_event_destroy(e0);
            return;
        }
        static const int bpw_cand[] = {1, 4, 8};
        float best = 1e30f;
        int bl = d->lanes, bv = d->variant, bt = 0, bw = 0;
        for (int c = 0; c < (int)(sizeof(cand) / sizeof(cand[0])); c++)
            for (int wi = 0; wi < (cand[c][0] ? 3 : 1); wi++) {
                rsvdt_spmm_desc t = *d;
                if (cand[c][0]) {
                    if (cand[c][0] > 8 && k <= cand[c][0] / 2) continue;
                    if (wi > 0 && bpw_cand[wi] > d->nb) continue;
                    t.lanes = cand[c][0];
                    t.variant = 9;
                    t.tile_bpw = bpw_cand[wi];
                    if (!tile_attach(&t, A, dir, cand[c][1])) continue;
                }
                if (rsvdt_hip_horner_stage(&t, x, x, NULL, NULL, 0.0, out, NULL, g_rsvdt.stream)) continue;
                rsvdt_hip_event_record(e0, g_rsvdt.stream);
                for (int rep = 0; rep < 3; rep++)
                    (void)rsvdt_hip_horner_stage(&t, x, x, NULL, NULL, 0.0, out, NULL, g_rsvdt.stream);
                rsvdt_hip_event_record(e1, g_rsvdt.stream);
                float ms = 0;
                rsvdt_hip_event_elapsed_ms(e0, e1, &ms);
                if (ms < best) { best = ms; bl = t.lanes; bv = t.variant; bt = cand[c][1]; bw = t.tile_bpw; }
            }
        rsvdt_hip_event_destroy(e0);
        rsvdt_hip_event_destroy(e1);
        A->mt_ok[q] = 1; A->mt_k[q] = k; A->mt_lanes[q] = bl; A->mt_variant[q] = bv; A->mt_tile[q] = bt;
        A->mt_bpw[q] = bw;
        if (g_rsvdt.opt.verbose)
            rsvdt_printf("[rsvdt] autotune (%s, FP64, %lld blocks of %lld columns): lanes %d, variant %d, tile %d rows, "
                         "%d blocks per workgroup: %.1f us per stage\n", dir ? "A" : "A^H", (long long)d->nb,
                         (long long)k, bl, bv, bt, bw, 1e3 * best / 3);
    }
    if (A->mt_variant[q] == 9) {
        rsvdt_spmm_desc t = *d;
        t.lanes = A->mt_lanes[q];
        t.variant = 9;
        t.tile_bpw = A->mt_bpw[q];
        if (tile_attach(&t, A, dir, A->mt_tile[q])) *d = t;
    } else {
        d->lanes = A->mt_lanes[q];
        d->variant = A->mt_variant[q];
    }
}

/* Forcing samples per chunk of S steps: S (form 0) or 2S+1 (form 1).  The
 * defaults, 64 and 2*31+1 = 63, match the GEMM's 64-row MFMA tile; with the
 * LDS-resident stepper (fused: one launch per chunk of a small state) the
 * chunk starts at 256 steps, since each chunk also costs the main stream a
 * wait on the forcing stream (~20 us, a quarter of a 64-step launch there).
 * S shrinks while the two chunk buffers would take more than 60 % of free
 * memory (an action frees large ones when its stepping ends, before the
 * transient removal's work). */
static PetscInt chunk_samples(int form, PetscInt S) { return form ? 2 * S + 1 : S; }

static PetscInt auto_chunk(int form, PetscInt N, PetscInt k, size_t esz, int fused)
{
    if (g_rsvdt.opt.chunk_steps > 0) return g_rsvdt.opt.chunk_steps;
    size_t fr = 0, tot = 0;
    rsvdt_hip_mem_info(&fr, &tot);
    /* the chunk buffers of an earlier action are reused (or replaced) */
    fr += rsvdt_scratch_size(SCR_FBUF) + rsvdt_scratch_size(SCR_FBUF2);
    double blk = (double)N * (double)k * (double)esz;
    PetscInt S = form ? 31 : (fused ? 256 : 64);
    while (S > 4 && 2.0 * (double)chunk_samples(form, S) * blk > 0.6 * (double)fr) S = form ? (S - 1) / 2 : S / 2;
    return S;
}

static PetscErrorCode stepper_init(stepper *st, Mat A, int dir, int prec, double h, PetscInt k, PetscInt Nwe,
                                   PetscInt Ns, const double *Fin, const double *idft, int fused)
{
    memset(st, 0, sizeof(*st));
    st->A = A; st->dir = dir; st->prec = prec; st->h = h;
    st->form = g_rsvdt.opt.rk4_form ? 1 : 0;
    st->N = A->rows; st->k = k; st->Nwe = Nwe; st->twoNs = 2 * Ns;
    st->esz = prec == 64 ? 16 : 8;
    st->Fin = Fin; st->idft = idft;
    st->S = auto_chunk(st->form, st->N, k, st->esz, fused);
    desc_for(&st->desc, A, dir, prec, k, 1);
    size_t blk = (size_t)st->N * (size_t)k * st->esz;
    const size_t nsamp = (size_t)chunk_samples(st->form, st->S);
    st->y = rsvdt_scratch(SCR_Y, blk);
    st->u1 = rsvdt_scratch(SCR_U1, blk);
    st->u2 = rsvdt_scratch(SCR_U2, blk);
    st->acc = st->form ? rsvdt_scratch(SCR_ACC, blk) : st->u2;
    st->fbuf[0] = rsvdt_scratch(SCR_FBUF, blk * nsamp);
    st->fbuf[1] = rsvdt_scratch(SCR_FBUF2, blk * nsamp);
    if (!st->y || !st->u1 || !st->u2 || !st->acc || !st->fbuf[0] || !st->fbuf[1])
        return rsvdt_err(RSVDT_ERR_MEM, "TSActionRK4", "cannot allocate the RK4 work buffers (%.2f GB each): %s",
                         blk / 1e9, rsvdt_hip_last_error());
    int hr = 0;
    for (int b = 0; b < 2 && !hr; b++) {
        hr = rsvdt_hip_event_create(&st->ev_frc[b]);
        if (!hr) hr = rsvdt_hip_event_create(&st->ev_rk[b]);
    }
    if (!hr) hr = rsvdt_hip_zero(st->y, blk, g_rsvdt.stream);
    if (hr) {
        stepper_free(st);
        return rsvdt_err(hr, "TSActionRK4", "device call failed: %s", rsvdt_hip_last_error());
    }
    return 0;
}

static void stepper_free(stepper *st)
{
    for (int b = 0; b < 2; b++) {
        rsvdt_hip_event_destroy(st->ev_frc[b]);
        rsvdt_hip_event_destroy(st->ev_rk[b]);
        st->ev_frc[b] = st->ev_rk[b] = NULL;
    }
}

static void *fslot(stepper *st, int buf, PetscInt j)
{
    return (char *)st->fbuf[buf] + (size_t)j * (size_t)st->N * (size_t)st->k * st->esz;
}

static rsvdt_stream_t frc_stream(void)
{
    return g_rsvdt.opt.overlap_forcing ? g_rsvdt.stream_frc : g_rsvdt.stream;
}

/* Forcing of the chunk of steps i0 .. i0+steps-1 (1-based) into buffer buf
 * (CreateForcingOnFly.c:17-19, positions of TSRK4.c:21-27; the adjoint reads
 * the iDFT table at -pos mod 2Ns):
 *   form 1: slot j = half-step sample 2(i0-1)+j, j = 0 .. 2 steps, from F_hat;
 *   form 0: slot j = g_{i0+j} = sum_m idft(m, 2(i0-1+j)) G_m. */
static PetscErrorCode stepper_forcing(stepper *st, int buf, PetscInt i0, PetscInt steps)
{
    const int64_t L = st->N * st->k;
    const int sgn = st->dir ? 1 : -1;
    const int64_t pos0 = 2 * (i0 - 1);
    const int64_t P = st->form ? 2 * steps + 1 : steps;
    const int stride = st->form ? sgn : 2 * sgn;
    RSVDT_HIPCHK("CreateForcingOnFly",
                 rsvdt_hip_coef_gemm(P, st->Nwe, L, st->idft, 1, st->Nwe, sgn * pos0, stride, st->twoNs, 1.0, st->Fin,
                                     L, fslot(st, buf, 0), L, st->prec, 0.0, frc_stream()));
    return 0;
}

/* One RK4 step on the whole block; s = step index within the chunk. */
static PetscErrorCode stepper_step(stepper *st, int buf, PetscInt s, double *snap)
{
    const double h = st->h;
    rsvdt_stream_t q = g_rsvdt.stream;
    if (st->form == 0) {
        const rsvdt_spmm_desc *d = &st->desc;
        RSVDT_HIPCHK("TSRK4", rsvdt_hip_horner_stage(d, st->y, st->y, NULL, NULL, h / 4, st->u1, NULL, q));
        RSVDT_HIPCHK("TSRK4", rsvdt_hip_horner_stage(d, st->u1, st->y, NULL, NULL, h / 3, st->u2, NULL, q));
        RSVDT_HIPCHK("TSRK4", rsvdt_hip_horner_stage(d, st->u2, st->y, NULL, NULL, h / 2, st->u1, NULL, q));
        RSVDT_HIPCHK("TSRK4", rsvdt_hip_horner_stage(d, st->u1, st->y, NULL, fslot(st, buf, s), h, st->y, snap, q));
        return 0;
    }
    void *F0 = fslot(st, buf, 2 * s), *Fh = fslot(st, buf, 2 * s + 1), *F1 = fslot(st, buf, 2 * s + 2);
    RSVDT_HIPCHK("TSRK4", rsvdt_hip_rk4_stage(&st->desc, 1, st->y, st->y, F0, st->acc, st->u1, h / 2, h / 6, NULL, q));
    RSVDT_HIPCHK("TSRK4", rsvdt_hip_rk4_stage(&st->desc, 2, st->u1, st->y, Fh, st->acc, st->u2, h / 2, h / 6 * 2, NULL, q));
    RSVDT_HIPCHK("TSRK4", rsvdt_hip_rk4_stage(&st->desc, 3, st->u2, st->y, Fh, st->acc, st->u1, h, h / 6 * 2, NULL, q));
    RSVDT_HIPCHK("TSRK4", rsvdt_hip_rk4_stage(&st->desc, 4, st->u1, NULL, F1, st->acc, st->y, 0.0, h / 6, snap, q));
    return 0;
}

/* The S steps of a full chunk from buffer buf (form 0) as one HIP graph: the
 * snapshot column of each step is read by the last stage from slots[t]
 * (filled per chunk before the graph runs), so the same graph serves every
 * chunk of its buffer. */
static PetscErrorCode stepper_capture(stepper *st, int buf, const int *slots, double *snap_base, void **exec)
{
    const double h = st->h;
    rsvdt_stream_t q = g_rsvdt.stream;
    const rsvdt_spmm_desc *d = &st->desc;
    const int64_t sstride = (int64_t)st->N * (int64_t)st->k;
    RSVDT_HIPCHK("TSActionRK4", rsvdt_hip_graph_begin(q));
    int rc = 0;
    for (PetscInt t = 0; t < st->S && !rc; t++) {
        rc = rsvdt_hip_horner_stage(d, st->y, st->y, NULL, NULL, h / 4, st->u1, NULL, q);
        if (!rc) rc = rsvdt_hip_horner_stage(d, st->u1, st->y, NULL, NULL, h / 3, st->u2, NULL, q);
        if (!rc) rc = rsvdt_hip_horner_stage(d, st->u2, st->y, NULL, NULL, h / 2, st->u1, NULL, q);
        if (!rc) rc = rsvdt_hip_horner_stage_ex(d, st->u1, st->y, NULL, fslot(st, buf, t), h, st->y, snap_base,
                                                snap_base ? slots + t : NULL, sstride, q);
    }
    void *e = NULL;
    int rc2 = rsvdt_hip_graph_end(q, &e);         /* always end the capture */
    if (rc || rc2) {
        if (e) rsvdt_hip_graph_destroy(e);
        return rsvdt_err(rc ? rc : rc2, "TSActionRK4", "graph capture failed: %s", rsvdt_hip_last_error());
    }
    *exec = e;
    return 0;
}

/* G_m = Q(z_m) F_hat_m for the Nwe frequency blocks (form 0), FP64, written
 * to G (which may be F itself: the last stage reads F row-locally only).
 * z_m = Ns * idft(m, +-1) is the iDFT's own half-step phase of frequency m.
 * Three Horner stages with per-block scales:
 *   w1 = (1+z)/2 F + (h/4) A F;  w2 = (1+2z) F + h A w1;
 *   G  = h/6 (1+4z+z^2) F + (h^2/6) A w2. */
static PetscErrorCode folded_forcing(Mat A, int dir, double h, PetscInt k, PetscInt Nwe, PetscInt Ns,
                                     const double *idft, const double *F, double *G)
{
    const PetscInt N = A->rows;
    const size_t blk = (size_t)N * (size_t)k * 16;
    double *zrow = malloc(sizeof(double) * 2 * (size_t)Nwe);
    double *sc = malloc(sizeof(double) * 6 * (size_t)Nwe);
    if (!zrow || !sc) { free(zrow); free(sc); return rsvdt_err(RSVDT_ERR_MEM, "TSActionRK4", "out of host memory"); }
    const PetscInt j1 = dir ? 1 : 2 * Ns - 1;
    int h1 = rsvdt_hip_memcpy_d2h(zrow, idft + 2 * (size_t)j1 * (size_t)Nwe, sizeof(double) * 2 * (size_t)Nwe);
    if (h1) { free(zrow); free(sc); return rsvdt_err(h1, "TSActionRK4", "device call failed: %s", rsvdt_hip_last_error()); }
    for (PetscInt m = 0; m < Nwe; m++) {
        const double zr = (double)Ns * zrow[2 * m], zi = (double)Ns * zrow[2 * m + 1];
        const double z2r = zr * zr - zi * zi, z2i = 2 * zr * zi;
        sc[2 * m] = 0.5 * (1 + zr);                       sc[2 * m + 1] = 0.5 * zi;               /* (1+z)/2 */
        sc[2 * (Nwe + m)] = 1 + 2 * zr;                   sc[2 * (Nwe + m) + 1] = 2 * zi;         /* 1+2z */
        sc[2 * (2 * Nwe + m)] = h / 6 * (1 + 4 * zr + z2r); sc[2 * (2 * Nwe + m) + 1] = h / 6 * (4 * zi + z2i);
    }
    free(zrow);
    double *dsc = rsvdt_scratch(SCR_FSCALE, sizeof(double) * 6 * (size_t)Nwe);
    if (!dsc) { free(sc); return rsvdt_err(RSVDT_ERR_MEM, "TSActionRK4", "cannot allocate forcing scales"); }
    int h2 = rsvdt_hip_memcpy_h2d(dsc, sc, sizeof(double) * 6 * (size_t)Nwe);
    free(sc);
    if (h2) return rsvdt_err(h2, "TSActionRK4", "device call failed: %s", rsvdt_hip_last_error());
    size_t fr = 0, tot = 0;
    rsvdt_hip_mem_info(&fr, &tot);
    PetscInt grp = Nwe;
    while (grp > 1 && 2.0 * (double)grp * (double)blk > 0.25 * (double)fr) grp = (grp + 1) / 2;
    double *w1 = rsvdt_scratch(SCR_W1, blk * (size_t)grp);
    double *w2 = rsvdt_scratch(SCR_W2, blk * (size_t)grp);
    if (!w1 || !w2) return rsvdt_err(RSVDT_ERR_MEM, "TSActionRK4", "cannot allocate forcing work buffers");
    rsvdt_stream_t q = g_rsvdt.stream;
    for (PetscInt b0 = 0; b0 < Nwe; b0 += grp) {
        const PetscInt g = grp < Nwe - b0 ? grp : Nwe - b0;
        rsvdt_spmm_desc d;
        desc_for(&d, A, dir, 64, k, g);
        const double *Fb = F + 2 * (size_t)b0 * (size_t)N * (size_t)k;
        double *Gb = G + 2 * (size_t)b0 * (size_t)N * (size_t)k;
        multi_block_tuned(&d, A, dir, k, Fb, w1);
        RSVDT_HIPCHK("TSActionRK4", rsvdt_hip_horner_stage(&d, Fb, Fb, dsc + 2 * b0, NULL, h / 4, w1, NULL, q));
        RSVDT_HIPCHK("TSActionRK4", rsvdt_hip_horner_stage(&d, w1, Fb, dsc + 2 * (Nwe + b0), NULL, h, w2, NULL, q));
        RSVDT_HIPCHK("TSActionRK4", rsvdt_hip_horner_stage(&d, w2, Fb, dsc + 2 * (2 * Nwe + b0), NULL, h * h / 6, Gb,
                                                           NULL, q));
    }
    return 0;
}

/* ------------------------------------------------------------------ event timing */
/* Pairs of HIP events recorded on the run's stream around each phase; the
 * elapsed times are read once at the end so that timing never stalls the
 * launch queue. */
typedef struct { void **ev; int n, cap; double ms; } phase_timer;
static void pt_begin_on(phase_timer *t, rsvdt_stream_t q)
{
    if (t->n + 2 > t->cap) {
        int nc = t->cap ? 2 * t->cap : 64;
        void **ne = realloc(t->ev, sizeof(void *) * (size_t)nc);
        if (!ne) return;
        for (int i = t->cap; i < nc; i++) ne[i] = NULL;
        t->ev = ne; t->cap = nc;
    }
    if (!t->ev[t->n]) rsvdt_hip_event_create(&t->ev[t->n]);
    if (!t->ev[t->n + 1]) rsvdt_hip_event_create(&t->ev[t->n + 1]);
    rsvdt_hip_event_record(t->ev[t->n], q);
}
static void pt_end_on(phase_timer *t, rsvdt_stream_t q)
{
    if (t->n + 2 > t->cap) return;
    rsvdt_hip_event_record(t->ev[t->n + 1], q);
    t->n += 2;
}
static void pt_begin(phase_timer *t) { pt_begin_on(t, g_rsvdt.stream); }
static void pt_end(phase_timer *t) { pt_end_on(t, g_rsvdt.stream); }
static void pt_collect(phase_timer *t)
{
    double ms = 0;
    for (int i = 0; i + 1 < t->n; i += 2) {
        float x = 0;
        rsvdt_hip_event_elapsed_ms(t->ev[i], t->ev[i + 1], &x);
        ms += x;
    }
    t->ms = ms;
}
static void pt_free(phase_timer *t)
{
    for (int i = 0; i < t->cap; i++) if (t->ev[i]) rsvdt_hip_event_destroy(t->ev[i]);
    free(t->ev);
    t->ev = NULL; t->n = t->cap = 0;
}

/* ------------------------------------------------------------------ SpMM autotune */
/* Picks the launch shape of the fused SpMM stage kernel of the current
 * stepping form for this operator, precision and direction -- lanes per row
 * (the column-block width a row's lanes cover), gather batch / streaming
 * variant and XCD row remap -- by timing three launches of every candidate
 * on scratch buffers (a Horner stage, or stage 2 of the stage form).  The
 * choice is kept on the Mat and reused by every later action. */
/* (columns per pass, tile height) of the tiled kernel the autotuner times */
static const int tile_cand[][2] = {{8, 32}, {16, 16}, {16, 32}, {32, 8}};

static PetscErrorCode autotune_lanes(Mat A, int dir, int prec, PetscInt k, void *a, void *b, void *c, void *d)
{
    const int p = prec == 32, q = dir ? 0 : 1, f = g_rsvdt.opt.rk4_form ? 1 : 0;
    if ((A->tuned_ok[f][p][q] && A->tuned_k[f][p][q] == k) || !g_rsvdt.opt.autotune || g_rsvdt.opt.spmm_lanes ||
        g_rsvdt.opt.spmm_variant >= 0)
        return 0;
    rsvdt_spmm_desc dsc;
    desc_for(&dsc, A, dir, prec, k, 1);
    void *e0, *e1;
    RSVDT_HIPCHK("autotune", rsvdt_hip_event_create(&e0));
    RSVDT_HIPCHK("autotune", rsvdt_hip_event_create(&e1));
    float best = 1e30f;
    int bl = 0, bv = 3, br = 0;
    for (int l = 4; l <= 64; l <<= 1) {
        if ((k + l - 1) / l > 4) continue;            /* at most 4 columns per lane */
        if (l >= 2 * k && l > 4) continue;            /* no more than half the lanes idle */
        const int cpl = (int)((k + l - 1) / l);
        /* gather-batch variants: 0-3 (the RK4 stage: with one column per
         * lane), 4-5 (12-nonzero batches) with 16 or more lanes and at most
         * two columns per lane, 6-7 (grouped rows, FP32 pairs) with one */
        int nvar = 1;
        if (f) nvar = cpl > 1 ? 1 : 4;
        else if (cpl == 1) nvar = l >= 16 ? 8 : 4;
        else if (cpl == 2) nvar = l >= 16 ? 6 : 4;
        else nvar = 4;
        for (int v = 0; v < nvar; v++)
            for (int r = 0; r < 4; r++) {
                dsc.lanes = l; dsc.variant = v; dsc.xcd_remap = r;
                /* all operands are scratch buffers: gathered a, row-local b (and d), out c */
#define RSVDT_TUNE_LAUNCH()                                                                              \
    (f ? rsvdt_hip_rk4_stage(&dsc, 2, a, b, NULL, d, c, 0.0, 0.0, NULL, g_rsvdt.stream)                   \
       : rsvdt_hip_horner_stage(&dsc, a, b, NULL, NULL, 0.0, c, NULL, g_rsvdt.stream))
                if (RSVDT_TUNE_LAUNCH()) continue;
                rsvdt_hip_event_record(e0, g_rsvdt.stream);
                for (int rep = 0; rep < 3; rep++) (void)RSVDT_TUNE_LAUNCH();
#undef RSVDT_TUNE_LAUNCH
                rsvdt_hip_event_record(e1, g_rsvdt.stream);
                float ms = 0;
                rsvdt_hip_event_elapsed_ms(e0, e1, &ms);
                if (ms < best) { best = ms; bl = l; bv = v; br = r; }
            }
    }
    /* FP32 column pairs: one pair per lane (variant 7, lanes = k/2) or two
     * (variant 8, lanes = k/4) */
    for (int v = 7; v <= 8 && !f && prec == 32; v++) {
        const PetscInt per = v == 7 ? 2 : 4, l = k / per;
        if (k % per || !(l == 4 || l == 8 || l == 16 || (v == 7 && l == 32))) continue;
        for (int r = 0; r < 4; r++) {
            dsc.lanes = (int)l; dsc.variant = v; dsc.xcd_remap = r;
            if (rsvdt_hip_horner_stage(&dsc, a, b, NULL, NULL, 0.0, c, NULL, g_rsvdt.stream)) continue;
            rsvdt_hip_event_record(e0, g_rsvdt.stream);
            for (int rep = 0; rep < 3; rep++) (void)rsvdt_hip_horner_stage(&dsc, a, b, NULL, NULL, 0.0, c, NULL, g_rsvdt.stream);
            rsvdt_hip_event_record(e1, g_rsvdt.stream);
            float ms = 0;
            rsvdt_hip_event_elapsed_ms(e0, e1, &ms);
            if (ms < best) { best = ms; bl = (int)l; bv = v; br = r; }
        }
    }
    /* the tiled kernel (Horner form; variant 10 prefetches the next pass's
     * rows into registers): columns per pass x tile height, on tiled forms
     * built here on first use */
    int bt = 0;
    for (int ci = 0; ci < (int)(sizeof(tile_cand) / sizeof(tile_cand[0])) && !f; ci++) {
        const int kc = tile_cand[ci][0], tr = tile_cand[ci][1];
        if (kc > 8 && k <= kc / 2) continue;          /* more than half the lanes of a pass idle */
        rsvdt_spmm_desc dt = dsc;
        dt.lanes = kc;
        if (!tile_attach(&dt, A, dir, tr)) continue;
        for (int tv = 9; tv <= 10; tv++)
            for (int r = 0; r < 4; r++) {
                dt.variant = tv;
                dt.xcd_remap = r;
                if (rsvdt_hip_horner_stage(&dt, a, b, NULL, NULL, 0.0, c, NULL, g_rsvdt.stream)) continue;
                rsvdt_hip_event_record(e0, g_rsvdt.stream);
                for (int rep = 0; rep < 3; rep++)
                    (void)rsvdt_hip_horner_stage(&dt, a, b, NULL, NULL, 0.0, c, NULL, g_rsvdt.stream);
                rsvdt_hip_event_record(e1, g_rsvdt.stream);
                float ms = 0;
                rsvdt_hip_event_elapsed_ms(e0, e1, &ms);
                if (ms < best) { best = ms; bl = kc; bv = tv; br = r; bt = tr; }
            }
    }
    rsvdt_hip_event_destroy(e0);
    rsvdt_hip_event_destroy(e1);
    A->tuned_lanes[f][p][q] = bl;
    A->tuned_variant[f][p][q] = bv;
    A->tuned_remap[f][p][q] = br;
    A->tuned_tile[f][p][q] = bt;
    A->tuned_ok[f][p][q] = 1;
    A->tuned_k[f][p][q] = k;
    if (g_rsvdt.opt.verbose)
        rsvdt_printf("[rsvdt] autotune (%s, FP%d, %s): %d lanes/row, variant %d, xcd remap %d, tile %d rows: %.1f us "
                     "per stage\n", dir ? "A" : "A^H", prec, f ? "stage form" : "Horner form", bl, bv, br, bt,
                     1e3 * best / 3);
    return 0;
}

/* ------------------------------------------------------------------ time stepping */
/* TSActionRK4.c:38-130: steps all k test vectors together from y = 0 over
 * rend = Nt + Ns (+RR) steps and stores Nstore snapshots, every RR steps
 * after Nt, into Y_all (block s = snapshot column s of every test vector;
 * the adjoint stores them reversed, SourceCode/SaveSnapshots.c:19-28). */
/* Steps i_first .. i_last (1-based, of 1 .. rend) of one stepping precision
 * on the stepper's state, from forcing chunks that start at i_first. */
static PetscErrorCode step_range(stepper *st, Mat A, RSVDt_vars *V, Mat Yall, PetscInt i_first, PetscInt i_last,
                                 int fused, phase_timer *t_rk4, phase_timer *t_frc, int64_t *launches)
{
    const int dir = st->dir;
    const PetscInt k = st->k, N = st->N;
    const PetscInt Nstore = V->TS.TransientRemoval ? V->RSVD.Nw + 1 : V->RSVD.Nw;
    const PetscInt nstep = i_last - i_first + 1;
    /* every failure is recorded in ierr and the loops stop, so that the tail
     * (device sync, graph destruction) always runs: a forcing GEMM may still
     * be writing a chunk buffer on the forcing stream when an error is met on
     * the main stream */
    PetscErrorCode ierr = 0;
#define TS_HIP(x)                                                                                      \
    do {                                                                                               \
        if (!ierr) {                                                                                   \
            int _h = (x);                                                                              \
            if (_h) ierr = rsvdt_err(_h, "TSActionRK4", "device call failed: %s", rsvdt_hip_last_error()); \
        }                                                                                              \
    } while (0)
    const PetscInt nch = (nstep + st->S - 1) / st->S;
    /* forcing of the first two chunks; later chunks are refilled as soon as
     * the steps that read their buffer are done */
    TS_HIP(rsvdt_hip_event_record(st->ev_rk[1], g_rsvdt.stream));   /* state ready */
    for (PetscInt c = 0; c < nch && c < 2 && !ierr; c++) {
        PetscInt i0 = i_first + c * st->S, steps = st->S < i_last - i0 + 1 ? st->S : i_last - i0 + 1;
        TS_HIP(rsvdt_hip_stream_wait_event(frc_stream(), st->ev_rk[1]));
        if (t_frc) pt_begin_on(t_frc, frc_stream());
        if (!ierr) ierr = stepper_forcing(st, (int)(c & 1), i0, steps);
        if (t_frc) pt_end_on(t_frc, frc_stream());
        TS_HIP(rsvdt_hip_event_record(st->ev_frc[c & 1], frc_stream()));
    }
    /* graph replay of full chunks (form 0): one graph per buffer */
    const int use_graph = g_rsvdt.opt.use_graph && st->form == 0 && !fused;
    void *gexec[2] = {NULL, NULL};
    int *slots = NULL;
    if (use_graph && !ierr) {
        slots = rsvdt_scratch(SCR_SLOTS, sizeof(int) * 2 * (size_t)st->S);
        if (!slots) ierr = rsvdt_err(RSVDT_ERR_MEM, "TSActionRK4", "cannot allocate slot table");
    }
    for (PetscInt c = 0; c < nch && !ierr; c++) {
        const int buf = (int)(c & 1);
        const PetscInt i0 = i_first + c * st->S, steps = st->S < i_last - i0 + 1 ? st->S : i_last - i0 + 1;
        TS_HIP(rsvdt_hip_stream_wait_event(g_rsvdt.stream, st->ev_frc[buf]));
        if (ierr) break;
        if (t_rk4) pt_begin(t_rk4);
        if (fused) {
            const double cf[4] = {st->h / 4, st->h / 3, st->h / 2, st->h};
            TS_HIP(rsvdt_hip_horner_fused(&st->desc, A->nnz, dir ? A->maxrow : A->maxrowH, st->y, fslot(st, buf, 0),
                                          (int64_t)N * k, cf, steps, Yall ? Yall->d : NULL, (int64_t)N * k, i0,
                                          V->TS.Nt, V->TS.ResRatio, Nstore, dir, g_rsvdt.stream));
            *launches += 1;
        } else if (use_graph && steps == st->S) {
            int *sl = slots + buf * st->S;
            if (Yall) TS_HIP(rsvdt_hip_snap_slots(sl, st->S, i0, V->TS.Nt, V->TS.ResRatio, Nstore, dir, g_rsvdt.stream));
            if (!ierr && !gexec[buf]) ierr = stepper_capture(st, buf, sl, Yall ? Yall->d : NULL, &gexec[buf]);
            TS_HIP(rsvdt_hip_graph_launch(gexec[buf], g_rsvdt.stream));
            *launches += 4 * steps;
        }
        for (PetscInt s = 0; s < steps && !ierr && !fused && !(use_graph && steps == st->S); s++) {
            PetscInt i = i0 + s;
            double *snap = NULL;
            if (Yall && i > V->TS.Nt && (i - V->TS.Nt) % V->TS.ResRatio == 0) {
                PetscInt pos = (i - V->TS.Nt) / V->TS.ResRatio - 1;
                pos %= Nstore;
                if (!dir) pos = Nstore - 1 - pos;
                snap = Yall->d + 2 * (size_t)pos * (size_t)N * (size_t)k;
            }
            ierr = stepper_step(st, buf, s, snap);
            *launches += 4;
        }
        if (t_rk4) pt_end(t_rk4);
        TS_HIP(rsvdt_hip_event_record(st->ev_rk[buf], g_rsvdt.stream));
        if (c + 2 < nch && !ierr) {
            const PetscInt j0 = i_first + (c + 2) * st->S, st2 = st->S < i_last - j0 + 1 ? st->S : i_last - j0 + 1;
            TS_HIP(rsvdt_hip_stream_wait_event(frc_stream(), st->ev_rk[buf]));
            if (t_frc) pt_begin_on(t_frc, frc_stream());
            if (!ierr) ierr = stepper_forcing(st, buf, j0, st2);
            if (t_frc) pt_end_on(t_frc, frc_stream());
            TS_HIP(rsvdt_hip_event_record(st->ev_frc[buf], frc_stream()));
        }
    }
#undef TS_HIP
    /* the main stream must not run ahead of a forcing GEMM still reading
     * F_hat, and no buffer may be reused while one is writing: always wait */
    int hs = rsvdt_hip_sync();
    if (hs && !ierr) ierr = rsvdt_err(hs, "TSActionRK4", "device call failed: %s", rsvdt_hip_last_error());
    for (int b = 0; b < 2; b++) rsvdt_hip_graph_destroy(gexec[b]);
    return ierr;
}

/* The time stepping of one action (TSActionRK4.c:38-130) over rend steps.
 * With the FP32 state and transient removal, the Nt steps before the first
 * snapshot run in FP32 and the snapshot period in FP64: the FP32 rounding
 * accumulated by then is a perturbation of the state at step Nt, i.e. part
 * of the transient, which the FP64 snapshots carry like the rest of it and
 * the removal takes out (kept in FP32, the removal's correction of the
 * transient would carry the rounding into the results). */
static PetscErrorCode time_stepping(Mat A, DFT_matrices *D, RSVDt_vars *V, Mat Fhat, int fhat_folded, Mat Yall,
                                    PetscInt max_steps, phase_timer *t_rk4, phase_timer *t_frc)
{
    RSVDT_CHK(rsvdt_aij_upload(A));
    const int dir = V->TS.DirAdj ? 1 : 0;
    const int prec = g_rsvdt.opt.precision;
    const PetscInt k = V->RSVD.k, Nwe = V->RSVD.Nw_eff, N = A->rows;
    PetscInt rend = V->TS.Nt + V->TS.Ns + (V->TS.TransientRemoval ? V->TS.ResRatio : 0);
    if (max_steps > 0 && max_steps < rend) rend = max_steps;
    const int split = prec == 32 && V->TS.TransientRemoval && Yall && V->TS.Nt > 0 && rend > V->TS.Nt;
    const PetscInt seg_last[2] = {split ? V->TS.Nt : rend, rend};
    const int seg_prec[2] = {prec, 64};
    int64_t launches = 0;
    PetscErrorCode ierr = 0;
    double *y64 = NULL;          /* the FP32 state at step Nt, widened */
    stepper st;
    memset(&st, 0, sizeof(st));
    for (int seg = 0; seg < (split ? 2 : 1) && !ierr; seg++) {
        const int p = seg_prec[seg];
        /* operators whose state fits in LDS: whole chunks of Horner steps per launch */
        const int fused = !g_rsvdt.opt.rk4_form && g_rsvdt.opt.fuse_small &&
                          rsvdt_hip_horner_fused_fits(N, A->nnz, A->is_real, p);
        ierr = stepper_init(&st, A, dir, p, V->TS.dt, k, Nwe, V->TS.Ns, Fhat->d, D->idft->d, fused);
        if (ierr) break;
        if (g_rsvdt.opt.autotune && !g_rsvdt.opt.spmm_lanes && !fused) {
            ierr = autotune_lanes(A, dir, p, k, st.u1, st.y, st.u2, st.form ? st.acc : st.u2);
            desc_for(&st.desc, A, dir, p, k, 1);
            int hz = ierr ? 0 : rsvdt_hip_zero(st.y, (size_t)N * k * st.esz, g_rsvdt.stream);
            if (hz) ierr = rsvdt_err(hz, "TSActionRK4", "device call failed: %s", rsvdt_hip_last_error());
        }
        if (!ierr && seg == 1) {
            int hc = rsvdt_hip_memcpy_d2d(st.y, y64, (size_t)N * (size_t)k * 16, g_rsvdt.stream);
            if (hc) ierr = rsvdt_err(hc, "TSActionRK4", "device call failed: %s", rsvdt_hip_last_error());
        }
        if (!ierr && seg == 0 && st.form == 0 && !fhat_folded) {   /* G_m = Q(z_m) F_hat_m in place of F_hat */
            if (t_frc) pt_begin(t_frc);
            ierr = folded_forcing(A, dir, V->TS.dt, k, Nwe, V->TS.Ns, D->idft->d, Fhat->d, Fhat->d);
            if (t_frc) pt_end(t_frc);
        }
        if (!ierr)
            ierr = step_range(&st, A, V, Yall, seg ? seg_last[0] + 1 : 1, seg_last[seg], fused, t_rk4, t_frc,
                              &launches);
        if (!ierr && split && seg == 0) {
            if (rsvdt_hip_malloc((void **)&y64, (size_t)N * (size_t)k * 16))
                ierr = rsvdt_err(RSVDT_ERR_MEM, "TSActionRK4", "cannot allocate the FP64 state: %s",
                                 rsvdt_hip_last_error());
            int hc = ierr ? 0 : rsvdt_hip_c2z((int64_t)N * k, (const float *)st.y, y64, g_rsvdt.stream);
            if (hc) ierr = rsvdt_err(hc, "TSActionRK4", "device call failed: %s", rsvdt_hip_last_error());
        }
        if (seg == 0 && split) stepper_free(&st);
    }
    if (y64) {
        rsvdt_hip_sync();
        rsvdt_hip_free(y64);
    }
    stepper_free(&st);
    /* in an action, chunk buffers of more than 1 GiB go back to the device:
     * the transient removal that follows sizes its groups to the free memory */
    if (Yall && rsvdt_scratch_size(SCR_FBUF) + rsvdt_scratch_size(SCR_FBUF2) > ((size_t)1 << 30)) {
        rsvdt_scratch_release(SCR_FBUF);
        rsvdt_scratch_release(SCR_FBUF2);
    }
    RSVDT_CHK(ierr);
    g_rsvdt.stats.steps = rend;
    g_rsvdt.stats.rk4_launches = launches;
    g_rsvdt.stats.chunk_steps = (int)st.S;
    g_rsvdt.stats.spmm_lanes = st.desc.lanes;
    g_rsvdt.stats.spmm_variant = st.desc.variant;
    g_rsvdt.stats.spmm_xcd_remap = st.desc.xcd_remap;
    g_rsvdt.stats.rk4_form = st.form;
    return 0;
}

/* Homogeneous RK4 steps on nb blocks of N x k in place (SourceCode/TSTransRK4.c:15-36
 * via TSExpADeltaT.c:25-34), FP64, in groups that bound the temporaries;
 * Horner stages (form 0) or the reference's stage order (form 1). */
PetscErrorCode rsvdt_homogeneous_steps(Mat A, int dir, double h, PetscInt nsteps, double *X, PetscInt N, PetscInt k,
                                       PetscInt nb)
{
    RSVDT_CHK(rsvdt_aij_upload(A));
    const int form = g_rsvdt.opt.rk4_form ? 1 : 0;
    size_t blk = (size_t)N * (size_t)k * 16;
    size_t fr = 0, tot = 0;
    rsvdt_hip_mem_info(&fr, &tot);
    const double ntmp = form ? 3.0 : 2.0;
    PetscInt G = nb;
    while (G > 1 && ntmp * (double)G * (double)blk > 0.4 * (double)fr) G = (G + 1) / 2;
    if (G > 65535) G = 65535;
    double *u1 = rsvdt_scratch(SCR_U1, blk * (size_t)G);
    double *u2 = rsvdt_scratch(SCR_U2, blk * (size_t)G);
    double *acc = form ? rsvdt_scratch(SCR_ACC, blk * (size_t)G) : u2;
    if (!u1 || !u2 || !acc) return rsvdt_err(RSVDT_ERR_MEM, "TSExpADeltaT", "cannot allocate work buffers");
    const int fused = form == 0 && g_rsvdt.opt.fuse_small && rsvdt_hip_horner_fused_fits(N, A->nnz, A->is_real, 64);
    for (PetscInt g0 = 0; g0 < nb; g0 += G) {
        PetscInt g = G < nb - g0 ? G : nb - g0;
        rsvdt_spmm_desc d;
        desc_for(&d, A, dir, 64, k, g);
        double *y = X + 2 * (size_t)g0 * (size_t)N * (size_t)k;
        rsvdt_stream_t q = g_rsvdt.stream;
        if (!fused && g > 1) multi_block_tuned(&d, A, dir, k, y, u1);
        else multi_block_shape(&d, k);
        if (fused) {
            const double cf[4] = {h / 4, h / 3, h / 2, h};
            RSVDT_HIPCHK("TSTransRK4", rsvdt_hip_horner_fused(&d, A->nnz, dir ? A->maxrow : A->maxrowH, y, NULL, 0, cf,
                                                              nsteps, NULL, 0, 0, 0, 1, 1, dir, q));
            continue;
        }
        for (PetscInt it = 0; it < nsteps; it++) {
            if (form == 0) {
                RSVDT_HIPCHK("TSTransRK4", rsvdt_hip_horner_stage(&d, y, y, NULL, NULL, h / 4, u1, NULL, q));
                RSVDT_HIPCHK("TSTransRK4", rsvdt_hip_horner_stage(&d, u1, y, NULL, NULL, h / 3, u2, NULL, q));
                RSVDT_HIPCHK("TSTransRK4", rsvdt_hip_horner_stage(&d, u2, y, NULL, NULL, h / 2, u1, NULL, q));
                RSVDT_HIPCHK("TSTransRK4", rsvdt_hip_horner_stage(&d, u1, y, NULL, NULL, h, y, NULL, q));
                continue;
            }
            RSVDT_HIPCHK("TSTransRK4", rsvdt_hip_rk4_stage(&d, 1, y, y, NULL, acc, u1, h / 2, h / 6, NULL, q));
            RSVDT_HIPCHK("TSTransRK4", rsvdt_hip_rk4_stage(&d, 2, u1, y, NULL, acc, u2, h / 2, h / 6 * 2, NULL, q));
            RSVDT_HIPCHK("TSTransRK4", rsvdt_hip_rk4_stage(&d, 3, u2, y, NULL, acc, u1, h, h / 6 * 2, NULL, q));
            RSVDT_HIPCHK("TSTransRK4", rsvdt_hip_rk4_stage(&d, 4, u1, NULL, NULL, acc, y, 0.0, h / 6, NULL, q));
        }
    }
    return 0;
}

/* ------------------------------------------------------------------ CholeskyQR */
/* Shifted CholeskyQR3 (Fukaya et al., SIAM J. Sci. Comput. 2020): a first
 * pass on X^H X + s I with s = 11(mn + n(n+1)) u ||X||_F^2, then CholeskyQR2.
 * Q is X R^{-1} with R = R3 R2 R1 upper triangular with a positive real
 * diagonal -- the unique factor that SLEPc's Gram-Schmidt BVOrthogonalize
 * also returns for a full-rank input (SourceCode/QRAllFreqs.c:26).
 *
 * npass passes on X (the first one shifted when shift_first); R_out
 * (optional) receives the product of the passes' factors.  T (optional): a
 * buffer like X of `bytes` bytes; the applies then alternate between X and
 * T (required for more than 64 columns with unit column strides, where the
 * apply cannot work in place) and the result is copied back into X. */
PetscErrorCode rsvdt_cholqr_bufs(PetscInt nb, PetscInt nrows, PetscInt n, double *X, int64_t xb, int64_t xr,
                                 int64_t xa, double *R_out, int npass, int shift_first, double drop_tol, double *T,
                                 size_t bytes, const rsvdt_cholqr_work *w, rsvdt_stream_t q, int report)
{
    const size_t mat = (size_t)nb * (size_t)n * (size_t)n * 16;
    RSVDT_HIPCHK("QRAllFreqs", rsvdt_hip_zero(w->info, 256, q));
    const double shift_rel = 11.0 * ((double)nrows * (double)n + (double)n * (double)(n + 1)) * 1.1102230246251565e-16;
    double *cur = X, *oth = T ? T : X;
    for (int pass = 0; pass < npass; pass++) {
        RSVDT_HIPCHK("QRAllFreqs", rsvdt_hip_bgram(nb, nrows, n, n, cur, xb, xr, xa, cur, xb, xr, xa, w->G, w->work, q));
        RSVDT_HIPCHK("QRAllFreqs", rsvdt_hip_bchol_inv(nb, n, w->G, (pass == 0 && shift_first) ? shift_rel : 0.0,
                                                       drop_tol, w->R, w->Ri, w->info, q));
        RSVDT_HIPCHK("QRAllFreqs", rsvdt_hip_bapply(nb, nrows, n, n, cur, xb, xr, xa, w->Ri, 0.0, NULL, 0, 0, 0, oth, xb,
                                                    xr, xa, q));
        if (T) { double *tmp = cur; cur = oth; oth = tmp; }
        if (R_out) {
            if (pass == 0) RSVDT_HIPCHK("QRAllFreqs", rsvdt_hip_memcpy_d2d(w->Ra, w->R, mat, q));
            else {
                RSVDT_HIPCHK("QRAllFreqs", rsvdt_hip_btrmm_upper(nb, n, w->R, w->Ra, w->Rt, q));
                RSVDT_HIPCHK("QRAllFreqs", rsvdt_hip_memcpy_d2d(w->Ra, w->Rt, mat, q));
            }
        }
    }
    if (cur != X) RSVDT_HIPCHK("QRAllFreqs", rsvdt_hip_memcpy_d2d(X, cur, bytes, q));
    if (R_out) RSVDT_HIPCHK("QRAllFreqs", rsvdt_hip_memcpy_d2d(R_out, w->Ra, mat, q));
    if (report) {
        int hinfo = 0;
        RSVDT_HIPCHK("QRAllFreqs", rsvdt_hip_memcpy_d2h(&hinfo, w->info, sizeof(int)));
        if (hinfo) rsvdt_printf("[rsvdt] warning: CholeskyQR met %d non-positive pivots (rank-deficient input)\n", hinfo);
    }
    return 0;
}

static PetscErrorCode cholqr_passes(PetscInt nb, PetscInt nrows, PetscInt n, double *X, int64_t xb, int64_t xr,
                                    int64_t xa, double *R_out, int npass, int shift_first, double drop_tol,
                                    double *T, size_t bytes)
{
    size_t mat = (size_t)nb * (size_t)n * (size_t)n * 16;
    rsvdt_cholqr_work w;
    w.G = rsvdt_scratch(SCR_G, mat);
    w.R = rsvdt_scratch(SCR_R, mat);
    w.Ri = rsvdt_scratch(SCR_RINV, mat);
    w.Ra = R_out ? rsvdt_scratch(SCR_RACC, mat) : NULL;
    w.Rt = R_out ? rsvdt_scratch(SCR_RTMP, mat) : NULL;
    w.work = rsvdt_scratch(SCR_WORK, rsvdt_hip_bgram_work_bytes(nb, nrows, n, n));
    w.info = rsvdt_scratch(SCR_INFO, 256);
    if (!w.G || !w.R || !w.Ri || (R_out && (!w.Ra || !w.Rt)) || !w.work || !w.info)
        return rsvdt_err(RSVDT_ERR_MEM, "QRAllFreqs", "cannot allocate CholeskyQR work");
    return rsvdt_cholqr_bufs(nb, nrows, n, X, xb, xr, xa, R_out, npass, shift_first, drop_tol, T, bytes, &w,
                             g_rsvdt.stream, 1);
}

PetscErrorCode rsvdt_cholqr(PetscInt nb, PetscInt nrows, PetscInt n, double *X, int64_t xb, int64_t xr, int64_t xa,
                            double *R_out, double drop_tol)
{
    double *T = NULL;
    size_t bytes = (size_t)nb * (size_t)nrows * (size_t)n * 16;
    if (n > 64 && xa == 1) {      /* the wide apply is out of place: a dense contiguous layout is required */
        const int64_t ext = (nb - 1) * xb + (nrows - 1) * xr + (n - 1) * xa + 1;
        if (ext != nb * nrows * n)
            return rsvdt_err(RSVDT_ERR_USER, "QRAllFreqs", "more than 64 columns need a contiguous layout");
        T = rsvdt_scratch(SCR_QRT, bytes);
        if (!T) return rsvdt_err(RSVDT_ERR_MEM, "QRAllFreqs", "cannot allocate CholeskyQR work");
    }
    return cholqr_passes(nb, nrows, n, X, xb, xr, xa, R_out, 3, 1, drop_tol, T, bytes);
}

PetscErrorCode rsvdt_span_basis(PetscInt nb, PetscInt nrows, PetscInt n, const double *X, double *Q, double *E,
                                double *T, int64_t xb, int64_t xr, int64_t xa, size_t bytes, double drop_tol)
{
    size_t mat = (size_t)nb * (size_t)n * (size_t)n * 16;
    double *C = rsvdt_scratch(SCR_PROJ, mat);
    double *D = rsvdt_scratch(SCR_MASK, mat);
    double *Rm = rsvdt_scratch(SCR_RMASK, mat);
    int *cnt = rsvdt_scratch(SCR_CNT, 256);
    void *work = rsvdt_scratch(SCR_WORK, rsvdt_hip_bgram_work_bytes(nb, nrows, n, n));
    if (!C || !D || !Rm || !cnt || !work) return rsvdt_err(RSVDT_ERR_MEM, "CreateMtilde", "cannot allocate basis work");
    rsvdt_stream_t q = g_rsvdt.stream;
    const int64_t nel = (int64_t)(bytes / 16);
    /* level 1 */
    RSVDT_HIPCHK("CreateMtilde", rsvdt_hip_memcpy_d2d(Q, X, bytes, q));
    RSVDT_CHK(cholqr_passes(nb, nrows, n, Q, xb, xr, xa, Rm, 3, 1, drop_tol, T, bytes));
    RSVDT_HIPCHK("CreateMtilde", rsvdt_hip_zero(cnt, 256, q));
    RSVDT_HIPCHK("CreateMtilde", rsvdt_hip_bdrop_mask(nb, n, Rm, D, cnt, q));
    int ndrop = 0;
    RSVDT_HIPCHK("CreateMtilde", rsvdt_hip_memcpy_d2h(&ndrop, cnt, sizeof(int)));
    if (ndrop == 0) return 0;
    /* level 2: E = ((I - Q Q^H) X) D, the residuals of the dropped columns */
    RSVDT_HIPCHK("CreateMtilde", rsvdt_hip_bgram(nb, nrows, n, n, Q, xb, xr, xa, X, xb, xr, xa, C, work, q));
    RSVDT_HIPCHK("CreateMtilde", rsvdt_hip_zscale((int64_t)nb * n * n, -1.0, C, C, q));
    RSVDT_HIPCHK("CreateMtilde", rsvdt_hip_bapply(nb, nrows, n, n, Q, xb, xr, xa, C, 1.0, X, xb, xr, xa, T, xb, xr, xa, q));
    RSVDT_HIPCHK("CreateMtilde", rsvdt_hip_bapply(nb, nrows, n, n, T, xb, xr, xa, D, 0.0, NULL, 0, 0, 0, E, xb, xr, xa, q));
    RSVDT_CHK(cholqr_passes(nb, nrows, n, E, xb, xr, xa, NULL, 3, 1, drop_tol, T, bytes));
    /* reorthogonalise against the level-1 basis (into T) and renormalise */
    RSVDT_HIPCHK("CreateMtilde", rsvdt_hip_bgram(nb, nrows, n, n, Q, xb, xr, xa, E, xb, xr, xa, C, work, q));
    RSVDT_HIPCHK("CreateMtilde", rsvdt_hip_zscale((int64_t)nb * n * n, -1.0, C, C, q));
    RSVDT_HIPCHK("CreateMtilde", rsvdt_hip_bapply(nb, nrows, n, n, Q, xb, xr, xa, C, 1.0, E, xb, xr, xa, T, xb, xr, xa, q));
    RSVDT_CHK(cholqr_passes(nb, nrows, n, T, xb, xr, xa, NULL, 2, 0, drop_tol, E, bytes));
    /* the level-2 columns take the dropped columns' places; one more pass */
    RSVDT_HIPCHK("CreateMtilde", rsvdt_hip_zaxpy(nel, 1.0, T, Q, q));
    RSVDT_CHK(cholqr_passes(nb, nrows, n, Q, xb, xr, xa, NULL, 1, 0, drop_tol, T, bytes));
    return 0;
}

/* Left singular vectors (descending) of nb tall N x n matrices in place:
 * CholeskyQR, one-sided Jacobi on R, Q <- Q U_R. */
PetscErrorCode rsvdt_svd_left(PetscInt nb, PetscInt nrows, PetscInt n, double *X, int64_t xb, int64_t xr, int64_t xa,
                              double *sig_dev)
{
    size_t mat = (size_t)nb * (size_t)n * (size_t)n * 16;
    double *Rf = rsvdt_scratch(SCR_MISC, mat);
    double *U = rsvdt_scratch(SCR_MISC2, mat);
    if (!Rf || !U) return rsvdt_err(RSVDT_ERR_MEM, "SVDAllFreqs", "cannot allocate SVD work");
    RSVDT_CHK(rsvdt_cholqr(nb, nrows, n, X, xb, xr, xa, Rf, 0.0));
    RSVDT_HIPCHK("SVDAllFreqs", rsvdt_hip_bjacobi_svd(nb, n, Rf, U, sig_dev, g_rsvdt.stream));
    RSVDT_HIPCHK("SVDAllFreqs", rsvdt_hip_bapply(nb, nrows, n, n, X, xb, xr, xa, U, 0.0, NULL, 0, 0, 0, X, xb, xr, xa,
                                                 g_rsvdt.stream));
    return 0;
}

/* ------------------------------------------------------------------ Permute */
/* SourceCode/PermuteMat.c:24-38 and ReversePermuteMat.c:24-38: column
 * ik + iw*k <-> ik*Nw + iw.  The blocked layout makes both a relabelling. */
PetscErrorCode PermuteMat(Mat F, RSVDt_vars *V)
{
    if (!F || F->kind != RSVDT_MAT_DENSE) return rsvdt_err(RSVDT_ERR_USER, "PermuteMat", "dense Mat required");
    PetscInt k = V->RSVD.k;
    if (F->cols % k) return rsvdt_err(RSVDT_ERR_USER, "PermuteMat", "columns not a multiple of k");
    if (F->order != 0 || F->kb != k) RSVDT_CHK(rsvdt_mat_to_layout(F, k, 0));
    F->order = 1;
    return 0;
}

PetscErrorCode ReversePermuteMat(Mat F, RSVDt_vars *V)
{
    if (!F || F->kind != RSVDT_MAT_DENSE) return rsvdt_err(RSVDT_ERR_USER, "ReversePermuteMat", "dense Mat required");
    PetscInt k = V->RSVD.k;
    if (F->cols % k) return rsvdt_err(RSVDT_ERR_USER, "ReversePermuteMat", "columns not a multiple of k");
    if (F->order != 1 || F->kb != k) RSVDT_CHK(rsvdt_mat_to_layout(F, k, 1));
    F->order = 0;
    return 0;
}

/* ------------------------------------------------------------------ DFT */
/* SourceCode/DFT.c:18-39: Y_hat = RR * Y_all * dft (per test vector), then the
 * reverse permutation to frequency-major: one GEMM over all test vectors
 * whose output blocks are already frequency-major. */
PetscErrorCode DFT(Mat Yall, DFT_matrices *D, RSVD_matrices *R, RSVDt_vars *V)
{
    PetscInt Nw = D->dft->rows, Nwe = V->RSVD.Nw_eff, k = Yall->cols / Nw, N = Yall->rows;
    if (Yall->cols % Nw) return rsvdt_err(RSVDT_ERR_USER, "DFT", "Y_all columns not a multiple of Nw");
    if (Yall->order != 1 || Yall->kb != k) RSVDT_CHK(rsvdt_mat_to_layout(Yall, k, 1));
    Mat Y;
    RSVDT_CHK(rsvdt_mat_alloc_dense(N, k * Nwe, k, 0, &Y));
    const int64_t L = N * k;
    RSVDT_HIPCHK("DFT", rsvdt_hip_coef_gemm(Nwe, Nw, L, D->dft->d, Nwe, 1, 0, 1, Nwe, (double)V->TS.ResRatio, Yall->d, L,
                                            Y->d, L, 64, 0.0, g_rsvdt.stream));
    RSVDT_MatDestroy(&Yall);
    R->Y_hat = Y;
    return 0;
}

/* ------------------------------------------------------------------ transient removal */
/* The removal is independent per test vector, so it works on the
 * per-test-vector row-major layout: X_ik (N x Ns1, row r = the Ns1
 * snapshots of row r) for a group of test vectors at a time, sized to the
 * free memory.  In that layout the snapshot basis, M_tilde and the Galerkin
 * sums are batched tall-skinny products with contiguous rows, and the RR
 * homogeneous steps of the basis are block SpMMs with Ns1 columns.
 * When one group holds all test vectors, the snapshots and the basis built
 * by CreateMtilde are kept for GalerkinProjAllModes (the reference computes
 * the same QR in both, SourceCode/CreateMtilde.c:43-46 and
 * GalerkinProjAllModes.c:31-33); otherwise each recomputes it per group. */
typedef struct {
    const void *key;             /* Y_all data the group in X, V was built from (NULL: none) */
    PetscInt b0, kb, N, n;
    size_t cap;                  /* bytes of each buffer */
    double *X, *V, *E, *T;       /* kb x N x n each; E and T are free between builds */
} tr_group;
static tr_group g_trg = {0};

/* phase times of the transient removal (ms), kept only with the Verbose
 * option, which synchronizes the device around each phase */
enum { TRT_BASIS = 0, TRT_STEPS, TRT_MTILDE, TRT_GALERKIN, TRT_SETUP, TRT_N };
static double g_trt[TRT_N];
static PetscInt g_trt_kb;
static double trt_mark(void)
{
    if (!g_rsvdt.opt.verbose) return 0;
    rsvdt_hip_sync();
    return RSVDT_Time();
}
static void trt_add(int phase, double t0)
{
    if (!g_rsvdt.opt.verbose) return;
    rsvdt_hip_sync();
    g_trt[phase] += 1e3 * (RSVDT_Time() - t0);
}

static void tr_group_free(tr_group *g)
{
    if (g->X || g->V || g->E || g->T) rsvdt_hip_sync();
    if (g->X) rsvdt_hip_free(g->X);
    if (g->V) rsvdt_hip_free(g->V);
    if (g->E) rsvdt_hip_free(g->E);
    if (g->T) rsvdt_hip_free(g->T);
    memset(g, 0, sizeof(*g));
}

/* The four group buffers are allocated once and reused by every group of
 * CreateMtilde and GalerkinProjAllModes (allocating and freeing buffers of
 * several GB per group costs more than the group's basis). */
static PetscErrorCode tr_group_reserve(tr_group *g, size_t bytes)
{
    if (g->cap >= bytes && g->X && g->V && g->E && g->T) return 0;
    tr_group_free(g);
    if (rsvdt_hip_malloc((void **)&g->X, bytes) || rsvdt_hip_malloc((void **)&g->V, bytes) ||
        rsvdt_hip_malloc((void **)&g->E, bytes) || rsvdt_hip_malloc((void **)&g->T, bytes)) {
        tr_group_free(g);
        return rsvdt_err(RSVDT_ERR_MEM, "CreateMtilde", "cannot allocate the snapshot basis (%.2f GB each)", bytes / 1e9);
    }
    g->cap = bytes;
    return 0;
}

/* test vectors per group: the four group buffers, the stepping temporaries
 * and the Galerkin output rows of a test vector must fit in half of the
 * free memory (buffers already held by the group count as free); the
 * TRGroup option caps the group (then split into balanced groups) */
static PetscInt tr_group_size(PetscInt k, PetscInt N, PetscInt n)
{
    size_t fr = 0, tot = 0;
    rsvdt_hip_mem_info(&fr, &tot);
    const double per = 6.0 * (double)N * (double)n * 16.0;
    PetscInt kb = (PetscInt)(0.5 * ((double)fr + 4.0 * (double)g_trg.cap) / per);
    if (g_rsvdt.opt.tr_group > 0 && kb > g_rsvdt.opt.tr_group) kb = g_rsvdt.opt.tr_group;
    if (kb < 1) kb = 1;
    if (kb >= k) return k;
    if (g_rsvdt.opt.tr_group > 0) {
        const PetscInt ng = (k + kb - 1) / kb;
        return (k + ng - 1) / ng;           /* balanced groups */
    }
    /* several groups: a power of two, so that a group's columns fill whole
     * 64-byte pieces of the k-column rows it is cut from and written back to
     * (measured on a 2M-row operator with k = 32: groups of 7 took 0.9 s
     * longer per removal than groups of 4 or 8) */
    PetscInt p2 = 1;
    while (2 * p2 <= kb) p2 *= 2;
    return p2;
}

void rsvdt_action_release(void) { tr_group_free(&g_trg); }

/* Snapshots (relaid) and orthonormal basis of test vectors b0 .. b0+kb-1. */
static PetscErrorCode tr_group_build(Mat Yall, PetscInt k, PetscInt b0, PetscInt kb, tr_group *g)
{
    const PetscInt N = Yall->rows, n = Yall->cols / k;
    if (g->key == (const void *)Yall->d && g->b0 == b0 && g->kb == kb && g->N == N && g->n == n) return 0;
    g->key = NULL;
    if (Yall->order != 1 || Yall->kb != k) RSVDT_CHK(rsvdt_mat_to_layout(Yall, k, 1));
    const size_t bytes = (size_t)kb * (size_t)N * (size_t)n * 16;
    const double tr0 = trt_mark();
    RSVDT_CHK(tr_group_reserve(g, bytes));
    trt_add(TRT_SETUP, tr0);
    g->b0 = b0; g->kb = kb; g->N = N; g->n = n;
    PetscErrorCode ierr = 0;
    double t0 = trt_mark();
    int h = rsvdt_hip_transpose_rows(N, n, k, b0, kb, Yall->d, g->X, 0, g_rsvdt.stream);
    if (h) ierr = rsvdt_err(h, "CreateMtilde", "device call failed: %s", rsvdt_hip_last_error());
    /* The snapshots of one period are a periodic response of rank <= Nw_eff
     * plus a decaying transient, so X_ik is often numerically rank-deficient
     * (singular values down to rounding level); the basis spans it to
     * rounding accuracy with the numerically dependent directions left out
     * (zero columns). */
    if (!ierr) ierr = rsvdt_span_basis(kb, N, n, g->X, g->V, g->E, g->T, N * n, n, 1, bytes, 1e-14);
    trt_add(TRT_BASIS, t0);
    if (ierr) return ierr;
    g->key = Yall->d;
    return 0;
}

/* M_tilde of the test vectors b0 .. b0+g-1, whose basis the group holds
 * (SourceCode/CreateMtilde.c:46-51). */
static PetscErrorCode mtilde_group(LNS_vars *LNS, RSVDt_vars *V, TS_removal_matrices *TSR, PetscInt N, PetscInt Ns1,
                                   PetscInt b0, PetscInt g)
{
    /* e^{A dT} V: the group's basis vectors as Ns1 blocks of N x g (the
     * snapshot layout), stepped by the same block SpMM as the actions.
     * Measured on a 2M-row operator this is faster than stepping each test
     * vector's N x Ns1 basis as one wide block (4.8 s against 6.6 s per
     * removal): a block of g columns stays in the last-level cache while its
     * rows are gathered.  The basis work arrays E and T are free once the
     * basis is built. */
    double *Vblk = g_trg.E, *MV = g_trg.T;
    double t0 = trt_mark();
    int h = rsvdt_hip_transpose_rows(N, Ns1, g, 0, g, g_trg.V, Vblk, 1, g_rsvdt.stream);
    if (h) return rsvdt_err(h, "CreateMtilde", "device call failed: %s", rsvdt_hip_last_error());
    RSVDT_CHK(rsvdt_homogeneous_steps(LNS->A, V->TS.DirAdj ? 1 : 0, V->TS.dt, V->TS.ResRatio, Vblk, N, g, Ns1));
    trt_add(TRT_STEPS, t0);
    t0 = trt_mark();
    h = rsvdt_hip_transpose_rows(N, Ns1, g, 0, g, Vblk, MV, 0, g_rsvdt.stream);
    if (h) return rsvdt_err(h, "CreateMtilde", "device call failed: %s", rsvdt_hip_last_error());
    void *work = rsvdt_scratch(SCR_WORK, rsvdt_hip_bgram_work_bytes(g, N, Ns1, Ns1));
    if (!work) return rsvdt_err(RSVDT_ERR_MEM, "CreateMtilde", "cannot allocate work");
    h = rsvdt_hip_bgram(g, N, Ns1, Ns1, g_trg.V, N * Ns1, Ns1, 1, MV, N * Ns1, Ns1, 1,
                        TSR->M_tilde_all->d + 2 * (size_t)b0 * (size_t)Ns1 * (size_t)Ns1, work, g_rsvdt.stream);
    if (h) return rsvdt_err(h, "CreateMtilde", "device call failed: %s", rsvdt_hip_last_error());
    trt_add(TRT_MTILDE, t0);
    return 0;
}

static PetscErrorCode mtilde_alloc(Mat Yall, LNS_vars *LNS, RSVDt_vars *V, TS_removal_matrices *TSR)
{
    PetscInt k = V->RSVD.k, Ns1 = V->RSVD.Nw + 1;
    if (Yall->cols != Ns1 * k) return rsvdt_err(RSVDT_ERR_USER, "CreateMtilde", "Y_all must have (Nw+1)*k columns");
    RSVDT_CHK(rsvdt_aij_upload(LNS->A));
    RSVDT_MatDestroy(&TSR->M_tilde_all);
    return rsvdt_mat_alloc_dense(Ns1, Ns1 * k, Ns1, 0, &TSR->M_tilde_all);
}

/* SourceCode/CreateMtilde.c:41-54: M_tilde_ik = V_ik^H e^{A dT} V_ik for every
 * test vector, e^{A dT} being RR homogeneous RK4 steps (TSExpADeltaT.c). */
PetscErrorCode CreateMtilde(Mat Yall, LNS_vars *LNS, RSVDt_vars *V, TS_removal_matrices *TSR)
{
    RSVDT_CHK(mtilde_alloc(Yall, LNS, V, TSR));
    PetscInt k = V->RSVD.k, Ns1 = V->RSVD.Nw + 1, N = Yall->rows;
    const PetscInt kb = tr_group_size(k, N, Ns1);
    g_trt_kb = kb;
    PetscErrorCode ierr = 0;
    for (PetscInt b0 = 0; b0 < k && !ierr; b0 += kb) {
        const PetscInt g = kb < k - b0 ? kb : k - b0;
        ierr = tr_group_build(Yall, k, b0, g, &g_trg);
        if (!ierr) ierr = mtilde_group(LNS, V, TSR, N, Ns1, b0, g);
    }
    rsvdt_hip_sync();
    if (ierr) tr_group_free(&g_trg);
    return ierr;
}

/* SourceCode/TSExpADeltaT.c:25-37: U <- e^{A dT} U column by column (RR
 * homogeneous RK4 steps); here all columns at once.  U must hold whole N x k
 * blocks (columns a multiple of k) in the snapshot layout, or a single N x c
 * matrix (stepped as one block of c columns). */
PetscErrorCode TSExpADeltaT(LNS_vars *LNS, RSVDt_vars *V, Mat U)
{
    if (U->kind != RSVDT_MAT_DENSE) return rsvdt_err(RSVDT_ERR_USER, "TSExpADeltaT", "dense Mat required");
    PetscInt kk = U->kb, nb = U->cols / U->kb;
    if (U->kb == 1 && U->cols > 1) {              /* column-major: relayout to one row-major block */
        RSVDT_CHK(rsvdt_mat_to_layout(U, U->cols, 0));
        kk = U->cols; nb = 1;
    }
    return rsvdt_homogeneous_steps(LNS->A, V->TS.DirAdj ? 1 : 0, V->TS.dt, V->TS.ResRatio, U->d, U->rows, kk, nb);
}

/* SourceCode/QRDecomposition.c:9-27: orthonormalize the columns of V in place. */
PetscErrorCode QRDecomposition(Mat Vm)
{
    if (Vm->kind != RSVDT_MAT_DENSE) return rsvdt_err(RSVDT_ERR_USER, "QRDecomposition", "dense Mat required");
    if (Vm->kb != Vm->cols) RSVDT_CHK(rsvdt_mat_to_layout(Vm, Vm->cols, 0));
    const size_t bytes = (size_t)Vm->rows * (size_t)Vm->cols * 16;
    double *X = NULL, *E = NULL, *T = NULL;
    PetscErrorCode ierr = 0;
    if (rsvdt_hip_malloc((void **)&X, bytes) || rsvdt_hip_malloc((void **)&E, bytes) ||
        rsvdt_hip_malloc((void **)&T, bytes))
        ierr = rsvdt_err(RSVDT_ERR_MEM, "QRDecomposition", "cannot allocate work");
    if (!ierr) {
        int h = rsvdt_hip_memcpy_d2d(X, Vm->d, bytes, g_rsvdt.stream);
        if (h) ierr = rsvdt_err(h, "QRDecomposition", "device call failed: %s", rsvdt_hip_last_error());
    }
    if (!ierr)
        ierr = rsvdt_span_basis(1, Vm->rows, Vm->cols, X, Vm->d, E, T, 0, Vm->cols, 1, bytes, 1e-14);
    rsvdt_hip_sync();
    if (X) rsvdt_hip_free(X);
    if (E) rsvdt_hip_free(E);
    if (T) rsvdt_hip_free(T);
    return ierr;
}

/* SourceCode/GalerkinProjAllModes.c:28-42 with GalerkinProj.c:26-74, for all
 * test vectors and frequencies at once:
 *   y1 = RR*Y[:,W1]*dft_m, y2 = RR*Y[:,W2]*dft_m  (W1 = 0..Nw-1, W2 = 1..Nw,
 *   swapped for the adjoint), ph = e^{+-i w m dT},
 *   a  = (ph I - M_tilde)^{-1} V^H (ph y1 - y2),   Y_hat_m = y1 - V a.
 * With the window coefficients C1 (y1 = X C1) and Ct (ph y1 - y2 = X Ct),
 * per test vector: B = (V^H X) Ct, a = solve, Y_hat = X C1 - V a.
 * The state below is shared by the groups of one projection. */
typedef struct {
    PetscInt k, N, Ns1, Nwe, kb;
    double *dC, *dph, *Ybrm;      /* window coefficients (C1, Ct), shifts, one group's output rows */
    double *B, *W;
    int *info;
    Mat Yup;
} gal_ctx;

static void gal_free(gal_ctx *c)
{
    rsvdt_hip_sync();
    if (c->dC) rsvdt_hip_free(c->dC);
    if (c->dph) rsvdt_hip_free(c->dph);
    if (c->Ybrm) rsvdt_hip_free(c->Ybrm);
    c->dC = c->dph = c->Ybrm = NULL;
}

static PetscErrorCode gal_begin(gal_ctx *c, Mat Yall, RSVDt_vars *V, DFT_matrices *D, PetscInt kb)
{
    memset(c, 0, sizeof(*c));
    PetscInt k = V->RSVD.k, Nw = V->RSVD.Nw, Ns1 = Nw + 1, Nwe = V->RSVD.Nw_eff, N = Yall->rows;
    c->k = k; c->N = N; c->Ns1 = Ns1; c->Nwe = Nwe; c->kb = kb;
    int dir = V->TS.DirAdj ? 1 : 0;
    double dT = V->TS.dt * (double)V->TS.ResRatio;
    double RR = (double)V->TS.ResRatio;
    /* host: dft (Nw x Nwe), the window coefficients for every test vector of a
     * group (C1 then Ct, kb x Ns1 x Nwe each) and the shifts ph (k x Nwe) */
    double *dft = malloc(sizeof(double) * 2 * (size_t)(Nw * Nwe));
    double *C = calloc((size_t)(2 * kb * Ns1 * Nwe) * 2, sizeof(double));
    double *ph = malloc(sizeof(double) * 2 * (size_t)(k * Nwe));
    if (!dft || !C || !ph) { free(dft); free(C); free(ph); return rsvdt_err(RSVDT_ERR_MEM, "GalerkinProjAllModes", "out of memory"); }
    int h = rsvdt_hip_memcpy_d2h(dft, D->dft->d, sizeof(double) * 2 * (size_t)(Nw * Nwe));
    if (h) { free(dft); free(C); free(ph); return rsvdt_err(h, "GalerkinProjAllModes", "device call failed: %s", rsvdt_hip_last_error()); }
    const PetscInt w1 = dir ? 0 : 1, w2 = dir ? 1 : 0;
    for (PetscInt m = 0; m < Nwe; m++) {
        double ang = (dir ? 1.0 : -1.0) * V->RSVD.w * (double)m * dT;   /* GalerkinProj.c:49-50 */
        double pr = cos(ang), pi = sin(ang);
        for (PetscInt ik = 0; ik < k; ik++) { ph[2 * (ik * Nwe + m)] = pr; ph[2 * (ik * Nwe + m) + 1] = pi; }
        for (PetscInt s = 0; s < Ns1; s++) {
            double d1r = 0, d1i = 0, d2r = 0, d2i = 0;
            PetscInt j1 = s - w1, j2 = s - w2;
            if (j1 >= 0 && j1 < Nw) { d1r = RR * dft[2 * (j1 * Nwe + m)]; d1i = RR * dft[2 * (j1 * Nwe + m) + 1]; }
            if (j2 >= 0 && j2 < Nw) { d2r = RR * dft[2 * (j2 * Nwe + m)]; d2i = RR * dft[2 * (j2 * Nwe + m) + 1]; }
            for (PetscInt b = 0; b < kb; b++) {
                double *c1 = C + 2 * (size_t)((b * Ns1 + s) * Nwe + m);
                double *ct = C + 2 * (size_t)(((kb + b) * Ns1 + s) * Nwe + m);
                c1[0] = d1r; c1[1] = d1i;
                ct[0] = pr * d1r - pi * d1i - d2r;
                ct[1] = pr * d1i + pi * d1r - d2i;
            }
        }
    }
    free(dft);
    const size_t csz = sizeof(double) * 2 * (size_t)(2 * kb * Ns1 * Nwe);
    const size_t ybytes = (size_t)kb * (size_t)N * (size_t)Nwe * 16;
    PetscErrorCode ierr = 0;
    if (rsvdt_hip_malloc((void **)&c->dC, csz) || rsvdt_hip_malloc((void **)&c->dph, sizeof(double) * 2 * (size_t)(k * Nwe)) ||
        rsvdt_hip_malloc((void **)&c->Ybrm, ybytes))
        ierr = rsvdt_err(RSVDT_ERR_MEM, "GalerkinProjAllModes", "cannot allocate work");
    if (!ierr && (rsvdt_hip_memcpy_h2d(c->dC, C, csz) || rsvdt_hip_memcpy_h2d(c->dph, ph, sizeof(double) * 2 * (size_t)(k * Nwe))))
        ierr = rsvdt_err(RSVDT_HIP_ERR_RUNTIME, "GalerkinProjAllModes", "device call failed: %s", rsvdt_hip_last_error());
    free(C); free(ph);
    if (!ierr) ierr = rsvdt_mat_alloc_dense(N, Nwe * k, k, 1, &c->Yup);
    size_t nb_b = (size_t)kb * (size_t)Ns1 * (size_t)Nwe * 16, nb_w = (size_t)kb * (size_t)Ns1 * (size_t)Ns1 * 16;
    c->B = rsvdt_scratch(SCR_MISC, nb_b);
    c->W = rsvdt_scratch(SCR_MISC2, nb_w);
    c->info = rsvdt_scratch(SCR_GINFO, 256);
    if (!ierr && (!c->B || !c->W || !c->info)) ierr = rsvdt_err(RSVDT_ERR_MEM, "GalerkinProj", "cannot allocate work");
    if (!ierr && rsvdt_hip_zero(c->info, 256, g_rsvdt.stream))
        ierr = rsvdt_err(RSVDT_HIP_ERR_RUNTIME, "GalerkinProj", "%s", rsvdt_hip_last_error());
    if (ierr) { gal_free(c); RSVDT_MatDestroy(&c->Yup); }
    return ierr;
}

/* The projection for test vectors b0 .. b0+g-1, whose snapshots and basis
 * the group holds. */
static PetscErrorCode gal_group(gal_ctx *c, TS_removal_matrices *TSR, PetscInt b0, PetscInt g)
{
    const PetscInt N = c->N, Ns1 = c->Ns1, Nwe = c->Nwe, kb = c->kb;
    const int64_t xs = N * Ns1;
    rsvdt_stream_t q = g_rsvdt.stream;
    const double t0 = trt_mark();
    /* work sizes follow the group's own batch count */
    void *work = rsvdt_scratch(SCR_WORK, rsvdt_hip_bgram_work_bytes(g, N, Ns1, Ns1));
    void *lw = rsvdt_scratch(SCR_G, rsvdt_hip_bshift_solve_work_bytes(g, Ns1, Nwe));
    if (!work || !lw) return rsvdt_err(RSVDT_ERR_MEM, "GalerkinProj", "cannot allocate work");
    const double *C1 = c->dC, *Ct = c->dC + 2 * (size_t)(kb * Ns1 * Nwe);
#define GCHK(x) do { int _h = (x); if (_h) return rsvdt_err(_h, "GalerkinProj", "device call failed: %s", rsvdt_hip_last_error()); } while (0)
    /* W = V^H X;  B = W Ct = V^H (ph y1 - y2) */
    GCHK(rsvdt_hip_bgram(g, N, Ns1, Ns1, g_trg.V, xs, Ns1, 1, g_trg.X, xs, Ns1, 1, c->W, work, q));
    GCHK(rsvdt_hip_bapply(g, Ns1, Ns1, Nwe, c->W, Ns1 * Ns1, Ns1, 1, Ct, 0.0, NULL, 0, 0, 0, c->B, Ns1 * Nwe, Nwe, 1, q));
    /* a = (ph I - M_tilde)^{-1} b, in place in B */
    GCHK(rsvdt_hip_bshift_solve(g, Ns1, Nwe, TSR->M_tilde_all->d + 2 * (size_t)b0 * Ns1 * Ns1,
                                c->dph + 2 * (size_t)b0 * Nwe, c->B, c->B, lw, c->info, q));
    GCHK(rsvdt_hip_zscale((int64_t)g * Ns1 * Nwe, -1.0, c->B, c->B, q));
    /* Y_hat = X C1 - V a  (per-test-vector rows), then to frequency blocks */
    GCHK(rsvdt_hip_bapply(g, N, Ns1, Nwe, g_trg.X, xs, Ns1, 1, C1, 0.0, NULL, 0, 0, 0, c->Ybrm, N * Nwe, Nwe, 1, q));
    GCHK(rsvdt_hip_bapply(g, N, Ns1, Nwe, g_trg.V, xs, Ns1, 1, c->B, 1.0, c->Ybrm, N * Nwe, Nwe, 1, c->Ybrm, N * Nwe,
                          Nwe, 1, q));
    GCHK(rsvdt_hip_transpose_rows(N, Nwe, c->k, b0, g, c->Ybrm, c->Yup->d, 1, q));
#undef GCHK
    trt_add(TRT_GALERKIN, t0);
    return 0;
}

/* Ends the projection: on success Y_hat_up takes the result and M_tilde_all
 * and Y_all are destroyed (GalerkinProjAllModes.c:41-42). */
static PetscErrorCode gal_end(gal_ctx *c, PetscErrorCode ierr, Mat Yall, TS_removal_matrices *TSR)
{
    int hinfo = 0;
    if (!ierr && rsvdt_hip_memcpy_d2h(&hinfo, c->info, sizeof(int)))
        ierr = rsvdt_err(RSVDT_HIP_ERR_RUNTIME, "GalerkinProj", "%s", rsvdt_hip_last_error());
    if (hinfo) rsvdt_printf("[rsvdt] warning: %d singular Galerkin systems\n", hinfo);
    gal_free(c);
    tr_group_free(&g_trg);
    if (ierr) { RSVDT_MatDestroy(&c->Yup); return ierr; }
    RSVDT_MatDestroy(&TSR->M_tilde_all);      /* GalerkinProjAllModes.c:41 */
    TSR->Y_hat_up = c->Yup;                    /* k-major, as the reference leaves it */
    c->Yup = NULL;
    RSVDT_MatDestroy(&Yall);                   /* GalerkinProjAllModes.c:42 */
    return 0;
}

PetscErrorCode GalerkinProjAllModes(Mat Yall, LNS_vars *LNS, RSVDt_vars *V, TS_removal_matrices *TSR, DFT_matrices *D)
{
    (void)LNS;
    if (!TSR->M_tilde_all) return rsvdt_err(RSVDT_ERR_USER, "GalerkinProjAllModes", "M_tilde_all is not set");
    const PetscInt k = V->RSVD.k, Ns1 = V->RSVD.Nw + 1;
    const PetscInt kb = tr_group_size(k, Yall->rows, Ns1);
    gal_ctx c;
    PetscErrorCode ierr = gal_begin(&c, Yall, V, D, kb);
    if (ierr) { tr_group_free(&g_trg); return ierr; }
    for (PetscInt b0 = 0; b0 < k && !ierr; b0 += kb) {
        const PetscInt g = kb < k - b0 ? kb : k - b0;
        ierr = tr_group_build(Yall, k, b0, g, &g_trg);
        if (!ierr) ierr = gal_group(&c, TSR, b0, g);
    }
    return gal_end(&c, ierr, Yall, TSR);
}

/* SourceCode/TransientRemovalStrategy.c:30-40 */
PetscErrorCode TransientRemovalStrategy(Mat Yall, LNS_vars *LNS, RSVD_matrices *R, RSVDt_vars *V,
                                        TS_removal_matrices *TSR, DFT_matrices *D)
{
    double t1 = RSVDT_Time();
    if (V->Display > 0)
        rsvdt_printf("*** Trasnient removal strategy begins! ***\nThe process takes about as long as one period integration\n");
    memset(g_trt, 0, sizeof(g_trt));
    /* CreateMtilde and GalerkinProjAllModes group by group, so that each
     * group's snapshot basis is built once (the reference computes the same
     * QR in both, CreateMtilde.c:43-46 and GalerkinProjAllModes.c:31-33) */
    double ts = trt_mark();
    RSVDT_CHK(mtilde_alloc(Yall, LNS, V, TSR));
    const PetscInt k = V->RSVD.k, Ns1 = V->RSVD.Nw + 1, N = Yall->rows;
    const PetscInt kb = tr_group_size(k, N, Ns1);
    g_trt_kb = kb;
    gal_ctx c;
    PetscErrorCode ierr = gal_begin(&c, Yall, V, D, kb);
    trt_add(TRT_SETUP, ts);
    if (ierr) { tr_group_free(&g_trg); return ierr; }
    for (PetscInt b0 = 0; b0 < k && !ierr; b0 += kb) {
        const PetscInt g = kb < k - b0 ? kb : k - b0;
        ierr = tr_group_build(Yall, k, b0, g, &g_trg);
        if (!ierr) ierr = mtilde_group(LNS, V, TSR, N, Ns1, b0, g);
        if (!ierr) ierr = gal_group(&c, TSR, b0, g);
    }
    ts = trt_mark();
    ierr = gal_end(&c, ierr, Yall, TSR);
    trt_add(TRT_SETUP, ts);
    RSVDT_CHK(ierr);
    RSVDT_CHK(ReversePermuteMat(TSR->Y_hat_up, V));
    R->Y_hat = TSR->Y_hat_up;
    TSR->Y_hat_up = NULL;
    if (g_rsvdt.opt.verbose)
        rsvdt_printf("[rsvdt] transient removal: %lld test vectors per group; basis %.1f ms, e^{A dT} steps %.1f ms, "
                     "M_tilde %.1f ms, Galerkin %.1f ms, set-up %.1f ms (of %.1f ms)\n",
                     (long long)g_trt_kb,
                     g_trt[TRT_BASIS], g_trt[TRT_STEPS], g_trt[TRT_MTILDE], g_trt[TRT_GALERKIN], g_trt[TRT_SETUP],
                     1e3 * (RSVDT_Time() - t1));
    if (V->Display > 0) rsvdt_printf("Trasnient removal strategy elapsed time = %.3f s\n", RSVDT_Time() - t1);
    return 0;
}

/* ------------------------------------------------------------------ TSActionRK4 */
PetscErrorCode TSActionRK4(RSVD_matrices *R, DFT_matrices *D, LNS_vars *LNS, RSVDt_vars *V, Directories *dirs,
                           TS_removal_matrices *TSR)
{
    (void)dirs;
    double t0 = RSVDT_Time();
    if (V->Display > 0) rsvdt_printf(V->TS.DirAdj ? "\n*** Direct action begins! ***\n" : "\n*** Adjoint action begins! ***\n");
    if (!R->Y_hat) return rsvdt_err(RSVDT_ERR_USER, "TSActionRK4", "Y_hat is not set");
    PetscInt k = V->RSVD.k, N = V->RSVD.N, Nwe = V->RSVD.Nw_eff;
    if (R->Y_hat->rows != N || R->Y_hat->cols != k * Nwe)
        return rsvdt_err(RSVDT_ERR_USER, "TSActionRK4", "Y_hat must be %lld x %lld, got %lld x %lld", (long long)N,
                         (long long)(k * Nwe), (long long)R->Y_hat->rows, (long long)R->Y_hat->cols);
    memset(&g_rsvdt.stats, 0, sizeof(g_rsvdt.stats));
    /* :49-50 Y_hat is moved into F_hat; the frequency-major blocks are what
     * the forcing GEMM reads, so the permutation of :76 is not materialized */
    R->F_hat = R->Y_hat;
    R->Y_hat = NULL;
    if (R->F_hat->order != 0 || R->F_hat->kb != k) RSVDT_CHK(rsvdt_mat_to_layout(R->F_hat, k, 0));
    PetscInt Nstore = V->TS.TransientRemoval ? V->RSVD.Nw + 1 : V->RSVD.Nw;
    Mat Yall;
    RSVDT_CHK(rsvdt_mat_alloc_dense(N, Nstore * k, k, 1, &Yall));
    phase_timer trk = {0}, tfr = {0}, tdf = {0};
    PetscErrorCode ierr = time_stepping(LNS->A, D, V, R->F_hat, 0, Yall, 0, &trk, &tfr);
    RSVDT_MatDestroy(&R->F_hat);                              /* :139 */
    if (ierr) { RSVDT_MatDestroy(&Yall); return ierr; }
    pt_begin(&tdf);
    ierr = V->TS.TransientRemoval ? TransientRemovalStrategy(Yall, LNS, R, V, TSR, D) : DFT(Yall, D, R, V);
    pt_end(&tdf);
    pt_collect(&trk); pt_collect(&tfr); pt_collect(&tdf);
    g_rsvdt.stats.rk4_ms = trk.ms;
    g_rsvdt.stats.forcing_ms = tfr.ms;
    g_rsvdt.stats.dft_ms = tdf.ms;
    g_rsvdt.stats.tr_ms = V->TS.TransientRemoval ? tdf.ms : 0;
    pt_free(&trk); pt_free(&tfr); pt_free(&tdf);
    RSVDT_CHK(ierr);
    RSVDT_HIPCHK("TSActionRK4", rsvdt_hip_sync());
    g_rsvdt.stats.total_ms = 1e3 * (RSVDT_Time() - t0);
    if (V->Display > 0)
        rsvdt_printf("*** %s action elapsed time = %.3f s (RK4 %.1f ms, forcing %.1f ms, %s %.1f ms, %lld steps) ***\n",
                     V->TS.DirAdj ? "Direct" : "Adjoint", RSVDT_Time() - t0, trk.ms, tfr.ms,
                     V->TS.TransientRemoval ? "transient removal" : "DFT", tdf.ms, (long long)g_rsvdt.stats.steps);
    return 0;
}

/* ------------------------------------------------------------------ ApplyWeightMats */
static PetscErrorCode apply_sparse(Mat M, int herm, Mat *Y, PetscInt k)
{
    RSVDT_CHK(rsvdt_aij_upload(M));
    PetscInt rows = herm ? M->cols : M->rows, cols = herm ? M->rows : M->cols;
    if ((*Y)->rows != cols)
        return rsvdt_err(RSVDT_ERR_USER, "ApplyWeightMats", "size mismatch: matrix has %lld columns, Y_hat %lld rows",
                         (long long)cols, (long long)(*Y)->rows);
    if ((*Y)->order != 0 || (*Y)->kb != k) RSVDT_CHK(rsvdt_mat_to_layout(*Y, k, 0));
    PetscInt nb = (*Y)->cols / k;
    Mat out;
    RSVDT_CHK(rsvdt_mat_alloc_dense(rows, (*Y)->cols, k, 0, &out));
    /* the block SpMM takes complex FP64 values; the device copies of a real
     * matrix hold real values, so its complex expansion is built once (from
     * the host copy) and kept on the Mat */
    const double *vz = herm ? M->d_valH64 : M->d_val64;
    if (M->is_real) {
        double **pz = herm ? &M->d_valzH : &M->d_valz;
        if (!*pz) {
            size_t nnz = (size_t)M->nnz;
            double *h = malloc(sizeof(double) * 2 * (nnz ? nnz : 1));
            if (!h) return rsvdt_err(RSVDT_ERR_MEM, "ApplyWeightMats", "out of memory");
            if (!herm) {
                for (size_t e = 0; e < nnz; e++) { h[2 * e] = M->h_val[2 * e]; h[2 * e + 1] = 0; }
            } else {       /* A^H of a real matrix: its transpose, in the order of the uploaded A^H */
                double *hr = malloc(sizeof(double) * (nnz ? nnz : 1));
                if (!hr) { free(h); return rsvdt_err(RSVDT_ERR_MEM, "ApplyWeightMats", "out of memory"); }
                int rc = rsvdt_hip_memcpy_d2h(hr, M->d_valH64, sizeof(double) * nnz);
                if (rc) { free(h); free(hr); return rsvdt_err(rc, "ApplyWeightMats", "%s", rsvdt_hip_last_error()); }
                for (size_t e = 0; e < nnz; e++) { h[2 * e] = hr[e]; h[2 * e + 1] = 0; }
                free(hr);
            }
            int rc = rsvdt_hip_malloc((void **)pz, sizeof(double) * 2 * (nnz ? nnz : 1));
            if (!rc) rc = rsvdt_hip_memcpy_h2d(*pz, h, sizeof(double) * 2 * nnz);
            free(h);
            if (rc) return rsvdt_err(rc, "ApplyWeightMats", "%s", rsvdt_hip_last_error());
        }
        vz = *pz;
    }
    RSVDT_HIPCHK("ApplyWeightMats",
                 rsvdt_hip_spmm_blocks(rows, k, nb, herm ? M->d_rowptrH : M->d_rowptr, herm ? M->d_colH : M->d_col, vz,
                                       (*Y)->d, cols * k, out->d, rows * k, g_rsvdt.stream));
    RSVDT_MatDestroy(Y);
    *Y = out;
    return 0;
}

/* SourceCode/ApplyWeightMats.c:16-80 (the Hermitian transposes use the
 * conjugate-transposed device copies built at load instead of transposing
 * the matrices in place). */
PetscErrorCode ApplyWeightMats(RSVD_matrices *R, RSVDt_vars *V, Weight_matrices *W, PetscBool before)
{
    PetscInt k = V->RSVD.k;
    if (!W) return 0;
    if (V->TS.DirAdj) {
        if (before) {
            if (W->InvInputWeightFlg) RSVDT_CHK(apply_sparse(W->W_f_sqrt_inv, 0, &R->Y_hat, k));
            if (W->InputMatrixFlg) RSVDT_CHK(apply_sparse(W->B, 0, &R->Y_hat, k));
        } else {
            if (W->OutputMatrixFlg) RSVDT_CHK(apply_sparse(W->C, 0, &R->Y_hat, k));
            if (W->OutputWeightFlg) RSVDT_CHK(apply_sparse(W->W_q_sqrt, 0, &R->Y_hat, k));
        }
    } else {
        if (before) {
            if (W->OutputWeightFlg) RSVDT_CHK(apply_sparse(W->W_q_sqrt, 1, &R->Y_hat, k));
            if (W->OutputMatrixFlg) RSVDT_CHK(apply_sparse(W->C, 1, &R->Y_hat, k));
        } else {
            if (W->InputMatrixFlg) RSVDT_CHK(apply_sparse(W->B, 1, &R->Y_hat, k));
            if (W->InvInputWeightFlg) RSVDT_CHK(apply_sparse(W->W_f_sqrt_inv, 1, &R->Y_hat, k));
        }
    }
    return 0;
}

/* ------------------------------------------------------------------ actions */
/* SourceCode/DirectActionRK4.c:24-28 */
PetscErrorCode DirectActionRK4(RSVD_matrices *R, RSVDt_vars *V, LNS_vars *LNS, DFT_matrices *D, Weight_matrices *W,
                               Directories *dirs, TS_removal_matrices *TSR)
{
    V->TS.DirAdj = PETSC_TRUE;
    RSVDT_CHK(ApplyWeightMats(R, V, W, PETSC_TRUE));
    RSVDT_CHK(TSActionRK4(R, D, LNS, V, dirs, TSR));
    RSVDT_CHK(ApplyWeightMats(R, V, W, PETSC_FALSE));
    return 0;
}

/* SourceCode/AdjointActionRK4.c:24-28 */
PetscErrorCode AdjointActionRK4(RSVD_matrices *R, RSVDt_vars *V, LNS_vars *LNS, DFT_matrices *D, Weight_matrices *W,
                                Directories *dirs, TS_removal_matrices *TSR)
{
    V->TS.DirAdj = PETSC_FALSE;
    RSVDT_CHK(ApplyWeightMats(R, V, W, PETSC_TRUE));
    RSVDT_CHK(TSActionRK4(R, D, LNS, V, dirs, TSR));
    RSVDT_CHK(ApplyWeightMats(R, V, W, PETSC_FALSE));
    return 0;
}

/* SourceCode/PowerIterationRK4.c:20-29 */
PetscErrorCode PowerIterationRK4(RSVD_matrices *R, RSVDt_vars *V, LNS_vars *LNS, DFT_matrices *D, Weight_matrices *W,
                                 Directories *dirs, TS_removal_matrices *TSR)
{
    for (PetscInt iq = 0; iq < V->RSVD.q; iq++) {
        if (V->Display > 0) rsvdt_printf("\n******** Inside power iteration, %d/%d *******\n", (int)iq + 1, (int)V->RSVD.q);
        RSVDT_CHK(QRAllFreqs(R, V));
        RSVDT_CHK(AdjointActionRK4(R, V, LNS, D, W, dirs, TSR));
        RSVDT_CHK(QRAllFreqs(R, V));
        RSVDT_CHK(DirectActionRK4(R, V, LNS, D, W, dirs, TSR));
    }
    if (V->RSVD.q > 0 && V->Display > 0) rsvdt_printf("\n******** Power iteration DONE! *************\n");
    return 0;
}

/* ------------------------------------------------------------------ bench */
PetscErrorCode RSVDT_BenchSteps(RSVD_matrices *R, RSVDt_vars *V, LNS_vars *LNS, DFT_matrices *D, PetscInt nsteps,
                                PetscInt warmup, double *ms_out, double *rk4_ms_out, double *forcing_ms_out)
{
    if (!R->Y_hat) return rsvdt_err(RSVDT_ERR_USER, "RSVDT_BenchSteps", "Y_hat is not set");
    Mat F = R->Y_hat;
    PetscInt k = V->RSVD.k;
    if (F->order != 0 || F->kb != k) RSVDT_CHK(rsvdt_mat_to_layout(F, k, 0));
    /* the folded forcing terms are computed once per action; here once,
     * outside the timed steps, into a copy so that Y_hat stays the sketch */
    Mat Fin = F;
    int folded = 0;
    if (!g_rsvdt.opt.rk4_form) {
        RSVDT_CHK(rsvdt_aij_upload(LNS->A));
        RSVDT_CHK(rsvdt_mat_alloc_dense(F->rows, F->cols, F->kb, F->order, &Fin));
        PetscErrorCode e = folded_forcing(LNS->A, V->TS.DirAdj ? 1 : 0, V->TS.dt, k, V->RSVD.Nw_eff, V->TS.Ns,
                                          D->idft->d, F->d, Fin->d);
        if (e) { RSVDT_MatDestroy(&Fin); return e; }
        folded = 1;
    }
    PetscErrorCode ierr = 0;
    if (warmup > 0) ierr = time_stepping(LNS->A, D, V, Fin, folded, NULL, warmup, NULL, NULL);
    if (!ierr) {
        int hs = rsvdt_hip_sync();
        if (hs) ierr = rsvdt_err(hs, "RSVDT_BenchSteps", "device call failed: %s", rsvdt_hip_last_error());
    }
    phase_timer trk = {0}, tfr = {0}, tall = {0};
    if (!ierr) {
        pt_begin(&tall);
        ierr = time_stepping(LNS->A, D, V, Fin, folded, NULL, nsteps, &trk, &tfr);
        pt_end(&tall);
        rsvdt_hip_sync();
        pt_collect(&trk); pt_collect(&tfr); pt_collect(&tall);
        if (ms_out) *ms_out = tall.ms;
        if (rk4_ms_out) *rk4_ms_out = trk.ms;
        if (forcing_ms_out) *forcing_ms_out = tfr.ms;
    }
    pt_free(&trk); pt_free(&tfr); pt_free(&tall);
    if (Fin != F) RSVDT_MatDestroy(&Fin);
    return ierr;
}
