/*
 * Output entry points: SourceCode/SaveResults.c:32-92 and SaveInputVarsCopy.c:26-38.
 */
#define _POSIX_C_SOURCE 200809L
#include "rsvdt_internal.h"

#include <stdlib.h>
#include <string.h>
#include <unistd.h>

static void put_be64o(unsigned char *p, uint64_t v)
{
    for (int i = 7; i >= 0; i--) { p[i] = (unsigned char)(v & 0xff); v >>= 8; }
}

/* Host buffers of a SaveResults call: two pinned staging buffers that the
 * device fills with row slices of the values, and one buffer that holds a
 * slice in file byte order. */
typedef struct { unsigned char *pin[2], *out; size_t cap; } wbufs;

static void wbufs_free(wbufs *w)
{
    rsvdt_hip_host_free(w->pin[0]);
    rsvdt_hip_host_free(w->pin[1]);
    free(w->out);
    memset(w, 0, sizeof(*w));
}

static PetscErrorCode wbufs_get(wbufs *w, size_t bytes)
{
    if (bytes <= w->cap) return 0;
    wbufs_free(w);
    if (rsvdt_hip_host_alloc((void **)&w->pin[0], bytes) || rsvdt_hip_host_alloc((void **)&w->pin[1], bytes) ||
        !(w->out = malloc(bytes))) {
        wbufs_free(w);
        return rsvdt_err(RSVDT_ERR_MEM, "SaveResults", "cannot allocate %zu-byte staging buffers", bytes);
    }
    w->cap = bytes;
    return 0;
}

/* n 64-bit words to big-endian */
static void to_be64(unsigned char *dst, const unsigned char *src, size_t n)
{
    const uint64_t *s = (const uint64_t *)src;
    uint64_t *d = (uint64_t *)dst;
    for (size_t i = 0; i < n; i++) d[i] = __builtin_bswap64(s[i]);
}

/* MatView of columns c0 .. c0+nc-1 of a dense Mat (PETSc's dense binary
 * layout: the AIJ one with every entry stored), streamed.  The values go to
 * the host in slices of rows (about 64 MB): while slice s is converted to
 * file byte order and written, slice s+1 is already in flight to the other
 * pinned buffer.  When the columns are one whole block of the Mat (a
 * frequency's k modes in the frequency-major layout) they are read in
 * place; otherwise they are first gathered into one rows x nc row-major
 * device buffer.  A result file costs its own size in transfers. */
static PetscErrorCode view_cols(Mat M, PetscInt c0, PetscInt nc, const char *path, wbufs *wb)
{
    const PetscInt rows = M->rows, kb = M->kb, nb = M->cols / M->kb;
    if (nc <= 0 || c0 < 0 || c0 + nc > M->cols) return rsvdt_err(RSVDT_ERR_USER, "SaveResults", "bad column range");
    rsvdt_stream_t q = g_rsvdt.stream;
    const double *src = NULL;
    double *T = NULL;
    PetscErrorCode ierr = 0;
    if (M->order == 0 && kb == nc && c0 % kb == 0) {
        src = M->d + 2 * (size_t)(c0 / kb) * (size_t)rows * (size_t)kb;
    } else {
        if (rsvdt_hip_malloc((void **)&T, (size_t)(rows * nc > 0 ? rows * nc : 1) * 16))
            return rsvdt_err(RSVDT_ERR_MEM, "SaveResults", "cannot allocate the column buffer: %s",
                             rsvdt_hip_last_error());
        for (PetscInt ci = 0; ci < nc && !ierr; ci++) {
            const PetscInt c = c0 + ci;
            const PetscInt b = M->order ? c % nb : c / kb, j = M->order ? c / nb : c % kb;
            int h = rsvdt_hip_zcopy3(1, rows, 1, M->d + 2 * ((size_t)b * (size_t)rows * (size_t)kb + (size_t)j), 0,
                                     kb, 0, T + 2 * (size_t)ci, 0, nc, 0, q);
            if (h) ierr = rsvdt_err(h, "SaveResults", "%s", rsvdt_hip_last_error());
        }
        src = T;
    }
    const PetscInt slice = (PetscInt)((64u << 20) / (16 * (size_t)nc)) + 1;   /* rows per slice: ~64 MB of values */
    const PetscInt srows = slice < rows ? slice : rows;
    if (!ierr) ierr = wbufs_get(wb, (size_t)srows * (size_t)nc * 16 + 64);
    FILE *f = ierr ? NULL : fopen(path, "wb");
    if (!ierr && !f) ierr = rsvdt_err(RSVDT_ERR_FILE, "SaveResults", "cannot create '%s'", path);
    if (!ierr) {
        unsigned char *buf = wb->out;
        int wr = 1;
        int64_t hdr[4] = {1211216, rows, nc, rows * nc};
        for (int i = 0; i < 4; i++) put_be64o(buf + 8 * i, (uint64_t)hdr[i]);
        wr &= fwrite(buf, 8, 4, f) == 4;
        for (PetscInt r = 0; r < srows; r++) put_be64o(buf + 8 * r, (uint64_t)nc);   /* row lengths */
        for (PetscInt r0 = 0; r0 < rows; r0 += slice) {
            const size_t nr = (size_t)(slice < rows - r0 ? slice : rows - r0);
            wr &= fwrite(buf, 8, nr, f) == nr;
        }
        for (PetscInt r = 0; r < srows; r++)                 /* column indices, row by row */
            for (PetscInt c = 0; c < nc; c++) put_be64o(buf + 8 * ((size_t)r * nc + c), (uint64_t)c);
        for (PetscInt r0 = 0; r0 < rows; r0 += slice) {
            const size_t nr = (size_t)(slice < rows - r0 ? slice : rows - r0);
            wr &= fwrite(buf, 8, nr * (size_t)nc, f) == nr * (size_t)nc;
        }
        /* values, row-major, double-buffered */
        const PetscInt nsl = (rows + slice - 1) / slice;
        int h = rsvdt_hip_memcpy_d2h_async(wb->pin[0], src, (size_t)srows * (size_t)nc * 16, q);
        for (PetscInt si = 0; si < nsl && !h && wr; si++) {
            const PetscInt r0 = si * slice;
            const size_t nr = (size_t)(slice < rows - r0 ? slice : rows - r0);
            h = rsvdt_hip_stream_sync(q);
            if (!h && si + 1 < nsl) {
                const PetscInt r1 = r0 + slice;
                const size_t nr1 = (size_t)(slice < rows - r1 ? slice : rows - r1);
                h = rsvdt_hip_memcpy_d2h_async(wb->pin[(si + 1) & 1], src + 2 * (size_t)r1 * (size_t)nc,
                                               nr1 * (size_t)nc * 16, q);
            }
            if (h) break;
            const size_t n = nr * (size_t)nc * 2;
            to_be64(buf, wb->pin[si & 1], n);
            wr &= fwrite(buf, 8, n, f) == n;
        }
        if (h) ierr = rsvdt_err(h, "SaveResults", "%s", rsvdt_hip_last_error());
        else if (!wr) ierr = rsvdt_err(RSVDT_ERR_FILE, "SaveResults", "write error on '%s'", path);
        rsvdt_hip_stream_sync(q);     /* no copy may still target a staging buffer */
    }
    if (f && fclose(f) != 0 && !ierr) ierr = rsvdt_err(RSVDT_ERR_FILE, "SaveResults", "write error on '%s'", path);
    rsvdt_hip_free(T);
    return ierr;
}

/* SourceCode/SaveResults.c:32-92: S_hat, then either k files of N x Nw_eff
 * (opt 1, 1-based U_hat_k<j>_allFreqs) or Nw_eff files of N x k (opt 2,
 * 0-based U_hat_Freq<m>_allK), and a copy of the inputs file. */
PetscErrorCode SaveResults(Resolvent_matrices *Res, RSVDt_vars *V, Directories *dirs)
{
    rsvdt_printf("\n********************************************\n************ Saving the output *************\n"
                 "********************************************\n");
    rsvdt_printf("\nThe results directory: %s\n\n", dirs->FolderDir);
    rsvdt_printf("One matrix of size %d (modes) x %d (frequencies) for gains\n", (int)V->RSVD.k, (int)V->RSVD.Nw_eff);
    snprintf(dirs->IO_dir, PETSC_MAX_PATH_LEN, "%s%s", dirs->FolderDir, "S_hat");
    RSVDT_CHK(RSVDT_MatView(Res->S_hat, dirs->IO_dir));
    PetscInt k = V->RSVD.k, Nwe = V->RSVD.Nw_eff;
    wbufs wb = {{NULL, NULL}, NULL, 0};
    PetscErrorCode ierr = 0;
    if (V->SaveResultsOpt == 1) {
        rsvdt_printf("%d matrices of size %d x %d for response modes\n%d matrices of size %d x %d for forcing modes\n\n",
                     (int)k, (int)Res->U_hat->rows, (int)Nwe, (int)k, (int)Res->V_hat->rows, (int)Nwe);
        RSVDT_CHK(PermuteMat(Res->U_hat, V));
        RSVDT_CHK(PermuteMat(Res->V_hat, V));
        for (PetscInt ik = 0; ik < k && !ierr; ik++) {
            snprintf(dirs->IO_dir, PETSC_MAX_PATH_LEN, "%s%s%d%s", dirs->FolderDir, "U_hat_k", (int)ik + 1, "_allFreqs");
            ierr = view_cols(Res->U_hat, ik * Nwe, Nwe, dirs->IO_dir, &wb);
            snprintf(dirs->IO_dir, PETSC_MAX_PATH_LEN, "%s%s%d%s", dirs->FolderDir, "V_hat_k", (int)ik + 1, "_allFreqs");
            if (!ierr) ierr = view_cols(Res->V_hat, ik * Nwe, Nwe, dirs->IO_dir, &wb);
        }
    } else {
        rsvdt_printf("%d matrices of size %d x %d for response modes\n%d matrices of size %d x %d for forcing modes\n\n",
                     (int)Nwe, (int)Res->U_hat->rows, (int)k, (int)Nwe, (int)Res->V_hat->rows, (int)k);
        for (PetscInt iw = 0; iw < Nwe && !ierr; iw++) {
            snprintf(dirs->IO_dir, PETSC_MAX_PATH_LEN, "%s%s%d%s", dirs->FolderDir, "U_hat_Freq", (int)iw, "_allK");
            ierr = view_cols(Res->U_hat, iw * k, k, dirs->IO_dir, &wb);
            snprintf(dirs->IO_dir, PETSC_MAX_PATH_LEN, "%s%s%d%s", dirs->FolderDir, "V_hat_Freq", (int)iw, "_allK");
            if (!ierr) ierr = view_cols(Res->V_hat, iw * k, k, dirs->IO_dir, &wb);
        }
    }
    wbufs_free(&wb);
    RSVDT_CHK(ierr);
    RSVDT_CHK(SaveInputVarsCopy(dirs));
    rsvdt_printf("DONE :))\n\n");
    return 0;
}

/* SourceCode/SaveInputVarsCopy.c:26-38: copy the -inputs file into the results folder. */
PetscErrorCode SaveInputVarsCopy(Directories *dirs)
{
    char fname[PETSC_MAX_PATH_LEN];
    if (!rsvdt_opt_get_string("inputs", fname, sizeof(fname)) || !fname[0]) return 0;
    const char *base = strrchr(fname, '/');
    base = base ? base + 1 : fname;
    char dst[2 * PETSC_MAX_PATH_LEN];
    snprintf(dst, sizeof(dst), "%s%s", dirs->FolderDir, base);
    FILE *in = fopen(fname, "rb");
    if (!in) return rsvdt_err(RSVDT_ERR_FILE, "SaveInputVarsCopy", "cannot open '%s'", fname);
    FILE *out = fopen(dst, "wb");
    if (!out) { fclose(in); return rsvdt_err(RSVDT_ERR_FILE, "SaveInputVarsCopy", "cannot create '%s'", dst); }
    char buf[1024];
    size_t n;
    while ((n = fread(buf, 1, sizeof(buf), in)) > 0) fwrite(buf, 1, n, out);
    fclose(in);
    fclose(out);
    return 0;
}
