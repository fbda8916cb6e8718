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

/* MatView of columns c0 .. c0+nc-1 of a dense Mat (PETSc's dense binary
 * layout: the AIJ one with every entry stored), streamed: the columns are
 * gathered on the device into one rows x nc row-major buffer, which goes to
 * the host and the file in slices of rows, so a result file costs its own
 * size in transfers and a slice of host memory. */
static PetscErrorCode view_cols(Mat M, PetscInt c0, PetscInt nc, const char *path)
{
    const PetscInt rows = M->rows, kb = M->kb, nb = M->cols / M->kb;
    if (nc <= 0 || c0 < 0 || c0 + nc > M->cols) return rsvdt_err(RSVDT_ERR_USER, "SaveResults", "bad column range");
    double *T = NULL;
    if (rsvdt_hip_malloc((void **)&T, (size_t)(rows * nc > 0 ? rows * nc : 1) * 16))
        return rsvdt_err(RSVDT_ERR_MEM, "SaveResults", "cannot allocate the column buffer: %s", rsvdt_hip_last_error());
    PetscErrorCode ierr = 0;
    for (PetscInt ci = 0; ci < nc && !ierr; ci++) {
        const PetscInt c = c0 + ci;
        const PetscInt b = M->order ? c % nb : c / kb, j = M->order ? c / nb : c % kb;
        int h = rsvdt_hip_zcopy3(1, rows, 1, M->d + 2 * ((size_t)b * (size_t)rows * (size_t)kb + (size_t)j), 0, kb, 0,
                                 T + 2 * (size_t)ci, 0, nc, 0, g_rsvdt.stream);
        if (h) ierr = rsvdt_err(h, "SaveResults", "%s", rsvdt_hip_last_error());
    }
    if (!ierr && rsvdt_hip_sync()) ierr = rsvdt_err(RSVDT_HIP_ERR_RUNTIME, "SaveResults", "%s", rsvdt_hip_last_error());
    FILE *f = ierr ? NULL : fopen(path, "wb");
    if (!ierr && !f) ierr = rsvdt_err(RSVDT_ERR_FILE, "SaveResults", "cannot create '%s'", path);
    const PetscInt slice = (PetscInt)((64u << 20) / (16 * (size_t)nc)) + 1;   /* rows per slice: ~64 MB of values */
    const size_t bufb = (size_t)(slice < rows ? slice : rows) * (size_t)nc * 16 + 64;
    unsigned char *buf = ierr ? NULL : malloc(bufb);
    double *hv = ierr ? NULL : malloc(bufb);
    if (!ierr && (!buf || !hv)) ierr = rsvdt_err(RSVDT_ERR_MEM, "SaveResults", "out of memory");
    if (!ierr) {
        int64_t hdr[4] = {1211216, rows, nc, rows * nc};
        for (int i = 0; i < 4; i++) put_be64o(buf + 8 * i, (uint64_t)hdr[i]);
        fwrite(buf, 8, 4, f);
        for (PetscInt r0 = 0; r0 < rows; r0 += slice) {            /* row lengths */
            const PetscInt nr = slice < rows - r0 ? slice : rows - r0;
            for (PetscInt r = 0; r < nr; r++) put_be64o(buf + 8 * r, (uint64_t)nc);
            fwrite(buf, 8, (size_t)nr, f);
        }
        for (PetscInt c = 0; c < nc; c++) put_be64o(buf + 8 * c, (uint64_t)c);   /* column indices, row by row */
        for (PetscInt r = 0; r < rows; r++) fwrite(buf, 8, (size_t)nc, f);
        for (PetscInt r0 = 0; r0 < rows && !ierr; r0 += slice) {   /* values, row-major */
            const PetscInt nr = slice < rows - r0 ? slice : rows - r0;
            const size_t n = (size_t)nr * (size_t)nc * 2;
            if (rsvdt_hip_memcpy_d2h(hv, T + 2 * (size_t)r0 * (size_t)nc, n * 8)) {
                ierr = rsvdt_err(RSVDT_HIP_ERR_RUNTIME, "SaveResults", "%s", rsvdt_hip_last_error());
                break;
            }
            for (size_t i = 0; i < n; i++) {
                uint64_t u;
                memcpy(&u, &hv[i], 8);
                put_be64o(buf + 8 * i, u);
            }
            fwrite(buf, 8, n, f);
        }
    }
    if (f && fclose(f) != 0 && !ierr) ierr = rsvdt_err(RSVDT_ERR_FILE, "SaveResults", "write error on '%s'", path);
    free(buf);
    free(hv);
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
    if (V->SaveResultsOpt == 1) {
        rsvdt_printf("%d matrices of size %d x %d for response modes\n%d matrices of size %d x %d for forcing modes\n\n",
                     (int)k, (int)Res->U_hat->rows, (int)Nwe, (int)k, (int)Res->V_hat->rows, (int)Nwe);
        RSVDT_CHK(PermuteMat(Res->U_hat, V));
        RSVDT_CHK(PermuteMat(Res->V_hat, V));
        for (PetscInt ik = 0; ik < k; ik++) {
            snprintf(dirs->IO_dir, PETSC_MAX_PATH_LEN, "%s%s%d%s", dirs->FolderDir, "U_hat_k", (int)ik + 1, "_allFreqs");
            RSVDT_CHK(view_cols(Res->U_hat, ik * Nwe, Nwe, dirs->IO_dir));
            snprintf(dirs->IO_dir, PETSC_MAX_PATH_LEN, "%s%s%d%s", dirs->FolderDir, "V_hat_k", (int)ik + 1, "_allFreqs");
            RSVDT_CHK(view_cols(Res->V_hat, ik * Nwe, Nwe, dirs->IO_dir));
        }
    } else {
        rsvdt_printf("%d matrices of size %d x %d for response modes\n%d matrices of size %d x %d for forcing modes\n\n",
                     (int)Nwe, (int)Res->U_hat->rows, (int)k, (int)Nwe, (int)Res->V_hat->rows, (int)k);
        for (PetscInt iw = 0; iw < Nwe; iw++) {
            snprintf(dirs->IO_dir, PETSC_MAX_PATH_LEN, "%s%s%d%s", dirs->FolderDir, "U_hat_Freq", (int)iw, "_allK");
            RSVDT_CHK(view_cols(Res->U_hat, iw * k, k, dirs->IO_dir));
            snprintf(dirs->IO_dir, PETSC_MAX_PATH_LEN, "%s%s%d%s", dirs->FolderDir, "V_hat_Freq", (int)iw, "_allK");
            RSVDT_CHK(view_cols(Res->V_hat, iw * k, k, dirs->IO_dir));
        }
    }
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
