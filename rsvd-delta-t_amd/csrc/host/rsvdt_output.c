/*
 * Output entry points: SourceCode/SaveResults.c:32-92 and SaveInputVarsCopy.c:26-38.
 */
#define _POSIX_C_SOURCE 200809L
#include "rsvdt_internal.h"

#include <stdlib.h>
#include <string.h>
#include <unistd.h>

static PetscErrorCode view_submat(Mat M, PetscInt c0, PetscInt nc, const char *path)
{
    /* columns c0 .. c0+nc-1 of a dense Mat, through the host */
    PetscInt rows = M->rows;
    double *all = malloc(sizeof(double) * 2 * (size_t)(rows * M->cols));
    if (!all) return rsvdt_err(RSVDT_ERR_MEM, "SaveResults", "out of memory");
    PetscErrorCode ierr = RSVDT_MatGetDense(M, all);
    Mat S = NULL;
    if (!ierr) ierr = RSVDT_MatCreateDense(rows, nc, all + 2 * (size_t)(c0 * rows), &S);
    free(all);
    if (!ierr) ierr = RSVDT_MatView(S, path);
    RSVDT_MatDestroy(&S);
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
            RSVDT_CHK(view_submat(Res->U_hat, ik * Nwe, Nwe, dirs->IO_dir));
            snprintf(dirs->IO_dir, PETSC_MAX_PATH_LEN, "%s%s%d%s", dirs->FolderDir, "V_hat_k", (int)ik + 1, "_allFreqs");
            RSVDT_CHK(view_submat(Res->V_hat, ik * Nwe, Nwe, dirs->IO_dir));
        }
    } else {
        rsvdt_printf("%d matrices of size %d x %d for response modes\n%d matrices of size %d x %d for forcing modes\n\n",
                     (int)Nwe, (int)Res->U_hat->rows, (int)k, (int)Nwe, (int)Res->V_hat->rows, (int)k);
        for (PetscInt iw = 0; iw < Nwe; iw++) {
            snprintf(dirs->IO_dir, PETSC_MAX_PATH_LEN, "%s%s%d%s", dirs->FolderDir, "U_hat_Freq", (int)iw, "_allK");
            RSVDT_CHK(view_submat(Res->U_hat, iw * k, k, dirs->IO_dir));
            snprintf(dirs->IO_dir, PETSC_MAX_PATH_LEN, "%s%s%d%s", dirs->FolderDir, "V_hat_Freq", (int)iw, "_allK");
            RSVDT_CHK(view_submat(Res->V_hat, iw * k, k, dirs->IO_dir));
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
