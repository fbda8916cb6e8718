/*
 * Per-frequency linear algebra entry points: QR of every N x k frequency
 * block (SourceCode/QRAllFreqs.c), the SVDs before and after the final
 * adjoint action (SVDAllFreqsBeforeAdjoint.c, SVDAllFreqs.c) and StoreU.
 * All Nw_eff frequencies are handled by one batched launch per step.
 */
#define _POSIX_C_SOURCE 200809L
#include "rsvdt_internal.h"

#include <stdlib.h>
#include <string.h>

static PetscErrorCode freq_blocks(Mat Y, PetscInt k, const char *fn)
{
    if (!Y || Y->kind != RSVDT_MAT_DENSE) return rsvdt_err(RSVDT_ERR_USER, fn, "Y_hat is not set");
    if (Y->cols % k) return rsvdt_err(RSVDT_ERR_USER, fn, "Y_hat columns not a multiple of k");
    if (Y->order != 0 || Y->kb != k) RSVDT_CHK(rsvdt_mat_to_layout(Y, k, 0));
    return 0;
}

/* SourceCode/QRAllFreqs.c:22-32: Y_hat_m <- Q_m for every frequency. */
PetscErrorCode QRAllFreqs(RSVD_matrices *R, RSVDt_vars *V)
{
    double t1 = RSVDT_Time();
    PetscInt k = V->RSVD.k;
    if (V->Display > 0) rsvdt_printf("\n*** QR decomposition begins! ***\n");
    RSVDT_CHK(freq_blocks(R->Y_hat, k, "QRAllFreqs"));
    Mat Y = R->Y_hat;
    PetscInt nb = Y->cols / k, N = Y->rows;
    RSVDT_CHK(rsvdt_cholqr(nb, N, k, Y->d, N * k, k, 1, NULL, 0.0));
    RSVDT_HIPCHK("QRAllFreqs", rsvdt_hip_sync());
    if (V->Display > 0) rsvdt_printf("*** QR decomposition elapsed time = %.3f s ***\n", RSVDT_Time() - t1);
    return 0;
}

/* SourceCode/SVDAllFreqsBeforeAdjoint.c:31-53: Y_hat_m <- its k left singular
 * vectors, in descending order of singular value. */
PetscErrorCode SVDAllFreqsBeforeAdjoint(RSVD_matrices *R, RSVDt_vars *V)
{
    double t1 = RSVDT_Time();
    PetscInt k = V->RSVD.k;
    if (V->Display > 0) rsvdt_printf("\n*** Reduced SVD begins! ***\n");
    RSVDT_CHK(freq_blocks(R->Y_hat, k, "SVDAllFreqsBeforeAdjoint"));
    Mat Y = R->Y_hat;
    PetscInt nb = Y->cols / k, N = Y->rows;
    double *sig;
    RSVDT_HIPCHK("SVDAllFreqsBeforeAdjoint", rsvdt_hip_malloc((void **)&sig, sizeof(double) * (size_t)(nb * k)));
    PetscErrorCode ierr = rsvdt_svd_left(nb, N, k, Y->d, N * k, k, 1, sig);
    rsvdt_hip_sync();
    rsvdt_hip_free(sig);
    RSVDT_CHK(ierr);
    if (V->Display > 0) rsvdt_printf("*** Reduced SVD elapsed time = %.3f s ***\n", RSVDT_Time() - t1);
    return 0;
}

/* SourceCode/StoreU.c:15 */
PetscErrorCode StoreU(RSVD_matrices *R, Resolvent_matrices *Res)
{
    if (!R->Y_hat) return rsvdt_err(RSVDT_ERR_USER, "StoreU", "Y_hat is not set");
    return rsvdt_mat_dup(R->Y_hat, &Res->U_hat);
}

/* SourceCode/SVDAllFreqs.c:34-64: V_hat_m <- left singular vectors of Y_hat_m,
 * S_hat(:, m) <- singular values, Y_hat destroyed, then the inverse weights
 * applied to U_hat and V_hat. */
PetscErrorCode SVDAllFreqs(RSVD_matrices *R, RSVDt_vars *V, Weight_matrices *W, Resolvent_matrices *Res)
{
    double t1 = RSVDT_Time();
    PetscInt k = V->RSVD.k, Nwe = V->RSVD.Nw_eff;
    if (V->Display > 0) rsvdt_printf("\n*** Reduced SVD begins! ***\n");
    RSVDT_CHK(freq_blocks(R->Y_hat, k, "SVDAllFreqs"));
    Mat Y = R->Y_hat;
    PetscInt nb = Y->cols / k, N = Y->rows;
    if (nb != Nwe) return rsvdt_err(RSVDT_ERR_USER, "SVDAllFreqs", "Y_hat has %lld frequency blocks, expected %lld",
                                    (long long)nb, (long long)Nwe);
    double *sig;
    RSVDT_HIPCHK("SVDAllFreqs", rsvdt_hip_malloc((void **)&sig, sizeof(double) * (size_t)(nb * k)));
    PetscErrorCode ierr = rsvdt_svd_left(nb, N, k, Y->d, N * k, k, 1, sig);
    if (ierr) { rsvdt_hip_free(sig); return ierr; }
    double *hs = malloc(sizeof(double) * (size_t)(nb * k));
    double *cm = calloc((size_t)(nb * k) * 2, sizeof(double));
    if (!hs || !cm) return rsvdt_err(RSVDT_ERR_MEM, "SVDAllFreqs", "out of memory");
    RSVDT_HIPCHK("SVDAllFreqs", rsvdt_hip_memcpy_d2h(hs, sig, sizeof(double) * (size_t)(nb * k)));
    rsvdt_hip_free(sig);
    /* S_hat is k x Nw_eff (column m = frequency m), real values stored complex */
    for (PetscInt m = 0; m < nb; m++)
        for (PetscInt j = 0; j < k; j++) cm[2 * (m * k + j)] = hs[m * k + j];
    ierr = RSVDT_MatCreateDense(k, nb, cm, &Res->S_hat);
    free(hs); free(cm);
    RSVDT_CHK(ierr);
    /* the left singular vectors were formed in place: Y_hat's storage becomes V_hat (:58 destroys Y_hat) */
    Res->V_hat = Y;
    R->Y_hat = NULL;
    if (W && W->InvOutputWeightFlg) {
        RSVD_matrices t = {NULL, Res->U_hat};
        RSVDt_vars v2 = *V;
        v2.TS.DirAdj = PETSC_TRUE;
        Weight_matrices w2 = {0};
        w2.InvInputWeightFlg = PETSC_TRUE;
        w2.W_f_sqrt_inv = W->W_q_sqrt_inv;
        RSVDT_CHK(ApplyWeightMats(&t, &v2, &w2, PETSC_TRUE));
        Res->U_hat = t.Y_hat;
    }
    if (W && W->InvInputWeightFlg) {
        RSVD_matrices t = {NULL, Res->V_hat};
        RSVDt_vars v2 = *V;
        v2.TS.DirAdj = PETSC_TRUE;
        Weight_matrices w2 = {0};
        w2.InvInputWeightFlg = PETSC_TRUE;
        w2.W_f_sqrt_inv = W->W_f_sqrt_inv;
        RSVDT_CHK(ApplyWeightMats(&t, &v2, &w2, PETSC_TRUE));
        Res->V_hat = t.Y_hat;
    }
    RSVDT_HIPCHK("SVDAllFreqs", rsvdt_hip_sync());
    if (V->Display > 0) rsvdt_printf("*** Reduced SVD elapsed time = %.3f s ***\n", RSVDT_Time() - t1);
    return 0;
}
