/*
 * rsvdt: command-line driver with the reference's call sequence
 * (SourceCode/main.c:107-150):  rsvdt -inputs variables.yaml [-Key value ...]
 */
#include "rsvdt.h"

#include <stdio.h>
#include <string.h>

#define CHK(x)                          \
    do {                                \
        PetscErrorCode _e = (x);        \
        if (_e) { RSVDT_Finalize(); return _e; } \
    } while (0)

int main(int argc, char **argv)
{
    Directories dirs;
    RSVDt_vars RSVDt;
    TransRun_vars TR;
    LNS_vars LNS;
    DFT_matrices DFTm;
    TS_removal_matrices TSR;
    Weight_matrices W;
    RSVD_matrices R;
    Resolvent_matrices Res;
    memset(&RSVDt, 0, sizeof(RSVDt)); memset(&TR, 0, sizeof(TR)); memset(&LNS, 0, sizeof(LNS));
    memset(&TSR, 0, sizeof(TSR)); memset(&W, 0, sizeof(W));
    CHK(RSVDT_Initialize(&argc, &argv));
    CHK(PreProcessing(&RSVDt, &W, &LNS, &R, &Res, &TR, &DFTm, &dirs));
    if (TR.TransRun) {   /* SourceCode/main.c:119-124: run and exit; the forward operator A */
        RSVDt.TS.DirAdj = PETSC_TRUE;
        CHK(TransientRunRK4(&TR, &RSVDt, &LNS, &DFTm, &dirs));
        RSVDT_MatDestroy(&LNS.A); RSVDT_MatDestroy(&DFTm.idft); RSVDT_MatDestroy(&DFTm.dft);
        return RSVDT_Finalize();
    }
    printf("\n********************************************\n*************** RSVD-\\Delta t **************\n"
           "********************************************\n");
    double t0 = RSVDT_Time();
    CHK(DirectActionRK4(&R, &RSVDt, &LNS, &DFTm, &W, &dirs, &TSR));
    CHK(PowerIterationRK4(&R, &RSVDt, &LNS, &DFTm, &W, &dirs, &TSR));
    CHK(SVDAllFreqsBeforeAdjoint(&R, &RSVDt));
    CHK(StoreU(&R, &Res));
    CHK(AdjointActionRK4(&R, &RSVDt, &LNS, &DFTm, &W, &dirs, &TSR));
    CHK(SVDAllFreqs(&R, &RSVDt, &W, &Res));
    printf("\n[rsvdt] time to modes: %.3f s\n", RSVDT_Time() - t0);
    CHK(SaveResults(&Res, &RSVDt, &dirs));
    RSVDT_MatDestroy(&Res.U_hat); RSVDT_MatDestroy(&Res.V_hat); RSVDT_MatDestroy(&Res.S_hat);
    RSVDT_MatDestroy(&LNS.A); RSVDT_MatDestroy(&DFTm.idft); RSVDT_MatDestroy(&DFTm.dft);
    return RSVDT_Finalize();
}
