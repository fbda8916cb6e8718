"""The FP32-state form (the Nt steps before the first snapshot in complex
FP32; the snapshot period, DFT, removal, QR and SVD in FP64) against the FP64
form on a reduced grid of the 2D flow operator: 3 fields on a 64 x 32
periodic grid (N = 6144, 11 nonzeros per row), shifted so that its rightmost
eigenvalue sits at -0.05 as on the full grid, k = 16, Nw = 32 -> 16
frequencies, q = 1, transient removal, the whole main.c sequence
(SourceCode/main.c:134-146) on the same sketch, at TL = T/2 and 6 T.

The FP32 rounding accumulated over the transient steps is a perturbation of
the state when the snapshots start, i.e. part of the transient, which the
removal takes out in FP64.  (With the snapshot period in FP32 as well, the
removal's correction of the transient carried the rounding into the gains:
4e-5 from FP64 at TL = T/2 on this grid.)  Required at every frequency:
sigma_1..3 within 1e-6 relative, the leading three response and forcing
subspaces within 1e-4 (sine of the largest principal angle)."""
import os
import sys

import numpy as np
import pytest
import scipy.sparse as sp
from scipy.sparse.linalg import eigs

from conftest import ROOT, crandn, subspace_gap

sys.path.insert(0, os.path.join(ROOT, "rsvd-delta-t_amd"))
import workloads  # noqa: E402

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("TLT", [0.5, 6.0])
def test_fp32_state_flow2d_reduced(pkg, TLT):
    nx, ny, k, Nw, nm = 64, 32, 16, 32, 3
    A0 = workloads._flow2d_unshifted(nx=nx, ny=ny)
    lam = float(np.max(eigs(A0, k=6, which="LR", return_eigenvectors=False, tol=1e-8, maxiter=50000,
                            ncv=60, v0=np.ones(A0.shape[0])).real))
    A = (A0 - (lam + 0.05) * sp.eye(A0.shape[0])).tocsr().astype(np.complex128)
    w = 2 * np.pi / 20
    T = 2 * np.pi / w
    dt = workloads.flow2d_dt(nx=nx, ny=ny, shift=0.05)
    N = A.shape[0]
    rng = np.random.default_rng(5)
    F = crandn(rng, N, k * Nw // 2)
    out = {}
    for prec in (64, 32):
        print(f"flow2d {nx}x{ny} FP{prec} TL={TLT}T", flush=True)
        s = pkg.Session(A, k=k, q=1, Nw=Nw, w=w, TL=TLT * T, dt=dt, transient_removal=True, precision=prec)
        out[prec] = s.run(F)
        s.close()
    pkg.set_options(precision=64)
    (U6, V6, S6), (U3, V3, S3) = out[64], out[32]
    nwe = S6.shape[1]
    assert nwe == Nw // 2
    es = np.abs(S3[:nm] - S6[:nm]) / S6[:nm]
    print(f"  sigma_1..{nm} largest relative difference {es.max():.2e}", flush=True)
    assert es.max() < 1e-6
    worst = 0.0
    for m in range(nwe):
        if (S6[nm - 1, m] - S6[nm, m]) / S6[nm - 1, m] < 1e-3:
            continue            # leading subspace not separated at this frequency
        sl = slice(m * k, m * k + nm)
        worst = max(worst, subspace_gap(U6[:, sl], U3[:, sl]), subspace_gap(V6[:, sl], V3[:, sl]))
    print(f"  leading {nm} subspaces: largest gap {worst:.2e}", flush=True)
    assert worst < 1e-4
