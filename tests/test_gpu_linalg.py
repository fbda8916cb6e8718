"""Parity of the batched per-frequency linear algebra (QRAllFreqs,
SVDAllFreqsBeforeAdjoint, SVDAllFreqs) with the oracle's Householder QR and
Jacobi SVD.  Q is unique (positive real diagonal of R) so it is compared
entrywise; singular vectors are compared up to a unit phase."""
import numpy as np
import pytest

import kat
import oracle
from conftest import crandn, rel, subspace_gap

pytestmark = pytest.mark.gpu


def session(pkg, N, k, Nw):
    A = kat.gl_operator(N=N, L=20.0)
    return A, pkg.Session(A, k=k, q=0, Nw=Nw, w=0.3, TL=1.0, dt=0.02, transient_removal=False)


@pytest.mark.parametrize("N,k,Nw,cond", [(2000, 16, 8, 1.0), (3000, 10, 4, 1e6), (1000, 32, 4, 1e10)])
def test_qr_allfreqs(pkg, N, k, Nw, cond):
    print(f"qr N={N} k={k} Nw={Nw} cond={cond}", flush=True)
    A, s = session(pkg, N, k, Nw)
    rng = np.random.default_rng(k)
    Y = crandn(rng, N, k * Nw)
    for m in range(Nw):
        Y[:, m * k:(m + 1) * k] = Y[:, m * k:(m + 1) * k] @ np.diag(np.logspace(0, -np.log10(cond), k))
    o = oracle.Oracle(A, k=k, q=0, Nw=Nw, w=0.3, TL=1.0, dt=0.02, transient_removal=False)
    Qo = o.qr_allfreqs(Y)
    s.set_sketch(Y)
    s.qr()
    Qg = s.get_Y()
    s.close()
    for m in range(Nw):
        Qm = Qg[:, m * k:(m + 1) * k]
        assert np.allclose(Qm.conj().T @ Qm, np.eye(k), atol=1e-12)
        # columns well determined up to cond * eps
        assert rel(Qm, Qo[:, m * k:(m + 1) * k]) < 1e-14 * cond * 100 + 1e-12


@pytest.mark.parametrize("N,k,Nw", [(2000, 16, 8), (700, 5, 3)])
def test_svd_allfreqs(pkg, N, k, Nw):
    print(f"svd N={N} k={k} Nw={Nw}", flush=True)
    A, s = session(pkg, N, k, Nw)
    rng = np.random.default_rng(N)
    Y = crandn(rng, N, k * Nw)
    for m in range(Nw):
        Y[:, m * k:(m + 1) * k] = Y[:, m * k:(m + 1) * k] @ (crandn(rng, k, k) @ np.diag(np.logspace(0, -4, k)))
    o = oracle.Oracle(A, k=k, q=0, Nw=Nw, w=0.3, TL=1.0, dt=0.02, transient_removal=False)
    Uo = o.svd_before_adjoint(Y)
    Vo, So = o.svd_allfreqs(Y)
    s.set_sketch(Y)
    s.svd_before_adjoint()
    Ug = s.get_Y()
    s.set_sketch(Y)
    s.store_u()
    s.svd_allfreqs()
    import rsvdt_amd  # noqa: F401  (loaded by the pkg fixture)
    Vg = pkg.mat_to_dense(pkg._vp(s.Res.V_hat))
    Sg = pkg.mat_to_dense(pkg._vp(s.Res.S_hat)).real
    s.close()
    assert np.allclose(Sg, So, rtol=1e-12)
    for m in range(Nw):
        for j in range(k):
            for X, Xo in ((Ug, Uo), (Vg, Vo)):
                c = np.vdot(X[:, m * k + j], Xo[:, m * k + j])
                assert abs(abs(c) - 1) < 1e-10
