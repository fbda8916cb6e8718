"""End-to-end pin of the oracle's main.c sequence (SourceCode/main.c:134-144):
with a full sketch (k = N) the randomized SVD is exact, so the gains must be
the singular values of the per-frequency RK4 transfer function G(A, w_m),
and for the tutorial operator the RK4 gains approach those of the continuous
resolvent (iwI - A)^{-1}."""
import numpy as np
import pytest

import kat
import oracle
from conftest import crandn, rel, subspace_gap


@pytest.mark.parametrize("q", [0, 1])
def test_full_sketch_gains_exact(q):
    N, Nw, w, dt = 12, 11, 0.4, 0.02
    A = kat.gl_operator(N=N, L=6.0)
    rng = np.random.default_rng(q)
    o = oracle.Oracle(A, k=N, q=q, Nw=Nw, w=w, TL=0.0, dt=dt, transient_removal=True)
    F = crandn(rng, N, N * o.Nw_eff)
    U, V, S = o.run(F)
    Sx = kat.exact_transfer_sigmas(A, Nw, w, dt)
    assert S.shape == Sx.shape
    assert np.allclose(S, Sx, rtol=1e-10)
    # forcing modes: right singular vectors of the direct transfer function.
    # (U_hat is the left basis of the direct response and is not rotated by
    # the final SVD -- SourceCode/StoreU.c:15, SVDAllFreqs.c:46 -- so with a
    # random full sketch it is not the optimal response; it is orthonormal.)
    h, Ns, Nt, RR = kat.grid(w, Nw, 0.0, dt)
    for m in range(o.Nw_eff):
        z = np.exp(2j * np.pi * kat.alias(m, o.Nw_eff, False) / (2 * Ns))
        Ug, sg, Vgh = np.linalg.svd(kat.transfer(A.toarray(), h, z))
        assert subspace_gap(V[:, m * N:m * N + 2], Vgh.conj().T[:, :2]) < 1e-8
        Um = U[:, m * N:(m + 1) * N]
        assert np.allclose(Um.conj().T @ Um, np.eye(N), atol=1e-10)


def test_tutorial_peak_gain():
    """A_GL at the tutorial's dt: sigma_1(w=-0.65) of the RK4 transfer function
    equals the dense resolvent's 1672.0219096 to 1e-9 (the published gain plot
    peaks there, Tutorial/Ginzburg-Landau/README.md:424)."""
    import petsc_bin, os
    from conftest import GOLDEN
    A = petsc_bin.read_mat(os.path.join(GOLDEN, "A_GL")).toarray()
    h, Ns, Nt, RR = kat.grid(0.05, 42, 10.0, 0.003)
    m = -13                                  # -0.65 / 0.05
    z = np.exp(2j * np.pi * m / (2 * Ns))
    s_rk4 = np.linalg.svd(kat.transfer(A, h, z), compute_uv=False)[0]
    s_ex = np.linalg.svd(np.linalg.inv(-0.65j * np.eye(500) - A), compute_uv=False)[0]
    assert abs(s_ex - 1672.02190959) < 1e-6
    assert abs(s_rk4 - s_ex) / s_ex < 1e-9
