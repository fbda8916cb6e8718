"""Pins the oracle's weighted main sequence (Oracle.run_weighted:
SourceCode/ApplyWeightMats.c:16-80 around each action, SVDAllFreqs.c:63-64 on
the modes) without its own formulas: with input and output matrices B (N x Nb)
and C (Nc x N), an output weight W_q^{1/2} (Nc x Nc) and an input weight
W_f^{-1/2} (Nb x Nb), a sketch of k = Nb columns spans every input, so the
gains must be the singular values of W_q^{1/2} C G(A, w_m) B W_f^{-1/2} for
the RK4 transfer function G (dense numpy), V_hat must be W_f^{-1/2} times its
right singular vectors and W_q^{1/2} U_hat must span its range."""
import numpy as np

import kat
import oracle
from conftest import crandn, subspace_gap


def _weights(rng, N, Nb, Nc):
    B = crandn(rng, N, Nb)
    C = crandn(rng, Nc, N)
    Lq = np.diag(rng.uniform(0.5, 2.0, Nc)) + 0.1 * np.triu(crandn(rng, Nc, Nc), 1)
    Wq = Lq @ Lq.conj().T                       # Hermitian positive definite "square root" of the output weight
    Wf_inv = np.diag(rng.uniform(0.5, 2.0, Nb)).astype(complex)
    return B, C, Wq, np.linalg.inv(Wq), Wf_inv


def test_weighted_full_input_sketch_exact():
    N, Nw, w, dt = 12, 11, 0.4, 0.02
    Nb, Nc = 3, 4
    k = Nb
    A = kat.gl_operator(N=N, L=6.0)
    rng = np.random.default_rng(21)
    B, C, Wq, Wq_inv, Wf_inv = _weights(rng, N, Nb, Nc)
    o = oracle.Oracle(A, k=k, q=1, Nw=Nw, w=w, TL=0.0, dt=dt, transient_removal=True)
    F = crandn(rng, Nb, k * o.Nw_eff)
    U, V, S = o.run_weighted(F, B=B, C=C, Wq_sqrt=Wq, Wq_sqrt_inv=Wq_inv, Wf_sqrt_inv=Wf_inv)
    assert U.shape == (Nc, k * o.Nw_eff) and V.shape == (Nb, k * o.Nw_eff) and S.shape == (k, o.Nw_eff)
    h, Ns, Nt, RR = kat.grid(w, Nw, 0.0, dt)
    for m in range(o.Nw_eff):
        z = np.exp(2j * np.pi * kat.alias(m, o.Nw_eff, False) / (2 * Ns))
        M = Wq @ C @ kat.transfer(A.toarray(), h, z) @ B @ Wf_inv
        Um, sm, Vmh = np.linalg.svd(M)
        assert np.allclose(S[:, m], sm[:k], rtol=1e-10)
        Vm = np.linalg.solve(Wf_inv, V[:, m * k:(m + 1) * k])
        for j in range(k):
            assert abs(abs(np.vdot(Vmh[j].conj(), Vm[:, j])) - 1) < 1e-9
        assert subspace_gap(Wq @ U[:, m * k:(m + 1) * k], Um[:, :k]) < 1e-9


def test_identity_weights_match_plain_run():
    """Identity B, C and weights reproduce the unweighted sequence exactly."""
    N, Nw = 20, 8
    A = kat.gl_operator(N=N, L=10.0)
    rng = np.random.default_rng(3)
    o = oracle.Oracle(A, k=3, q=1, Nw=Nw, w=0.3, TL=1.0, dt=0.02, transient_removal=True)
    F = crandn(rng, N, 3 * o.Nw_eff)
    U0, V0, S0 = o.run(F)
    I = np.eye(N)
    U1, V1, S1 = o.run_weighted(F, B=I, C=I, Wq_sqrt=I, Wq_sqrt_inv=I, Wf_sqrt_inv=I)
    assert np.allclose(S1, S0, rtol=1e-12)
    assert np.allclose(np.abs(np.sum(U1.conj() * U0, axis=0)), 1, atol=1e-10)
    assert np.allclose(np.abs(np.sum(V1.conj() * V0, axis=0)), 1, atol=1e-10)
