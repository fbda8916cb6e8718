"""Pins the oracle's dense kernels against LAPACK (numpy): Householder QR with
a positive real diagonal (the unique Q that Gram-Schmidt also returns),
Jacobi SVD (singular values and phase-invariant vectors) and the LU solve
that replaces the PGMRES of SourceCode/GalerkinProj.c:57-62."""
import numpy as np
import pytest

import oracle
from conftest import crandn, phase_align


@pytest.mark.parametrize("m,n", [(50, 5), (200, 16), (33, 33), (65, 64)])
def test_qr_matches_lapack(m, n):
    rng = np.random.default_rng(m * n)
    A = crandn(rng, m, n)
    Q, R = oracle.qr(A)
    Ql, Rl = np.linalg.qr(A)
    ph = np.diag(Rl) / np.abs(np.diag(Rl))
    Ql = Ql * ph
    Rl = Rl * ph.conj()[:, None]
    assert np.allclose(Q, Ql, atol=1e-12)
    assert np.allclose(R, Rl, atol=1e-11 * np.abs(Rl).max())
    assert np.all(np.abs(np.diag(R).imag) < 1e-13) and np.all(np.diag(R).real > 0)
    assert np.allclose(Q @ R, A, atol=1e-12 * np.abs(A).max())


def test_qr_ill_conditioned():
    rng = np.random.default_rng(2)
    A = crandn(rng, 300, 10)
    A[:, 9] = A[:, 8] + 1e-9 * crandn(rng, 300)
    Q, R = oracle.qr(A)
    assert np.allclose(Q.conj().T @ Q, np.eye(10), atol=1e-13)
    assert np.allclose(Q @ R, A, atol=1e-12 * np.abs(A).max())


@pytest.mark.parametrize("m,n", [(40, 4), (300, 16), (64, 32)])
def test_svd_matches_lapack(m, n):
    rng = np.random.default_rng(m + n)
    A = crandn(rng, m, n) @ np.diag(np.logspace(0, -6, n))
    U, s = oracle.svd_left(A)
    Ul, sl, _ = np.linalg.svd(A, full_matrices=False)
    assert np.allclose(s, sl, rtol=1e-12, atol=1e-14 * sl[0])
    for j in range(n):
        assert np.allclose(phase_align(Ul[:, j], U[:, j]), Ul[:, j], atol=1e-8 * sl[0] / sl[j])


def test_lu_solve():
    rng = np.random.default_rng(9)
    for n in (1, 5, 65):
        A = crandn(rng, n, n)
        b = crandn(rng, n)
        x = oracle.lu_solve(A, b)
        assert np.allclose(x, np.linalg.solve(A, b), rtol=1e-10)
