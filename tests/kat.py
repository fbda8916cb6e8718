"""Closed-form known answers for the RSVD-Delta-t time-stepped actions.

These are derived independently of the step-by-step algorithm: the RK4
recursion y_i = P y_{i-1} + (forcing term) is linear with a forcing that is a
sum of discrete exponentials, so the discrete solution from y_0 = 0 is the sum
of a periodic part, (z^2 I - P)^{-1} Q(z) F / Ns per frequency, and the
homogeneous transient -P^i y_per(0).  Dense numpy linear algebra on small
operators evaluates it exactly (to rounding), including the transient.

Conventions pinned here (reference file:line):
  * time grid: SourceCode/SetupTimeFreqGrid.c:16-21
  * iDFT samples idft[m, j] = (1/Ns) exp(+i w_m j h/2), aliased frequencies
    for complex operators: SourceCode/CreateDFTiDFTMats.c:42-57,93
  * stage forcings F0, F1/2, F1 at half-step positions 2i-2, 2i-1, 2i (direct)
    and their negatives (adjoint): SourceCode/TSRK4.c:18-47
  * snapshots every RR steps after Nt, reversed for the adjoint:
    SourceCode/SaveSnapshots.c:19-28
  * Y_hat = RR * Y_all * dft, dft[s, m] = exp(-2 pi i m s / Nw): SourceCode/DFT.c:29,38
"""
import numpy as np


def grid(w, Nw, TL, dt):
    T = 2 * np.pi / w
    dt_w = T / Nw
    h = dt_w / np.ceil(dt_w / dt)
    Ns = int(round(T / h))
    Nt = int(round(TL / h))
    return h, Ns, Nt, Ns // Nw


def alias(m, Nw_eff, real_op):
    """Signed frequency index of column m (MATLAB FFT ordering for complex A)."""
    if real_op:
        return m
    half = Nw_eff // 2 - 1 if Nw_eff % 2 == 0 else (Nw_eff - 1) // 2
    return m if m <= half else m - Nw_eff


def rk4_P(A, h):
    n = A.shape[0]
    hA = h * A
    I = np.eye(n)
    return I + hA + hA @ hA / 2 + hA @ hA @ hA / 6 + hA @ hA @ hA @ hA / 24


def rk4_Q(A, h, z):
    """One RK4 step from y=0 with stage forcings (1, z, z, z^2) * F (F = identity)."""
    n = A.shape[0]
    I = np.eye(n)
    k1 = I
    k2 = A @ (h / 2 * k1) + z * I
    k3 = A @ (h / 2 * k2) + z * I
    k4 = A @ (h * k3) + z * z * I
    return h / 6 * (k1 + 2 * k2 + 2 * k3 + k4)


def transfer(A, h, z):
    """G(A, w) = (z^2 I - P)^{-1} Q(z), z = exp(i w h / 2)."""
    P = rk4_P(A, h)
    return np.linalg.solve(z * z * np.eye(A.shape[0]) - P, rk4_Q(A, h, z))


def exact_action(A, F, k, Nw, w, TL, dt, direct=True, transient_removal=False, include_transient=True):
    """Exact discrete Y_hat (N x k*Nw_eff, frequency-major) of one action.

    With transient_removal=True this returns what the Galerkin strategy must
    produce when the snapshot space is the whole state space (Nw+1 >= N):
    the periodic part alone."""
    A = np.asarray(A.toarray() if hasattr(A, "toarray") else A, dtype=np.complex128)
    N = A.shape[0]
    real_op = np.linalg.norm(A.imag) < 1e-12
    Nw_eff = Nw // 2 if real_op else Nw
    h, Ns, Nt, RR = grid(w, Nw, TL, dt)
    Aop = A if direct else A.conj().T
    P = rk4_P(Aop, h)
    Nstore = Nw + 1 if transient_removal else Nw
    Y_all = np.zeros((Nstore, N, k), dtype=np.complex128)
    # periodic parts per frequency
    X = np.zeros((Nw_eff, N, k), dtype=np.complex128)
    zs = np.zeros(Nw_eff, dtype=np.complex128)
    for m in range(Nw_eff):
        ma = alias(m, Nw_eff, real_op)
        z = np.exp(2j * np.pi * ma / (2 * Ns))
        if not direct:
            z = 1 / z
        zs[m] = z
        Fm = F[:, m * k:(m + 1) * k]
        X[m] = transfer(Aop, h, z) @ Fm / Ns
    y0 = X.sum(axis=0)
    Pi = np.linalg.matrix_power(P, Nt + RR)
    PR = np.linalg.matrix_power(P, RR)
    for s in range(Nstore):
        i = Nt + (s + 1) * RR
        yper = np.einsum("m,mnk->nk", zs ** (2 * i), X)
        y = yper - (Pi @ y0 if include_transient else 0)
        col = s if direct else Nstore - 1 - s
        Y_all[col] = y
        Pi = PR @ Pi
    dft = np.exp(-2j * np.pi * np.outer(np.arange(Nw), np.arange(Nw_eff)) / Nw)
    Yh = np.zeros((N, k * Nw_eff), dtype=np.complex128)
    if transient_removal:
        # periodic part of the windowed DFT y1 (window 0..Nw-1 direct, 1..Nw adjoint)
        for m in range(Nw_eff):
            p = 2 * (Nt + RR) if direct else 2 * Nt
            Yh[:, m * k:(m + 1) * k] = Ns * X[m] * zs[m] ** p
        return Yh, dict(h=h, Ns=Ns, Nt=Nt, RR=RR, Nw_eff=Nw_eff, real_op=real_op)
    for m in range(Nw_eff):
        Yh[:, m * k:(m + 1) * k] = RR * np.einsum("s,snk->nk", dft[:, m], Y_all[:Nw])
    return Yh, dict(h=h, Ns=Ns, Nt=Nt, RR=RR, Nw_eff=Nw_eff, real_op=real_op)


def exact_transfer_sigmas(A, Nw, w, dt, direct=True):
    """Singular values of the per-frequency RK4 transfer function G(A, w_m)."""
    A = np.asarray(A.toarray() if hasattr(A, "toarray") else A, dtype=np.complex128)
    real_op = np.linalg.norm(A.imag) < 1e-12
    Nw_eff = Nw // 2 if real_op else Nw
    h, Ns, Nt, RR = grid(w, Nw, 0.0, dt)
    out = []
    for m in range(Nw_eff):
        z = np.exp(2j * np.pi * alias(m, Nw_eff, real_op) / (2 * Ns))
        out.append(np.linalg.svd(transfer(A, h, z), compute_uv=False))
    return np.array(out).T   # N x Nw_eff


def gl_operator(N=500, U=2 + 0.2j, gamma=1 - 1j, mu0=0.23, mu2=-0.01, L=85.0):
    """Linearised Ginzburg-Landau: A = -U d/dx + gamma d2/dx2 + mu(x), central
    differences on N uniform points of [-L, L], Dirichlet (truncated stencil)."""
    import scipy.sparse as sp
    dx = 2 * L / (N - 1)
    x = -L + dx * np.arange(N)
    mu = (mu0 - U ** 2 / (4 * gamma)) + mu2 * x ** 2 / 2
    main = -2 * gamma / dx ** 2 + mu
    up = -U / (2 * dx) + gamma / dx ** 2
    lo = U / (2 * dx) + gamma / dx ** 2
    return sp.diags([lo * np.ones(N - 1), main, up * np.ones(N - 1)], [-1, 0, 1]).tocsr().astype(np.complex128)
