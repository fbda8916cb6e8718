"""Synthetic operators of the benchmark workloads, written as PETSc-binary
files or built in memory (scipy CSR).

* gl(N): the linearised Ginzburg-Landau operator (complex tridiagonal),
  A = -U d/dx + gamma d2/dx2 + mu(x), U = 2+0.2i, gamma = 1-i,
  mu(x) = (mu0 - U^2/(4 gamma)) + mu2 x^2/2, mu0 = 0.23, mu2 = -0.01,
  N uniform points of [-85, 85], central differences, Dirichlet.
* flow2d(): 3 coupled real fields on a 512 x 256 periodic grid (N = 393,216):
  advection by u = 1 + 0.5 sin(2 pi y/Ly) in x, diffusion nu = 1e-3 with the
  compact 9-point (Mehrstellen) Laplacian, and a seeded dense 3x3 reaction
  per grid point with N(0, 0.1^2) entries; Lx = 2, Ly = 1 (h = 1/256).
  11 nonzeros per row.  The operator is shifted by a multiple of I so that its
  spectral abscissa is -0.05.
* flow3d(): 5 coupled fields on a 128 x 64 x 48 grid (N = 1,966,080),
  7-point diffusion nu = 5e-4, advection by u = (1 + 0.3 tanh(y), 0, 0),
  per-point seeded 5x5 coupling, periodic in x and z, Dirichlet in y, h = 1/64,
  shifted to abscissa -0.02.
Field index is fastest in the row ordering (row = nf*point + field), so the
per-point couplings are dense 3x3/5x5 diagonal blocks and the stencil
neighbours of a row are +-nf, +-nf*nx, ... rows away.
"""
import numpy as np
import scipy.sparse as sp


def gl(N=500, U=2 + 0.2j, gamma=1 - 1j, mu0=0.23, mu2=-0.01, L=85.0):
    dx = 2 * L / (N - 1)
    x = -L + dx * np.arange(N)
    mu = (mu0 - U ** 2 / (4 * gamma)) + mu2 * x ** 2 / 2
    main = -2 * gamma / dx ** 2 + mu
    up = -U / (2 * dx) + gamma / dx ** 2
    lo = U / (2 * dx) + gamma / dx ** 2
    return sp.diags([lo * np.ones(N - 1), main, up * np.ones(N - 1)], [-1, 0, 1]).tocsr().astype(np.complex128)


def _flow2d_unshifted(nx=512, ny=256, nf=3, nu=1e-3, Lx=2.0, Ly=1.0, sigma=0.1, seed=7):
    hx, hy = Lx / nx, Ly / ny
    assert abs(hx - hy) < 1e-15, "the 9-point stencil assumes a square grid"
    h = hx
    ii, jj = np.meshgrid(np.arange(nx), np.arange(ny), indexing="xy")   # shape (ny, nx)
    pt = (jj * nx + ii).ravel()
    y = (jj.ravel() + 0.5) * hy
    u = 1.0 + 0.5 * np.sin(2 * np.pi * y / Ly)
    rows, cols, vals = [], [], []
    # compact 9-point Laplacian: (1/(6h^2)) [[1,4,1],[4,-20,4],[1,4,1]]
    w9 = {(0, 0): -20.0, (1, 0): 4.0, (-1, 0): 4.0, (0, 1): 4.0, (0, -1): 4.0,
          (1, 1): 1.0, (1, -1): 1.0, (-1, 1): 1.0, (-1, -1): 1.0}
    for (di, dj), wgt in w9.items():
        nb = (((jj + dj) % ny) * nx + ((ii + di) % nx)).ravel()
        v = nu * wgt / (6 * h * h) * np.ones(pt.size)
        if dj == 0 and di != 0:                         # central advection -u d/dx
            v = v - u * di / (2 * h)
        for f in range(nf):
            rows.append(nf * pt + f); cols.append(nf * nb + f); vals.append(v)
    rng = np.random.default_rng(seed)
    R = rng.normal(0.0, sigma, size=(pt.size, nf, nf))
    for f in range(nf):
        for g in range(nf):
            rows.append(nf * pt + f); cols.append(nf * pt + g); vals.append(R[:, f, g])
    N = nf * nx * ny
    A = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(N, N))
    A.sum_duplicates()
    A.sort_indices()
    return A


def _flow3d_unshifted(nx=128, ny=64, nz=48, nf=5, nu=5e-4, h=1.0 / 64, sigma=0.1, seed=11):
    ii, jj, kk = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    ii, jj, kk = ii.ravel(), jj.ravel(), kk.ravel()
    pt = (kk * ny + jj) * nx + ii
    y = (jj + 1) * h - 0.5 * (ny + 1) * h          # centred, Dirichlet walls just outside the grid
    u = 1.0 + 0.3 * np.tanh(y)
    rows, cols, vals = [], [], []
    for (di, dj, dk), wgt in {(0, 0, 0): -6.0, (1, 0, 0): 1.0, (-1, 0, 0): 1.0, (0, 1, 0): 1.0, (0, -1, 0): 1.0,
                              (0, 0, 1): 1.0, (0, 0, -1): 1.0}.items():
        j2 = jj + dj
        keep = (j2 >= 0) & (j2 < ny)                  # Dirichlet in y
        nb = (((kk + dk) % nz) * ny + j2) * nx + ((ii + di) % nx)
        v = nu * wgt / (h * h) * np.ones(pt.size)
        if di != 0:
            v = v - u * di / (2 * h)
        for f in range(nf):
            rows.append((nf * pt + f)[keep]); cols.append((nf * nb + f)[keep]); vals.append(v[keep])
    rng = np.random.default_rng(seed)
    R = rng.normal(0.0, sigma, size=(pt.size, nf, nf))
    for f in range(nf):
        for g in range(nf):
            rows.append(nf * pt + f); cols.append(nf * pt + g); vals.append(R[:, f, g])
    N = nf * nx * ny * nz
    A = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(N, N))
    A.sum_duplicates()
    A.sort_indices()
    return A


def rightmost_real_part(A, tol=1e-6):
    """Real part of the rightmost eigenvalue (ARPACK, which='LR')."""
    from scipy.sparse.linalg import eigs
    vals = eigs(A.astype(np.float64) if np.isrealobj(A.data) else A, k=4, which="LR", tol=tol,
                maxiter=20000, return_eigenvectors=False, ncv=40)
    return float(np.max(vals.real))


def shifted(A, abscissa, lam=None):
    lam = rightmost_real_part(A) if lam is None else lam
    return (A - (lam - abscissa) * sp.eye(A.shape[0], format="csr")).tocsr(), lam


def rightmost_eigs_gpu(pkg, A, dt, block=16, total_time=800.0, qr_every=4, seed=3):
    """Rightmost eigenvalues of A by block subspace iteration with the RK4
    propagator P = e^{A RR dt} applied on the GPU (TSExpADeltaT through the
    engine), re-orthonormalised by QRDecomposition every `qr_every`
    applications, and a Rayleigh-Ritz step at the end: lambda = log(mu) /
    (RR dt) for the Ritz values mu of P on the final block.  Converges for
    the eigenvalues whose decay gap to the block's last one, times
    total_time, is large; returns them sorted by real part, largest first."""
    import ctypes
    s = pkg.Session(A, k=block, q=0, Nw=64, w=2 * np.pi / 20, TL=0.0, dt=dt, transient_removal=False)
    L = pkg.lib()
    rng = np.random.default_rng(seed)
    N = A.shape[0]
    X = rng.standard_normal((N, block)) + 1j * rng.standard_normal((N, block))
    U = pkg.mat_from_dense(X, block)
    s.V.TS.DirAdj = 1
    m = s.RR
    napply = max(1, int(total_time / (m * s.dt)))
    for it in range(napply):
        pkg.check(L.TSExpADeltaT(ctypes.byref(s.LNS), ctypes.byref(s.V), U), "TSExpADeltaT")
        if (it + 1) % qr_every == 0:
            pkg.check(L.QRDecomposition(U), "QRDecomposition")
    pkg.check(L.QRDecomposition(U), "QRDecomposition")
    Q = pkg.mat_to_dense(U)
    pkg.check(L.TSExpADeltaT(ctypes.byref(s.LNS), ctypes.byref(s.V), U), "TSExpADeltaT")
    W = pkg.mat_to_dense(U)
    pkg.destroy(U)
    s.close()
    mu = np.linalg.eigvals(Q.conj().T @ W)
    lam = np.log(mu.astype(complex)) / (m * s.dt)
    return lam[np.argsort(-lam.real)]


# rightmost real parts of the unshifted operators (computed once; kept here
# so that building a workload is fast)
FLOW2D_LAMBDA = 0.0003284555252514587   # shift-invert ARPACK about sigma = 0.3
# rightmost_eigs_gpu(pkg, _flow3d_unshifted(), flow3d_dt(), block=16) with 300
# and 600 time units agree to 8 digits; the same call returns FLOW2D_LAMBDA
# for the 2D operator
FLOW3D_LAMBDA = -0.00433013


def flow2d(abscissa=-0.05):
    A0 = _flow2d_unshifted()
    A, lam = shifted(A0, abscissa, FLOW2D_LAMBDA)
    return A.astype(np.complex128)


def flow3d(abscissa=-0.02):
    A0 = _flow3d_unshifted()
    A, lam = shifted(A0, abscissa, FLOW3D_LAMBDA)
    return A.astype(np.complex128)


def rk4_stable_dt(A, factor=0.5):
    """factor x the largest RK4-stable step for the spectrum of A, estimated
    from the largest-magnitude eigenvalues (the RK4 region reaches 2.785 on
    the negative real axis and 2.828 on the imaginary axis)."""
    from scipy.sparse.linalg import eigs
    vals = eigs(A, k=6, which="LM", tol=1e-6, return_eigenvectors=False, ncv=40)

    def stable(z):
        r = 1 + z + z * z / 2 + z ** 3 / 6 + z ** 4 / 24
        return abs(r) <= 1.0
    hi = 10.0 / max(abs(vals))
    lo = 0.0
    for _ in range(60):
        mid = 0.5 * (lo + hi)
        if all(stable(mid * v) for v in vals):
            lo = mid
        else:
            hi = mid
    return factor * lo


def _rk4_limit(lams):
    """Largest h with |R(h lam)| <= 1 for every lam (R the RK4 amplification
    polynomial 1 + z + z^2/2 + z^3/6 + z^4/24), by bisection."""
    lams = np.asarray(lams)
    lo, hi = 0.0, 10.0 / np.max(np.abs(lams))
    for _ in range(80):
        mid = 0.5 * (lo + hi)
        z = mid * lams
        r = np.abs(1 + z + z * z / 2 + z ** 3 / 6 + z ** 4 / 24)
        if np.all(r <= 1.0 + 1e-12):
            lo = mid
        else:
            hi = mid
    return lo


def flow2d_dt(factor=0.5, nx=512, ny=256, nu=1e-3, Lx=2.0, shift=0.05):
    """factor x the RK4 stability limit of the 2D flow operator, from the
    Fourier symbol of its constant-coefficient part (9-point diffusion plus
    central advection at every velocity the field takes, u in [0.5, 1.5]);
    the O(0.1) reaction terms do not move the limit."""
    h = Lx / nx
    a = 2 * np.pi * np.arange(nx) / nx
    b = 2 * np.pi * np.arange(ny) / ny
    A, B = np.meshgrid(a, b, indexing="ij")
    diff = nu * (-20 + 8 * np.cos(A) + 8 * np.cos(B) + 4 * np.cos(A) * np.cos(B)) / (6 * h * h)
    lams = [diff - 1j * u * np.sin(A) / h - shift for u in np.linspace(0.5, 1.5, 11)]
    return factor * _rk4_limit(np.concatenate([l.ravel() for l in lams]))


def flow3d_dt(factor=0.5, nx=128, ny=64, nz=48, nu=5e-4, h=1.0 / 64, shift=0.02):
    """factor x the RK4 stability limit of the 3D flow operator (7-point
    diffusion + central advection with u in [0.7, 1.3]), from its symbol."""
    a = 2 * np.pi * np.arange(nx) / nx
    c = 2 * np.pi * np.arange(nz) / nz
    b = np.pi * np.arange(1, ny + 1) / (ny + 1)        # Dirichlet modes in y
    A, B, C = np.meshgrid(a, b, c, indexing="ij")
    diff = nu * (-6 + 2 * np.cos(A) + 2 * np.cos(B) + 2 * np.cos(C)) / (h * h)
    lams = [diff - 1j * u * np.sin(A) / h - shift for u in np.linspace(0.7, 1.3, 7)]
    return factor * _rk4_limit(np.concatenate([l.ravel() for l in lams]))
