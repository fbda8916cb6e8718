"""Kernel-level parity of the gfx950 HIP kernels, called through the C ABI
(include/rsvdt_hip.h), against dense numpy FP64 references."""
import ctypes

import numpy as np
import pytest
import scipy.sparse as sp

import kat
from conftest import crandn, rel
from gpu_util import Desc, Dev, grouped_form, init, tiled_form

pytestmark = pytest.mark.gpu
V = ctypes.c_void_p
I64 = ctypes.c_int64


@pytest.fixture(scope="module")
def L(pkg):
    return init(pkg)


@pytest.mark.parametrize("P,K,Lc,p0,sgn,cmod,oprec", [(64, 64, 5000, 3, 1, 200, 64), (17, 9, 777, 5, -1, 40, 64),
                                                     (1, 64, 300, 0, 1, 100, 64), (129, 129, 1000, 0, 1, 129, 64),
                                                     (33, 16, 4096, 7, -1, 64, 32), (64, 32, 300001, 9, -1, 8192, 64),
                                                     (64, 32, 600003, 4, 1, 8192, 32)])
def test_coef_gemm(L, pkg, P, K, Lc, p0, sgn, cmod, oprec):
    """out[p] = 0.7 sum_q coef(q, col_p) X[q] against numpy; the last two
    cases have outputs larger than the last-level cache (~300 MB)."""
    print(f"coef_gemm P={P} K={K} L={Lc} p0={p0} sgn={sgn} cmod={cmod} prec={oprec}", flush=True)
    rng = np.random.default_rng(P + K)
    coef = crandn(rng, K, cmod)          # coef[q][c], cp = 1, cq = cmod
    X = crandn(rng, K, Lc)
    cols = (p0 + sgn * np.arange(P)) % cmod
    ref = 0.7 * (coef[:, cols].T @ X)
    dc, dx = Dev(pkg, coef), Dev(pkg, X)
    dout = Dev(pkg, nbytes=P * Lc * (16 if oprec == 64 else 8))
    L.rsvdt_hip_coef_gemm.argtypes = [I64, I64, I64, V, I64, I64, I64, ctypes.c_int, I64, ctypes.c_double, V, I64,
                                      V, I64, ctypes.c_int, ctypes.c_double, V]
    rc = L.rsvdt_hip_coef_gemm(P, K, Lc, dc.p, cmod, 1, p0, sgn, cmod, 0.7, dx.p, Lc, dout.p, Lc, oprec, 0.0, None)
    assert rc == 0
    out = dout.get(np.complex128 if oprec == 64 else np.complex64, (P, Lc))
    assert rel(out, ref) < (1e-14 if oprec == 64 else 1e-6)


@pytest.mark.parametrize("nb,nr,na,nc,layout", [(4, 3000, 16, 16, "freq"), (3, 5000, 65, 65, "snap"),
                                                (2, 100, 5, 33, "freq"), (7, 999, 10, 3, "snap"),
                                                (3, 4001, 129, 129, "freq"), (2, 777, 129, 64, "freq"),
                                                (5, 1001, 33, 48, "freq"), (3, 4096, 1, 7, "freq"),
                                                (64, 2000, 16, 16, "freq"), (2, 3, 20, 20, "freq")])
def test_bgram_bapply(L, pkg, nb, nr, na, nc, layout):
    print(f"bgram nb={nb} nr={nr} na={na} nc={nc} {layout}", flush=True)
    rng = np.random.default_rng(na * nc)
    if layout == "freq":   # X[b][r][a]
        X = crandn(rng, nb, nr, na); Z = crandn(rng, nb, nr, nc)
        xs = (nr * na, na, 1); zs = (nr * nc, nc, 1)
        Xb = X; Zb = Z
    else:                  # X[a][r][b]
        X = crandn(rng, na, nr, nb); Z = crandn(rng, nc, nr, nb)
        xs = (1, nb, nr * nb); zs = (1, nb, nr * nb)
        Xb = X.transpose(2, 1, 0); Zb = Z.transpose(2, 1, 0)
    G = np.einsum("bra,brc->bac", Xb.conj(), Zb)
    L.rsvdt_hip_bgram_work_bytes.restype = ctypes.c_size_t
    L.rsvdt_hip_bgram_work_bytes.argtypes = [I64] * 4
    dX, dZ = Dev(pkg, X), Dev(pkg, Z)
    dG = Dev(pkg, nbytes=nb * na * nc * 16)
    dw = Dev(pkg, nbytes=L.rsvdt_hip_bgram_work_bytes(nb, nr, na, nc))
    L.rsvdt_hip_bgram.argtypes = [I64] * 4 + [V, I64, I64, I64, V, I64, I64, I64, V, V, V]
    assert L.rsvdt_hip_bgram(nb, nr, na, nc, dX.p, *xs, dZ.p, *zs, dG.p, dw.p, None) == 0
    assert rel(dG.get(np.complex128, (nb, na, nc)), G) < 1e-13
    # bapply: Out = 0.5*Z + X @ C, with C nb x na x nc
    C = crandn(rng, nb, na, nc)
    ref = 0.5 * Zb + np.einsum("bra,bac->brc", Xb, C)
    dC = Dev(pkg, C)
    dO = Dev(pkg, nbytes=Z.nbytes)
    L.rsvdt_hip_bapply.argtypes = [I64] * 4 + [V, I64, I64, I64, V, ctypes.c_double, V, I64, I64, I64, V, I64, I64,
                                               I64, V]
    assert L.rsvdt_hip_bapply(nb, nr, na, nc, dX.p, *xs, dC.p, 0.5, dZ.p, *zs, dO.p, *zs, None) == 0
    out = dO.get(np.complex128, Z.shape)
    if layout != "freq":
        out = out.transpose(2, 1, 0)
    assert rel(out, ref) < 1e-13


@pytest.mark.parametrize("nb,n", [(5, 16), (3, 20), (3, 33), (2, 40), (3, 65), (2, 70), (2, 81), (2, 96), (2, 129),
                                  (8, 1)])
def test_chol_jacobi_solve(L, pkg, nb, n):
    """Cholesky + inverse (in registers for n <= 80, global memory beyond),
    Jacobi SVD, shifted solves (registers for n <= 79, LDS for n <= 96,
    global beyond) against numpy."""
    print(f"dense nb={nb} n={n}", flush=True)
    rng = np.random.default_rng(n)
    X = crandn(rng, nb, 4 * n + 3, n)
    G = np.einsum("bra,brc->bac", X.conj(), X)
    dG = Dev(pkg, G)
    dR, dRi = Dev(pkg, nbytes=G.nbytes), Dev(pkg, nbytes=G.nbytes)
    info = Dev(pkg, np.zeros(4, dtype=np.int32))
    L.rsvdt_hip_bchol_inv.argtypes = [I64, I64, V, ctypes.c_double, ctypes.c_double, V, V, V, V]
    assert L.rsvdt_hip_bchol_inv(nb, n, dG.p, 0.0, 0.0, dR.p, dRi.p, info.p, None) == 0
    R = dR.get(np.complex128, G.shape)
    Ri = dRi.get(np.complex128, G.shape)
    for b in range(nb):
        assert np.allclose(np.tril(R[b], -1), 0)
        assert rel(R[b].conj().T @ R[b], G[b]) < 1e-13
        assert np.allclose(Ri[b] @ R[b], np.eye(n), atol=1e-10)
        assert np.all(np.diag(R[b]).real > 0)
    if n <= 64:
        sig_dev, dU = Dev(pkg, nbytes=nb * n * 8), Dev(pkg, nbytes=G.nbytes)
        L.rsvdt_hip_bjacobi_svd.argtypes = [I64, I64, V, V, V, V]
        assert L.rsvdt_hip_bjacobi_svd(nb, n, dR.p, dU.p, sig_dev.p, None) == 0
        U = dU.get(np.complex128, G.shape)
        sg = sig_dev.get(np.float64, (nb, n))
        for b in range(nb):
            Ul, sl, _ = np.linalg.svd(R[b])
            assert np.allclose(sg[b], sl, rtol=1e-12)
            assert np.allclose(U[b].conj().T @ U[b], np.eye(n), atol=1e-12)
            for j in range(n):
                c = np.vdot(U[b][:, j], Ul[:, j])
                assert abs(abs(c) - 1) < 1e-9
    # shifted solves (ph I - M) x = b for nrhs shifts
    nrhs = 3
    M = crandn(rng, nb, n, n) / np.sqrt(n)
    B = crandn(rng, nb, n, nrhs)
    sh = np.exp(1j * rng.uniform(0, 6, size=(nb, nrhs))) * 3
    dM, dB, dsh = Dev(pkg, M), Dev(pkg, B), Dev(pkg, sh)
    L.rsvdt_hip_bshift_solve_work_bytes.restype = ctypes.c_size_t
    L.rsvdt_hip_bshift_solve_work_bytes.argtypes = [I64] * 3
    dw = Dev(pkg, nbytes=L.rsvdt_hip_bshift_solve_work_bytes(nb, n, nrhs))
    L.rsvdt_hip_bshift_solve.argtypes = [I64, I64, I64, V, V, V, V, V, V, V]
    assert L.rsvdt_hip_bshift_solve(nb, n, nrhs, dM.p, dsh.p, dB.p, dB.p, dw.p, info.p, None) == 0
    Xs = dB.get(np.complex128, B.shape)
    for b in range(nb):
        for m in range(nrhs):
            x = np.linalg.solve(sh[b, m] * np.eye(n) - M[b], B[b, :, m])
            assert rel(Xs[b, :, m], x) < 1e-11


def rk4_ref(A, y, F0, Fh, F1, h):
    k1 = A @ y + F0
    k2 = A @ (y + h / 2 * k1) + Fh
    k3 = A @ (y + h / 2 * k2) + Fh
    k4 = A @ (y + h * k3) + F1
    return y + h / 6 * k1 + h / 3 * k2 + h / 3 * k3 + h / 6 * k4




@pytest.mark.parametrize("k,real,prec,nb,lanes,variant", [
    (1, False, 64, 1, 0, 0), (3, True, 64, 1, 0, 0), (10, False, 64, 1, 0, 0), (16, True, 64, 1, 0, 0),
    (16, False, 64, 1, 8, 0), (32, False, 64, 3, 0, 0), (40, True, 64, 1, 0, 0), (100, False, 64, 1, 0, 0),
    (16, True, 32, 1, 0, 0), (10, False, 32, 2, 4, 0),
    (16, True, 64, 2, 16, 1), (16, True, 64, 1, 16, 2), (16, False, 64, 1, 16, 3), (16, True, 32, 1, 16, 3),
    (8, True, 64, 2, 32, 3)])
def test_rk4_step(L, pkg, k, real, prec, nb, lanes, variant):
    print(f"rk4 k={k} real={real} prec={prec} nb={nb} lanes={lanes} variant={variant}", flush=True)
    rng = np.random.default_rng(k + nb)
    N = 3001
    A = sp.random(N, N, density=5.0 / N, random_state=k, format="csr", dtype=np.float64)
    if not real:
        A = A + 1j * sp.random(N, N, density=3.0 / N, random_state=k + 1, format="csr")
    A = (A + sp.eye(N) * (-1.0)).tocsr().astype(np.complex128)
    A.sort_indices()
    h = 0.05
    y = crandn(rng, nb, N, k); F = crandn(rng, 3, nb, N, k)
    ref = np.stack([rk4_ref(A, y[b], F[0, b], F[1, b], F[2, b], h) for b in range(nb)])
    cdt = np.complex128 if prec == 64 else np.complex64
    vals = A.data.real.astype(np.float64) if real else A.data.astype(np.complex128)   # FP64 values for both state precisions
    drp = Dev(pkg, A.indptr.astype(np.int32)); dci = Dev(pkg, A.indices.astype(np.int32)); dv = Dev(pkg, vals)
    dy = Dev(pkg, y.astype(cdt)); dF = [Dev(pkg, F[i].astype(cdt)) for i in range(3)]
    du1, du2, dacc = (Dev(pkg, nbytes=y.astype(cdt).nbytes) for _ in range(3))
    dsnap = Dev(pkg, nbytes=y.nbytes)
    d = Desc(N, k, nb, drp.p, dci.p, dv.p, int(real), prec, lanes, 1, variant)
    L.rsvdt_hip_rk4_stage.argtypes = [ctypes.POINTER(Desc), ctypes.c_int, V, V, V, V, V, ctypes.c_double,
                                      ctypes.c_double, V, V]
    f = L.rsvdt_hip_rk4_stage
    assert f(ctypes.byref(d), 1, dy.p, dy.p, dF[0].p, dacc.p, du1.p, h / 2, h / 6, None, None) == 0
    assert f(ctypes.byref(d), 2, du1.p, dy.p, dF[1].p, dacc.p, du2.p, h / 2, h / 6 * 2, None, None) == 0
    assert f(ctypes.byref(d), 3, du2.p, dy.p, dF[1].p, dacc.p, du1.p, h, h / 6 * 2, None, None) == 0
    assert f(ctypes.byref(d), 4, du1.p, None, dF[2].p, dacc.p, dy.p, 0.0, h / 6, dsnap.p, None) == 0
    out = dy.get(cdt, y.shape)
    snap = dsnap.get(np.complex128, y.shape)
    tol = 1e-13 if prec == 64 else 3e-6
    assert rel(out, ref) < tol
    assert np.array_equal(snap, out.astype(np.complex128))


@pytest.mark.parametrize("k,real,prec,nb,lanes,variant,scaled,forced", [
    (1, False, 64, 1, 0, 0, False, True), (16, True, 64, 1, 0, 3, False, True), (16, True, 64, 1, 16, 1, True, False),
    (10, False, 64, 3, 0, 0, True, True), (16, False, 64, 2, 8, 0, False, False), (32, True, 64, 1, 0, 2, False, True),
    (40, False, 64, 1, 0, 0, True, False), (16, True, 32, 1, 0, 3, False, True), (16, True, 32, 1, 8, 0, False, True),
    (10, False, 32, 2, 4, 0, True, True), (16, True, 64, 1, 16, 4, False, True), (16, False, 64, 2, 16, 5, True, True),
    (32, True, 32, 1, 32, 4, False, True), (12, True, 64, 1, 16, 5, False, False),
    (16, True, 32, 1, 8, 7, False, True), (32, False, 32, 2, 16, 7, True, True), (8, True, 32, 3, 4, 7, True, False),
    (64, False, 32, 1, 32, 7, False, True), (32, True, 64, 1, 16, 4, False, True), (32, False, 64, 2, 16, 5, True, True),
    (24, True, 64, 1, 16, 5, False, False), (32, True, 64, 1, 16, 2, False, True), (32, False, 64, 2, 16, 3, True, False),
    (16, True, 64, 1, 8, 1, False, True), (16, True, 64, 1, 4, 3, False, True), (32, False, 64, 2, 8, 2, True, True),
    (13, True, 32, 1, 4, 1, False, False), (16, True, 32, 1, 4, 8, False, True), (32, False, 32, 2, 8, 8, True, True),
    (64, True, 32, 1, 16, 8, True, False)])
def test_horner_stage(L, pkg, k, real, prec, nb, lanes, variant, scaled, forced):
    """out_b = ys_b y_b + c A vin_b + g_b, with the snapshot copy, against numpy."""
    print(f"horner k={k} real={real} prec={prec} nb={nb} lanes={lanes} variant={variant} scaled={scaled} "
          f"forced={forced}", flush=True)
    rng = np.random.default_rng(7 * k + nb)
    N = 2999
    A = sp.random(N, N, density=6.0 / N, random_state=k + 3, format="csr", dtype=np.float64)
    if not real:
        A = A + 1j * sp.random(N, N, density=3.0 / N, random_state=k + 4, format="csr")
    A = (A - sp.eye(N)).tocsr().astype(np.complex128)
    A.sort_indices()
    c = 0.037
    vin, y, g = crandn(rng, nb, N, k), crandn(rng, nb, N, k), crandn(rng, nb, N, k)
    ys = crandn(rng, nb) if scaled else np.ones(nb, complex)
    ref = np.stack([ys[b] * y[b] + c * (A @ vin[b]) + (g[b] if forced else 0) for b in range(nb)])
    cdt = np.complex128 if prec == 64 else np.complex64
    vals = A.data.real.astype(np.float64) if real else A.data.astype(np.complex128)   # FP64 values for both state precisions
    drp = Dev(pkg, A.indptr.astype(np.int32)); dci = Dev(pkg, A.indices.astype(np.int32)); dv = Dev(pkg, vals)
    dvin, dy, dg = Dev(pkg, vin.astype(cdt)), Dev(pkg, y.astype(cdt)), Dev(pkg, g.astype(cdt))
    dys = Dev(pkg, ys.astype(np.complex128))
    dsnap = Dev(pkg, nbytes=y.nbytes)
    d = Desc(N, k, nb, drp.p, dci.p, dv.p, int(real), prec, lanes, 1, variant)
    L.rsvdt_hip_horner_stage.argtypes = [ctypes.POINTER(Desc), V, V, V, V, ctypes.c_double, V, V, V]
    # out aliases y (the in-place last stage of a step)
    rc = L.rsvdt_hip_horner_stage(ctypes.byref(d), dvin.p, dy.p, dys.p if scaled else None, dg.p if forced else None,
                                  c, dy.p, dsnap.p, None)
    assert rc == 0
    out = dy.get(cdt, y.shape)
    snap = dsnap.get(np.complex128, y.shape)
    assert rel(out, ref) < (1e-14 if prec == 64 else 2e-6)
    assert np.array_equal(snap, out.astype(np.complex128))


@pytest.mark.parametrize("variant", [9, 10])
@pytest.mark.parametrize("k,real,prec,nb,kc,tr,remap,scaled,forced,bpw", [
    (32, True, 64, 1, 8, 32, 3, False, True, 1), (32, False, 64, 3, 8, 32, 0, True, True, 1),
    (16, True, 64, 2, 16, 16, 3, True, False, 1), (12, False, 64, 1, 8, 32, 1, False, True, 1),
    (40, True, 64, 1, 16, 48, 2, True, True, 1), (32, True, 32, 1, 16, 64, 3, False, True, 1),
    (16, False, 32, 2, 8, 32, 0, True, True, 1), (32, True, 64, 1, 32, 8, 3, False, False, 1),
    (64, False, 64, 1, 32, 32, 0, True, True, 1), (8, True, 64, 4, 8, 32, 0, True, True, 1),
    (8, True, 64, 7, 8, 32, 0, True, False, 3), (32, False, 64, 5, 16, 32, 3, True, True, 4),
    (12, True, 32, 4, 8, 32, 0, False, True, 4)])
def test_horner_tile(L, pkg, k, real, prec, nb, kc, tr, remap, scaled, forced, bpw, variant):
    """The Horner stage on the tiled form (variant 9; 10 with the next pass's
    rows prefetched into registers): out_b = ys_b y_b + c A vin_b + g_b with
    the snapshot copy, against numpy, on a banded multi-field-like operator
    whose tiles share most of their columns (and k not a multiple of the
    columns per pass), with one or several blocks per workgroup (the last
    workgroup's share short when bpw does not divide nb)."""
    print(f"horner tile variant={variant} k={k} real={real} prec={prec} nb={nb} kc={kc} tr={tr} remap={remap} "
          f"bpw={bpw}", flush=True)
    rng = np.random.default_rng(11 * k + nb + variant)
    N = 4001
    offs = [-61, -7, -1, 0, 1, 2, 7, 61]
    A = sp.diags([rng.standard_normal(N - abs(o)) for o in offs], offs, format="csr")
    if not real:
        A = A + 1j * sp.diags([rng.standard_normal(N - 3)], [3], format="csr")
    A = A.tocsr().astype(np.complex128)
    A.sort_indices()
    c = 0.029
    vin, y, g = crandn(rng, nb, N, k), crandn(rng, nb, N, k), crandn(rng, nb, N, k)
    ys = crandn(rng, nb) if scaled else np.ones(nb, complex)
    ref = np.stack([ys[b] * y[b] + c * (A @ vin[b]) + (g[b] if forced else 0) for b in range(nb)])
    cdt = np.complex128 if prec == 64 else np.complex64
    vals = A.data.real.astype(np.float64) if real else A.data.astype(np.complex128)
    tptr, tcol, lidx, hmax, emax = tiled_form(A, tr)
    drp = Dev(pkg, A.indptr.astype(np.int32)); dci = Dev(pkg, A.indices.astype(np.int32)); dv = Dev(pkg, vals)
    dtp, dtc, dli = Dev(pkg, tptr), Dev(pkg, tcol), Dev(pkg, lidx)
    dvin, dy, dg = Dev(pkg, vin.astype(cdt)), Dev(pkg, y.astype(cdt)), Dev(pkg, g.astype(cdt))
    dys = Dev(pkg, ys.astype(np.complex128))
    dsnap = Dev(pkg, nbytes=y.astype(np.complex128).nbytes)
    d = Desc(N, k, nb, drp.p, dci.p, dv.p, int(real), prec, kc, remap, variant, 0, None, None, None, tr, hmax, emax,
             dtp.p, dtc.p, dli.p, bpw)
    L.rsvdt_hip_horner_stage.argtypes = [ctypes.POINTER(Desc), V, V, V, V, ctypes.c_double, V, V, V]
    # out aliases y (the in-place last stage of a step)
    rc = L.rsvdt_hip_horner_stage(ctypes.byref(d), dvin.p, dy.p, dys.p if scaled else None, dg.p if forced else None,
                                  c, dy.p, dsnap.p, None)
    assert rc == 0, L.rsvdt_hip_last_error()
    out = dy.get(cdt, y.shape)
    snap = dsnap.get(np.complex128, y.shape)
    assert rel(out, ref) < (1e-14 if prec == 64 else 2e-6)
    assert np.array_equal(snap, out.astype(np.complex128))
    # the first stage gathers y itself (vin aliases y, no nontemporal hint on it)
    dy2 = Dev(pkg, y.astype(cdt))
    dout = Dev(pkg, nbytes=y.astype(cdt).nbytes)
    rc = L.rsvdt_hip_horner_stage(ctypes.byref(d), dy2.p, dy2.p, None, None, c, dout.p, None, None)
    assert rc == 0, L.rsvdt_hip_last_error()
    ref1 = np.stack([y[b] + c * (A @ y[b]) for b in range(nb)])
    assert rel(dout.get(cdt, y.shape), ref1) < (1e-14 if prec == 64 else 2e-6)


@pytest.mark.parametrize("k,lanes", [(16, 16), (32, 32), (4, 4)])
def test_horner_stage_xcd_remaps(L, pkg, k, lanes):
    """Every workgroup-to-rows map (identity, one range per XCD, chunks of 64
    or 256 row blocks dealt to the XCDs, with a tail that fills no whole
    round) computes every row once: identical results, equal to numpy."""
    print(f"horner remaps k={k} lanes={lanes}", flush=True)
    rng = np.random.default_rng(k)
    N = 40000 + 7
    A = sp.random(N, N, density=5.0 / N, random_state=k, format="csr", dtype=np.float64)
    A = (A - sp.eye(N)).tocsr()
    A.sort_indices()
    vin, y = crandn(rng, N, k), crandn(rng, N, k)
    ref = y + 0.05 * (A @ vin)
    drp = Dev(pkg, A.indptr.astype(np.int32)); dci = Dev(pkg, A.indices.astype(np.int32))
    dv = Dev(pkg, A.data.astype(np.float64))
    dvin, dy = Dev(pkg, vin), Dev(pkg, y)
    outs = []
    for remap in (0, 1, 2, 3):
        dout = Dev(pkg, nbytes=y.nbytes)
        d = Desc(N, k, 1, drp.p, dci.p, dv.p, 1, 64, lanes, remap, 3)
        L.rsvdt_hip_horner_stage.argtypes = [ctypes.POINTER(Desc), V, V, V, V, ctypes.c_double, V, V, V]
        assert L.rsvdt_hip_horner_stage(ctypes.byref(d), dvin.p, dy.p, None, None, 0.05, dout.p, None, None) == 0
        outs.append(dout.get(np.complex128, y.shape))
    assert rel(outs[0], ref) < 1e-14
    for o in outs[1:]:
        assert np.array_equal(o, outs[0])


@pytest.mark.parametrize("N,k,real,prec,nb,snaps,band", [
    (700, 10, False, 64, 1, True, False), (500, 16, True, 64, 1, True, False), (333, 3, False, 64, 4, False, False),
    (700, 10, False, 32, 1, True, False), (1200, 2, True, 64, 1, True, False), (90, 40, False, 32, 2, False, False),
    (500, 10, False, 64, 1, True, True), (1024, 4, True, 32, 2, False, True), (1000, 3, False, 64, 1, True, True)])
def test_horner_fused(L, pkg, N, k, real, prec, nb, snaps, band):
    """Whole runs of Horner steps in one launch (LDS-resident state) against
    numpy: y' = y + hA(y + h/2 A(y + h/3 A(y + h/4 Ay))) + g_s, with the FP64
    snapshots of the steps that store one; band: a periodic tridiagonal
    operator (at most 3 nonzeros per row, the register-resident variant)."""
    print(f"fused N={N} k={k} real={real} prec={prec} nb={nb} snaps={snaps} band={band}", flush=True)
    rng = np.random.default_rng(N + k)
    if band:
        d = [rng.standard_normal(N) + (0 if real else 1j * rng.standard_normal(N)) for _ in range(3)]
        A = sp.diags([d[0][1:], d[1] - 4, d[2][:-1], d[0][:1], d[2][-1:]], [-1, 0, 1, N - 1, -(N - 1)])
    else:
        A = sp.random(N, N, density=4.0 / N, random_state=k, format="csr", dtype=np.float64)
        if not real:
            A = A + 1j * sp.random(N, N, density=2.0 / N, random_state=k + 1, format="csr")
    A = (A - 2 * sp.eye(N)).tocsr().astype(np.complex128)
    A.sort_indices()
    h, nsteps = 0.05, 23
    Nt, RR, Nstore = 9, 4, 3
    y0 = crandn(rng, nb, N, k)
    g = crandn(rng, nsteps, N, k) if nb == 1 else None
    cdt = np.complex128 if prec == 64 else np.complex64
    vals = A.data.real.astype(np.float64) if real else A.data.astype(np.complex128)
    drp = Dev(pkg, A.indptr.astype(np.int32)); dci = Dev(pkg, A.indices.astype(np.int32)); dv = Dev(pkg, vals)
    dg = Dev(pkg, g.astype(cdt)) if g is not None else None
    coef = np.array([h / 4, h / 3, h / 2, h])
    d = Desc(N, k, nb, drp.p, dci.p, dv.p, int(real), prec, 0, 0, 0)
    L.rsvdt_hip_horner_fused.argtypes = [ctypes.POINTER(Desc), I64, I64, V, V, I64, V, I64, V, I64, I64, I64, I64,
                                         I64, ctypes.c_int, V]
    L.rsvdt_hip_horner_fused_fits.argtypes = [I64, I64, ctypes.c_int, ctypes.c_int]
    assert L.rsvdt_hip_horner_fused_fits(N, A.nnz, int(real), prec) == 1
    maxrow = int(np.max(np.diff(A.indptr)))
    tol = 1e-12 if prec == 64 else 3e-5
    # two snapshot schedules: the run starts before the first snapshot (direct
    # slot order), and inside the snapshot period off the snapshot grid
    # (adjoint slot order, reversed); both wrap around the Nstore slots
    for i0, dr in ((5, 0), (15, 1)):
        y = y0.astype(cdt).astype(np.complex128)
        snap_ref = np.zeros((Nstore, N, k), complex)
        for st in range(nsteps):
            for b in range(nb):
                v = y[b]
                for c in (h / 4, h / 3, h / 2):
                    v = (y[b] + c * (A @ v)).astype(cdt).astype(np.complex128) if prec == 32 else y[b] + c * (A @ v)
                yn = y[b] + h * (A @ v) + (g[st].astype(cdt) if g is not None else 0)
                y[b] = yn.astype(cdt).astype(np.complex128)
            i = i0 + st
            if snaps and i > Nt and (i - Nt) % RR == 0:
                pos = ((i - Nt) // RR - 1) % Nstore
                pos = Nstore - 1 - pos if not dr else pos
                snap_ref[pos] = y[0]
        # maxrow -1: nonzeros staged in LDS; the true maxrow: in registers when the rows are short enough
        for mr in (-1, maxrow):
            dy = Dev(pkg, y0.astype(cdt))
            dsnap = Dev(pkg, np.zeros((Nstore, N, k), complex))
            rc = L.rsvdt_hip_horner_fused(ctypes.byref(d), A.nnz, mr, dy.p, dg.p if dg else None, N * k,
                                          coef.ctypes.data_as(V), nsteps, dsnap.p if snaps else None, N * k, i0, Nt,
                                          RR, Nstore, dr, None)
            assert rc == 0, L.rsvdt_hip_last_error()
            out = dy.get(cdt, y.shape)
            assert rel(out, y) < tol, (mr, i0)
            if snaps:
                assert rel(dsnap.get(np.complex128, snap_ref.shape), snap_ref) < tol, (mr, i0)
    # an operator too large for LDS is refused before any launch
    nbig = 4000
    assert L.rsvdt_hip_horner_fused_fits(nbig, 10 * nbig, int(real), prec) == 0
    big = Desc(nbig, k, 1, drp.p, dci.p, dv.p, int(real), prec, 0, 0, 0)
    assert L.rsvdt_hip_horner_fused(ctypes.byref(big), 10 * nbig, 10, dy.p, None, 0, coef.ctypes.data_as(V), 1, None,
                                    0, 0, 0, 1, 1, 1, None) != 0


def _point_operator(P, F, rng, cplx):
    """F fields per grid point: dense F x F coupling at the point plus a
    diagonal (same-field) coupling to the periodic neighbours p+-1, p+-7, and a
    few random entries; rows of one point share their columns."""
    S = sp.diags([1.0, 1.0, 1.0, 1.0], [-7, -1, 1, 7], shape=(P, P), format="csr")
    S = (S + sp.diags([1.0, 1.0, 1.0, 1.0], [P - 1, P - 7, -(P - 1), -(P - 7)], shape=(P, P))).tocsr()
    B = rng.standard_normal((F, F)) + (1j * rng.standard_normal((F, F)) if cplx else 0)
    A = sp.kron(sp.eye(P), B) + sp.kron(S, sp.diags(rng.standard_normal(F)))
    R = sp.random(P * F, P * F, density=0.5 / (P * F), random_state=int(rng.integers(1000)), format="csr")
    return (A + R - 2 * sp.eye(P * F)).tocsr().astype(np.complex128)


@pytest.mark.parametrize("F,G,k,real,prec,nb,scaled", [(5, 5, 32, True, 64, 1, False), (3, 3, 16, True, 64, 2, True),
                                                       (5, 8, 32, False, 64, 1, True), (4, 4, 10, False, 32, 1, False),
                                                       (2, 2, 64, True, 32, 3, False), (3, 5, 16, False, 64, 1, True)])
def test_horner_grouped(L, pkg, F, G, k, real, prec, nb, scaled):
    """The Horner stage on the row-grouped operator (variant 6) against numpy,
    with the forcing term, block scales and the snapshot copy."""
    print(f"grouped F={F} G={G} k={k} real={real} prec={prec} nb={nb}", flush=True)
    rng = np.random.default_rng(F * 10 + G + k)
    A = _point_operator(811, F, rng, not real)
    if real:
        A = A.real.astype(np.complex128)
    N = A.shape[0]
    A.sort_indices()
    c = 0.021
    vin, y, g = crandn(rng, nb, N, k), crandn(rng, nb, N, k), crandn(rng, nb, N, k)
    ys = crandn(rng, nb) if scaled else np.ones(nb, complex)
    cdt = np.complex128 if prec == 64 else np.complex64
    vin, y, g = vin.astype(cdt), y.astype(cdt), g.astype(cdt)
    ref = np.stack([ys[b] * y[b] + c * (A @ vin[b].astype(complex)) + g[b] for b in range(nb)])
    gptr, gword, gval = grouped_form(A, G)
    vals = A.data.real.astype(np.float64) if real else A.data.astype(np.complex128)
    gv = gval.real.astype(np.float64) if real else gval.astype(np.complex128)
    drp = Dev(pkg, A.indptr.astype(np.int32)); dci = Dev(pkg, A.indices.astype(np.int32)); dv = Dev(pkg, vals)
    dgp, dgw, dgv = Dev(pkg, gptr), Dev(pkg, gword), Dev(pkg, gv)
    dvin, dy, dg, dys = Dev(pkg, vin), Dev(pkg, y), Dev(pkg, g), Dev(pkg, ys.astype(np.complex128))
    dout, dsnap = Dev(pkg, nbytes=y.nbytes), Dev(pkg, nbytes=nb * N * k * 16)
    d = Desc(N, k, nb, drp.p, dci.p, dv.p, int(real), prec, 0, 0, 6, G, dgp.p, dgw.p, dgv.p)
    L.rsvdt_hip_horner_stage.argtypes = [ctypes.POINTER(Desc), V, V, V, V, ctypes.c_double, V, V, V]
    rc = L.rsvdt_hip_horner_stage(ctypes.byref(d), dvin.p, dy.p, dys.p if scaled else None, dg.p, c, dout.p,
                                  dsnap.p, None)
    assert rc == 0, L.rsvdt_hip_last_error()
    out = dout.get(cdt, y.shape)
    assert rel(out, ref) < (1e-14 if prec == 64 else 2e-6)
    assert np.array_equal(dsnap.get(np.complex128, y.shape), out.astype(np.complex128))
    # variant 6 without the grouped arrays is refused
    d0 = Desc(N, k, nb, drp.p, dci.p, dv.p, int(real), prec, 0, 0, 6, 0, None, None, None)
    assert L.rsvdt_hip_horner_stage(ctypes.byref(d0), dvin.p, dy.p, None, None, c, dout.p, None, None) != 0


def test_horner_stage_rejects_bad_desc(L, pkg):
    d = Desc(0, 16, 1, None, None, None, 1, 64, 0, 1, 0)
    L.rsvdt_hip_horner_stage.argtypes = [ctypes.POINTER(Desc), V, V, V, V, ctypes.c_double, V, V, V]
    assert L.rsvdt_hip_horner_stage(ctypes.byref(d), None, None, None, None, 1.0, None, None, None) != 0


@pytest.mark.parametrize("nb,nr,n", [(4, 3001, 16), (2, 1500, 64), (3, 777, 40)])
def test_bapply_in_place(L, pkg, nb, nr, n):
    """X <- X C in place (the CholeskyQR apply), unit column strides."""
    rng = np.random.default_rng(nr + n)
    X = crandn(rng, nb, nr, n)
    C = crandn(rng, nb, n, n)
    ref = np.einsum("bra,bac->brc", X, C)
    dX, dC = Dev(pkg, X), Dev(pkg, C)
    L.rsvdt_hip_bapply.argtypes = [I64] * 4 + [V, I64, I64, I64, V, ctypes.c_double, V, I64, I64, I64, V, I64, I64,
                                               I64, V]
    assert L.rsvdt_hip_bapply(nb, nr, n, n, dX.p, nr * n, n, 1, dC.p, 0.0, None, 0, 0, 0, dX.p, nr * n, n, 1,
                              None) == 0
    assert rel(dX.get(np.complex128, X.shape), ref) < 1e-13


def test_bapply_in_place_wide_rejected(L, pkg):
    """More than 64 output columns are split over workgroups: in place is refused."""
    X = np.zeros((1, 64, 129), complex)
    dX, dC = Dev(pkg, X), Dev(pkg, np.zeros((1, 129, 129), complex))
    L.rsvdt_hip_bapply.argtypes = [I64] * 4 + [V, I64, I64, I64, V, ctypes.c_double, V, I64, I64, I64, V, I64, I64,
                                               I64, V]
    assert L.rsvdt_hip_bapply(1, 64, 129, 129, dX.p, 64 * 129, 129, 1, dC.p, 0.0, None, 0, 0, 0, dX.p, 64 * 129, 129,
                              1, None) != 0


@pytest.mark.parametrize("n,N,kt,b0,kb", [(129, 3001, 16, 0, 16), (65, 1000, 32, 8, 8), (7, 50, 3, 1, 2),
                                          (64, 2048, 16, 0, 16)])
def test_transpose_rows(L, pkg, n, N, kt, b0, kb):
    """[s][r][b] <-> [j][r][s] for the batch range b0 .. b0+kb-1."""
    rng = np.random.default_rng(n + kb)
    Y = crandn(rng, n, N, kt)
    dY = Dev(pkg, Y)
    dT = Dev(pkg, nbytes=kb * N * n * 16)
    L.rsvdt_hip_transpose_rows.argtypes = [I64, I64, I64, I64, I64, V, V, ctypes.c_int, V]
    assert L.rsvdt_hip_transpose_rows(N, n, kt, b0, kb, dY.p, dT.p, 0, None) == 0
    T = dT.get(np.complex128, (kb, N, n))
    assert np.array_equal(T, Y[:, :, b0:b0 + kb].transpose(2, 1, 0))
    back = np.zeros_like(Y)
    dB = Dev(pkg, back)
    assert L.rsvdt_hip_transpose_rows(N, n, kt, b0, kb, dT.p, dB.p, 1, None) == 0
    B = dB.get(np.complex128, Y.shape)
    assert np.array_equal(B[:, :, b0:b0 + kb], Y[:, :, b0:b0 + kb])
    assert not np.any(np.delete(B, np.s_[b0:b0 + kb], axis=2))


def test_allocation_cache(L, pkg):
    """Freed device blocks of at most 256 MB are kept and handed to a later
    request of their size (up to 2x); larger ones are freed; the cache counts
    as free memory and release_cached empties it.  Garbage collection is held
    off so that no earlier test's buffer is freed into the cache meanwhile."""
    import gc
    gc.collect()
    gc.disable()
    try:
        _allocation_cache_checks(L)
    finally:
        gc.enable()


def _allocation_cache_checks(L):
    V = ctypes.c_void_p
    L.rsvdt_hip_malloc.argtypes = [ctypes.POINTER(V), ctypes.c_size_t]
    L.rsvdt_hip_free.argtypes = [V]
    L.rsvdt_hip_cached_bytes.restype = ctypes.c_size_t
    L.rsvdt_hip_mem_info.argtypes = [ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ctypes.c_size_t)]
    assert L.rsvdt_hip_release_cached() == 0
    assert L.rsvdt_hip_cached_bytes() == 0
    p = V()
    assert L.rsvdt_hip_malloc(ctypes.byref(p), 3 << 20) == 0
    first = p.value
    # the block is usable: a device round trip through it
    x = np.arange(1024, dtype=np.float64)
    assert L.rsvdt_hip_memcpy_h2d(p, x.ctypes.data_as(V), ctypes.c_size_t(x.nbytes)) == 0
    assert L.rsvdt_hip_free(p) == 0
    assert L.rsvdt_hip_cached_bytes() == 3 << 20
    f, t = ctypes.c_size_t(), ctypes.c_size_t()
    assert L.rsvdt_hip_mem_info(ctypes.byref(f), ctypes.byref(t)) == 0 and f.value <= t.value
    q = V()
    assert L.rsvdt_hip_malloc(ctypes.byref(q), (2 << 20) + 5) == 0      # within 2x: the cached block
    assert q.value == first and L.rsvdt_hip_cached_bytes() == 0
    y = np.zeros(1024)
    assert L.rsvdt_hip_memcpy_h2d(q, x.ctypes.data_as(V), ctypes.c_size_t(x.nbytes)) == 0
    assert L.rsvdt_hip_memcpy_d2h(y.ctypes.data_as(V), q, ctypes.c_size_t(y.nbytes)) == 0
    assert np.array_equal(x, y)
    assert L.rsvdt_hip_free(q) == 0
    r = V()
    assert L.rsvdt_hip_malloc(ctypes.byref(r), 1 << 20) == 0            # 3 MB is over 2x of 1 MB: a new block
    assert r.value != first and L.rsvdt_hip_cached_bytes() == 3 << 20
    assert L.rsvdt_hip_free(r) == 0
    big = V()
    assert L.rsvdt_hip_malloc(ctypes.byref(big), 300 << 20) == 0
    assert L.rsvdt_hip_free(big) == 0
    assert L.rsvdt_hip_cached_bytes() == (3 << 20) + (1 << 20)            # the 300 MB block was freed
    assert L.rsvdt_hip_release_cached() == 0
    assert L.rsvdt_hip_cached_bytes() == 0
