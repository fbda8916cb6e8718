"""Parity of the resolvent actions (DirectActionRK4 / AdjointActionRK4 with
the snapshot DFT or the transient removal) between the HIP engine and the
FP64 CPU oracle, on the same seeded inputs.  Tolerance: 1e-10 relative
(FP64 state); the FP32-state form is checked at 2e-5."""
import numpy as np
import pytest

import kat
import oracle
from conftest import crandn, rel

pytestmark = pytest.mark.gpu

CASES = [
    # N, Nw, complex, TR, TL, k
    (64, 16, True, False, 3.0, 3),
    (64, 16, True, True, 1.0, 3),
    (64, 16, False, False, 3.0, 2),
    (64, 16, False, True, 1.0, 2),
    (50, 7, True, False, 2.0, 5),
    (65, 64, True, True, 0.0, 16),     # Nstore = N: exact removal, 16 columns
]


@pytest.mark.parametrize("N,Nw,cplx,tr,TL,k", CASES)
@pytest.mark.parametrize("direct", [True, False])
@pytest.mark.parametrize("form,fuse,variant", [(0, 1, -1), (0, 0, -1), (0, 0, 6), (1, 0, -1)])
def test_action_vs_oracle(pkg, N, Nw, cplx, tr, TL, k, direct, form, fuse, variant):
    """form 0: Horner stages with the folded forcing, whole chunks per launch
    in LDS (fuse 1) or one launch per stage (fuse 0; variant 6: on the
    row-grouped operator); form 1: the reference's stage order."""
    print(f"action N={N} Nw={Nw} complex={cplx} TR={tr} TL={TL} k={k} direct={direct} form={form} fuse={fuse} "
          f"variant={variant}", flush=True)
    A = kat.gl_operator(N=N, L=20.0)
    if not cplx:
        A = A.real.astype(complex)
    rng = np.random.default_rng(N * 7 + Nw)
    o = oracle.Oracle(A, k=k, q=0, Nw=Nw, w=0.3, TL=TL, dt=0.02, transient_removal=tr)
    F = crandn(rng, N, k * o.Nw_eff)
    Yo = o.action(F, direct=direct)
    s = pkg.Session(A, k=k, q=0, Nw=Nw, w=0.3, TL=TL, dt=0.02, transient_removal=tr, rk4_form=form, fuse_small=fuse,
                    spmm_variant=variant)
    assert (s.Ns, s.Nt, s.RR) == (o.Ns, o.Nt, o.RR)
    s.set_sketch(F)
    s.direct() if direct else s.adjoint()
    Yg = s.get_Y()
    s.close()
    assert rel(Yg, Yo) < 1e-10


@pytest.mark.parametrize("direct", [True, False])
@pytest.mark.parametrize("fuse", [1, 0])
def test_removal_groups(pkg, direct, fuse):
    """The transient removal in groups of test vectors (3, 3, 3, 1 of k = 10,
    as a large operator is processed): the group buffers are reused across
    groups and the basis is rebuilt for the Galerkin step."""
    print(f"removal groups direct={direct} fuse={fuse}", flush=True)
    N, Nw, k = 64, 16, 10
    A = kat.gl_operator(N=N, L=20.0)
    rng = np.random.default_rng(321)
    o = oracle.Oracle(A, k=k, q=0, Nw=Nw, w=0.3, TL=1.0, dt=0.02, transient_removal=True)
    F = crandn(rng, N, k * o.Nw_eff)
    Yo = o.action(F, direct=direct)
    s = pkg.Session(A, k=k, q=0, Nw=Nw, w=0.3, TL=1.0, dt=0.02, transient_removal=True, fuse_small=fuse, tr_group=3)
    s.set_sketch(F)
    s.direct() if direct else s.adjoint()
    Yg = s.get_Y()
    s.close()
    pkg.set_options(tr_group=0)
    assert rel(Yg, Yo) < 1e-10


@pytest.mark.parametrize("direct", [True, False])
@pytest.mark.parametrize("group", [0, 3])
def test_removal_entry_points(pkg, direct, group):
    """CreateMtilde then GalerkinProjAllModes called through the C ABI on the
    oracle's snapshots (SourceCode/TransientRemovalStrategy.c:30-36 without
    the final permutation): Y_hat_up, k-major, against the oracle's action."""
    import ctypes
    print(f"removal entry points direct={direct} group={group}", flush=True)
    N, Nw, k = 64, 16, 7
    A = kat.gl_operator(N=N, L=20.0)
    rng = np.random.default_rng(99)
    o = oracle.Oracle(A, k=k, q=0, Nw=Nw, w=0.3, TL=1.0, dt=0.02, transient_removal=True)
    F = crandn(rng, N, k * o.Nw_eff)
    Ysnap = o.snapshots(F, direct=direct)
    Yo = o.action(F, direct=direct)
    Nwe = o.Nw_eff
    Yo_k = Yo.reshape(N, Nwe, k).transpose(0, 2, 1).reshape(N, k * Nwe)     # column ik*Nwe + m
    s = pkg.Session(A, k=k, q=0, Nw=Nw, w=0.3, TL=1.0, dt=0.02, transient_removal=True, tr_group=group)
    s.V.TS.DirAdj = int(direct)
    L = pkg.lib()
    Ym = pkg.mat_from_dense(np.asarray(Ysnap))
    pkg.check(L.CreateMtilde(Ym, ctypes.byref(s.LNS), ctypes.byref(s.V), ctypes.byref(s.TSR)), "CreateMtilde")
    pkg.check(L.GalerkinProjAllModes(Ym, ctypes.byref(s.LNS), ctypes.byref(s.V), ctypes.byref(s.TSR),
                                     ctypes.byref(s.D)), "GalerkinProjAllModes")       # destroys Y_all
    assert not s.TSR.M_tilde_all                                                     # destroyed as well
    Yup = pkg.mat_to_dense(ctypes.c_void_p(s.TSR.Y_hat_up))
    pkg.destroy(s.TSR.Y_hat_up)
    s.TSR.Y_hat_up = None
    s.close()
    pkg.set_options(tr_group=0)
    assert rel(Yup, Yo_k) < 1e-10


def test_chunk_boundaries_and_lanes(pkg):
    """Forcing-chunk sizes that do not divide the step count, and every lane width."""
    A = kat.gl_operator(N=48, L=20.0)
    rng = np.random.default_rng(11)
    o = oracle.Oracle(A, k=4, q=0, Nw=8, w=0.3, TL=1.3, dt=0.02, transient_removal=False)
    F = crandn(rng, 48, 4 * o.Nw_eff)
    Yo = o.action(F)
    for chunk, lanes, form in [(1, 4, 0), (3, 8, 0), (7, 16, 0), (32, 32, 0), (5, 64, 0), (1, 4, 1), (3, 8, 1),
                               (7, 16, 1), (32, 32, 1), (5, 64, 1)]:
        print(f"chunk={chunk} lanes={lanes} form={form}", flush=True)
        s = pkg.Session(A, k=4, q=0, Nw=8, w=0.3, TL=1.3, dt=0.02, transient_removal=False, chunk_steps=chunk,
                        spmm_lanes=lanes, rk4_form=form, fuse_small=0)
        s.set_sketch(F)
        s.direct()
        assert rel(s.get_Y(), Yo) < 1e-10
        s.close()
    pkg.set_options(chunk_steps=0, spmm_lanes=0)


def test_config0_ginzburg_landau(pkg):
    """The CPU-runnable Ginzburg-Landau case: N = 500, k = 10, Nw = 64,
    w = 0.05, dt = T/4096, TL = 10, transient removal on; one direct action."""
    print("config0 GL N=500 k=10 Nw=64 TR", flush=True)
    A = kat.gl_operator(N=500)
    rng = np.random.default_rng(2024)
    k, Nw, w, TL, dt = 10, 64, 0.05, 10.0, 0.03068
    o = oracle.Oracle(A, k=k, q=0, Nw=Nw, w=w, TL=TL, dt=dt, transient_removal=True)
    assert o.Ns == 4096 and o.RR == 64
    F = crandn(rng, 500, k * o.Nw_eff)
    Yo = o.action(F)
    for fuse in (1, 0):
        s = pkg.Session(A, k=k, q=0, Nw=Nw, w=w, TL=TL, dt=dt, transient_removal=True, fuse_small=fuse)
        s.set_sketch(F)
        s.direct()
        Yg = s.get_Y()
        s.close()
        assert rel(Yg, Yo) < 1e-10, fuse


def test_fp32_state_form(pkg):
    """Complex-FP32 RK4 state and SpMM (DFT, QR, SVD and transient removal in FP64)."""
    A = kat.gl_operator(N=64, L=20.0)
    rng = np.random.default_rng(5)
    o = oracle.Oracle(A, k=3, q=0, Nw=16, w=0.3, TL=1.0, dt=0.02, transient_removal=True)
    F = crandn(rng, 64, 3 * o.Nw_eff)
    Yo = o.action(F)
    s = pkg.Session(A, k=3, q=0, Nw=16, w=0.3, TL=1.0, dt=0.02, transient_removal=True, precision=32)
    s.set_sketch(F)
    s.direct()
    e = rel(s.get_Y(), Yo)
    s.close()
    pkg.set_options(precision=64)
    print(f"fp32 relative error {e:.2e}")
    assert e < 2e-5


@pytest.mark.parametrize("TL", [5.0, 20.0])
def test_rank_deficient_snapshot_basis(pkg, TL):
    """Real operator, Nw_eff = 16 < Nw + 1 = 33 snapshots and a transient that
    has decayed: the snapshot basis is numerically rank-deficient (singular
    values down to 1e-17 of the largest).  The removal must still return the
    periodic response (exact discrete answer) and agree with the oracle."""
    print(f"rank-deficient snapshots TL={TL}", flush=True)
    A = kat.gl_operator(N=120, L=30.0).real.astype(complex)
    rng = np.random.default_rng(int(TL))
    o = oracle.Oracle(A, k=2, q=0, Nw=32, w=0.3, TL=TL, dt=0.02, transient_removal=True)
    F = crandn(rng, 120, 2 * o.Nw_eff)
    Yo = o.action(F)
    Yp, _ = kat.exact_action(A, F, 2, 32, 0.3, TL, 0.02, transient_removal=True)
    s = pkg.Session(A, k=2, q=0, Nw=32, w=0.3, TL=TL, dt=0.02, transient_removal=True)
    s.set_sketch(F)
    s.direct()
    Yg = s.get_Y()
    s.close()
    assert np.all(np.isfinite(Yg))
    assert rel(Yg, Yp) < 1e-9
    assert rel(Yg, Yo) < 1e-9


@pytest.mark.parametrize("tr,direct,prec", [(False, True, 64), (True, False, 64), (True, True, 32)])
def test_graph_replay(pkg, tr, direct, prec):
    """Full forcing chunks replayed from a captured HIP graph (one per forcing
    buffer, snapshot slots refreshed per chunk) give the same result bit for
    bit as direct launches; the tail chunk runs direct.  A power iteration
    (q = 1) re-captures per action, with eager QR and transient removal work
    between the replays."""
    A = kat.gl_operator(N=64, L=20.0)
    rng = np.random.default_rng(31)
    o = oracle.Oracle(A, k=3, q=1, Nw=16, w=0.3, TL=1.3, dt=0.02, transient_removal=tr)
    F = crandn(rng, 64, 3 * o.Nw_eff)
    out = {}
    for g in (0, 1):
        print(f"graph={g} tr={tr} direct={direct} prec={prec}", flush=True)
        # a fixed launch shape (no autotuning), so both runs sum in the same order
        s = pkg.Session(A, k=3, q=1, Nw=16, w=0.3, TL=1.3, dt=0.02, transient_removal=tr, chunk_steps=7,
                        use_graph=g, precision=prec, fuse_small=0, autotune=0)
        s.set_sketch(F)
        s.direct() if direct else s.adjoint()
        Y1 = s.get_Y()
        U, Vh, S = s.run(F)
        out[g] = (Y1, U, Vh, S)
        s.close()
    for a, b in zip(out[0], out[1]):
        assert np.array_equal(a, b)
    if prec == 64:
        assert rel(out[1][0], o.action(F, direct=direct)) < 1e-10


@pytest.mark.parametrize("real,Nw,Ns", [(False, 16, 512), (True, 64, 4096), (False, 15, 300)])
def test_idft_table_on_device(pkg, real, Nw, Ns):
    """The iDFT table CreateDFTiDFTMats builds on the device against the
    closed form idft(m, j) = exp(+2 pi i m_alias j / (2 Ns)) 2/(2 Ns)
    (SourceCode/CreateDFTiDFTMats.c:42-57,93), m_alias the aliased frequency
    index of a complex operator (odd Nw included)."""
    import ctypes
    print(f"idft table real={real} Nw={Nw} Ns={Ns}", flush=True)
    L = pkg.lib()
    V = pkg.RSVDt_vars()
    Nwe = Nw // 2 if real else Nw
    V.RSVD.Nw, V.RSVD.Nw_eff, V.TS.Ns, V.TS.RealOperator = Nw, Nwe, Ns, int(real)
    D, LNS = pkg.DFT_matrices(), pkg.LNS_vars()
    pkg.check(L.RSVDT_Initialize(None, None), "RSVDT_Initialize")
    pkg.check(L.CreateDFTiDFTMats(ctypes.byref(V), ctypes.byref(D), ctypes.byref(LNS)), "CreateDFTiDFTMats")
    idft = pkg.mat_to_dense(pkg._vp(D.idft))                        # Nw_eff x 2Ns
    pkg.destroy(D.idft)
    pkg.destroy(D.dft)
    nt2 = 2 * Ns
    m = np.arange(Nwe)
    if not real:
        half = Nwe // 2 - 1 if Nwe % 2 == 0 else (Nwe - 1) // 2
        m = np.where(m > half, m - Nwe + nt2, m)
    p = np.outer(m.astype(np.int64), np.arange(nt2, dtype=np.int64)) % nt2     # exact phase index
    ref = np.exp(2j * np.pi * p / nt2) * (2.0 / nt2)
    assert idft.shape == (Nwe, nt2)
    assert np.max(np.abs(idft - ref)) < 1e-15 * (2.0 / nt2) * 8
