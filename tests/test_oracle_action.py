"""Pins the oracle's resolvent actions (TSActionRK4 + DFT / transient removal)
against the exact discrete solution of the RK4 recursion, evaluated with
dense linear algebra in numpy (tests/kat.py): periodic part
(z^2 I - P)^{-1} Q(z) F / Ns per frequency plus the homogeneous transient
-P^i y_per(0).  A sign, index or window error in the oracle changes the
result at O(1) and fails these checks."""
import numpy as np
import pytest

import kat
import oracle
from conftest import crandn, rel

CASES = [
    # (N, Nw, complex, direct, TR, TL)
    (40, 8, True, True, False, 5.0),
    (40, 8, True, False, False, 5.0),
    (40, 7, True, True, False, 5.0),
    (40, 8, False, True, False, 5.0),
    (40, 8, False, False, False, 5.0),
    (40, 39, True, True, True, 0.0),     # Nw + 1 = N: Galerkin projection on the whole space is exact
    (40, 39, True, False, True, 0.0),
    (41, 40, False, True, True, 2.0),    # real operator, Nw_eff = 20, Nstore = 41 = N
    (41, 40, False, False, True, 2.0),
]


@pytest.mark.parametrize("N,Nw,cplx,direct,tr,TL", CASES)
def test_action_exact(N, Nw, cplx, direct, tr, TL):
    A = kat.gl_operator(N=N, L=20.0)
    if not cplx:
        A = A.real.astype(complex)
    k = 3
    rng = np.random.default_rng(N + Nw + int(direct))
    o = oracle.Oracle(A, k=k, q=0, Nw=Nw, w=0.3, TL=TL, dt=0.02, transient_removal=tr)
    F = crandn(rng, N, k * o.Nw_eff)
    Y = o.action(F, direct=direct)
    Yk, info = kat.exact_action(A, F, k, Nw, 0.3, TL, 0.02, direct=direct, transient_removal=tr)
    assert info["Ns"] == o.Ns and info["Nt"] == o.Nt and info["RR"] == o.RR
    assert rel(Y, Yk) < 1e-11


def test_transient_is_included():
    """Without transient removal and with a short TL the exact answer differs
    from the periodic part; the oracle follows the former (not the latter)."""
    A = kat.gl_operator(N=40, L=20.0)
    rng = np.random.default_rng(5)
    o = oracle.Oracle(A, k=2, q=0, Nw=8, w=0.3, TL=0.5, dt=0.02, transient_removal=False)
    F = crandn(rng, 40, 2 * o.Nw_eff)
    Y = o.action(F)
    Yfull, _ = kat.exact_action(A, F, 2, 8, 0.3, 0.5, 0.02)
    Yper, _ = kat.exact_action(A, F, 2, 8, 0.3, 0.5, 0.02, include_transient=False)
    assert rel(Y, Yfull) < 1e-11
    assert rel(Y, Yper) > 1e-4


@pytest.mark.parametrize("direct", [True, False])
def test_transient_removal_reduces_error(direct):
    """With Nw + 1 < N the projection is not exact, but it must remove most of
    the transient (SourceCode/TransientRemovalStrategy.c)."""
    A = kat.gl_operator(N=64, L=20.0)
    rng = np.random.default_rng(7)
    k, Nw, w, dt, TL = 2, 16, 0.3, 0.02, 1.0
    F = crandn(rng, 64, k * Nw)
    o_tr = oracle.Oracle(A, k=k, q=0, Nw=Nw, w=w, TL=TL, dt=dt, transient_removal=True)
    o_no = oracle.Oracle(A, k=k, q=0, Nw=Nw, w=w, TL=TL, dt=dt, transient_removal=False)
    Yper, _ = kat.exact_action(A, F, k, Nw, w, TL, dt, direct=direct, transient_removal=True)
    e_tr = rel(o_tr.action(F, direct), Yper)
    e_no = rel(o_no.action(F, direct), Yper)
    assert e_tr < e_no / 5


def test_snapshot_store_layout():
    """Y_all column ik*Nstore + s holds y at step Nt + (s+1) RR (direct) and the
    reversed order for the adjoint (SourceCode/SaveSnapshots.c:19-28)."""
    A = kat.gl_operator(N=24, L=10.0)
    rng = np.random.default_rng(3)
    o = oracle.Oracle(A, k=2, q=0, Nw=4, w=0.5, TL=0.3, dt=0.05, transient_removal=True)
    F = crandn(rng, 24, 2 * o.Nw_eff)
    Yd = o.snapshots(F, direct=True)
    Ya = o.snapshots(F, direct=False)
    assert Yd.shape == (24, 5 * 2)
    # the direct store's first snapshot is the state after Nt + RR steps: run max_steps
    o.set_max_steps(o.Nt + o.RR)
    Yp = o.snapshots(F, direct=True)
    assert np.allclose(Yp[:, 0], Yd[:, 0], rtol=0, atol=1e-14 * np.abs(Yd).max())
    assert np.allclose(Yp[:, 1:5], 0)
    o.set_max_steps(o.Nt + o.RR)
    Yq = o.snapshots(F, direct=False)
    assert np.allclose(Yq[:, 4], Ya[:, 4], rtol=0, atol=1e-14 * np.abs(Ya).max())


@pytest.mark.parametrize("TL", [5.0, 20.0, 60.0])
def test_rank_deficient_snapshots(TL):
    """Snapshots of a real operator with Nw_eff = 16 frequencies and a decayed
    transient are numerically rank-deficient (17 of 33 singular values below
    1e-10 of the largest); the Householder basis completes them with
    rounding-noise directions and the removal still returns the periodic part."""
    A = kat.gl_operator(N=120, L=30.0).real.astype(complex)
    rng = np.random.default_rng(0)
    o = oracle.Oracle(A, k=2, q=0, Nw=32, w=0.3, TL=TL, dt=0.02, transient_removal=True)
    F = crandn(rng, 120, 2 * o.Nw_eff)
    Ys = o.snapshots(F)
    sv = np.linalg.svd(Ys[:, :33], compute_uv=False)
    assert sv[-1] / sv[0] < 1e-15
    Yp, _ = kat.exact_action(A, F, 2, 32, 0.3, TL, 0.02, transient_removal=True)
    assert rel(o.action(F), Yp) < 1e-9
