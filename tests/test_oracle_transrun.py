"""Pins the oracle's transient-run restatement (orc_trans_run / orc_trans_est,
SourceCode/TransientRunRK4.c:39-190 and SourceCode/TransientRemovalEst.c:26-118)
without its own formulas:

* the homogeneous RK4 trajectory against dense powers of the RK4 step
  polynomial P(hA) = I + hA + (hA)^2/2 + (hA)^3/6 + (hA)^4/24 (and of P(hA^H)
  for the adjoint);
* the stop rules against those closed-form norms (the run stops at the first
  monitored norm below TransConVal / above TransDivVal, before recording it);
* the removal estimate's initial norms against numpy's FFT of the snapshots,
  and its updated norms against least-squares residuals (numpy lstsq) of
  the synthetic snapshot sets, which define the projection residual without
  any QR;
* the period-end estimates of a run against the estimate on closed-form
  snapshots of the period (the circular snapshot store)."""
import numpy as np
import pytest

import kat
import oracle

W, NW, DT = 0.3, 8, 0.02


def _setup(N=48, direct=True):
    A = kat.gl_operator(N=N, L=20.0)
    o = oracle.Oracle(A, k=1, q=0, Nw=NW, w=W, TL=0.0, dt=DT, transient_removal=False)
    return A.toarray(), o


def _step_poly(A, h, direct=True):
    M = A if direct else A.conj().T
    Z = h * M
    I = np.eye(A.shape[0])
    return I + Z + Z @ Z / 2 + Z @ Z @ Z / 6 + Z @ Z @ Z @ Z / 24


def _q0(N, seed=0):
    rng = np.random.default_rng(seed)
    q = rng.random(N) + 1j * rng.random(N)
    return q / np.linalg.norm(q)


@pytest.mark.parametrize("direct", [True, False])
def test_trajectory_and_norms_closed_form(direct):
    A, o = _setup()
    q0 = _q0(A.shape[0])
    r = o.trans_run(q0, periods=2, save_mod=100, div_val=1e3, con_val=1e-12, direct=direct, save=True)
    T = 2 * np.pi / W
    dT = T / NW
    h = dT / np.ceil(dT / o.dt)
    assert r["dt"] == h and r["Nt_period"] == round(T / h)
    assert r["reason"] == 0 and r["stop"] == 2 * r["Nt_period"]
    P = _step_poly(A, h, direct)
    y = q0.copy()
    ref_norms = [np.linalg.norm(y)]
    snaps = [y.copy()]
    for i in range(1, r["stop"] + 1):
        y = P @ y
        if i % 100 == 0:
            ref_norms.append(np.linalg.norm(y))
            snaps.append(y.copy())
    assert np.max(np.abs(r["norms"] - ref_norms) / np.array(ref_norms)) < 1e-12
    assert np.linalg.norm(r["last"] - y) / np.linalg.norm(y) < 1e-12
    assert np.linalg.norm(r["snaps"] - np.array(snaps)) / np.linalg.norm(snaps) < 1e-12


def test_stop_rules():
    A, o = _setup()
    q0 = _q0(A.shape[0], 1)
    base = o.trans_run(q0, periods=3, save_mod=50, con_val=0.0)
    norms = base["norms"]
    # convergence: threshold between two monitored norms of the decaying run
    j = 5
    assert norms[j] < norms[j - 1]
    cv = 0.5 * (norms[j] + norms[j - 1])
    r = o.trans_run(q0, periods=3, save_mod=50, con_val=cv)
    assert r["reason"] == 3 and r["stop"] == 50 * j
    assert np.array_equal(r["norms"][:j], norms[:j]) and np.all(r["norms"][j:] == 0)
    # divergence: a shifted (unstable) operator grows past the threshold
    As = A + 0.8 * np.eye(A.shape[0])      # spectral abscissa of A is -0.58
    o2 = oracle.Oracle(As, k=1, q=0, Nw=NW, w=W, TL=0.0, dt=DT, transient_removal=False)
    free = o2.trans_run(q0, periods=3, save_mod=50, div_val=np.inf, con_val=0.0)
    grow = free["norms"]
    jd = int(np.argmax(grow > 3.0))
    assert jd > 0
    r = o2.trans_run(q0, periods=3, save_mod=50, div_val=3.0, con_val=0.0)
    assert r["reason"] == 2 and r["stop"] == 50 * jd
    assert np.all(r["norms"][jd:] == 0)


def _synthetic(Qt, qss, p, w, dT):
    skip = 1 if p == 1 else 0
    ns = Qt.shape[1] - skip
    out = []
    for iw in range(NW // 2):
        ph = np.exp(1j * (w * iw) * np.arange(ns) * dT)
        out.append(Qt[:, skip:] + np.outer(qss, ph))
    return out


@pytest.mark.parametrize("p", [1, 2])
def test_removal_estimate_fft_and_lstsq(p):
    A, o = _setup()
    N = A.shape[0]
    rng = np.random.default_rng(p)
    Qt = (rng.standard_normal((N, NW)) + 1j * rng.standard_normal((N, NW))) * 0.1
    qss = -_q0(N, 7)
    ini, upd = o.trans_est(Qt, qss, p)
    Qhat = np.fft.fft(Qt, axis=1) / NW
    assert np.max(np.abs(ini - np.linalg.norm(Qhat[:, :NW // 2], axis=0))) < 1e-13
    dT = 2 * np.pi / W / NW
    for iw, S in enumerate(_synthetic(Qt, qss, p, W, dT)):
        x = np.linalg.lstsq(S, qss, rcond=None)[0]
        res = np.linalg.norm(S @ x - qss) / np.linalg.norm(qss)
        assert abs(upd[iw] - res) < 1e-12 * max(1.0, res) + 1e-14


def _slow_operator(N, seed=5):
    """Normal operator with slowly decaying modes (-0.02 .. -0.12) at
    frequencies in [-2, 2]: its trajectories keep many directions, so the
    synthetic snapshot sets leave a residual well above rounding."""
    rng = np.random.default_rng(seed)
    Qm, _ = np.linalg.qr(rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N)))
    lam = -0.02 - 0.1 * rng.random(N) + 1j * rng.uniform(-2, 2, N)
    return (Qm * lam) @ Qm.conj().T


def test_run_estimates_use_the_last_period():
    import scipy.sparse as sp
    A = _slow_operator(40)
    o = oracle.Oracle(sp.csr_matrix(A), k=1, q=0, Nw=NW, w=W, TL=0.0, dt=DT, transient_removal=False)
    q0 = _q0(40, 3)
    r = o.trans_run(q0, periods=2, save_mod=10**9, est=True, con_val=0.0)
    assert r["n_est"] == 2
    P = _step_poly(A, r["dt"])
    ntp, delta = r["Nt_period"], r["Nt_period"] // NW
    traj = {0: q0}
    y = q0.copy()
    for i in range(1, 2 * ntp + 1):
        y = P @ y
        if i % delta == 0:
            traj[i] = y.copy()
    for p in (1, 2):
        Qt = np.stack([traj[(p - 1) * ntp + j * delta] for j in range(NW)], axis=1)
        ini, upd = o.trans_est(Qt, -q0, p)
        assert np.min(upd) > 1e-6
        assert np.max(np.abs(r["init"][p - 1] - ini) / ini) < 1e-10
        assert np.max(np.abs(r["upd"][p - 1] - upd) / upd) < 1e-9
