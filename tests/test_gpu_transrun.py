"""The transient-run mode (TransientRunRK4 / TransientRemovalEst / TSTransRK4,
SourceCode/TransientRunRK4.c:39-227, SourceCode/TransientRemovalEst.c:26-127,
SourceCode/TSTransRK4.c:15-36) on the GPU against the oracle, through the C
ABI: the files it writes (norm history, saved and last snapshots, per-period
removal estimates) at 1e-10 relative, the stop rules, and the initial-vector
file path.  Files are read back with the tests' own PETSc-binary reader."""
import os

import numpy as np
import pytest
import scipy.sparse as sp

import kat
import oracle
from petsc_bin import read_vec

pytestmark = pytest.mark.gpu
W, NW, DT = 0.3, 8, 0.02


def _slow_operator(N, seed=5):
    rng = np.random.default_rng(seed)
    Qm, _ = np.linalg.qr(rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N)))
    lam = -0.02 - 0.1 * rng.random(N) + 1j * rng.uniform(-2, 2, N)
    return sp.csr_matrix((Qm * lam) @ Qm.conj().T)


def _q0(N, seed):
    rng = np.random.default_rng(seed)
    q = rng.random(N) + 1j * rng.random(N)
    return q / np.linalg.norm(q)


def _rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b))


@pytest.mark.parametrize("direct", [True, False])
@pytest.mark.parametrize("form", [0, 1])
def test_transient_run_vs_oracle(pkg, tmp_path, direct, form):
    print(f"transient run direct={direct} form={form}", flush=True)
    A = _slow_operator(40)
    q0 = _q0(40, 3)
    o = oracle.Oracle(A, k=1, q=0, Nw=NW, w=W, TL=0.0, dt=DT, transient_removal=False)
    ref = o.trans_run(q0, periods=2, save_mod=100, est=True, save=True, con_val=1e-8, direct=direct)
    s = pkg.Session(A, k=1, q=0, Nw=NW, w=W, TL=0.0, dt=DT, transient_removal=False, rk4_form=form)
    dt = s.transient_run(tmp_path, periods=2, save_mod=100, est=True, save=True, q0=q0, direct=direct)
    s.close()
    assert dt == ref["dt"]
    norms = read_vec(tmp_path / "q_transient_norms")
    assert np.all(norms.imag == 0)
    assert _rel(norms.real, ref["norms"]) < 1e-10
    assert _rel(read_vec(tmp_path / "q_transient_last_snapshot"), ref["last"]) < 1e-10
    for pos in range(ref["Nt_saved"]):
        assert _rel(read_vec(tmp_path / f"q_transient_{pos}"), ref["snaps"][pos]) < 1e-10
    for p in (1, 2):
        ini = read_vec(tmp_path / f"Initial_transient_norm_period_{p}").real
        upd = read_vec(tmp_path / f"Updated_transient_norm_period_{p}").real
        assert _rel(ini, ref["init"][p - 1]) < 1e-10
        assert _rel(upd, ref["upd"][p - 1]) < 1e-9
    assert not os.path.exists(tmp_path / "Initial_transient_norm_period_3")


def test_transient_run_stops_and_random_start(pkg, tmp_path):
    """Convergence stop on the Ginzburg-Landau operator, and the seeded random
    start (uniform entries, normalised: the first recorded norm is 1)."""
    A = kat.gl_operator(N=48, L=20.0)
    o = oracle.Oracle(A, k=1, q=0, Nw=NW, w=W, TL=0.0, dt=DT, transient_removal=False)
    s = pkg.Session(A, k=1, q=0, Nw=NW, w=W, TL=0.0, dt=DT, transient_removal=False)
    q0 = _q0(48, 9)
    free = o.trans_run(q0, periods=3, save_mod=50, con_val=0.0)
    cv = 0.5 * (free["norms"][5] + free["norms"][4])
    ref = o.trans_run(q0, periods=3, save_mod=50, con_val=cv)
    assert ref["reason"] == 3
    s.transient_run(tmp_path / "a", periods=3, save_mod=50, con_val=cv, q0=q0)
    g = read_vec(tmp_path / "a" / "q_transient_norms").real
    assert np.array_equal(g == 0, ref["norms"] == 0)
    assert _rel(g, ref["norms"]) < 1e-10
    assert _rel(read_vec(tmp_path / "a" / "q_transient_last_snapshot"), ref["last"]) < 1e-10
    s.transient_run(tmp_path / "b", periods=1, save_mod=50, seed=11)
    g = read_vec(tmp_path / "b" / "q_transient_norms").real
    assert abs(g[0] - 1) < 1e-14 and np.all(np.isfinite(g)) and 0 < g[-1] < 1e-3
    s.close()


def test_trans_removal_est_entry_point(pkg, tmp_path):
    """TransientRemovalEst called directly on random snapshots, both period kinds."""
    A = _slow_operator(40)
    o = oracle.Oracle(A, k=1, q=0, Nw=NW, w=W, TL=0.0, dt=DT, transient_removal=False)
    s = pkg.Session(A, k=1, q=0, Nw=NW, w=W, TL=0.0, dt=DT, transient_removal=False)
    rng = np.random.default_rng(4)
    Qt = 0.1 * (rng.standard_normal((40, NW)) + 1j * rng.standard_normal((40, NW)))
    qss = -_q0(40, 8)
    for p in (1, 2):
        ini, upd = s.trans_removal_est(Qt, qss, p, tmp_path)
        ri, ru = o.trans_est(Qt, qss, p)
        assert _rel(ini, ri) < 1e-12 and _rel(upd, ru) < 1e-10
    s.close()
