"""End-to-end parity of the main.c sequence (SourceCode/main.c:134-146) on
the GPU with the oracle, through the Python mirror of the entry points and
through the rsvdt command-line driver (variables.yaml + PETSc-binary I/O)."""
import os
import subprocess

import numpy as np
import pytest

import kat
import oracle
import petsc_bin
from conftest import crandn, rel, subspace_gap

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("q,tr,cplx", [(0, True, True), (1, True, True), (1, False, False)])
def test_run_vs_oracle(pkg, q, tr, cplx):
    print(f"e2e q={q} TR={tr} complex={cplx}", flush=True)
    N, k, Nw = 80, 4, 16
    A = kat.gl_operator(N=N, L=20.0)
    if not cplx:
        A = A.real.astype(complex)
    rng = np.random.default_rng(q + 3)
    o = oracle.Oracle(A, k=k, q=q, Nw=Nw, w=0.3, TL=2.0, dt=0.02, transient_removal=tr)
    F = crandn(rng, N, k * o.Nw_eff)
    Uo, Vo, So = o.run(F)
    s = pkg.Session(A, k=k, q=q, Nw=Nw, w=0.3, TL=2.0, dt=0.02, transient_removal=tr)
    Ug, Vg, Sg = s.run(F)
    s.close()
    assert np.allclose(Sg, So, rtol=1e-10)
    for m in range(o.Nw_eff):
        for j in range(2):     # leading modes are well separated
            for X, Xo in ((Ug, Uo), (Vg, Vo)):
                c = np.vdot(X[:, m * k + j], Xo[:, m * k + j])
                assert abs(abs(c) - 1) < 1e-9


def test_full_sketch_exact_gains(pkg):
    N, Nw, w, dt = 12, 11, 0.4, 0.02
    A = kat.gl_operator(N=N, L=6.0)
    rng = np.random.default_rng(1)
    s = pkg.Session(A, k=N, q=0, Nw=Nw, w=w, TL=0.0, dt=dt, transient_removal=True)
    U, V, S = s.run(crandn(rng, N, N * s.Nw_eff))
    s.close()
    assert np.allclose(S, kat.exact_transfer_sigmas(A, Nw, w, dt), rtol=1e-10)


@pytest.mark.parametrize("opt", [2, 1])
def test_cli_drop_in(pkg, tmp_path, opt):
    """rsvdt -inputs variables.yaml with InputForcingFlg, PETSc-binary in/out;
    results per frequency (SaveResultsOpt 2: U_hat_Freq<m>_allK, 0-based) or
    per mode (1: U_hat_k<j>_allFreqs, 1-based), SourceCode/SaveResults.c:32-92."""
    N, k, Nw = 60, 3, 8
    A = kat.gl_operator(N=N, L=20.0)
    rng = np.random.default_rng(4)
    F = crandn(rng, N, k * Nw)
    petsc_bin.write_mat(str(tmp_path / "A"), A)
    petsc_bin.write_mat(str(tmp_path / "F_hat"), F)
    (tmp_path / "res").mkdir()
    yaml = f"""RootDir: {tmp_path}/
ResultsDir: res/
OperatorDir: A
k: 99
q: 1
Nw: {Nw}
w: 0.3
TwoPI: false
TransientLength: 2
dt: 0.02
TransientRemoval: true
Display: 1
SaveResultsOpt: {opt}
InputForcingFlg: true
InputForcingDir: F_hat
TransICDir: none
"""
    (tmp_path / "vars.yaml").write_text(yaml)
    r = subprocess.run([pkg.BINPATH, "-inputs", "vars.yaml"], cwd=str(tmp_path), capture_output=True, text=True,
                       timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    assert "Based on the input forcing: k = 3" in r.stdout
    out = tmp_path / "res" / "ResolventModes_0"
    S = petsc_bin.read_dense(str(out / "S_hat")).real
    o = oracle.Oracle(A, k=k, q=1, Nw=Nw, w=0.3, TL=2.0, dt=0.02, transient_removal=True)
    Uo, Vo, So = o.run(F)
    assert S.shape == (k, Nw)
    assert np.allclose(S, So, rtol=1e-10)
    Nwe = o.Nw_eff
    if opt == 2:
        for m in (0, Nwe - 1):
            Um = petsc_bin.read_dense(str(out / f"U_hat_Freq{m}_allK"))
            Vm = petsc_bin.read_dense(str(out / f"V_hat_Freq{m}_allK"))
            assert Um.shape == (N, k) and Vm.shape == (N, k)
            for j in range(k):
                assert abs(abs(np.vdot(Um[:, j], Uo[:, m * k + j])) - 1) < 1e-9
                assert abs(abs(np.vdot(Vm[:, j], Vo[:, m * k + j])) - 1) < 1e-9
    else:
        for j in range(k):
            Uj = petsc_bin.read_dense(str(out / f"U_hat_k{j + 1}_allFreqs"))
            Vj = petsc_bin.read_dense(str(out / f"V_hat_k{j + 1}_allFreqs"))
            assert Uj.shape == (N, Nwe) and Vj.shape == (N, Nwe)
            for m in range(Nwe):
                assert abs(abs(np.vdot(Uj[:, m], Uo[:, m * k + j])) - 1) < 1e-9
                assert abs(abs(np.vdot(Vj[:, m], Vo[:, m * k + j])) - 1) < 1e-9
    assert (out / "vars.yaml").read_text() == yaml


def test_cli_transient_run(pkg, tmp_path):
    """rsvdt -inputs variables.yaml with TransRun (SourceCode/main.c:119-124):
    the initial vector read from TransICDir, norm monitor every TransSaveMod
    steps, saved snapshots and the removal estimate of each period, against the
    oracle; the outputs land in TransientSnapshots_0 with the inputs copy."""
    N, Nw = 60, 8
    A = kat.gl_operator(N=N, L=20.0)
    rng = np.random.default_rng(12)
    q0 = rng.random(N) + 1j * rng.random(N)
    petsc_bin.write_mat(str(tmp_path / "A"), A)
    (tmp_path / "res").mkdir()
    petsc_bin.write_vec(str(tmp_path / "res" / "q0"), q0)
    yaml = f"""RootDir: {tmp_path}/
ResultsDir: res/
OperatorDir: A
k: 2
q: 0
Nw: {Nw}
w: 0.3
TransientLength: 2
dt: 0.02
TransRun: true
TransICFlg: true
TransICDir: q0
TransSave: true
TransSaveMod: 200
TransPeriods: 2
TransRemovalEst: true
TransConVal: 1e-30
"""
    (tmp_path / "vars.yaml").write_text(yaml)
    r = subprocess.run([pkg.BINPATH, "-inputs", "vars.yaml"], cwd=str(tmp_path), capture_output=True, text=True,
                       timeout=120)
    print(r.stdout[-3000:], r.stderr)
    assert r.returncode == 0
    out = tmp_path / "res" / "TransientSnapshots_0"
    o = oracle.Oracle(A, k=2, q=0, Nw=Nw, w=0.3, TL=2.0, dt=0.02, transient_removal=False)
    ref = o.trans_run(q0, periods=2, save_mod=200, con_val=1e-30, est=True, save=True)
    norms = petsc_bin.read_vec(str(out / "q_transient_norms")).real
    assert np.max(np.abs(norms - ref["norms"]) / ref["norms"]) < 1e-10
    last = petsc_bin.read_vec(str(out / "q_transient_last_snapshot"))
    assert np.linalg.norm(last - ref["last"]) / np.linalg.norm(ref["last"]) < 1e-10
    for pos in range(ref["Nt_saved"]):
        assert (out / f"q_transient_{pos}").exists()
    for p in (1, 2):
        upd = petsc_bin.read_vec(str(out / f"Updated_transient_norm_period_{p}")).real
        assert np.allclose(upd, ref["upd"][p - 1], rtol=1e-8, atol=1e-12)
    assert "Total integration time" in r.stdout and "DONE" in r.stdout
    assert (out / "vars.yaml").read_text() == yaml
