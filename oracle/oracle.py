"""ctypes front-end of the CPU FP64 oracle (oracle/rsvdt_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and
by bench.py's cpu_baseline leg, always as the checker or the CPU baseline and
never as a stand-in for the HIP path.

Array conventions follow the reference's PETSc MATDENSE (column-major):
a sketch/response matrix is numpy complex128 of shape (N, k*Nw_eff) whose
column m*k + ik is frequency m of test vector ik (frequency-major, as
SourceCode/CreateRandomMat.c and SourceCode/DFT.c leave it).
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle_rsvdt.so")
_lib = None

_i64 = ctypes.c_int64
_dbl = ctypes.c_double
_int = ctypes.c_int
_vp = ctypes.c_void_p


class _Ctx(ctypes.Structure):
    _fields_ = [
        ("A_n", _i64), ("A_nnz", _i64), ("A_rowptr", _vp), ("A_col", _vp), ("A_val", _vp),
        ("AH_n", _i64), ("AH_nnz", _i64), ("AH_rowptr", _vp), ("AH_col", _vp), ("AH_val", _vp),
        ("N", _i64), ("k", _i64), ("q", _i64), ("Nw", _i64), ("Nw_eff", _i64),
        ("w", _dbl), ("real_op", _int), ("transient_removal", _int),
        ("dt", _dbl), ("TL", _dbl), ("Ns", _i64), ("Nt", _i64), ("RR", _i64),
        ("idft", _vp), ("dft", _vp), ("nthreads", _int), ("max_steps", _i64),
    ]


def build():
    """Compile the oracle with its Makefile (gcc + OpenMP)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = ctypes.CDLL(_LIB)
        P = ctypes.POINTER
        L.orc_init.argtypes = [P(_Ctx), _i64, _i64, _vp, _vp, _vp, _i64, _i64, _i64, _dbl, _dbl, _dbl, _int, _int]
        L.orc_init.restype = _int
        L.orc_free.argtypes = [P(_Ctx)]
        L.orc_action.argtypes = [P(_Ctx), _int, _vp, _vp]
        L.orc_action.restype = _int
        L.orc_action_snapshots.argtypes = [P(_Ctx), _int, _vp, _vp]
        L.orc_action_snapshots.restype = _int
        L.orc_qr_allfreqs.argtypes = [P(_Ctx), _vp]
        L.orc_qr_allfreqs.restype = _int
        L.orc_svd_before_adjoint.argtypes = [P(_Ctx), _vp]
        L.orc_svd_before_adjoint.restype = _int
        L.orc_svd_allfreqs.argtypes = [P(_Ctx), _vp, _vp, _vp]
        L.orc_svd_allfreqs.restype = _int
        L.orc_rsvdt_run.argtypes = [P(_Ctx), _vp, _vp, _vp, _vp]
        L.orc_rsvdt_run.restype = _int
        L.orc_qr.argtypes = [_i64, _i64, _vp, _vp]
        L.orc_qr.restype = _int
        L.orc_svd_left.argtypes = [_i64, _i64, _vp, _vp]
        L.orc_svd_left.restype = _int
        L.orc_lu_solve.argtypes = [_i64, _vp, _vp]
        L.orc_lu_solve.restype = _int
        L.orc_setup_grid.argtypes = [_dbl, _i64, _dbl, _dbl, P(_dbl), P(_i64), P(_i64), P(_i64)]
        L.orc_spmv_add.argtypes = [P(_Ctx), _int, _vp, _vp, _vp]
        L.orc_get_dft.argtypes = [P(_Ctx), _vp, _vp]
        L.orc_time_steps.argtypes = [P(_Ctx), _int, _vp, _i64, _i64]
        L.orc_time_steps.restype = _dbl
        L.orc_trans_est.argtypes = [P(_Ctx), _vp, _vp, _i64, _vp, _vp]
        L.orc_trans_est.restype = _int
        L.orc_trans_run.argtypes = [P(_Ctx), _int, _i64, _i64, _dbl, _dbl, _int, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                                    P(_dbl)]
        L.orc_trans_run.restype = _int
        _lib = L
    return _lib


def _ptr(a):
    return a.ctypes.data_as(_vp)


def setup_grid(w, Nw, TL, dt):
    """SetupTimeFreqGrid.c:16-21 -> (dt, Ns, Nt, RR)."""
    d = _dbl(); Ns = _i64(); Nt = _i64(); RR = _i64()
    lib().orc_setup_grid(w, Nw, TL, dt, ctypes.byref(d), ctypes.byref(Ns), ctypes.byref(Nt), ctypes.byref(RR))
    return d.value, Ns.value, Nt.value, RR.value


class Oracle:
    """One operator + run configuration (the reference's RSVDt_vars / LNS_vars)."""

    def __init__(self, A, k, q, Nw, w, TL, dt, transient_removal, nthreads=0):
        import scipy.sparse as sp
        A = sp.csr_matrix(A)
        A.sort_indices()
        self._rowptr = np.ascontiguousarray(A.indptr, dtype=np.int64)
        self._col = np.ascontiguousarray(A.indices, dtype=np.int64)
        self._val = np.ascontiguousarray(A.data, dtype=np.complex128)
        self.c = _Ctx()
        rc = lib().orc_init(ctypes.byref(self.c), A.shape[0], A.nnz, _ptr(self._rowptr), _ptr(self._col),
                            _ptr(self._val), k, q, Nw, w, TL, dt, int(bool(transient_removal)), nthreads)
        if rc:
            raise RuntimeError("orc_init failed")

    def __del__(self):
        if _lib is not None and hasattr(self, "c"):
            _lib.orc_free(ctypes.byref(self.c))

    # scalar info
    N = property(lambda s: s.c.N)
    k = property(lambda s: s.c.k)
    Nw = property(lambda s: s.c.Nw)
    Nw_eff = property(lambda s: s.c.Nw_eff)
    Ns = property(lambda s: s.c.Ns)
    Nt = property(lambda s: s.c.Nt)
    RR = property(lambda s: s.c.RR)
    dt = property(lambda s: s.c.dt)
    real_op = property(lambda s: bool(s.c.real_op))

    def set_max_steps(self, n):
        self.c.max_steps = int(n)

    def dft_mats(self):
        idft = np.zeros((self.Nw_eff, 2 * self.Ns), dtype=np.complex128, order="F")
        dft = np.zeros((self.Nw, self.Nw_eff), dtype=np.complex128, order="F")
        lib().orc_get_dft(ctypes.byref(self.c), _ptr(idft), _ptr(dft))
        return idft, dft

    def _check(self, F):
        F = np.asfortranarray(F, dtype=np.complex128)
        if F.shape != (self.N, self.k * self.Nw_eff):
            raise ValueError(f"sketch must be {self.N} x {self.k * self.Nw_eff}, got {F.shape}")
        return F

    def action(self, F, direct=True):
        F = self._check(F)
        Y = np.zeros_like(F)
        if lib().orc_action(ctypes.byref(self.c), int(direct), _ptr(F), _ptr(Y)):
            raise RuntimeError("orc_action failed")
        return Y

    def snapshots(self, F, direct=True):
        F = self._check(F)
        nstore = self.Nw + 1 if self.c.transient_removal else self.Nw
        Y = np.zeros((self.N, nstore * self.k), dtype=np.complex128, order="F")
        lib().orc_action_snapshots(ctypes.byref(self.c), int(direct), _ptr(F), _ptr(Y))
        return Y

    def qr_allfreqs(self, Y):
        Y = np.array(self._check(Y), order="F", copy=True)
        if lib().orc_qr_allfreqs(ctypes.byref(self.c), _ptr(Y)):
            raise RuntimeError("orc_qr_allfreqs failed")
        return Y

    def svd_before_adjoint(self, Y):
        Y = np.array(self._check(Y), order="F", copy=True)
        lib().orc_svd_before_adjoint(ctypes.byref(self.c), _ptr(Y))
        return Y

    def svd_allfreqs(self, Y):
        Y = self._check(Y)
        V = np.zeros_like(Y)
        S = np.zeros((self.k, self.Nw_eff), dtype=np.float64, order="F")
        lib().orc_svd_allfreqs(ctypes.byref(self.c), _ptr(Y), _ptr(V), _ptr(S))
        return V, S

    def run(self, F):
        """main.c sequence: returns (U_hat, V_hat, S_hat)."""
        F = self._check(F)
        U = np.zeros_like(F); V = np.zeros_like(F)
        S = np.zeros((self.k, self.Nw_eff), dtype=np.float64, order="F")
        if lib().orc_rsvdt_run(ctypes.byref(self.c), _ptr(F), _ptr(U), _ptr(V), _ptr(S)):
            raise RuntimeError("orc_rsvdt_run failed")
        return U, V, S

    # ---- weights and input/output matrices (numpy products around the C actions)
    def _weights(self, Y, direct, before, W):
        """ApplyWeightMats.c:16-80: direct before W_f^{-1/2} then B, after C then
        W_q^{1/2}; adjoint before W_q^{1/2,H} then C^H, after B^H then W_f^{-1/2,H}."""
        H = lambda M: M.conj().T
        if direct:
            seq = [("Wf_sqrt_inv", False), ("B", False)] if before else [("C", False), ("Wq_sqrt", False)]
        else:
            seq = [("Wq_sqrt", True), ("C", True)] if before else [("B", True), ("Wf_sqrt_inv", True)]
        for name, herm in seq:
            M = W.get(name)
            if M is not None:
                M = np.asarray(M.toarray() if hasattr(M, "toarray") else M, dtype=np.complex128)
                Y = (H(M) if herm else M) @ Y
        return np.asfortranarray(Y)

    def _blocks(self, Y, fn):
        k = self.k
        out = np.array(Y, dtype=np.complex128, order="F", copy=True)
        for m in range(self.Nw_eff):
            out[:, m * k:(m + 1) * k] = fn(Y[:, m * k:(m + 1) * k])
        return out

    def _act(self, Y, W, direct):
        Y = self._weights(Y, direct, True, W)
        Y = self.action(Y, direct=direct)
        return self._weights(Y, direct, False, W)

    def run_weighted(self, F, **W):
        """main.c:134-144 with ApplyWeightMats (W: any of B (N x Nb), C (Nc x N),
        Wq_sqrt and Wq_sqrt_inv (Nc x Nc), Wf_sqrt_inv (Nb x Nb)): F is
        Nb x k*Nw_eff; returns (U_hat Nc x k*Nw_eff, V_hat Nb x k*Nw_eff, S_hat).
        QRAllFreqs / SVDAllFreqs* act per frequency block with orc_qr /
        orc_svd_left; SVDAllFreqs.c:63-64 scales U by W_q^{-1/2} and V by W_f^{-1/2}."""
        Y = self._act(np.asfortranarray(F, dtype=np.complex128), W, True)
        for _ in range(self.c.q):
            Y = self._act(self._blocks(Y, lambda b: qr(b)[0]), W, False)
            Y = self._act(self._blocks(Y, lambda b: qr(b)[0]), W, True)
        U = self._blocks(Y, lambda b: svd_left(b)[0])                  # SVDAllFreqsBeforeAdjoint, StoreU
        Z = self._act(U, W, False)
        S = np.zeros((self.k, self.Nw_eff))
        V = np.array(Z, order="F", copy=True)
        for m in range(self.Nw_eff):
            Vm, sm = svd_left(Z[:, m * self.k:(m + 1) * self.k])
            V[:, m * self.k:(m + 1) * self.k] = Vm
            S[:, m] = sm
        if W.get("Wq_sqrt_inv") is not None:
            U = np.asarray(W["Wq_sqrt_inv"].toarray() if hasattr(W["Wq_sqrt_inv"], "toarray") else W["Wq_sqrt_inv"]) @ U
        if W.get("Wf_sqrt_inv") is not None:
            V = np.asarray(W["Wf_sqrt_inv"].toarray() if hasattr(W["Wf_sqrt_inv"], "toarray") else W["Wf_sqrt_inv"]) @ V
        return U, V, S

    def time_steps(self, Fk, nvec, nsteps, direct=True):
        """Wall seconds of nvec x nsteps reference-structured RK4 steps (CPU baseline sample)."""
        Fk = np.asfortranarray(Fk, dtype=np.complex128)
        assert Fk.shape == (self.N, self.Nw_eff)
        return lib().orc_time_steps(ctypes.byref(self.c), int(direct), _ptr(Fk), nvec, nsteps)

    def trans_est(self, Qt, qss, period_index):
        """TransientRemovalEst.c:26-118 -> (initial norms, updated norms), Nw/2 each."""
        Qt = np.asfortranarray(Qt, dtype=np.complex128)
        assert Qt.shape == (self.N, self.Nw)
        qss = np.ascontiguousarray(qss, dtype=np.complex128)
        a, b = np.zeros(self.Nw // 2), np.zeros(self.Nw // 2)
        if lib().orc_trans_est(ctypes.byref(self.c), _ptr(Qt), _ptr(qss), period_index, _ptr(a), _ptr(b)):
            raise RuntimeError("orc_trans_est failed")
        return a, b

    def trans_run(self, q0, periods=1, save_mod=100, div_val=1e3, con_val=1e-8, est=False, save=False,
                  direct=True):
        """TransientRunRK4.c:39-190.  Returns a dict: norms (Nt_saved), last (N),
        snaps (Nt_saved x N, when save), init/upd (periods x Nw/2, when est),
        stop, n_est, Nt_saved, Nt_period, reason (0 none, 1 NaN, 2 diverged,
        3 converged), dt."""
        q0 = np.ascontiguousarray(q0, dtype=np.complex128)
        T = 2 * np.pi / self.c.w
        dT = T / self.Nw
        dt2 = dT / np.ceil(dT / self.dt)
        ntp = int(round(T / dt2))
        nsv = (ntp * periods) // save_mod + 1
        norms = np.zeros(nsv)
        last = np.zeros(self.N, np.complex128)
        snaps = np.zeros((nsv, self.N), np.complex128) if save else None
        ini = np.zeros((periods, self.Nw // 2)) if est else None
        upd = np.zeros((periods, self.Nw // 2)) if est else None
        info = np.zeros(5, np.int64)
        d = _dbl()
        rc = lib().orc_trans_run(ctypes.byref(self.c), int(direct), periods, save_mod, div_val, con_val, int(est),
                                 _ptr(q0), _ptr(norms), _ptr(last), _ptr(snaps) if save else None,
                                 _ptr(ini) if est else None, _ptr(upd) if est else None, _ptr(info), ctypes.byref(d))
        if rc:
            raise RuntimeError("orc_trans_run failed")
        assert info[2] == nsv and info[3] == ntp
        return dict(norms=norms, last=last, snaps=snaps, init=ini, upd=upd, stop=int(info[0]), n_est=int(info[1]),
                    Nt_saved=nsv, Nt_period=ntp, reason=int(info[4]), dt=d.value)

    def spmv_add(self, x, f=None, direct=True):
        x = np.ascontiguousarray(x, dtype=np.complex128)
        y = np.zeros_like(x)
        fp = _ptr(np.ascontiguousarray(f, dtype=np.complex128)) if f is not None else None
        lib().orc_spmv_add(ctypes.byref(self.c), int(direct), _ptr(x), fp, _ptr(y))
        return y


def qr(a):
    a = np.array(a, dtype=np.complex128, order="F", copy=True)
    m, n = a.shape
    r = np.zeros((n, n), dtype=np.complex128, order="F")
    if lib().orc_qr(m, n, _ptr(a), _ptr(r)):
        raise RuntimeError("orc_qr failed")
    return a, r


def svd_left(a):
    a = np.array(a, dtype=np.complex128, order="F", copy=True)
    m, n = a.shape
    s = np.zeros(n)
    if lib().orc_svd_left(m, n, _ptr(a), _ptr(s)):
        raise RuntimeError("orc_svd_left failed")
    return a, s


def lu_solve(a, b):
    a = np.array(a, dtype=np.complex128, order="F", copy=True)
    b = np.array(b, dtype=np.complex128, copy=True)
    if lib().orc_lu_solve(a.shape[0], _ptr(a), _ptr(b)):
        raise RuntimeError("orc_lu_solve: singular")
    return b
