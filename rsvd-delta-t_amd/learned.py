"""Learned sketch for RSVD-Delta-t on a parameterised operator family.

A coordinate MLP maps (x, w, mu0, Re U) -- a grid point, a frequency and the
family parameters of a linearised Ginzburg-Landau operator -- to 8 complex
forcing values; evaluated on the grid at every frequency it gives 8 columns
per frequency block of the initial sketch, and the remaining k - 8 columns
stay random (CreateRandomMat, SourceCode/CreateRandomMat.c:33-39, through
RSVDT_SetSketchColumns).  The network is trained on the leading forcing
modes (right singular vectors, V_hat of SVDAllFreqs) that plain RSVD-Delta-t
computes for sampled family members, with a phase- and scale-invariant
alignment loss, by the hand-written BF16 MFMA kernels of csrc/hip/mlp.hip
(forward, backward, loss gradient, Adam).

Family: A = -U d/dx + gamma d2/dx2 + mu(x) on N = 500 points of [-85, 85],
U = Re U + 0.2i, gamma = 1 - i, mu(x) = mu0 - U^2/(4 gamma) + mu2 x^2/2,
mu2 = -0.01, with mu0 in [0.1, 0.4] and Re U in [1.5, 2.5]; Nw = 64,
w = 0.05, dt = T/4096, TransientLength 10, transient removal on.

  python learned.py --train 200 --eval 50      (the full evaluation)
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

LAYERS = [(64, 256), (256, 256), (256, 256), (256, 16)]
NCOL = 8                    # learned columns per frequency block
XS, KMAX = 85.0, 2.0        # x scale and largest Fourier wavenumber of the features
FAMILY = dict(N=500, Nw=64, w=0.05, TL=10.0, mu0=(0.1, 0.4), Ur=(1.5, 2.5))


def _pkg():
    import importlib.util
    if "rsvdt_amd" in sys.modules:
        return sys.modules["rsvdt_amd"]
    spec = importlib.util.spec_from_file_location("rsvdt_amd", os.path.join(HERE, "__init__.py"))
    mod = importlib.util.module_from_spec(spec)
    sys.modules["rsvdt_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


class _Buf:
    """Device buffer through the engine's C ABI."""

    def __init__(self, L, nbytes):
        self.L = L
        self.p = ctypes.c_void_p()
        self.nbytes = int(nbytes)
        if L.rsvdt_hip_malloc(ctypes.byref(self.p), ctypes.c_size_t(max(self.nbytes, 16))) != 0:
            raise MemoryError(f"device allocation of {self.nbytes} bytes failed")

    def put(self, arr, offset_bytes=0):
        a = np.ascontiguousarray(arr)
        assert offset_bytes + a.nbytes <= self.nbytes
        dst = ctypes.c_void_p(self.p.value + int(offset_bytes))
        _check(self.L.rsvdt_hip_memcpy_h2d(dst, a.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(a.nbytes)))

    def get(self, dtype, shape, offset_bytes=0):
        out = np.empty(shape, dtype=dtype)
        src = ctypes.c_void_p(self.p.value + int(offset_bytes))
        _check(self.L.rsvdt_hip_memcpy_d2h(out.ctypes.data_as(ctypes.c_void_p), src, ctypes.c_size_t(out.nbytes)))
        return out

    def at(self, offset_bytes):
        return ctypes.c_void_p(self.p.value + int(offset_bytes))

    def __del__(self):
        try:
            self.L.rsvdt_hip_free(self.p)
        except Exception:
            pass


def _check(rc):
    if rc:
        raise RuntimeError(f"device call failed with code {rc}: {_pkg().lib().rsvdt_hip_last_error().decode()}")


def _bind(L):
    V, I64, D, F = ctypes.c_void_p, ctypes.c_int64, ctypes.c_double, ctypes.c_float
    L.rsvdt_hip_malloc.argtypes = [ctypes.POINTER(V), ctypes.c_size_t]
    L.rsvdt_hip_memcpy_h2d.argtypes = [V, V, ctypes.c_size_t]
    L.rsvdt_hip_memcpy_d2h.argtypes = [V, V, ctypes.c_size_t]
    L.rsvdt_hip_free.argtypes = [V]
    L.rsvdt_hip_last_error.restype = ctypes.c_char_p
    L.rsvdt_hip_mlp_features.argtypes = [I64, V, D, D, D, V, V]
    L.rsvdt_hip_gemm_bf16.argtypes = [I64, I64, I64, V, I64, I64, V, I64, I64, V, I64, ctypes.c_int, V, V, I64, V]
    L.rsvdt_hip_colsum.argtypes = [I64, I64, V, I64, V, V]
    L.rsvdt_hip_sketch_loss.argtypes = [I64, I64, I64, V, V, V, V, D, V]
    L.rsvdt_hip_adam.argtypes = [I64, V, V, V, V, F, F, F, F, F, F, V]
    L.RSVDT_SetSketchColumns.argtypes = [I64, I64, I64, V]
    L.RSVDT_SetSketchColumnsDevice.argtypes = [I64, I64, I64, V, I64]


def family_operator(mu0, Ur, N=500):
    sys.path.insert(0, HERE)
    import workloads
    return workloads.gl(N, U=Ur + 0.2j, mu0=mu0)


def omegas(Nw, w, real_op):
    """Angular frequency of each frequency block (SetupTimeFreqGrid aliasing:
    blocks m >= Nw/2 of a complex operator are the negative frequencies)."""
    m = np.arange(Nw // 2 if real_op else Nw)
    return np.where(m < Nw // 2, m, m - Nw) * w


class SketchNet:
    """The coordinate MLP, with FP32 master weights on the device."""

    def __init__(self, seed=0, batch_points=16000):
        self.pkg = _pkg()
        self.L = self.pkg.lib()
        _bind(self.L)
        self.L.rsvdt_hip_init(0)
        rng = np.random.default_rng(seed)
        self.offsets = []
        off = 0
        host = []
        for i, o in LAYERS:
            W = (rng.standard_normal((o, i)) * np.sqrt(2.0 / i)).astype(np.float32)
            b = np.zeros(o, np.float32)
            self.offsets.append((off, off + W.size))
            off += W.size + o
            host += [W.ravel(), b]
        self.nparam = off
        flat = np.concatenate(host).astype(np.float32)
        self.w = _Buf(self.L, 4 * off)
        self.g = _Buf(self.L, 4 * off)
        self.m = _Buf(self.L, 4 * off)
        self.v = _Buf(self.L, 4 * off)
        self.w.put(flat)
        self.m.put(np.zeros(off, np.float32))
        self.v.put(np.zeros(off, np.float32))
        self.t = 0
        self._alloc(batch_points)

    def _alloc(self, B):
        self.B = B
        self.q = _Buf(self.L, 32 * B)
        self.feat = _Buf(self.L, 4 * 64 * B)
        self.Z = [_Buf(self.L, 4 * 256 * B) for _ in range(3)]
        self.A = [_Buf(self.L, 4 * 256 * B) for _ in range(3)]
        self.out = _Buf(self.L, 4 * 16 * B)
        self.dout = _Buf(self.L, 4 * 16 * B)
        self.dZ = [_Buf(self.L, 4 * 256 * B) for _ in range(3)]
        self.loss = _Buf(self.L, 4 * max(B, 1))

    def _wptr(self, li, buf=None):
        buf = buf or self.w
        return buf.at(4 * self.offsets[li][0]), buf.at(4 * self.offsets[li][1])

    def params_host(self):
        flat = self.w.get(np.float32, (self.nparam,))
        out = []
        for (i, o), (a, b) in zip(LAYERS, self.offsets):
            out.append((flat[a:a + o * i].reshape(o, i).astype(np.float64), flat[b:b + o].astype(np.float64)))
        return out

    def save(self, path):
        """FP32 weights of the trained network (numpy .npz, no pickled objects)."""
        np.savez(path, w=self.w.get(np.float32, (self.nparam,)), layers=np.array(LAYERS, np.int64))

    def load(self, path):
        with np.load(path, allow_pickle=False) as z:
            if not np.array_equal(z["layers"], np.array(LAYERS, np.int64)) or z["w"].shape != (self.nparam,):
                raise ValueError(f"{path}: weights do not match the network layout")
            self.w.put(np.ascontiguousarray(z["w"], dtype=np.float32))
        return self

    def grads_host(self):
        flat = self.g.get(np.float32, (self.nparam,))
        out = []
        for (i, o), (a, b) in zip(LAYERS, self.offsets):
            out.append((flat[a:a + o * i].reshape(o, i).astype(np.float64), flat[b:b + o].astype(np.float64)))
        return out

    # -------------------------------------------------------------- passes
    def forward_dev(self, n, q_dev):
        """Network outputs (n x 16 FP32, device) for n points q (n x 4 FP64, device)."""
        L = self.L
        assert n <= self.B
        _check(L.rsvdt_hip_mlp_features(n, q_dev, XS, 1.0, KMAX, self.feat.p, None))
        a_in, ld_in = self.feat.p, 64
        for li, (i, o) in enumerate(LAYERS):
            W, b = self._wptr(li)
            last = li == len(LAYERS) - 1
            C = self.out.p if last else self.A[li].p
            Z = None if last else self.Z[li].p
            _check(L.rsvdt_hip_gemm_bf16(n, o, i, a_in, ld_in, 1, W, 1, i, C, o, 2 if last else 1, b, Z, o, None))
            a_in, ld_in = C, o
        return self.out

    def backward_dev(self, n):
        """Gradients into self.g from self.dout (n x 16)."""
        L = self.L
        dz, ldz = self.dout.p, 16
        for li in range(len(LAYERS) - 1, -1, -1):
            i, o = LAYERS[li]
            a_in = self.feat.p if li == 0 else self.A[li - 1].p
            gW, gb = self._wptr(li, self.g)
            W, _ = self._wptr(li)
            # dW = dz^T a_in (o x i, K = n); db = column sums of dz
            _check(L.rsvdt_hip_gemm_bf16(o, i, n, dz, 1, ldz, a_in, i, 1, gW, i, 0, None, None, 0, None))
            _check(L.rsvdt_hip_colsum(n, o, dz, ldz, gb, None))
            if li > 0:
                # dz_prev = (dz W) * silu'(z_prev)
                _check(L.rsvdt_hip_gemm_bf16(n, i, o, dz, ldz, 1, W, i, 1, self.dZ[li - 1].p, i, 3, None,
                                             self.Z[li - 1].p, i, None))
                dz, ldz = self.dZ[li - 1].p, i

    def adam_step(self, lr, b1=0.9, b2=0.999, eps=1e-8):
        self.t += 1
        _check(self.L.rsvdt_hip_adam(self.nparam, self.w.p, self.g.p, self.m.p, self.v.p, lr, b1, b2, eps,
                                     1 - b1 ** self.t, 1 - b2 ** self.t, None))

    # -------------------------------------------------------------- training
    def train(self, data, epochs=20, groups_per_batch=32, lr=1e-3, seed=0, log=None):
        """data: dict with q (G x npts x 4), V (G x npts x NCOL complex), groups
        in random order.  Returns the per-epoch mean loss."""
        Q, Vt = data["q"], data["V"]
        G, npts = Q.shape[0], Q.shape[1]
        gb = min(groups_per_batch, G)
        B = gb * npts
        if B > self.B:
            self._alloc(B)
        qd = _Buf(self.L, Q.nbytes)
        qd.put(Q.astype(np.float64))
        Vd = _Buf(self.L, Vt.nbytes)
        Vd.put(Vt.astype(np.complex128))
        nb = G // gb
        rng = np.random.default_rng(seed)
        hist = []
        for ep in range(epochs):
            tot = 0.0
            for bi in rng.permutation(nb):
                g0 = int(bi) * gb
                qp = qd.at(g0 * npts * 32)
                vp = Vd.at(g0 * npts * NCOL * 16)
                self.forward_dev(B, qp)
                _check(self.L.rsvdt_hip_sketch_loss(gb, npts, NCOL, self.out.p, vp, self.dout.p, self.loss.p,
                                                    1.0 / (gb * NCOL), None))
                self.backward_dev(B)
                lr_t = lr * 0.5 * (1 + np.cos(np.pi * (ep + 0.5) / epochs))
                self.adam_step(lr_t)
                tot += float(self.loss.get(np.float32, (gb,)).sum())
            hist.append(tot / nb)
            if log:
                log(f"epoch {ep}: alignment loss {hist[-1]:.4f}")
        return hist

    # -------------------------------------------------------------- inference
    def _infer(self, mu0, Ur, x, om):
        n = len(x) * len(om)
        if n > self.B:
            self._alloc(n)
        q = np.zeros((len(om), len(x), 4))
        q[:, :, 0] = x[None, :]
        q[:, :, 1] = om[:, None]
        q[:, :, 2] = mu0
        q[:, :, 3] = Ur
        self.q.put(q.reshape(n, 4))
        return n, self.forward_dev(n, self.q.p)

    def columns(self, mu0, Ur, x, om):
        """Learned columns (len(om) x len(x) x NCOL complex) for one family member."""
        n, out = self._infer(mu0, Ur, x, om)
        o = out.get(np.float32, (n, 16)).astype(np.float64)
        c = o[:, 0::2] + 1j * o[:, 1::2]
        return c.reshape(len(om), len(x), NCOL)

    def columns_dev(self, mu0, Ur, x, om):
        """The same left on the device in the network's output layout (FP32
        interleaved, point-major): (pointer, ld, Nw_eff, N) for
        RSVDT_SetSketchColumnsDevice; valid until the next forward pass."""
        _, out = self._infer(mu0, Ur, x, om)
        return out.p, 16, len(om), len(x)


# ------------------------------------------------------------------ family runs
def _session(pkg, A, k, q, **opts):
    Nw, w = FAMILY["Nw"], FAMILY["w"]
    T = 2 * np.pi / w
    return pkg.Session(A, k=k, q=q, Nw=Nw, w=w, TL=FAMILY["TL"], dt=T / 4096, transient_removal=True, **opts)


def rsvd_run(pkg, A, k, q, seed, learned=None, learned_dev=None, **opts):
    """Plain RSVD-Delta-t (main.c sequence) on A; learned: optional host array
    (Nw_eff x N x ncol complex) of sketch columns, or learned_dev: the same on
    the device as SketchNet.columns_dev returns it; opts: engine options
    (precision, rk4_form, fuse_small, ...).  Returns (S_hat (k x Nw_eff),
    V_hat (N x k*Nw_eff), wall seconds)."""
    L = pkg.lib()
    _bind(L)
    t0 = time.time()
    s = _session(pkg, A, k, q, **opts)
    if learned is not None:
        nwe, N, kl = learned.shape
        host = np.asfortranarray(learned.transpose(1, 0, 2).reshape(N, nwe * kl))
        pkg.check(L.RSVDT_SetSketchColumns(N, kl, nwe, host.ctypes.data_as(ctypes.c_void_p)), "RSVDT_SetSketchColumns")
    elif learned_dev is not None:
        ptr, ld, nwe, N = learned_dev
        pkg.check(L.RSVDT_SetSketchColumnsDevice(N, NCOL, nwe, ptr, ld), "RSVDT_SetSketchColumnsDevice")
    try:
        s.random_sketch(seed=seed, gaussian=True)
    finally:
        if learned is not None or learned_dev is not None:
            L.RSVDT_SetSketchColumns(0, 0, 0, None)
    s.direct()
    s.power_iteration()
    s.svd_before_adjoint()
    s.store_u()
    s.adjoint()
    s.svd_allfreqs()
    L.rsvdt_hip_sync()
    wall = time.time() - t0
    S = pkg.mat_to_dense(pkg._vp(s.Res.S_hat)).real
    Vh = pkg.mat_to_dense(pkg._vp(s.Res.V_hat))
    s.close()
    return S, Vh, wall


def family_sample(rng, n):
    return [(rng.uniform(*FAMILY["mu0"]), rng.uniform(*FAMILY["Ur"])) for _ in range(n)]


def training_data(pkg, members, k=16, q=1, seed=1373, log=None):
    """Leading NCOL forcing modes of each member at every frequency from plain
    RSVD-Delta-t; groups (member, frequency) in random order."""
    N = FAMILY["N"]
    x = np.linspace(-85, 85, N)
    Qs, Vs = [], []
    for i, (mu0, Ur) in enumerate(members):
        A = family_operator(mu0, Ur, N)
        S, Vh, wall = rsvd_run(pkg, A, k, q, seed + i)
        nwe = S.shape[1]
        om = omegas(FAMILY["Nw"], FAMILY["w"], False)[:nwe]
        V = Vh.reshape(N, nwe, k)[:, :, :NCOL].transpose(1, 0, 2)       # nwe x N x NCOL
        q4 = np.zeros((nwe, N, 4))
        q4[:, :, 0] = x[None, :]
        q4[:, :, 1] = om[:, None]
        q4[:, :, 2] = mu0
        q4[:, :, 3] = Ur
        Qs.append(q4)
        Vs.append(V)
        if log and (i % 20 == 0 or i == len(members) - 1):
            log(f"member {i}: mu0={mu0:.3f} Re U={Ur:.3f} sigma1 max {S[0].max():.1f} ({wall:.2f} s)")
    Q = np.concatenate(Qs)
    V = np.concatenate(Vs)
    perm = np.random.default_rng(seed).permutation(Q.shape[0])
    return {"q": Q[perm], "V": V[perm]}


def sigma1_error(S, Sref):
    return float(np.max(np.abs(S[0] - Sref[0]) / Sref[0]))


def evaluate(pkg, net, members, qs=(0, 1, 2), k=16, q_ref=4, seed=9001, tol=1e-6, log=None):
    """sigma_1 error of plain (k random columns) and learned (NCOL learned + k -
    NCOL random) RSVD-Delta-t against a converged reference (k random, q_ref
    power iterations), per member and power-iteration count."""
    N = FAMILY["N"]
    x = np.linspace(-85, 85, N)
    rows = []
    for i, (mu0, Ur) in enumerate(members):
        A = family_operator(mu0, Ur, N)
        Sref, _, _ = rsvd_run(pkg, A, k, q_ref, seed + 7919 * i)
        nwe = Sref.shape[1]
        om = omegas(FAMILY["Nw"], FAMILY["w"], False)[:nwe]
        t0 = time.time()
        cols = net.columns_dev(mu0, Ur, x, om)      # read in place by each learned run below
        t_inf = time.time() - t0
        row = {"mu0": mu0, "Ur": Ur, "plain": {}, "learned": {}, "inference_s": t_inf}
        for q in qs:
            S, _, wall = rsvd_run(pkg, A, k, q, seed + i)
            row["plain"][q] = (sigma1_error(S, Sref), wall)
            S, _, wall = rsvd_run(pkg, A, k, q, seed + i, learned_dev=cols)
            row["learned"][q] = (sigma1_error(S, Sref), wall + t_inf)
        rows.append(row)
        if log:
            log(f"member {i}: " + " ".join(f"q={q}: plain {row['plain'][q][0]:.1e} learned {row['learned'][q][0]:.1e}"
                                            for q in qs))
    summary = {}
    for form in ("plain", "learned"):
        need, times = [], []
        for r in rows:
            ok = [q for q in qs if r[form][q][0] <= tol]
            qn = ok[0] if ok else None
            need.append(qn)
            times.append(r[form][qn][1] if qn is not None else None)
        summary[form] = {
            "power_iterations_to_tol": need,
            "mean_power_iterations": float(np.mean([n for n in need if n is not None])) if any(n is not None for n in need) else None,
            "fraction_reaching_tol": float(np.mean([n is not None for n in need])),
            "mean_wall_s_to_tol": float(np.mean([t for t in times if t is not None])) if any(t is not None for t in times) else None,
            "median_sigma1_err": {q: float(np.median([r[form][q][0] for r in rows])) for q in qs},
        }
    return rows, summary


WEIGHTS = os.path.join(HERE, "learned_gl_sketch.npz")

# the straightforward GPU form of the reference path: FP64 state, the
# reference's RK4 stage order, one launch per stage, random sketch
PLAIN = dict(precision=64, rk4_form=1, fuse_small=0)
# the optimised form: Horner stages with the folded forcing, whole forcing
# chunks per launch with the state in LDS, learned sketch columns
FAST = dict(precision=64, rk4_form=0, fuse_small=1)
# the same with the FP32 state (matrix values, DFT, Gram, QR, SVD and the
# transient removal stay FP64)
FAST32 = dict(precision=32, rk4_form=0, fuse_small=1)


def time_to_modes(pkg, net, members, k=16, q_ref=4, tol=1e-6, qmax=4, seed=4242, fast=FAST, plain=PLAIN,
                  extra=(("fast_fp32", FAST32, True), ("fast_random", FAST, False)), log=None):
    """Wall time to sigma_1 relative error <= tol (against a converged
    reference: q_ref power iterations, optimised form) for the plain form
    with a random sketch and for the optimised form with the learned sketch,
    each with the fewest power iterations that reach tol; `extra` adds forms
    (name, options, learned columns or not): by default the optimised form
    with the FP32 state, and the optimised FP64 form with a random sketch --
    the engine's best form without the network, which isolates what the
    learned columns buy.  log: optional callable given one line per
    member.  Returns per-member rows and the mean times."""
    N = FAMILY["N"]
    x = np.linspace(-85, 85, N)
    rows = []
    for i, (mu0, Ur) in enumerate(members):
        A = family_operator(mu0, Ur, N)
        Sref, _, _ = rsvd_run(pkg, A, k, q_ref, seed + 7919 * i, **fast)
        om = omegas(FAMILY["Nw"], FAMILY["w"], False)[:Sref.shape[1]]
        row = {"mu0": mu0, "Ur": Ur}
        for name, opts, lrn in (("plain", plain, False), ("fast", fast, True)) + tuple(extra):
            for q in range(qmax + 1):
                t0 = time.time()
                cols = net.columns_dev(mu0, Ur, x, om) if lrn else None
                S, _, _ = rsvd_run(pkg, A, k, q, seed + i, learned_dev=cols, **opts)
                wall = time.time() - t0
                err = sigma1_error(S, Sref)
                if err <= tol or q == qmax:
                    row[name] = {"q": q, "wall_s": wall, "sigma1_err": err}
                    break
        rows.append(row)
        if log:
            log(f"member {i + 1}/{len(members)}: " + ", ".join(
                f"{n} q={row[n]['q']} {1e3 * row[n]['wall_s']:.1f} ms" for n in row if isinstance(row[n], dict)))
    ok = all(r["plain"]["sigma1_err"] <= tol and r["fast"]["sigma1_err"] <= tol for r in rows)
    tp = float(np.mean([r["plain"]["wall_s"] for r in rows]))
    tf = float(np.mean([r["fast"]["wall_s"] for r in rows]))
    summ = {"plain_s": tp, "fast_s": tf, "speedup": tp / tf, "all_reach_tol": ok,
            "plain_q": [r["plain"]["q"] for r in rows], "fast_q": [r["fast"]["q"] for r in rows]}
    for n, _, _ in extra:
        te = float(np.mean([r[n]["wall_s"] for r in rows]))
        summ[n] = {"s": te, "speedup": tp / te, "speedup_of_fast": te / tf, "q": [r[n]["q"] for r in rows],
                   "max_sigma1_err": float(max(r[n]["sigma1_err"] for r in rows)),
                   "all_reach_tol": all(r[n]["sigma1_err"] <= tol for r in rows)}
    return rows, summ


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--train", type=int, default=200)
    ap.add_argument("--eval", type=int, default=50)
    ap.add_argument("--epochs", type=int, default=1000)
    ap.add_argument("--target-q", type=int, default=3,
                    help="power iterations of the plain runs whose forcing modes are the training targets")
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--save", default=None, help="write the trained weights here (.npz)")
    args = ap.parse_args()
    pkg = _pkg()
    log = lambda s: print(s, flush=True)
    rng = np.random.default_rng(args.seed)
    train_m = family_sample(rng, args.train)
    eval_m = family_sample(rng, args.eval)
    t0 = time.time()
    data = training_data(pkg, train_m, q=args.target_q, log=log)
    t_data = time.time() - t0
    net = SketchNet(seed=args.seed)
    t0 = time.time()
    hist = net.train(data, epochs=args.epochs, log=log)
    t_train = time.time() - t0
    if args.save:
        net.save(args.save)
    if not eval_m:
        print(json.dumps({"learned_sketch": {"train_members": args.train, "epochs": args.epochs,
                                             "target_q": args.target_q,
                                             "final_loss": hist[-1], "train_s": round(t_train, 1)}}))
        return
    rows, summary = evaluate(pkg, net, eval_m, qs=(0, 1, 2, 3), log=log)
    print(json.dumps({"learned_sketch": {"train_members": args.train, "eval_members": args.eval,
                                         "target_q": args.target_q,
                                         "epochs": args.epochs, "final_loss": hist[-1],
                                         "data_s": round(t_data, 1), "train_s": round(t_train, 1),
                                         "summary": summary}}))


if __name__ == "__main__":
    main()
