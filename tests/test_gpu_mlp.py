"""The learned-sketch network's HIP kernels (BF16 MFMA GEMM with its
epilogues, features, alignment-loss gradient, Adam) and its full forward and
backward pass on the GPU against the numpy restatement (oracle/mlp_oracle.py),
which the CPU tests pin by finite differences."""
import ctypes
import os
import sys

import numpy as np
import pytest

import mlp_oracle as M
from conftest import rel
from gpu_util import Dev, init

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "rsvd-delta-t_amd"))
import learned  # noqa: E402

pytestmark = pytest.mark.gpu
V = ctypes.c_void_p
I64 = ctypes.c_int64


@pytest.fixture(scope="module")
def L(pkg):
    L = init(pkg)
    learned._bind(L)
    return L


def test_features(L, pkg):
    rng = np.random.default_rng(0)
    q = np.stack([rng.uniform(-85, 85, 999), rng.uniform(-3, 3, 999), rng.uniform(0.1, 0.4, 999),
                  rng.uniform(1.5, 2.5, 999)], axis=1)
    dq, df = Dev(pkg, q), Dev(pkg, nbytes=999 * 64 * 4)
    assert L.rsvdt_hip_mlp_features(999, dq.p, 85.0, 2.0, 2.0, df.p, None) == 0
    f = df.get(np.float32, (999, 64)).astype(np.float64)
    assert np.max(np.abs(f - M.features(q, 85.0, 2.0, 2.0))) < 2e-6


@pytest.mark.parametrize("M_,N_,K_,epi,trans,bkn", [(100, 256, 64, 1, False, False), (37, 16, 256, 2, False, False),
                                                    (256, 64, 1000, 0, True, False), (513, 200, 256, 3, False, False),
                                                    (16000, 256, 256, 3, False, True), (256, 256, 16000, 0, True, True),
                                                    (64, 16, 7000, 0, True, True), (77, 45, 33, 2, True, True)])
def test_gemm_bf16_epilogues(L, pkg, M_, N_, K_, epi, trans, bkn):
    """C = A B with BF16 operands (FP32 accumulation) and each epilogue, for
    A k- or m-contiguous (trans) and B k- or n-contiguous (bkn) in memory --
    the forward, weight-gradient and input-gradient layouts of the network --
    including reductions long enough to be split (K = 7000, 16000)."""
    rng = np.random.default_rng(M_ + K_)
    A = rng.standard_normal((M_, K_)).astype(np.float32)
    W = rng.standard_normal((N_, K_)).astype(np.float32)        # B(k, n) = W[n, k]
    bias = rng.standard_normal(N_).astype(np.float32)
    Z = rng.standard_normal((M_, N_)).astype(np.float32)
    if trans:   # A given transposed in memory: A(m, k) = At[k, m]
        At = np.ascontiguousarray(A.T)
        dA, sam, sak = Dev(pkg, At), 1, M_
    else:
        dA, sam, sak = Dev(pkg, A), K_, 1
    if bkn:     # B given k-major in memory: B(k, n) = Bkn[k, n]
        dW, sbk, sbn = Dev(pkg, np.ascontiguousarray(W.T)), N_, 1
    else:
        dW, sbk, sbn = Dev(pkg, W), 1, K_
    db, dZ = Dev(pkg, bias), Dev(pkg, Z)
    dC = Dev(pkg, nbytes=M_ * N_ * 4)
    assert L.rsvdt_hip_gemm_bf16(M_, N_, K_, dA.p, sam, sak, dW.p, sbk, sbn, dC.p, N_, epi, db.p, dZ.p, N_,
                                 None) == 0
    acc = M.bf16(A).astype(np.float64) @ M.bf16(W).astype(np.float64).T
    if epi == 1:
        z = acc + bias
        ref, zref = z / (1 + np.exp(-z)), z
        assert rel(dZ.get(np.float32, (M_, N_)), zref) < 1e-6
    elif epi == 2:
        ref = acc + bias
    elif epi == 3:
        s = 1 / (1 + np.exp(-Z.astype(np.float64)))
        ref = acc * s * (1 + Z * (1 - s))
    else:
        ref = acc
    # FP32 accumulation: the rounding error grows like sqrt(K)
    assert rel(dC.get(np.float32, (M_, N_)), ref) < 1e-5 * max(1.0, np.sqrt(K_ / 1000))


def test_sketch_loss_gradient(L, pkg):
    rng = np.random.default_rng(3)
    g, npts, ncol = 5, 123, 8
    F = rng.standard_normal((g * npts, 2 * ncol)).astype(np.float32)
    Vt = rng.standard_normal((g, npts, ncol)) + 1j * rng.standard_normal((g, npts, ncol))
    dF_, dV = Dev(pkg, F), Dev(pkg, Vt)
    dG, dl = Dev(pkg, nbytes=F.nbytes), Dev(pkg, nbytes=4 * g)
    assert L.rsvdt_hip_sketch_loss(g, npts, ncol, dF_.p, dV.p, dG.p, dl.p, 0.25, None) == 0
    Lref, Gref = M.loss_grad(F.astype(np.float64), Vt, 0.25)
    assert rel(dl.get(np.float32, (g,)), Lref) < 1e-6
    assert rel(dG.get(np.float32, F.shape), Gref) < 1e-5


def test_adam(L, pkg):
    rng = np.random.default_rng(4)
    n = 1000
    w, g, m, v = (rng.standard_normal(n).astype(np.float32) for _ in range(4))
    v = np.abs(v)
    dw, dg, dm, dv = Dev(pkg, w), Dev(pkg, g), Dev(pkg, m), Dev(pkg, v)
    assert L.rsvdt_hip_adam(n, dw.p, dg.p, dm.p, dv.p, 1e-2, 0.9, 0.999, 1e-8, 1 - 0.9 ** 3, 1 - 0.999 ** 3, None) == 0
    wr, mr, vr = M.adam(w.astype(np.float64), g, m, v, 1e-2, 0.9, 0.999, 1e-8, 3)
    assert rel(dw.get(np.float32, (n,)), wr) < 1e-6
    assert rel(dm.get(np.float32, (n,)), mr) < 1e-6


def test_network_gradients_vs_oracle(L, pkg):
    """One forward + loss + backward of the whole network on the GPU against
    the numpy restatement with the same BF16 rounding of the GEMM operands."""
    rng = np.random.default_rng(5)
    net = learned.SketchNet(seed=11, batch_points=2 * 500)
    g, npts = 2, 500
    q = np.zeros((g, npts, 4))
    q[:, :, 0] = np.linspace(-85, 85, npts)[None]
    q[:, :, 1] = rng.uniform(-1, 1, (g, 1))
    q[:, :, 2] = rng.uniform(0.1, 0.4, (g, 1))
    q[:, :, 3] = rng.uniform(1.5, 2.5, (g, 1))
    Vt = rng.standard_normal((g, npts, learned.NCOL)) + 1j * rng.standard_normal((g, npts, learned.NCOL))
    net.q.put(q.reshape(-1, 4))
    dV = Dev(pkg, Vt)
    out = net.forward_dev(g * npts, net.q.p)
    scale = 1.0 / (g * learned.NCOL)
    assert L.rsvdt_hip_sketch_loss(g, npts, learned.NCOL, out.p, dV.p, net.dout.p, net.loss.p, scale, None) == 0
    net.backward_dev(g * npts)
    params = net.params_host()
    feat = M.features(q.reshape(-1, 4), learned.XS, 1.0, learned.KMAX)
    oref, cache = M.forward(params, feat, use_bf16=True)
    assert rel(out.get(np.float32, (g * npts, 16)), oref) < 2e-3
    Lref, dref = M.loss_grad(oref, Vt, scale)
    assert abs(net.loss.get(np.float32, (g,)).sum() - Lref.sum()) < 1e-3 * abs(Lref.sum())
    gref = M.backward(params, cache, dref, use_bf16=True)
    ggpu = net.grads_host()
    for li, ((gW, gb), (rW, rb)) in enumerate(zip(ggpu, gref)):
        print(f"layer {li}: dW rel {rel(gW, rW):.2e}  db rel {rel(gb, rb):.2e}", flush=True)
        assert rel(gW, rW) < 3e-2
        assert rel(gb, rb) < 3e-2


def test_training_reduces_alignment_loss(pkg):
    """A few epochs on fixed smooth targets bring the alignment loss down."""
    rng = np.random.default_rng(6)
    g, npts = 16, 500
    x = np.linspace(-85, 85, npts)
    q = np.zeros((g, npts, 4))
    q[:, :, 0] = x[None]
    q[:, :, 1] = rng.uniform(-1, 1, (g, 1))
    q[:, :, 2] = rng.uniform(0.1, 0.4, (g, 1))
    q[:, :, 3] = rng.uniform(1.5, 2.5, (g, 1))
    Vt = np.zeros((g, npts, learned.NCOL), complex)
    for j in range(learned.NCOL):
        Vt[:, :, j] = np.exp(-((x[None] - 10 * j + 30) / (15 + 5 * q[:, :, 2])) ** 2) * np.exp(1j * 0.5 * x[None] * q[:, :, 3])
    net = learned.SketchNet(seed=1, batch_points=8 * npts)
    hist = net.train({"q": q, "V": Vt}, epochs=40, groups_per_batch=8, lr=2e-3)
    print("loss history", hist[0], hist[-1], flush=True)
    assert hist[-1] < 0.5 * hist[0]
