"""Pins the numpy restatement of the learned-sketch network (oracle/mlp_oracle.py):
its backward pass against central finite differences of its forward pass
(FP64, no BF16 rounding), and its alignment loss against the loss's
invariances and extreme values."""
import numpy as np

import mlp_oracle as M


def _params(rng, scale=0.3):
    return [(rng.standard_normal((o, i)) * scale / np.sqrt(i), rng.standard_normal(o) * 0.1) for i, o in M.LAYERS]


def _batch(rng, g=3, npts=7, ncol=8):
    q = np.stack([rng.uniform(-85, 85, g * npts), rng.uniform(-2, 2, g * npts), rng.uniform(0.1, 0.4, g * npts),
                  rng.uniform(1.5, 2.5, g * npts)], axis=1)
    V = rng.standard_normal((g, npts, ncol)) + 1j * rng.standard_normal((g, npts, ncol))
    return q, V


def _total_loss(params, q, V):
    feat = M.features(q, 85.0, 2.0, 2.0)
    out, _ = M.forward(params, feat, use_bf16=False)
    L, _ = M.loss_grad(out, V, 1.0)
    return L.sum()


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(0)
    params = _params(rng)
    q, V = _batch(rng)
    feat = M.features(q, 85.0, 2.0, 2.0)
    out, cache = M.forward(params, feat, use_bf16=False)
    _, dout = M.loss_grad(out, V, 1.0)
    grads = M.backward(params, cache, dout, use_bf16=False)
    eps = 1e-6
    for li in range(len(params)):
        W, b = params[li]
        for (r, c) in [(0, 0), (W.shape[0] - 1, W.shape[1] - 1), (W.shape[0] // 2, 3)]:
            Wp = W.copy(); Wp[r, c] += eps
            Wm = W.copy(); Wm[r, c] -= eps
            pp = list(params); pp[li] = (Wp, b)
            pm = list(params); pm[li] = (Wm, b)
            fd = (_total_loss(pp, q, V) - _total_loss(pm, q, V)) / (2 * eps)
            assert abs(fd - grads[li][0][r, c]) <= 1e-6 * max(1.0, abs(fd)), (li, r, c, fd, grads[li][0][r, c])
        j = b.shape[0] - 1
        bp = b.copy(); bp[j] += eps
        bm = b.copy(); bm[j] -= eps
        pp = list(params); pp[li] = (W, bp)
        pm = list(params); pm[li] = (W, bm)
        fd = (_total_loss(pp, q, V) - _total_loss(pm, q, V)) / (2 * eps)
        assert abs(fd - grads[li][1][j]) <= 1e-6 * max(1.0, abs(fd))


def test_loss_gradient_matches_finite_differences():
    rng = np.random.default_rng(1)
    g, npts, ncol = 2, 5, 3
    F = rng.standard_normal((g * npts, 2 * ncol))
    V = rng.standard_normal((g, npts, ncol)) + 1j * rng.standard_normal((g, npts, ncol))
    L, dF = M.loss_grad(F, V, 0.5)
    eps = 1e-7
    for idx in [(0, 0), (3, 5), (9, 2)]:
        Fp = F.copy(); Fp[idx] += eps
        Fm = F.copy(); Fm[idx] -= eps
        fd = (M.loss_grad(Fp, V, 0.5)[0].sum() - M.loss_grad(Fm, V, 0.5)[0].sum()) / (2 * eps)
        assert abs(fd - dF[idx]) < 1e-6


def test_loss_invariances_and_extremes():
    rng = np.random.default_rng(2)
    g, npts, ncol = 2, 6, 4
    V = rng.standard_normal((g, npts, ncol)) + 1j * rng.standard_normal((g, npts, ncol))
    # f = c * v (any complex c) is a perfect alignment: loss 0, gradient 0
    c = 3.0 * np.exp(0.7j)
    fc = c * V
    F = np.stack([fc.real, fc.imag], axis=-1).reshape(g * npts, 2 * ncol)
    L, dF = M.loss_grad(F, V, 1.0)
    assert np.all(np.abs(L) < 1e-14) and np.max(np.abs(dF)) < 1e-12
    # f orthogonal to v: loss 1 per column
    w = rng.standard_normal((g, npts, ncol)) + 1j * rng.standard_normal((g, npts, ncol))
    w -= V * (np.einsum("gxj,gxj->gj", V.conj(), w) / np.sum(np.abs(V) ** 2, axis=1))[:, None, :]
    F2 = np.stack([w.real, w.imag], axis=-1).reshape(g * npts, 2 * ncol)
    L2, _ = M.loss_grad(F2, V, 1.0)
    assert np.allclose(L2, ncol)
    # invariance to a per-column phase and scale of f
    fr = rng.standard_normal((g, npts, ncol)) + 1j * rng.standard_normal((g, npts, ncol))
    ph = 2.5 * np.exp(1j * rng.uniform(0, 6, (g, 1, ncol)))
    Fa = np.stack([fr.real, fr.imag], axis=-1).reshape(g * npts, 2 * ncol)
    Fb = np.stack([(ph * fr).real, (ph * fr).imag], axis=-1).reshape(g * npts, 2 * ncol)
    assert np.allclose(M.loss_grad(Fa, V, 1.0)[0], M.loss_grad(Fb, V, 1.0)[0])


def test_bf16_rounding():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, -3.14159, 65504.0, 1e-3])
    r = M.bf16(x)
    assert r[0] == 1.0
    assert r[1] == 1.0                       # tie, rounds to even
    assert r[2] == 1.0 + 2 ** -7             # above the tie, rounds up
    assert abs(r[3] + 3.140625) < 1e-12
    assert np.all(np.abs(r - x) <= np.abs(x) * 2 ** -8)
