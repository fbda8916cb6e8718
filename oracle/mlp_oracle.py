"""CPU restatement of the learned-sketch network (numpy).

TEST INFRASTRUCTURE ONLY: the checker for the HIP kernels of
rsvd-delta-t_amd/csrc/hip/mlp.hip (features, BF16-MFMA GEMMs with their
epilogues, the alignment loss and its gradient, Adam).  Only tests/ may
import it.  The network has no counterpart in the reference; it proposes the
columns CreateRandomMat (SourceCode/CreateRandomMat.c:33-39) would draw at
random.

Pinning: test_oracle_mlp.py checks the backward pass against central finite
differences of the forward pass in FP64 (bf16=False), and the loss against
its invariances (column phase and scale) and its closed-form extremes.
"""
import numpy as np

LAYERS = [(64, 256), (256, 256), (256, 256), (256, 16)]   # (in, out) of the four linear layers


def bf16(a):
    """Round to the nearest BF16 value (ties to even), returned as float64."""
    a32 = np.asarray(a, dtype=np.float32)
    b = a32.view(np.uint32).astype(np.uint64)
    r = ((b + 0x7FFF + ((b >> 16) & 1)) & 0xFFFF0000).astype(np.uint32)
    return r.view(np.float32).astype(np.float64)


def features(q, xs, ws, kmax):
    """[x/xs, w/ws, mu0, U, sin(k_j x), cos(k_j x) (k_j = kmax (j+1)/28, j < 28), 0 x 4]."""
    q = np.asarray(q, dtype=np.float64)
    x, w, mu, u = q[:, 0], q[:, 1], q[:, 2], q[:, 3]
    out = np.zeros((q.shape[0], 64))
    out[:, 0], out[:, 1], out[:, 2], out[:, 3] = x / xs, w / ws, mu, u
    for j in range(28):
        a = kmax * (j + 1) / 28.0 * x
        out[:, 4 + 2 * j] = np.sin(a)
        out[:, 5 + 2 * j] = np.cos(a)
    return out.astype(np.float32).astype(np.float64)


def _sig(z):
    return 1.0 / (1.0 + np.exp(-z))


def _mm(a, b, use_bf16):
    return (bf16(a) @ bf16(b)) if use_bf16 else (a @ b)


def forward(params, feat, use_bf16=True):
    """params: list of (W (out x in), b (out)).  Returns (out, cache)."""
    a = feat
    cache = {"a": [a], "z": []}
    for li, (W, b) in enumerate(params):
        z = _mm(a, W.T, use_bf16) + b
        if li < len(params) - 1:
            cache["z"].append(z)
            a = z * _sig(z)
            cache["a"].append(a)
        else:
            a = z
    return a, cache


def loss_grad(F, V, scale):
    """F: (g*npts) x 2ncol real (re, im interleaved); V: g x npts x ncol complex.
    L_g = scale sum_j (1 - |<v,f>|^2/(|f|^2|v|^2)); returns (L per group, dF)."""
    g, npts, ncol = V.shape
    f = F.reshape(g, npts, ncol, 2)
    fc = f[..., 0] + 1j * f[..., 1]
    s = np.einsum("gxj,gxj->gj", V.conj(), fc)
    n = np.sum(np.abs(fc) ** 2, axis=1)
    nv = np.sum(np.abs(V) ** 2, axis=1)
    L = scale * np.sum(1.0 - np.abs(s) ** 2 / (n * nv), axis=1)
    gc = (2.0 / (n * nv))[:, None, :] * ((np.abs(s) ** 2 / n)[:, None, :] * fc - s[:, None, :] * V) * scale
    dF = np.stack([gc.real, gc.imag], axis=-1).reshape(F.shape)
    return L, dF


def backward(params, cache, dout, use_bf16=True):
    """Gradients (dW, db) per layer for dL/dout = dout."""
    grads = [None] * len(params)
    dz = dout
    for li in range(len(params) - 1, -1, -1):
        W, _ = params[li]
        a_in = cache["a"][li]
        dW = _mm(dz.T, a_in, use_bf16)
        db = dz.sum(axis=0)
        grads[li] = (dW, db)
        if li > 0:
            da = _mm(dz, W, use_bf16)
            z = cache["z"][li - 1]
            s = _sig(z)
            dz = da * s * (1.0 + z * (1.0 - s))
    return grads


def adam(w, g, m, v, lr, b1, b2, eps, t):
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    w = w - lr * (m / (1 - b1 ** t)) / (np.sqrt(v / (1 - b2 ** t)) + eps)
    return w, m, v
