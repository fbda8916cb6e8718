"""The no-allocation QR entry point (rsvdt_qr_blocks_noalloc) against the
existing QRAllFreqs entry point on the same blocks: two calls in a row, the
second with a smaller input (the static scratch keeps the first call's
contents), and the size limit."""
import ctypes

import numpy as np
import pytest

import kat
from conftest import crandn, rel
from gpu_util import Dev, init

pytestmark = pytest.mark.gpu
V = ctypes.c_void_p
I64 = ctypes.c_int64


def qr_allfreqs(pkg, F, k):
    """QRAllFreqs (the reference-compatible entry point) on the frequency-major
    sketch F (N x k*nfreq, column m*k + j), through a session as a driver
    makes the call."""
    N, nf = F.shape[0], F.shape[1] // k
    s = pkg.Session(kat.gl_operator(N=N, L=20.0), k=k, q=0, Nw=nf, w=0.3, TL=1.0, dt=0.02,
                    transient_removal=False)
    assert s.Nw_eff == nf
    s.set_sketch(F)
    s.qr()
    Q = s.get_Y()
    s.close()
    return Q


def to_blocks(F, k):
    """N x k*nfreq (column m*k + j) -> nfreq blocks of N x k row-major."""
    N = F.shape[0]
    return np.ascontiguousarray(F.reshape(N, -1, k).transpose(1, 0, 2))


def from_blocks(B):
    nf, N, k = B.shape
    return B.transpose(1, 0, 2).reshape(N, nf * k)


@pytest.mark.parametrize("calls", [[(3000, 16, 8), (1000, 8, 3)], [(20000, 32, 32), (4096, 12, 5)]])
def test_noalloc_qr_matches_qrallfreqs(pkg, calls):
    L = init(pkg)
    L.rsvdt_qr_blocks_noalloc.argtypes = [V, I64, I64, I64, V]
    L.rsvdt_qr_blocks_noalloc_bytes.argtypes = [I64, I64, I64]
    L.rsvdt_qr_blocks_noalloc_bytes.restype = ctypes.c_size_t
    for i, (N, k, nf) in enumerate(calls):
        print(f"no-alloc QR call {i}: N={N} k={k} nfreq={nf}, work {L.rsvdt_qr_blocks_noalloc_bytes(nf, N, k)} "
              "bytes", flush=True)
        rng = np.random.default_rng(N + k)
        F = crandn(rng, N, k * nf)
        Qref = qr_allfreqs(pkg, F, k)
        d = Dev(pkg, to_blocks(F, k))
        assert L.rsvdt_qr_blocks_noalloc(d.p, N, k, nf, None) == 0, L.rsvdt_hip_last_error()
        Q = from_blocks(d.get(np.complex128, (nf, N, k)))
        assert rel(Q, Qref) < 1e-13
        # orthonormal columns per block, R = Q^H F upper triangular with a positive diagonal
        for m in range(nf):
            Qm, Fm = Q[:, m * k:(m + 1) * k], F[:, m * k:(m + 1) * k]
            assert np.abs(Qm.conj().T @ Qm - np.eye(k)).max() < 1e-12
            Rm = Qm.conj().T @ Fm
            assert np.abs(np.tril(Rm, -1)).max() < 1e-10 * np.abs(Rm).max()
            assert np.all(np.diag(Rm).real > 0)


def test_noalloc_qr_rejects_oversize(pkg):
    L = init(pkg)
    L.rsvdt_qr_blocks_noalloc.argtypes = [V, I64, I64, I64, V]
    d = Dev(pkg, nbytes=16 * 64 * 65)
    # k > 64 and work beyond the static scratch are rejected before any launch
    assert L.rsvdt_qr_blocks_noalloc(d.p, 64, 65, 1, None) == 102
    assert b"at most 64 columns" in L.rsvdt_hip_last_error()
    assert L.rsvdt_qr_blocks_noalloc(d.p, 1 << 26, 64, 4096, None) == 102
    assert b"static scratch" in L.rsvdt_hip_last_error()
