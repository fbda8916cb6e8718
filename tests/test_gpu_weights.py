"""Input/output matrices and weights (ReadWeightMats / ApplyWeightMats /
SVDAllFreqs, SourceCode/ApplyWeightMats.c:16-80, SVDAllFreqs.c:63-64) on the
GPU against the oracle's weighted main sequence, with Nb != N != Nc:
gains at 1e-10, modes up to a unit phase per column."""
import numpy as np
import pytest
import scipy.sparse as sp

import kat
import oracle
from conftest import crandn

pytestmark = pytest.mark.gpu


def _weights(rng, N, Nb, Nc):
    B = sp.random(N, Nb, density=0.3, random_state=int(rng.integers(100)), format="csr")
    B = (B + sp.csr_matrix((np.ones(Nb), (np.arange(Nb) * (N // Nb), np.arange(Nb))), shape=(N, Nb))).astype(complex)
    C = crandn(rng, Nc, N) * (rng.random((Nc, N)) < 0.2)
    Lq = np.diag(rng.uniform(0.5, 2.0, Nc)) + 0.1 * np.triu(crandn(rng, Nc, Nc), 1)
    Wq = Lq @ Lq.conj().T
    Wf_inv = np.diag(rng.uniform(0.5, 2.0, Nb)).astype(complex)
    return B, C, Wq, np.linalg.inv(Wq), Wf_inv


@pytest.mark.parametrize("N,Nb,Nc,k,Nw,tr,fuse", [(12, 3, 4, 3, 11, True, 1), (80, 20, 30, 4, 16, True, 0),
                                                 (80, 25, 10, 3, 16, False, 1)])
def test_weighted_run_vs_oracle(pkg, N, Nb, Nc, k, Nw, tr, fuse):
    print(f"weighted N={N} Nb={Nb} Nc={Nc} k={k} Nw={Nw} tr={tr} fuse={fuse}", flush=True)
    A = kat.gl_operator(N=N, L=6.0 if N == 12 else 20.0)
    rng = np.random.default_rng(N + Nb + Nc)
    B, C, Wq, Wq_inv, Wf_inv = _weights(rng, N, Nb, Nc)
    TL = 0.0 if N == 12 else 1.0
    o = oracle.Oracle(A, k=k, q=1, Nw=Nw, w=0.4, TL=TL, dt=0.02, transient_removal=tr)
    F = crandn(rng, Nb, k * o.Nw_eff)
    Uo, Vo, So = o.run_weighted(F, B=B, C=C, Wq_sqrt=Wq, Wq_sqrt_inv=Wq_inv, Wf_sqrt_inv=Wf_inv)
    s = pkg.Session(A, k=k, q=1, Nw=Nw, w=0.4, TL=TL, dt=0.02, transient_removal=tr, fuse_small=fuse)
    s.set_weights(B=B, C=C, Wq_sqrt=Wq, Wq_sqrt_inv=Wq_inv, Wf_sqrt_inv=Wf_inv)
    U, V, S = s.run(F)
    s.close()
    assert U.shape == Uo.shape and V.shape == Vo.shape
    assert np.allclose(S, So, rtol=1e-10)
    for X, Xo in ((U, Uo), (V, Vo)):
        ph = np.abs(np.sum(X.conj() * Xo, axis=0)) / (np.linalg.norm(X, axis=0) * np.linalg.norm(Xo, axis=0))
        assert np.max(np.abs(ph - 1)) < 1e-9
