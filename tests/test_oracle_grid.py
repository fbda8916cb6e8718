"""Pins the oracle's time grid and DFT/iDFT tables against closed forms.

SourceCode/SetupTimeFreqGrid.c:16-21, SourceCode/CreateDFTiDFTMats.c:18-93."""
import numpy as np
import pytest

import kat
import oracle
from conftest import crandn


@pytest.mark.parametrize("w,Nw,TL,dt", [(0.05, 64, 10.0, 0.03068), (0.05, 42, 10.0, 0.003), (0.3, 7, 5.0, 0.02),
                                        (2 * np.pi / 20, 128, 10.0, 0.002)])
def test_grid_closed_form(w, Nw, TL, dt):
    h, Ns, Nt, RR = oracle.setup_grid(w, Nw, TL, dt)
    T = 2 * np.pi / w
    assert h <= dt * (1 + 1e-15)
    assert Ns == Nw * RR                      # an integer number of steps per snapshot interval
    assert abs(Ns * h - T) < 1e-9 * T
    assert Nt == round(TL / h)
    assert (h, Ns, Nt, RR) == kat.grid(w, Nw, TL, dt)


def test_tutorial_grid():
    # Tutorial/Ginzburg-Landau inputs: w = 0.05, Nw = 42, dt = 0.003, TL = 10
    h, Ns, Nt, RR = oracle.setup_grid(0.05, 42, 10.0, 0.003)
    assert RR == 998 and Ns == 41916
    assert abs(h - 0.0029979890) < 1e-9


@pytest.mark.parametrize("complex_op,Nw", [(True, 8), (True, 7), (False, 8)])
def test_dft_tables_closed_form(complex_op, Nw):
    rng = np.random.default_rng(1)
    A = kat.gl_operator(N=16, L=10.0)
    if not complex_op:
        A = A.real.astype(complex)
    o = oracle.Oracle(A, k=2, q=0, Nw=Nw, w=0.3, TL=1.0, dt=0.05, transient_removal=False)
    idft, dft = o.dft_mats()
    Ns, Nwe = o.Ns, o.Nw_eff
    assert Nwe == (Nw // 2 if not complex_op else Nw)
    j = np.arange(2 * Ns)
    for m in range(Nwe):
        ma = kat.alias(m, Nwe, not complex_op)
        # (1/Ns) e^{+i w_m j h/2}, w_m = w*ma and w*Ns*h = 2 pi
        ref = np.exp(2j * np.pi * ma * j / (2 * Ns)) / Ns
        assert np.allclose(idft[m], ref, atol=1e-14, rtol=0)
        s = np.arange(Nw)
        assert np.allclose(dft[:, m], np.exp(-2j * np.pi * m * s / Nw), atol=1e-13, rtol=0)
    # the forcing samples form an inverse DFT: sum_j idft[m, j] idft[m', j]^* = 2/Ns delta
    G = idft @ idft.conj().T
    assert np.allclose(G, 2.0 / Ns * np.eye(Nwe), atol=1e-12)
