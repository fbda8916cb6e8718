"""The SpMM stage kernels at the benchmark operators' full sizes (2D flow,
N = 393,216, k = 16; 3D flow, N = 1,966,080, k = 32), through the C ABI,
against scipy in FP64: one Horner stage (with forcing term) and one RK4 step
in the reference stage order, in both state precisions, with the default
launch shape and an XCD-remapped one."""
import ctypes
import os
import sys

import numpy as np
import pytest

from conftest import crandn, rel
from gpu_util import Desc, Dev, grouped_form, init

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "rsvd-delta-t_amd"))
import workloads  # noqa: E402

pytestmark = pytest.mark.gpu
V = ctypes.c_void_p
I64 = ctypes.c_int64




@pytest.fixture(scope="module")
def L(pkg):
    return init(pkg)


_OPS = {}


def operator(name):
    if name not in _OPS:
        _OPS.clear()
        A = workloads.flow2d() if name == "flow2d" else workloads.flow3d()
        _OPS[name] = (A, 16 if name == "flow2d" else 32)
    return _OPS[name]


@pytest.mark.parametrize("name", ["flow2d", "flow3d"])
@pytest.mark.parametrize("prec,variant,remap", [(64, 3, 1), (64, 0, 0), (32, 2, 1), (64, 6, 0), (32, 6, 3), (32, 7, 0)])
def test_horner_stage_full_size(L, pkg, name, prec, variant, remap):
    print(f"horner full size {name} prec={prec} variant={variant} remap={remap}", flush=True)
    A, k = operator(name)
    N = A.shape[0]
    rng = np.random.default_rng(prec + variant)
    vin, y, g = crandn(rng, N, k), crandn(rng, N, k), crandn(rng, N, k)
    c = 0.013
    cdt = np.complex128 if prec == 64 else np.complex64
    vin, y, g = vin.astype(cdt), y.astype(cdt), g.astype(cdt)
    ref = y.astype(np.complex128) + c * (A @ vin.astype(np.complex128)) + g.astype(np.complex128)
    drp, dci = Dev(pkg, A.indptr.astype(np.int32)), Dev(pkg, A.indices.astype(np.int32))
    dv = Dev(pkg, A.data.real.astype(np.float64))
    dvin, dy, dg, dout = Dev(pkg, vin), Dev(pkg, y), Dev(pkg, g), Dev(pkg, nbytes=y.nbytes)
    d = Desc(N, k, 1, drp.p, dci.p, dv.p, 1, prec, 0, remap, variant)
    if variant == 6:    # the row-grouped form, one group per grid point (3 or 5 fields)
        G = 3 if name == "flow2d" else 5
        gptr, gword, gval = grouped_form(A, G)
        dgp, dgw, dgv = Dev(pkg, gptr), Dev(pkg, gword), Dev(pkg, gval.real.astype(np.float64))
        d.group, d.gptr, d.gword, d.gval = G, dgp.p, dgw.p, dgv.p
    L.rsvdt_hip_horner_stage.argtypes = [ctypes.POINTER(Desc), V, V, V, V, ctypes.c_double, V, V, V]
    assert L.rsvdt_hip_horner_stage(ctypes.byref(d), dvin.p, dy.p, None, dg.p, c, dout.p, None, None) == 0
    out = dout.get(cdt, y.shape)
    assert rel(out, ref) < (1e-14 if prec == 64 else 1e-6)


@pytest.mark.parametrize("name", ["flow2d", "flow3d"])
def test_rk4_step_full_size(L, pkg, name):
    """One step of the reference stage order, FP64, default launch shape."""
    print(f"rk4 step full size {name}", flush=True)
    A, k = operator(name)
    N = A.shape[0]
    rng = np.random.default_rng(7)
    y = crandn(rng, N, k)
    F = [crandn(rng, N, k) for _ in range(3)]
    h = 0.01
    k1 = A @ y + F[0]
    k2 = A @ (y + h / 2 * k1) + F[1]
    k3 = A @ (y + h / 2 * k2) + F[1]
    k4 = A @ (y + h * k3) + F[2]
    ref = y + h / 6 * k1 + h / 3 * k2 + h / 3 * k3 + h / 6 * k4
    drp, dci = Dev(pkg, A.indptr.astype(np.int32)), Dev(pkg, A.indices.astype(np.int32))
    dv = Dev(pkg, A.data.real.astype(np.float64))
    dy = Dev(pkg, y)
    dF = [Dev(pkg, f) for f in F]
    du1, du2, dacc = (Dev(pkg, nbytes=y.nbytes) for _ in range(3))
    d = Desc(N, k, 1, drp.p, dci.p, dv.p, 1, 64, 0, 1, 3)
    L.rsvdt_hip_rk4_stage.argtypes = [ctypes.POINTER(Desc), ctypes.c_int, V, V, V, V, V, ctypes.c_double,
                                      ctypes.c_double, V, V]
    f = L.rsvdt_hip_rk4_stage
    assert f(ctypes.byref(d), 1, dy.p, dy.p, dF[0].p, dacc.p, du1.p, h / 2, h / 6, None, None) == 0
    assert f(ctypes.byref(d), 2, du1.p, dy.p, dF[1].p, dacc.p, du2.p, h / 2, h / 6 * 2, None, None) == 0
    assert f(ctypes.byref(d), 3, du2.p, dy.p, dF[1].p, dacc.p, du1.p, h, h / 6 * 2, None, None) == 0
    assert f(ctypes.byref(d), 4, du1.p, None, dF[2].p, dacc.p, dy.p, 0.0, h / 6, None, None) == 0
    assert rel(dy.get(np.complex128, y.shape), ref) < 1e-13
