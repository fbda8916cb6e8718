import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "tests"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def load_pkg():
    import importlib.util
    if "rsvdt_amd" in sys.modules:
        return sys.modules["rsvdt_amd"]
    path = os.path.join(ROOT, "rsvd-delta-t_amd", "__init__.py")
    spec = importlib.util.spec_from_file_location("rsvdt_amd", path)
    mod = importlib.util.module_from_spec(spec)
    sys.modules["rsvdt_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


@pytest.fixture(scope="session")
def pkg():
    return load_pkg()


def crandn(rng, *shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


def subspace_gap(U1, U2):
    """sin of the largest principal angle between span(U1) and span(U2)
    (orthonormal columns), as ||(I - U1 U1^H) U2||_2, accurate to rounding."""
    R = U2 - U1 @ (U1.conj().T @ U2)
    return float(np.linalg.norm(R, 2))


def phase_align(u, v):
    """v rotated by the unit phase that best matches u."""
    c = np.vdot(v, u)
    return v * (c / abs(c)) if abs(c) > 0 else v
