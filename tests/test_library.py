"""CPU checks of the C-ABI library: it loads, exports every function the
headers in include/ declare, and its host logic (PETSc-binary reader/writer,
variables.yaml reader, time grid) behaves like the reference's.  No GPU."""
import ctypes
import os
import re

import numpy as np
import pytest

import kat
import petsc_bin
from conftest import GOLDEN, ROOT


def declared_functions():
    names = set()
    for h in ("rsvdt.h", "rsvdt_hip.h"):
        txt = open(os.path.join(ROOT, "include", h)).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        for m in re.finditer(r"^[A-Za-z_][\w \*]*?\b([A-Za-z_]\w*)\s*\(([^;{]*?)\)\s*;", txt, flags=re.M):
            if not m.group(0).startswith("typedef"):
                names.add(m.group(1))
    return names


def test_exports_every_declared_symbol(pkg):
    L = pkg.lib()
    names = declared_functions()
    assert len(names) > 60
    missing = [n for n in sorted(names) if not hasattr(L, n)]
    assert not missing, missing
    assert set(pkg.ENTRY_POINTS + pkg.HIP_ENTRY_POINTS) == names


def test_reader_matches_reference_file(pkg):
    A, is_real = pkg.load_aij(os.path.join(GOLDEN, "A_GL"))
    B = petsc_bin.read_mat(os.path.join(GOLDEN, "A_GL"))
    assert A.shape == (500, 500) and A.nnz == 2000 and not is_real
    assert (A != B).nnz == 0
    # header of the reference fixture: int64 big-endian [1211216, 500, 500, 2000]
    hdr = np.frombuffer(open(os.path.join(GOLDEN, "A_GL"), "rb").read(32), dtype=">i8")
    assert list(hdr) == [1211216, 500, 500, 2000]


def test_real_operator_detection(pkg, tmp_path):
    A = kat.gl_operator(N=30, L=5.0).real.astype(complex)
    p = str(tmp_path / "A_real")
    petsc_bin.write_mat(p, A)
    B, is_real = pkg.load_aij(p)
    assert is_real and (B != A).nnz == 0
    A2 = A.copy(); A2[3, 3] += 2e-12j
    petsc_bin.write_mat(p, A2)
    assert pkg.load_aij(p)[1] is False     # ||Im A||_F = 2e-12 >= 1e-12 (SourceCode/LNSStructure.c:40)


def test_writer_roundtrip(pkg, tmp_path):
    A = kat.gl_operator(N=17, L=5.0)
    p1, p2 = str(tmp_path / "a"), str(tmp_path / "b")
    petsc_bin.write_mat(p1, A)
    L = pkg.lib()
    m = ctypes.c_void_p()
    assert L.RSVDT_MatLoad(p1.encode(), 1, ctypes.byref(m)) == 0
    assert L.RSVDT_MatView(m, p2.encode()) == 0
    pkg.destroy(m)
    assert open(p1, "rb").read() == open(p2, "rb").read()


def test_reader_32bit_indices_and_native_dense(pkg, tmp_path):
    A = kat.gl_operator(N=9, L=5.0).tocsr()
    p = str(tmp_path / "a32")
    with open(p, "wb") as f:
        f.write(np.array([1211216, 9, 9, A.nnz], dtype=">i4").tobytes())
        f.write(np.diff(A.indptr).astype(">i4").tobytes())
        f.write(A.indices.astype(">i4").tobytes())
        f.write(A.data.astype(">c16").tobytes())
    B, _ = pkg.load_aij(p)
    assert (A != B).nnz == 0
    D = np.arange(12).reshape(3, 4) + 1j
    with open(p, "wb") as f:
        f.write(np.array([1211216, 3, 4, -1], dtype=">i8").tobytes())
        f.write(D.astype(">c16").tobytes())
    B, _ = pkg.load_aij(p)
    assert np.array_equal(B.toarray(), D)


@pytest.mark.parametrize("mutate,frag", [
    (lambda b: b[:20], "truncated"),
    (lambda b: b"\x00" * 8 + b[8:], "classid"),
    (lambda b: b[:-100], "truncated"),
])
def test_reader_rejects_bad_files(pkg, tmp_path, capfd, mutate, frag):
    raw = open(os.path.join(GOLDEN, "A_GL"), "rb").read()
    raw = raw[:32 + 500 * 8 + 2000 * 8 + 2000 * 16]      # drop the trailing Vec object
    p = str(tmp_path / "bad")
    open(p, "wb").write(mutate(raw))
    m = ctypes.c_void_p()
    rc = pkg.lib().RSVDT_MatLoad(p.encode(), 1, ctypes.byref(m))
    assert rc != 0
    assert frag in capfd.readouterr().err


def test_reader_rejects_nan(pkg, tmp_path, capfd):
    A = kat.gl_operator(N=9, L=5.0).tocsr()
    A.data[4] = np.nan
    p = str(tmp_path / "nan")
    petsc_bin.write_mat(p, A)
    rc = pkg.lib().RSVDT_MatLoad(p.encode(), 1, ctypes.byref(ctypes.c_void_p()))
    assert rc == 83 and "non-finite" in capfd.readouterr().err


BASE_YAML = """RootDir: {root}/
ResultsDir: res/
OperatorDir: A
k: 5
q: 1
Nw: 16
w: 0.5
TransientLength: 3   # comment
dt: 0.02
TransientRemoval: true
TransICDir: ic
"""


def read_user_input(pkg, path):
    L = pkg.lib()
    L.RSVDT_Finalize()
    L.RSVDT_Initialize(None, None)
    L.RSVDT_OptionsSetValue(b"inputs", path.encode())
    V = pkg.RSVDt_vars(); W = pkg.Weight_matrices(); LNS = pkg.LNS_vars(); Res = pkg.Resolvent_matrices()
    TR = pkg.TransRun_vars(); D = pkg.Directories()
    L.ReadUserInput.argtypes = [ctypes.c_void_p] * 6
    rc = L.ReadUserInput(ctypes.byref(V), ctypes.byref(W), ctypes.byref(LNS), ctypes.byref(Res), ctypes.byref(TR),
                         ctypes.byref(D))
    return rc, V, W, LNS, TR, D


def test_read_user_input_values_and_defaults(pkg, tmp_path, capfd):
    p = str(tmp_path / "v.yaml")
    open(p, "w").write(BASE_YAML.format(root=tmp_path) + "TwoPI: true\n")
    rc, V, W, LNS, TR, D = read_user_input(pkg, p)
    assert rc == 0
    assert V.RSVD.k == 5 and V.RSVD.q == 1 and V.RSVD.Nw == 16
    assert abs(V.RSVD.w - 0.5 * 2 * np.pi) < 1e-15        # TwoPI (SourceCode/ReadUserInput.c:82)
    assert V.TS.TransientLength == 3 and V.TS.dt == 0.02 and V.TS.TransientRemoval == 1
    assert V.RSVD.RandSeed == 1373 and V.Display == 2 and V.SaveResultsOpt == 1   # defaults
    assert D.OperatorDir == b"A"
    out = capfd.readouterr().out
    assert "Warning: 'RandSeed' variable not found" in out


@pytest.mark.parametrize("drop,msg", [("TransientLength", "Must specify 'TransientLength'"),
                                      ("dt:", "Must specify 'dt'"), ("w:", "Must specify 'dt'"),
                                      ("Nw", "Must specify 'Nw'"), ("RootDir", "Must specify 'RootDir'"),
                                      ("TransICDir", "Must specify 'TransICDir'")])
def test_read_user_input_required(pkg, tmp_path, capfd, drop, msg):
    txt = "".join(l + "\n" for l in BASE_YAML.format(root=tmp_path).splitlines() if not l.startswith(drop))
    p = str(tmp_path / "v.yaml")
    open(p, "w").write(txt)
    rc = read_user_input(pkg, p)[0]
    assert rc == 83
    assert msg in capfd.readouterr().err


def test_read_user_input_flag_dirs(pkg, tmp_path, capfd):
    p = str(tmp_path / "v.yaml")
    open(p, "w").write(BASE_YAML.format(root=tmp_path) + "InputForcingFlg: 1\n")
    assert read_user_input(pkg, p)[0] == 83
    assert "Must specify 'InputForcingDir'" in capfd.readouterr().err
    open(p, "w").write(BASE_YAML.format(root=tmp_path) + "DiscFlg: yes\n")
    assert read_user_input(pkg, p)[0] == 83
    assert "specify 'beta'" in capfd.readouterr().err
    open(p, "w").write(BASE_YAML.format(root=tmp_path) + "DiscFlg: yes\nbeta: 0.1\nPrecision: fp32\n")
    rc, V, W, LNS, TR, D = read_user_input(pkg, p)
    assert rc == 0 and LNS.RSVDt.Disc.DiscFlg == 1 and LNS.RSVDt.Disc.beta == 0.1
    assert pkg.get_options().precision == 32
    assert "'Precision: fp32' with transient removal: the steps before the first snapshot run in FP32" in \
        capfd.readouterr().out                                                      # TransientRemoval is on
    pkg.set_options(precision=64)


@pytest.mark.parametrize("w,Nw,TL,dt", [(0.05, 64, 10.0, 0.03068), (0.05, 42, 10.0, 0.003)])
def test_setup_grid_matches(pkg, w, Nw, TL, dt):
    V = pkg.RSVDt_vars()
    V.RSVD.w, V.RSVD.Nw, V.TS.TransientLength, V.TS.dt = w, Nw, TL, dt
    TR = pkg.TransRun_vars()
    assert pkg.lib().SetupTimeFreqGrid(ctypes.byref(V), ctypes.byref(TR)) == 0
    assert (V.TS.dt, V.TS.Ns, V.TS.Nt, V.TS.ResRatio) == kat.grid(w, Nw, TL, dt)


def test_matview_reproduces_reference_file(pkg, tmp_path):
    """MatLoad then MatView of the reference's own operator file writes its
    first object (the Mat: header, row lengths, columns, values -- explicit
    zeros included) byte for byte; the trailing Vec object of the file is not
    part of the Mat (SourceCode/LNSStructure.c:26 reads only the Mat)."""
    L = pkg.lib()
    m = ctypes.c_void_p()
    src = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "A_GL")
    assert L.RSVDT_MatLoad(src.encode(), 1, ctypes.byref(m)) == 0
    out = str(tmp_path / "A_view")
    assert L.RSVDT_MatView(m, out.encode()) == 0
    pkg.destroy(m)
    ref = open(src, "rb").read()
    hdr = np.frombuffer(ref[:32], dtype=">i8")
    M, nz = int(hdr[1]), int(hdr[3])
    nbytes = 32 + 8 * M + 8 * nz + 16 * nz
    got = open(out, "rb").read()
    assert len(got) == nbytes and got == ref[:nbytes]


def test_device_sketch_columns_arguments(pkg, capfd):
    """RSVDT_SetSketchColumnsDevice checks its sizes on the host (the device
    array is not read until CreateRandomMat): ld below 2 ncols or an empty
    grid is an error with the entry point's name; a null array clears any
    columns set before."""
    L = pkg.lib()
    fake = ctypes.c_void_p(0x1000)          # never dereferenced by the checks
    assert L.RSVDT_SetSketchColumnsDevice(500, 8, 32, fake, 15) != 0
    assert "RSVDT_SetSketchColumnsDevice" in capfd.readouterr().err
    assert L.RSVDT_SetSketchColumnsDevice(0, 8, 32, fake, 16) != 0
    capfd.readouterr()
    assert L.RSVDT_SetSketchColumnsDevice(500, 8, 32, None, 16) == 0
    assert L.RSVDT_SetSketchColumns(0, 0, 0, None) == 0
