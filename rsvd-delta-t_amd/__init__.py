"""ctypes binding of librsvdt (the MI355X RSVD-Delta-t engine).

The C library exports the reference's entry points with its struct types
(SourceCode/Variables.h:5-110); this module mirrors those structs with
ctypes so that tests and benchmarks drive the engine through exactly the
calls a reference driver makes (SourceCode/main.c:113-146):

    s = Session(A, k=10, q=0, Nw=64, w=0.05, TL=10, dt=0.0307, transient_removal=True)
    s.set_sketch(F)          # N x k*Nw_eff, frequency-major (column m*k + ik)
    s.direct()               # DirectActionRK4
    Y = s.get_Y()

The package directory name is not a Python identifier; load it with
importlib (see load()).  Every compute call goes to the HIP library; there
is no CPU fallback, and loading fails loudly if librsvdt.so is missing.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "lib", "librsvdt.so")
BINPATH = os.path.join(HERE, "bin", "rsvdt")
PETSC_MAX_PATH_LEN = 4096

_i64 = ctypes.c_int64
_int = ctypes.c_int
_dbl = ctypes.c_double
_vp = ctypes.c_void_p


class RSVD_vars(ctypes.Structure):
    _fields_ = [("N", _i64), ("Nb", _i64), ("Nc", _i64), ("k", _i64), ("q", _i64), ("Nw", _i64),
                ("Nw_eff", _i64), ("w", _dbl), ("TwoPI", _int), ("InputForcingFlg", _int), ("RandSeed", _i64)]


class TS_vars(ctypes.Structure):
    _fields_ = [("Ns", _i64), ("Nt", _i64), ("ResRatio", _i64), ("TransientLength", _dbl), ("dt", _dbl),
                ("RealOperator", _int), ("TransientRemoval", _int), ("DirAdj", _int)]


class Discounting(ctypes.Structure):
    _fields_ = [("DiscFlg", _int), ("beta", _dbl)]


class RSVDt_vars(ctypes.Structure):
    _fields_ = [("RSVD", RSVD_vars), ("TS", TS_vars), ("Disc", Discounting), ("Display", _i64),
                ("SaveResultsOpt", _i64)]


class TransRun_vars(ctypes.Structure):
    _fields_ = [("TransRun", _int), ("TransSave", _int), ("TransRemovalEst", _int), ("TransICFlg", _int),
                ("TransSaveMod", _i64), ("TransPeriods", _i64), ("TransDivVal", _dbl), ("TransConVal", _dbl)]


class DFT_matrices(ctypes.Structure):
    _fields_ = [("idft", _vp), ("dft", _vp)]


class LNS_vars(ctypes.Structure):
    _fields_ = [("A", _vp), ("RSVDt", RSVDt_vars)]


class Weight_matrices(ctypes.Structure):
    _fields_ = [("W_q_sqrt", _vp), ("W_q_sqrt_inv", _vp), ("W_f_sqrt_inv", _vp), ("B", _vp), ("C", _vp),
                ("InvInputWeightFlg", _int), ("OutputWeightFlg", _int), ("InvOutputWeightFlg", _int),
                ("InputMatrixFlg", _int), ("OutputMatrixFlg", _int)]


class RSVD_matrices(ctypes.Structure):
    _fields_ = [("F_hat", _vp), ("Y_hat", _vp)]


class TS_removal_matrices(ctypes.Structure):
    _fields_ = [("M_tilde_all", _vp), ("Y_hat_up", _vp)]


class Resolvent_matrices(ctypes.Structure):
    _fields_ = [("U_hat", _vp), ("V_hat", _vp), ("S_hat", _vp)]


class Directories(ctypes.Structure):
    _fields_ = [(n, ctypes.c_char * PETSC_MAX_PATH_LEN) for n in
                ("RootDir", "ResultsDir", "OperatorDir", "IO_dir", "FolderDir", "InputForcingDir",
                 "InvInputWeightDir", "OutputWeightDir", "InvOutputWeightDir", "InputMatrixDir",
                 "OutputMatrixDir", "TransICDir")]


class RSVDT_Options(ctypes.Structure):
    _fields_ = [("precision", _int), ("chunk_steps", _int), ("spmm_lanes", _int), ("xcd_remap", _int),
                ("use_graph", _int), ("autotune", _int), ("verbose", _int), ("overlap_forcing", _int), ("rk4_form", _int),
                ("fuse_small", _int), ("spmm_variant", _int), ("tr_group", _int)]


class RSVDT_ActionStats(ctypes.Structure):
    _fields_ = [("rk4_ms", _dbl), ("forcing_ms", _dbl), ("dft_ms", _dbl), ("tr_ms", _dbl), ("total_ms", _dbl),
                ("steps", _i64), ("rk4_launches", _i64), ("chunk_steps", _int), ("spmm_lanes", _int),
                ("spmm_variant", _int), ("spmm_xcd_remap", _int), ("rk4_form", _int)]


# Every symbol the C headers (include/rsvdt.h, include/rsvdt_hip.h) declare.
ENTRY_POINTS = [
    "PreProcessing", "ReadUserInput", "SetupTimeFreqGrid", "LNSStructure", "CreateDFTiDFTMats",
    "ReadWeightMats", "CreateRandomMat", "CreateResultsDir", "DirectActionRK4", "AdjointActionRK4",
    "PowerIterationRK4", "ApplyWeightMats", "TSActionRK4", "DFT", "PermuteMat", "ReversePermuteMat",
    "TransientRemovalStrategy", "CreateMtilde", "GalerkinProjAllModes", "TSExpADeltaT", "QRDecomposition",
    "QRAllFreqs", "SVDAllFreqsBeforeAdjoint", "StoreU", "SVDAllFreqs", "SaveResults", "SaveInputVarsCopy",
    "RSVDT_Initialize", "RSVDT_Finalize", "RSVDT_OptionsSetValue", "RSVDT_OptionsClear", "RSVDT_Time",
    "RSVDT_MatCreateDense", "RSVDT_MatCreateDenseBlocked", "RSVDT_MatGetDense", "RSVDT_MatCreateAIJ",
    "RSVDT_MatGetSize", "RSVDT_MatGetAIJ", "RSVDT_MatDestroy", "RSVDT_MatLoad", "RSVDT_MatView", "RSVDT_GetOptions",
    "RSVDT_SetOptions", "RSVDT_GetActionStats", "RSVDT_BenchSteps", "RSVDT_SetSketchColumns",
    "RSVDT_SetSketchColumnsDevice",
    "TransientRunRK4", "TransientRemovalEst", "TSTransRK4", "RSVDT_VecCreate", "RSVDT_VecGetValues",
    "RSVDT_VecGetSize", "RSVDT_VecDestroy", "RSVDT_VecLoad", "RSVDT_VecView",
]
HIP_ENTRY_POINTS = [
    "rsvdt_hip_init", "rsvdt_hip_last_error", "rsvdt_hip_malloc", "rsvdt_hip_free", "rsvdt_hip_release_cached",
    "rsvdt_hip_cached_bytes", "rsvdt_hip_sketch_cols_f32", "rsvdt_hip_idft_table", "rsvdt_hip_memcpy_h2d",
    "rsvdt_hip_memcpy_d2h", "rsvdt_hip_memcpy_d2d", "rsvdt_hip_zero", "rsvdt_hip_sync", "rsvdt_hip_mem_info",
    "rsvdt_hip_stream_create", "rsvdt_hip_stream_destroy", "rsvdt_hip_stream_wait_event", "rsvdt_hip_stream_sync",
    "rsvdt_hip_host_alloc", "rsvdt_hip_host_free", "rsvdt_hip_memcpy_d2h_async",
    "rsvdt_hip_event_create", "rsvdt_hip_event_destroy", "rsvdt_hip_event_record", "rsvdt_hip_event_elapsed_ms",
    "rsvdt_hip_rk4_stage", "rsvdt_hip_horner_stage", "rsvdt_hip_horner_fused", "rsvdt_hip_horner_fused_fits", "rsvdt_hip_horner_stage_ex", "rsvdt_hip_snap_slots",
    "rsvdt_hip_graph_begin", "rsvdt_hip_graph_end", "rsvdt_hip_graph_launch", "rsvdt_hip_graph_destroy", "rsvdt_hip_spmm_blocks", "rsvdt_hip_coef_gemm", "rsvdt_hip_bgram_work_bytes",
    "rsvdt_hip_bgram", "rsvdt_hip_bapply", "rsvdt_hip_bchol_inv", "rsvdt_hip_bdrop_mask", "rsvdt_hip_btrmm_upper",
    "rsvdt_hip_bjacobi_svd", "rsvdt_hip_bshift_solve_work_bytes", "rsvdt_hip_bshift_solve", "rsvdt_hip_zscale", "rsvdt_hip_zaxpy",
    "rsvdt_hip_z2c", "rsvdt_hip_c2z", "rsvdt_hip_zcopy3", "rsvdt_hip_transpose_rows", "rsvdt_hip_fill_random", "rsvdt_hip_znorm2", "rsvdt_hip_mlp_features", "rsvdt_hip_gemm_bf16",
    "rsvdt_hip_colsum", "rsvdt_hip_sketch_loss", "rsvdt_hip_adam", "rsvdt_hip_trans_synth",
    "rsvdt_qr_blocks_noalloc_bytes", "rsvdt_qr_blocks_noalloc",
]

_lib = None


def build():
    """Compile librsvdt.so and the rsvdt driver (hipcc for gfx950 + gcc)."""
    subprocess.run(["make", "-s", "-C", HERE, "-j8"], check=True)


def lib():
    """Load librsvdt.so; raises if it has not been built (no fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIBPATH):
        raise RuntimeError(f"librsvdt.so not built ({LIBPATH}); run __graft_entry__.build()")
    L = ctypes.CDLL(LIBPATH)
    P = ctypes.POINTER
    for n in ENTRY_POINTS + HIP_ENTRY_POINTS:
        getattr(L, n)   # raises AttributeError if a declared symbol is missing
    L.RSVDT_Time.restype = _dbl
    L.rsvdt_hip_last_error.restype = ctypes.c_char_p
    L.RSVDT_Initialize.argtypes = [_vp, _vp]
    L.RSVDT_OptionsSetValue.argtypes = [ctypes.c_char_p, ctypes.c_char_p]
    L.RSVDT_MatCreateDense.argtypes = [_i64, _i64, _vp, P(_vp)]
    L.RSVDT_MatCreateDenseBlocked.argtypes = [_i64, _i64, _i64, _vp, P(_vp)]
    L.RSVDT_MatGetDense.argtypes = [_vp, _vp]
    L.RSVDT_MatCreateAIJ.argtypes = [_i64, _i64, _vp, _vp, _vp, P(_vp)]
    L.RSVDT_MatGetSize.argtypes = [_vp, P(_i64), P(_i64)]
    L.RSVDT_MatGetAIJ.argtypes = [_vp, P(_i64), _vp, _vp, _vp, P(_int)]
    L.RSVDT_MatDestroy.argtypes = [P(_vp)]
    L.RSVDT_MatLoad.argtypes = [ctypes.c_char_p, _int, P(_vp)]
    L.RSVDT_MatView.argtypes = [_vp, ctypes.c_char_p]
    L.RSVDT_GetOptions.argtypes = [P(RSVDT_Options)]
    L.RSVDT_SetOptions.argtypes = [P(RSVDT_Options)]
    L.RSVDT_GetActionStats.argtypes = [P(RSVDT_ActionStats)]
    L.RSVDT_SetSketchColumns.argtypes = [_i64, _i64, _i64, _vp]
    L.RSVDT_SetSketchColumnsDevice.argtypes = [_i64, _i64, _i64, _vp, _i64]
    act = [P(RSVD_matrices), P(RSVDt_vars), P(LNS_vars), P(DFT_matrices), P(Weight_matrices), P(Directories),
           P(TS_removal_matrices)]
    for n in ("DirectActionRK4", "AdjointActionRK4", "PowerIterationRK4"):
        getattr(L, n).argtypes = act
    L.SetupTimeFreqGrid.argtypes = [P(RSVDt_vars), P(TransRun_vars)]
    L.CreateDFTiDFTMats.argtypes = [P(RSVDt_vars), P(DFT_matrices), P(LNS_vars)]
    L.QRAllFreqs.argtypes = [P(RSVD_matrices), P(RSVDt_vars)]
    L.TSExpADeltaT.argtypes = [P(LNS_vars), P(RSVDt_vars), _vp]
    L.QRDecomposition.argtypes = [_vp]
    L.SVDAllFreqsBeforeAdjoint.argtypes = [P(RSVD_matrices), P(RSVDt_vars)]
    L.StoreU.argtypes = [P(RSVD_matrices), P(Resolvent_matrices)]
    L.SVDAllFreqs.argtypes = [P(RSVD_matrices), P(RSVDt_vars), P(Weight_matrices), P(Resolvent_matrices)]
    L.TSActionRK4.argtypes = [P(RSVD_matrices), P(DFT_matrices), P(LNS_vars), P(RSVDt_vars), P(Directories),
                              P(TS_removal_matrices)]
    L.ApplyWeightMats.argtypes = [P(RSVD_matrices), P(RSVDt_vars), P(Weight_matrices), _int]
    L.PermuteMat.argtypes = [_vp, P(RSVDt_vars)]
    L.ReversePermuteMat.argtypes = [_vp, P(RSVDt_vars)]
    L.DFT.argtypes = [_vp, P(DFT_matrices), P(RSVD_matrices), P(RSVDt_vars)]
    L.TransientRemovalStrategy.argtypes = [_vp, P(LNS_vars), P(RSVD_matrices), P(RSVDt_vars),
                                           P(TS_removal_matrices), P(DFT_matrices)]
    L.RSVDT_BenchSteps.argtypes = [P(RSVD_matrices), P(RSVDt_vars), P(LNS_vars), P(DFT_matrices), _i64, _i64,
                                   P(_dbl), P(_dbl), P(_dbl)]
    L.PreProcessing.argtypes = [P(RSVDt_vars), P(Weight_matrices), P(LNS_vars), P(RSVD_matrices),
                                P(Resolvent_matrices), P(TransRun_vars), P(DFT_matrices), P(Directories)]
    L.SaveResults.argtypes = [P(Resolvent_matrices), P(RSVDt_vars), P(Directories)]
    L.TransientRunRK4.argtypes = [P(TransRun_vars), P(RSVDt_vars), P(LNS_vars), P(DFT_matrices), P(Directories)]
    L.TransientRemovalEst.argtypes = [_vp, _vp, _i64, P(RSVDt_vars), P(DFT_matrices), P(Directories)]
    L.TSTransRK4.argtypes = [P(LNS_vars), _vp, P(RSVDt_vars), _i64, _vp]
    L.RSVDT_VecCreate.argtypes = [_i64, _vp, P(_vp)]
    L.RSVDT_VecGetValues.argtypes = [_vp, _vp]
    L.RSVDT_VecGetSize.argtypes = [_vp, P(_i64)]
    L.RSVDT_VecDestroy.argtypes = [P(_vp)]
    L.RSVDT_VecLoad.argtypes = [ctypes.c_char_p, P(_vp)]
    L.RSVDT_VecView.argtypes = [_vp, ctypes.c_char_p]
    _lib = L
    return L


class RSVDTError(RuntimeError):
    pass


def check(rc, what):
    if rc:
        raise RSVDTError(f"{what} failed with code {rc}")


def _blocked_host(F):
    return np.asfortranarray(F, dtype=np.complex128)


def mat_from_dense(F, kb=None):
    """Device Mat from a host matrix (column-major semantics, kb-column blocks)."""
    F = _blocked_host(F)
    m = _vp()
    rows, cols = F.shape
    check(lib().RSVDT_MatCreateDenseBlocked(rows, cols, kb or 1, F.ctypes.data_as(_vp), ctypes.byref(m)),
          "RSVDT_MatCreateDenseBlocked")
    return m


def mat_to_dense(m):
    r, c = _i64(), _i64()
    check(lib().RSVDT_MatGetSize(m, ctypes.byref(r), ctypes.byref(c)), "RSVDT_MatGetSize")
    out = np.zeros((r.value, c.value), dtype=np.complex128, order="F")
    check(lib().RSVDT_MatGetDense(m, out.ctypes.data_as(_vp)), "RSVDT_MatGetDense")
    return out


def mat_from_csr(A):
    import scipy.sparse as sp
    A = sp.csr_matrix(A)
    A.sort_indices()
    rp = np.ascontiguousarray(A.indptr, dtype=np.int64)
    ci = np.ascontiguousarray(A.indices, dtype=np.int64)
    v = np.ascontiguousarray(A.data, dtype=np.complex128)
    m = _vp()
    check(lib().RSVDT_MatCreateAIJ(A.shape[0], A.shape[1], rp.ctypes.data_as(_vp), ci.ctypes.data_as(_vp),
                                   v.ctypes.data_as(_vp), ctypes.byref(m)), "RSVDT_MatCreateAIJ")
    return m


def load_aij(path):
    """MatLoad through the C reader (no GPU needed): returns scipy CSR and the real-operator flag."""
    import scipy.sparse as sp
    m = _vp()
    check(lib().RSVDT_MatLoad(path.encode(), 1, ctypes.byref(m)), "RSVDT_MatLoad")
    r, c, nnz, isr = _i64(), _i64(), _i64(), _int()
    lib().RSVDT_MatGetSize(m, ctypes.byref(r), ctypes.byref(c))
    lib().RSVDT_MatGetAIJ(m, ctypes.byref(nnz), None, None, None, ctypes.byref(isr))
    rp = np.zeros(r.value + 1, dtype=np.int64)
    ci = np.zeros(max(nnz.value, 1), dtype=np.int64)
    v = np.zeros(max(nnz.value, 1), dtype=np.complex128)
    lib().RSVDT_MatGetAIJ(m, None, rp.ctypes.data_as(_vp), ci.ctypes.data_as(_vp), v.ctypes.data_as(_vp), None)
    destroy(m)
    return sp.csr_matrix((v[:nnz.value], ci[:nnz.value], rp), shape=(r.value, c.value)), bool(isr.value)


def destroy(m):
    if m:
        h = _vp(m if isinstance(m, int) else m.value)
        lib().RSVDT_MatDestroy(ctypes.byref(h))


def vec_from_numpy(v):
    """A device Vec holding the complex vector v."""
    v = np.ascontiguousarray(v, dtype=np.complex128)
    h = _vp()
    check(lib().RSVDT_VecCreate(v.shape[0], v.ctypes.data_as(_vp), ctypes.byref(h)), "RSVDT_VecCreate")
    return h


def vec_to_numpy(h):
    n = _i64()
    check(lib().RSVDT_VecGetSize(h, ctypes.byref(n)), "RSVDT_VecGetSize")
    v = np.zeros(n.value, np.complex128)
    check(lib().RSVDT_VecGetValues(h, v.ctypes.data_as(_vp)), "RSVDT_VecGetValues")
    return v


def vec_destroy(h):
    lib().RSVDT_VecDestroy(ctypes.byref(h))


def vec_load(path):
    """VecLoad of a PETSc binary Vec file -> numpy complex vector."""
    h = _vp()
    check(lib().RSVDT_VecLoad(str(path).encode(), ctypes.byref(h)), f"RSVDT_VecLoad({path})")
    try:
        return vec_to_numpy(h)
    finally:
        vec_destroy(h)


def vec_view(v, path):
    """VecView of a numpy vector into a PETSc binary file."""
    h = vec_from_numpy(v)
    try:
        check(lib().RSVDT_VecView(h, str(path).encode()), f"RSVDT_VecView({path})")
    finally:
        vec_destroy(h)


def get_options():
    o = RSVDT_Options()
    check(lib().RSVDT_GetOptions(ctypes.byref(o)), "RSVDT_GetOptions")
    return o


# engine option defaults (RSVDT_Initialize's); a Session starts from these
OPTION_DEFAULTS = dict(precision=64, chunk_steps=0, spmm_lanes=0, xcd_remap=1, use_graph=0, autotune=1, verbose=0,
                       overlap_forcing=1, rk4_form=0, fuse_small=1, spmm_variant=-1, tr_group=0)


def set_options(**kw):
    o = get_options()
    for k, v in kw.items():
        if k not in OPTION_DEFAULTS:
            raise KeyError(f"unknown engine option {k!r}")
        setattr(o, k, int(v))
    check(lib().RSVDT_SetOptions(ctypes.byref(o)), "RSVDT_SetOptions")


def action_stats():
    s = RSVDT_ActionStats()
    lib().RSVDT_GetActionStats(ctypes.byref(s))
    return {f: getattr(s, f) for f, _ in RSVDT_ActionStats._fields_}


class Session:
    """One operator and run configuration, driven through the reference entry points."""

    def __init__(self, A, k, q, Nw, w, TL, dt, transient_removal, display=0, precision=64, **opts):
        L = lib()
        check(L.RSVDT_Initialize(None, None), "RSVDT_Initialize")
        set_options(**dict(OPTION_DEFAULTS, precision=precision, **opts))
        import scipy.sparse as sp
        A = sp.csr_matrix(A)
        self.V = RSVDt_vars()
        self.V.RSVD.k, self.V.RSVD.q, self.V.RSVD.Nw, self.V.RSVD.w = k, q, Nw, w
        self.V.TS.TransientLength, self.V.TS.dt = TL, dt
        self.V.TS.TransientRemoval = int(bool(transient_removal))
        self.V.Display = display
        self.V.SaveResultsOpt = 1
        self.TR = TransRun_vars()
        check(L.SetupTimeFreqGrid(ctypes.byref(self.V), ctypes.byref(self.TR)), "SetupTimeFreqGrid")
        # LNSStructure (SourceCode/LNSStructure.c:22-56) with the operator given in memory
        self.LNS = LNS_vars()
        self.LNS.A = mat_from_csr(A).value
        N = A.shape[0]
        real = np.linalg.norm(np.asarray(A.data).imag) < 1e-12
        self.V.RSVD.N = self.V.RSVD.Nb = self.V.RSVD.Nc = N
        self.V.TS.RealOperator = int(real)
        self.V.RSVD.Nw_eff = Nw // 2 if real else Nw
        self.LNS.RSVDt = self.V
        self.D = DFT_matrices()
        check(L.CreateDFTiDFTMats(ctypes.byref(self.V), ctypes.byref(self.D), ctypes.byref(self.LNS)),
              "CreateDFTiDFTMats")
        self.W = Weight_matrices()
        self.R = RSVD_matrices()
        self.TSR = TS_removal_matrices()
        self.Res = Resolvent_matrices()
        self.dirs = Directories()

    N = property(lambda s: s.V.RSVD.N)
    k = property(lambda s: s.V.RSVD.k)
    Nw_eff = property(lambda s: s.V.RSVD.Nw_eff)
    Ns = property(lambda s: s.V.TS.Ns)
    Nt = property(lambda s: s.V.TS.Nt)
    RR = property(lambda s: s.V.TS.ResRatio)
    dt = property(lambda s: s.V.TS.dt)

    def _args(self):
        return (ctypes.byref(self.R), ctypes.byref(self.V), ctypes.byref(self.LNS), ctypes.byref(self.D),
                ctypes.byref(self.W), ctypes.byref(self.dirs), ctypes.byref(self.TSR))

    def set_weights(self, B=None, C=None, Wq_sqrt=None, Wq_sqrt_inv=None, Wf_sqrt_inv=None):
        """ReadWeightMats (SourceCode/ReadWeightMats.c:26-99) with the matrices
        held in memory (dense or scipy sparse): input matrix B (N x Nb), output
        matrix C (Nc x N), output weight W_q^{1/2} and its inverse (Nc x Nc),
        inverse input weight W_f^{-1/2} (Nb x Nb).  Sets Nb and Nc."""
        import scipy.sparse as sp
        W = self.W
        for name, M, flag in (("B", B, "InputMatrixFlg"), ("C", C, "OutputMatrixFlg"),
                              ("W_q_sqrt", Wq_sqrt, "OutputWeightFlg"), ("W_q_sqrt_inv", Wq_sqrt_inv, "InvOutputWeightFlg"),
                              ("W_f_sqrt_inv", Wf_sqrt_inv, "InvInputWeightFlg")):
            destroy(getattr(W, name))
            setattr(W, name, None)
            setattr(W, flag, 0)
            if M is not None:
                setattr(W, name, mat_from_csr(sp.csr_matrix(M, dtype=np.complex128)).value)
                setattr(W, flag, 1)
        self.V.RSVD.Nb = B.shape[1] if B is not None else self.N
        self.V.RSVD.Nc = C.shape[0] if C is not None else self.N
        self.LNS.RSVDt = self.V

    def set_sketch(self, F):
        F = np.asarray(F)
        Nb = self.V.RSVD.Nb or self.N
        if F.shape != (Nb, self.k * self.Nw_eff):
            raise ValueError(f"sketch must be {Nb} x {self.k * self.Nw_eff}")
        destroy(self.R.Y_hat)
        self.R.Y_hat = mat_from_dense(F, self.k).value

    def random_sketch(self, seed=1373, gaussian=True):
        """CreateRandomMat (SourceCode/CreateRandomMat.c:33-39): a seeded sketch generated on the device."""
        destroy(self.R.Y_hat)
        self.R.Y_hat = None
        self.V.RSVD.InputForcingFlg = 0
        self.V.RSVD.RandSeed = seed
        lib().RSVDT_OptionsSetValue(b"RandDist", b"gaussian" if gaussian else b"uniform")
        L = lib()
        L.CreateRandomMat.argtypes = [ctypes.POINTER(RSVD_matrices), ctypes.POINTER(RSVDt_vars),
                                      ctypes.POINTER(Directories)]
        check(L.CreateRandomMat(ctypes.byref(self.R), ctypes.byref(self.V), ctypes.byref(self.dirs)),
              "CreateRandomMat")

    def get_Y(self):
        return mat_to_dense(_vp(self.R.Y_hat))

    def direct(self):
        check(lib().DirectActionRK4(*self._args()), "DirectActionRK4")

    def adjoint(self):
        check(lib().AdjointActionRK4(*self._args()), "AdjointActionRK4")

    def power_iteration(self):
        check(lib().PowerIterationRK4(*self._args()), "PowerIterationRK4")

    def qr(self):
        check(lib().QRAllFreqs(ctypes.byref(self.R), ctypes.byref(self.V)), "QRAllFreqs")

    def svd_before_adjoint(self):
        check(lib().SVDAllFreqsBeforeAdjoint(ctypes.byref(self.R), ctypes.byref(self.V)), "SVDAllFreqsBeforeAdjoint")

    def store_u(self):
        check(lib().StoreU(ctypes.byref(self.R), ctypes.byref(self.Res)), "StoreU")

    def svd_allfreqs(self):
        check(lib().SVDAllFreqs(ctypes.byref(self.R), ctypes.byref(self.V), ctypes.byref(self.W),
                                ctypes.byref(self.Res)), "SVDAllFreqs")

    def run(self, F):
        """The main.c sequence (SourceCode/main.c:134-144); returns (U_hat, V_hat, S_hat)."""
        self.set_sketch(F)
        self.direct()
        self.power_iteration()
        self.svd_before_adjoint()
        self.store_u()
        self.adjoint()
        self.svd_allfreqs()
        U = mat_to_dense(_vp(self.Res.U_hat))
        Vh = mat_to_dense(_vp(self.Res.V_hat))
        S = mat_to_dense(_vp(self.Res.S_hat)).real
        return U, Vh, S

    def bench_steps(self, nsteps, warmup=0, direct=True):
        self.V.TS.DirAdj = int(direct)
        ms, rk, fr = _dbl(), _dbl(), _dbl()
        check(lib().RSVDT_BenchSteps(ctypes.byref(self.R), ctypes.byref(self.V), ctypes.byref(self.LNS),
                                     ctypes.byref(self.D), nsteps, warmup, ctypes.byref(ms), ctypes.byref(rk),
                                     ctypes.byref(fr)), "RSVDT_BenchSteps")
        return ms.value, rk.value, fr.value

    def transient_run(self, folder, periods=1, save_mod=100, div_val=1e3, con_val=1e-8, est=False, save=False,
                      q0=None, seed=1373, direct=True):
        """TransientRunRK4 (SourceCode/TransientRunRK4.c:39-227) with this
        session's operator, w, Nw and dt: files go to `folder`; q0 (optional)
        is written there as the initial vector (TransICFlg), else a seeded
        random vector is used.  Returns the dt the run used."""
        L = lib()
        folder = os.path.join(os.path.abspath(folder), "")
        os.makedirs(folder, exist_ok=True)
        TR = TransRun_vars()
        TR.TransRun, TR.TransSave, TR.TransRemovalEst = 1, int(bool(save)), int(bool(est))
        TR.TransSaveMod, TR.TransPeriods, TR.TransDivVal, TR.TransConVal = save_mod, periods, div_val, con_val
        V = RSVDt_vars.from_buffer_copy(self.V)
        V.RSVD.RandSeed = seed
        V.TS.DirAdj = int(bool(direct))
        dirs = Directories()
        dirs.FolderDir = folder.encode()
        if q0 is not None:
            TR.TransICFlg = 1
            vec_view(q0, os.path.join(folder, "q0_input"))
            dirs.RootDir = folder.encode()
            dirs.TransICDir = b"q0_input"
        check(L.TransientRunRK4(ctypes.byref(TR), ctypes.byref(V), ctypes.byref(self.LNS), ctypes.byref(self.D),
                                ctypes.byref(dirs)), "TransientRunRK4")
        return V.TS.dt

    def trans_removal_est(self, Qt, qss, period_index, folder):
        """TransientRemovalEst (SourceCode/TransientRemovalEst.c:26-127) on
        N x Nw snapshots Qt; returns (initial norms, updated norms) read back
        from the files it writes into `folder`."""
        folder = os.path.join(os.path.abspath(folder), "")
        os.makedirs(folder, exist_ok=True)
        dirs = Directories()
        dirs.FolderDir = folder.encode()
        m = mat_from_dense(np.asarray(Qt, dtype=np.complex128), kb=1)
        v = vec_from_numpy(qss)
        try:
            check(lib().TransientRemovalEst(m, v, period_index, ctypes.byref(self.V), ctypes.byref(self.D),
                                            ctypes.byref(dirs)), "TransientRemovalEst")
        finally:
            destroy(m)
            vec_destroy(v)
        a = vec_load(os.path.join(folder, f"Initial_transient_norm_period_{period_index}"))
        b = vec_load(os.path.join(folder, f"Updated_transient_norm_period_{period_index}"))
        return a.real, b.real

    def close(self):
        for h in (self.R.Y_hat, self.R.F_hat, self.Res.U_hat, self.Res.V_hat, self.Res.S_hat, self.LNS.A,
                  self.D.idft, self.D.dft):
            destroy(h)
        self.R.Y_hat = self.R.F_hat = self.Res.U_hat = self.Res.V_hat = self.Res.S_hat = None
        self.LNS.A = self.D.idft = self.D.dft = None
        for name in ("B", "C", "W_q_sqrt", "W_q_sqrt_inv", "W_f_sqrt_inv"):
            destroy(getattr(self.W, name))
            setattr(self.W, name, None)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def load():
    """Import this package by path (its directory name is not an identifier)."""
    import importlib.util
    import sys
    if "rsvdt_amd" in sys.modules:
        return sys.modules["rsvdt_amd"]
    spec = importlib.util.spec_from_file_location("rsvdt_amd", os.path.join(HERE, "__init__.py"))
    mod = importlib.util.module_from_spec(spec)
    sys.modules["rsvdt_amd"] = mod
    spec.loader.exec_module(mod)
    return mod
