"""Device buffers through the library's own C ABI (no torch)."""
import ctypes

import numpy as np


class Dev:
    def __init__(self, pkg, arr=None, nbytes=None):
        self.L = pkg.lib()
        self.L.rsvdt_hip_malloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.L.rsvdt_hip_memcpy_h2d.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t]
        self.L.rsvdt_hip_memcpy_d2h.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t]
        self.L.rsvdt_hip_free.argtypes = [ctypes.c_void_p]
        self.p = ctypes.c_void_p()
        self.nbytes = int(arr.nbytes if arr is not None else nbytes)
        assert self.L.rsvdt_hip_malloc(ctypes.byref(self.p), max(self.nbytes, 16)) == 0
        if arr is not None:
            a = np.ascontiguousarray(arr)
            assert self.L.rsvdt_hip_memcpy_h2d(self.p, a.ctypes.data_as(ctypes.c_void_p), self.nbytes) == 0

    def get(self, dtype, shape):
        out = np.zeros(shape, dtype=dtype)
        assert out.nbytes <= self.nbytes
        assert self.L.rsvdt_hip_sync() == 0
        assert self.L.rsvdt_hip_memcpy_d2h(out.ctypes.data_as(ctypes.c_void_p), self.p, out.nbytes) == 0
        return out

    def __del__(self):
        try:
            self.L.rsvdt_hip_free(self.p)
        except Exception:
            pass


def init(pkg):
    L = pkg.lib()
    rc = L.rsvdt_hip_init(0)
    assert rc == 0, L.rsvdt_hip_last_error()
    return L


V_ = ctypes.c_void_p
I64_ = ctypes.c_int64


class Desc(ctypes.Structure):
    """Mirror of rsvdt_spmm_desc (include/rsvdt_hip.h)."""
    _fields_ = [("nrows", I64_), ("k", I64_), ("nb", I64_), ("rowptr", V_), ("col", V_), ("val", V_),
                ("a_real", ctypes.c_int), ("prec", ctypes.c_int), ("lanes", ctypes.c_int), ("xcd_remap", ctypes.c_int),
                ("variant", ctypes.c_int), ("group", ctypes.c_int), ("gptr", V_), ("gword", V_), ("gval", V_),
                ("tile_rows", ctypes.c_int), ("tile_hmax", ctypes.c_int), ("tile_emax", ctypes.c_int),
                ("tptr", V_), ("tcol", V_), ("lidx", V_), ("tile_bpw", ctypes.c_int)]


def tiled_form(A, TR):
    """Tiled form of a CSR matrix (rsvdt_spmm_desc.tile_rows/tptr/tcol/lidx):
    per tile of TR consecutive rows its distinct columns, ascending, and for
    each nonzero the position of its column in its tile's list -- built here
    with numpy, independently of the library's own builder.  Returns
    (tptr, tcol, lidx, hmax, emax)."""
    A = A.tocsr()
    A.sort_indices()
    N = A.shape[0]
    nt = (N + TR - 1) // TR
    rows = np.repeat(np.arange(N, dtype=np.int64), np.diff(A.indptr))
    tile = rows // TR
    key = tile * N + A.indices.astype(np.int64)
    uniq, inv = np.unique(key, return_inverse=True)
    counts = np.bincount(uniq // N, minlength=nt)
    tptr = np.zeros(nt + 1, np.int32)
    np.cumsum(counts, out=tptr[1:])
    lidx = (inv - tptr[tile]).astype(np.uint16)
    tcol = (uniq % N).astype(np.int32)
    ends = np.minimum(np.arange(1, nt + 1) * TR, N)
    emax = int(np.max(A.indptr[ends] - A.indptr[np.arange(nt) * TR]))
    return tptr, tcol, lidx, int(counts.max()), emax


def grouped_form(A, G):
    """Row-grouped form of a CSR matrix (rsvdt_spmm_desc.group/gptr/gword/gval):
    per group of G rows its distinct columns, ascending, as (col << 8 | row mask),
    and the values in (column, row) order -- built here independently of the
    library's own builder."""
    A = A.tocsr()
    A.sort_indices()
    N = A.shape[0]
    rows = np.repeat(np.arange(N), np.diff(A.indptr))
    grp, r = rows // G, rows % G
    order = np.lexsort((r, A.indices, grp))
    g_s, c_s, r_s = grp[order], A.indices[order].astype(np.int64), r[order]
    vals = A.data[order]
    new = np.ones(len(order), bool)
    new[1:] = (g_s[1:] != g_s[:-1]) | (c_s[1:] != c_s[:-1])
    starts = np.flatnonzero(new)
    words = (c_s[starts] << 8) | np.bitwise_or.reduceat(np.left_shift(1, r_s), starts)
    ng = (N + G - 1) // G
    counts = np.bincount(g_s[starts], minlength=ng)
    gptr = np.zeros(ng + 1, np.int32)
    np.cumsum(counts, out=gptr[1:])
    return gptr, words.astype(np.uint32), vals
