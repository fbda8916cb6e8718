"""Minimal PETSc binary format reader/writer used by the tests and fixture scripts.

Format (PETSc MatView/MatLoad, binary viewer, 64-bit indices, complex scalars):
big-endian int64 header [MAT_FILE_CLASSID=1211216, M, N, nz], then M int64
row lengths, nz int64 column indices, nz complex128 values (re, im).  A dense
matrix is written in the same AIJ layout (every entry stored).  nz == -1 marks
PETSc's native dense layout: M*N values in row-major order follow the header.
Vectors: [VEC_FILE_CLASSID=1211214, n] then n values.
"""
import numpy as np
import scipy.sparse as sp

MAT_FILE_CLASSID = 1211216
VEC_FILE_CLASSID = 1211214


def read_mat(path, index_bytes=8, scalar="complex"):
    it = np.dtype(">i8" if index_bytes == 8 else ">i4")
    st = np.dtype(">c16" if scalar == "complex" else ">f8")
    with open(path, "rb") as f:
        hdr = np.frombuffer(f.read(4 * it.itemsize), dtype=it)
        if hdr[0] != MAT_FILE_CLASSID:
            raise ValueError(f"{path}: not a PETSc Mat (classid {hdr[0]})")
        M, N, nz = int(hdr[1]), int(hdr[2]), int(hdr[3])
        if nz == -1:
            vals = np.frombuffer(f.read(M * N * st.itemsize), dtype=st)
            return sp.csr_matrix(vals.astype(np.complex128).reshape(M, N))
        rl = np.frombuffer(f.read(M * it.itemsize), dtype=it).astype(np.int64)
        cols = np.frombuffer(f.read(nz * it.itemsize), dtype=it).astype(np.int64)
        vals = np.frombuffer(f.read(nz * st.itemsize), dtype=st).astype(np.complex128)
    indptr = np.zeros(M + 1, dtype=np.int64)
    np.cumsum(rl, out=indptr[1:])
    return sp.csr_matrix((vals, cols, indptr), shape=(M, N))


def write_mat(path, A):
    """Write a sparse or dense matrix in PETSc AIJ binary layout (complex, int64)."""
    if sp.issparse(A):
        A = sp.csr_matrix(A)
        A.sort_indices()
        indptr, cols, vals = A.indptr, A.indices, A.data
        M, N = A.shape
    else:
        A = np.asarray(A)
        M, N = A.shape
        indptr = np.arange(M + 1, dtype=np.int64) * N
        cols = np.tile(np.arange(N, dtype=np.int64), M)
        vals = A.reshape(-1)
    rl = np.diff(indptr).astype(">i8")
    with open(path, "wb") as f:
        f.write(np.array([MAT_FILE_CLASSID, M, N, len(vals)], dtype=">i8").tobytes())
        f.write(rl.tobytes())
        f.write(np.asarray(cols, dtype=">i8").tobytes())
        f.write(np.asarray(vals, dtype=np.complex128).astype(">c16").tobytes())


def read_dense(path):
    return read_mat(path).toarray()


def write_vec(path, v):
    v = np.asarray(v, dtype=np.complex128)
    with open(path, "wb") as f:
        f.write(np.array([VEC_FILE_CLASSID, len(v)], dtype=">i8").tobytes())
        f.write(v.astype(">c16").tobytes())


def read_vec(path):
    """PETSc binary Vec (64-bit header, complex scalars) -> complex vector."""
    with open(path, "rb") as f:
        hdr = np.frombuffer(f.read(16), dtype=">i8")
        if hdr[0] != VEC_FILE_CLASSID:
            raise ValueError(f"{path}: not a PETSc Vec (classid {hdr[0]})")
        n = int(hdr[1])
        v = np.frombuffer(f.read(16 * n), dtype=">c16").astype(np.complex128)
        if v.shape[0] != n or f.read(1):
            raise ValueError(f"{path}: size mismatch")
    return v
