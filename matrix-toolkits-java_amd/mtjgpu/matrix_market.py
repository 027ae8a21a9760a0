"""Matrix Market input (mtjg_mm_*): MTJ's CompRowMatrix(MatrixVectorReader)
and DenseVector(MatrixVectorReader) constructors (S/CompRowMatrix.java:86-166,
M/DenseVector.java:52-96).  Parsing runs in the host library; the results are
host arrays or device objects."""
import ctypes as C
import os

import numpy as np

from . import _lib as L
from . import CompRowMatrix, DenseVector, _ck, default_context

# MatrixInfo.MatrixSymmetry of a parsed matrix (MTJG_MM_* in the C header)
GENERAL, SYMMETRIC, SKEW_SYMMETRIC = 0, 1, 2


class _Parsed:
    def __init__(self, handle):
        self._lib = L.load()
        self.handle = handle

    def shape(self):
        r, c, z, s = C.c_int32(), C.c_int32(), C.c_int64(), C.c_int32()
        _ck(self._lib.mtjg_mm_shape(self.handle, C.byref(r), C.byref(c), C.byref(z), C.byref(s)))
        return r.value, c.value, z.value, s.value

    def close(self):
        if self.handle:
            self._lib.mtjg_mm_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def _open_matrix(source):
    lib = L.load()
    h = C.c_void_p()
    if isinstance(source, (bytes, bytearray)):
        buf = bytes(source)
        _ck(lib.mtjg_mm_parse_matrix(buf, len(buf), C.byref(h)))
    else:
        _ck(lib.mtjg_mm_read_matrix(os.fsencode(source), C.byref(h)))
    return _Parsed(h)


def _open_vector(source):
    lib = L.load()
    h = C.c_void_p()
    if isinstance(source, (bytes, bytearray)):
        buf = bytes(source)
        _ck(lib.mtjg_mm_parse_vector(buf, len(buf), C.byref(h)))
    else:
        _ck(lib.mtjg_mm_read_vector(os.fsencode(source), C.byref(h)))
    return _Parsed(h)


def read_matrix(source):
    """Parses a Matrix Market matrix (a path, or the file's bytes).  Returns
    (rows, cols, row_pointer, column_index, data, symmetry) exactly as MTJ's
    CompRowMatrix(MatrixVectorReader) stores it."""
    with _open_matrix(source) as m:
        r, c, z, s = m.shape()
        rp = np.empty(r + 1, np.int32)
        ci = np.empty(z, np.int32)
        va = np.empty(z, np.float64)
        _ck(m._lib.mtjg_mm_copy_csr(m.handle, rp.ctypes.data, ci.ctypes.data if z else None,
                                     va.ctypes.data if z else None))
        return r, c, rp, ci, va, s


def read_vector(source):
    """Parses a Matrix Market vector (MTJ's "%%MatrixMarket vector ..."
    format) into a dense float64 array."""
    with _open_vector(source) as m:
        n = m.shape()[0]
        x = np.empty(n, np.float64)
        if n:
            _ck(m._lib.mtjg_mm_copy_vector(m.handle, x.ctypes.data))
        return x


def load_matrix(source, ctx=None):
    """new CompRowMatrix(new MatrixVectorReader(...)) on the device."""
    ctx = ctx or default_context()
    with _open_matrix(source) as m:
        h = C.c_void_p()
        _ck(m._lib.mtjg_mm_to_csr(ctx.handle, m.handle, C.byref(h)))
    A = CompRowMatrix.__new__(CompRowMatrix)
    A.ctx = ctx
    A._lib = ctx._lib
    A.handle = h
    return A


def load_vector(source, ctx=None):
    """new DenseVector(new MatrixVectorReader(...)) on the device."""
    ctx = ctx or default_context()
    x = read_vector(source)
    return DenseVector(x, ctx)
