"""Seeded synthetic systems for the CSR solve path (libmtjproblems.so).

Each generator returns (n_rows, n_cols, row_pointer, column_index, data) as
numpy arrays (int32, int32, float64) with sorted, unique columns and a stored
diagonal in every row.  The same seed always produces the same bits.
"""
import ctypes as C
import os

import numpy as np

from ._lib import LIB_DIR

_PATH = os.path.join(LIB_DIR, "libmtjproblems.so")
_lib = None


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(_PATH):
            raise ImportError("libmtjproblems.so is not built; run `make -C matrix-toolkits-java_amd`")
        lib = C.CDLL(_PATH)
        P = C.c_void_p
        lib.mtjp_fill_uniform.argtypes = [C.c_int64, C.c_uint64, P]
        lib.mtjp_fill_normal.argtypes = [C.c_int64, C.c_uint64, P]
        lib.mtjp_stencil_nnz.restype = C.c_int64
        lib.mtjp_stencil_nnz.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int]
        lib.mtjp_poisson.restype = C.c_int64
        lib.mtjp_poisson.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, P, P, P]
        lib.mtjp_varcoef_poisson.restype = C.c_int64
        lib.mtjp_varcoef_poisson.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_uint64, P, P, P]
        lib.mtjp_convdiff3d.restype = C.c_int64
        lib.mtjp_convdiff3d.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, C.c_uint64, P, P, P, P]
        lib.mtjp_powerlaw_lengths.restype = C.c_int64
        lib.mtjp_powerlaw_lengths.argtypes = [C.c_int32, C.c_double, C.c_int32, C.c_uint64, P]
        lib.mtjp_powerlaw_fill.argtypes = [C.c_int32, P, C.c_uint64, P, P, P]
        lib.mtjp_symmetrize.restype = C.c_int64
        lib.mtjp_symmetrize.argtypes = [C.c_int32, P, P, P, P, P, P]
        lib.mtjp_write_mm.argtypes = [C.c_char_p, C.c_int32, C.c_int32, P, P, P, C.c_int]
        _lib = lib
    return _lib


def uniform(n, seed):
    """n values uniform in [-1, 1) from splitmix64(seed)."""
    out = np.empty(n, np.float64)
    _load().mtjp_fill_uniform(n, seed, out.ctypes.data)
    return out


def normal(n, seed):
    out = np.empty(n, np.float64)
    _load().mtjp_fill_normal(n, seed, out.ctypes.data)
    return out


def _alloc(dim, nx, ny, nz):
    n = nx * ny * (nz if dim == 3 else 1)
    nnz = int(_load().mtjp_stencil_nnz(dim, nx, ny, nz if dim == 3 else 1))
    return n, np.empty(n + 1, np.int32), np.empty(nnz, np.int32), np.empty(nnz, np.float64)


def poisson2d(nx, ny=None):
    """5-point Laplacian on nx*ny (diagonal 4, off-diagonal -1)."""
    ny = ny or nx
    n, rp, ci, va = _alloc(2, nx, ny, 1)
    _load().mtjp_poisson(2, nx, ny, 1, rp.ctypes.data, ci.ctypes.data, va.ctypes.data)
    return n, n, rp, ci, va


def poisson3d(nx, ny=None, nz=None):
    """7-point Laplacian on nx*ny*nz (diagonal 6, off-diagonal -1)."""
    ny = ny or nx
    nz = nz or nx
    n, rp, ci, va = _alloc(3, nx, ny, nz)
    _load().mtjp_poisson(3, nx, ny, nz, rp.ctypes.data, ci.ctypes.data, va.ctypes.data)
    return n, n, rp, ci, va


def varcoef_poisson(dim, nx, ny=None, nz=None, seed=1):
    """Cell-centred Poisson with kappa ~ lognormal(0,1) per cell and
    harmonic-mean face coefficients (SPD)."""
    ny = ny or nx
    nz = (nz or nx) if dim == 3 else 1
    n, rp, ci, va = _alloc(dim, nx, ny, nz)
    _load().mtjp_varcoef_poisson(dim, nx, ny, nz, seed, rp.ctypes.data, ci.ctypes.data,
                                 va.ctypes.data)
    return n, n, rp, ci, va


def convdiff3d(nx, ny=None, nz=None, peclet=10.0, seed=1):
    """3-D 7-point first-order upwind convection-diffusion, unit velocity drawn
    from the seed, diffusion 1/peclet, scaled by h^2 (nonsymmetric)."""
    ny = ny or nx
    nz = nz or nx
    n, rp, ci, va = _alloc(3, nx, ny, nz)
    vel = np.empty(3, np.float64)
    _load().mtjp_convdiff3d(nx, ny, nz, peclet, seed, rp.ctypes.data, ci.ctypes.data,
                            va.ctypes.data, vel.ctypes.data)
    return n, n, rp, ci, va


def powerlaw(n, alpha=2.1, kmax=2000, seed=0):
    """Random CSR with off-diagonal row lengths P(k) ~ k^-alpha on [1, kmax],
    distinct uniform columns, N(0,1) values and diagonal 1 + sum |a_ij|."""
    lib = _load()
    lens = np.empty(n, np.int32)
    nnz = int(lib.mtjp_powerlaw_lengths(n, alpha, kmax, seed, lens.ctypes.data))
    rp = np.empty(n + 1, np.int32)
    ci = np.empty(nnz, np.int32)
    va = np.empty(nnz, np.float64)
    lib.mtjp_powerlaw_fill(n, lens.ctypes.data, seed, rp.ctypes.data, ci.ctypes.data, va.ctypes.data)
    return n, n, rp, ci, va


def powerlaw_spd(n, alpha=2.1, kmax=2000, seed=0):
    """SPD companion of powerlaw(): S = B_off + B_off^T with every diagonal
    entry 1 + sum_j |s_ij| (strictly diagonally dominant, positive diagonal),
    where B_off is powerlaw(n, alpha, kmax, seed) without its diagonal.  Row i
    of S holds the union of row i and column i of B_off plus the diagonal."""
    n, _, rp, ci, va = powerlaw(n, alpha, kmax, seed)
    lib = _load()
    srp = np.empty(n + 1, np.int32)
    nnz = int(lib.mtjp_symmetrize(n, rp.ctypes.data, ci.ctypes.data, va.ctypes.data, srp.ctypes.data, None, None))
    if nnz < 0:
        raise MemoryError("symmetrize: out of host memory")
    sci = np.empty(nnz, np.int32)
    sva = np.empty(nnz, np.float64)
    lib.mtjp_symmetrize(n, rp.ctypes.data, ci.ctypes.data, va.ctypes.data, srp.ctypes.data, sci.ctypes.data,
                        sva.ctypes.data)
    return n, n, srp, sci, sva


def spmv_bytes(n_rows, nnz, value_bytes=8, index_bytes=4):
    """Algorithmic HBM bytes of one y = A x: values and column indices once,
    row pointers once, x read once, y written once."""
    return (value_bytes + index_bytes) * nnz + index_bytes * (n_rows + 1) + 2 * value_bytes * n_rows


def write_matrix_market(path, n_rows, n_cols, rp, ci, va, lower_only=False):
    """Writes CSR arrays as a Matrix Market file in MTJ's MatrixVectorWriter
    layout (lower triangle with a "symmetric" header if lower_only)."""
    rp = np.ascontiguousarray(rp, np.int32)
    ci = np.ascontiguousarray(ci, np.int32)
    va = np.ascontiguousarray(va, np.float64)
    if _load().mtjp_write_mm(os.fsencode(path), int(n_rows), int(n_cols), rp.ctypes.data,
                             ci.ctypes.data, va.ctypes.data, int(bool(lower_only))) != 0:
        raise OSError("cannot write %s" % path)
