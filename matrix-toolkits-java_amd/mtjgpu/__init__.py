"""MTJ's sparse solve API over the MI355X C ABI (libmtjgpu.so).

Class and method names follow no.uib.cipr.matrix(.sparse): DenseVector,
CompRowMatrix (mult / multAdd / transMult / transMultAdd), DiagonalPreconditioner,
ILU, DefaultIterationMonitor, CG, BiCGstab, GMRES, IR, and the exceptions
IterativeSolverNotConvergedException (with NotConvergedException.Reason),
IndexOutOfBoundsException, IllegalArgumentException and
MatrixSingularException.  Every operand lives in device memory; numpy arrays
cross the boundary only through explicit uploads and downloads.

Reference citations: S/ = src/main/java/no/uib/cipr/matrix/sparse/,
M/ = src/main/java/no/uib/cipr/matrix/.
"""
import ctypes as C
import enum
import sys

import numpy as np

from . import _lib as L
from ._lib import load

__all__ = [
    "Context", "DenseVector", "CompRowMatrix", "DiagonalPreconditioner", "ILU", "ICC", "SSOR", "AMG",
    "DefaultIterationMonitor", "CG", "BiCGstab", "BiCG", "CGS", "QMR", "GMRES", "IR", "MixedCG",
    "Norm", "Reason", "IterativeSolverNotConvergedException", "IndexOutOfBoundsException",
    "IllegalArgumentException", "MatrixSingularException", "MtjgError",
    "UnsupportedOperationException", "IOException", "ZeroDiagonalException",
    "MissingDiagonalException", "ZeroPivotException", "default_context",
]


# ---------------------------------------------------------------------------
# exceptions (MTJ's taxonomy)

class MtjgError(RuntimeError):
    """Any failure reported by the library (RuntimeException in MTJ)."""

    def __init__(self, code, msg):
        super().__init__(msg)
        self.code = code


class IndexOutOfBoundsException(MtjgError, IndexError):
    pass


class IllegalArgumentException(MtjgError, ValueError):
    pass


class UnsupportedOperationException(MtjgError):
    pass


class IOException(MtjgError, OSError):
    """java.io.IOException / EOFException (unreadable or malformed input)."""


class MatrixSingularException(MtjgError):
    pass


class ZeroDiagonalException(MtjgError):
    """RuntimeException("Zero diagonal on row i") (S/DiagonalPreconditioner.java:71-72)."""

    def __init__(self, code, msg, row):
        super().__init__(code, msg)
        self.row = row


class MissingDiagonalException(ZeroDiagonalException):
    """RuntimeException("Missing diagonal entry on row k") (S/ILU.java:134-136)."""


class ZeroPivotException(ZeroDiagonalException):
    """RuntimeException("Zero pivot encountered ...") (S/ILU.java:105-107)."""


class Reason(enum.Enum):
    """NotConvergedException.Reason (M/NotConvergedException.java:33-49)."""
    Iterations = L.NOT_CONVERGED_ITERATIONS
    Divergence = L.NOT_CONVERGED_DIVERGENCE
    Breakdown = L.NOT_CONVERGED_BREAKDOWN


class IterativeSolverNotConvergedException(MtjgError):
    """S/IterativeSolverNotConvergedException.java:29-91."""

    def __init__(self, reason, iterations, residual, msg=""):
        super().__init__(reason.value, msg or "Solver did not converge: %s" % reason.name)
        self.reason = reason
        self.iterations = iterations
        self.residual = residual

    def getReason(self):
        return self.reason

    def getIterations(self):
        return self.iterations

    def getResidual(self):
        return self.residual


class Norm(enum.Enum):
    """Vector.Norm (M/Vector.java:159-180)."""
    One = L.NORM_ONE
    Two = L.NORM_TWO
    TwoRobust = L.NORM_TWO_ROBUST
    Infinity = L.NORM_INFINITY


def _raise(code, row=None):
    lib = load()
    msg = (lib.mtjg_last_error() or b"").decode() or lib.mtjg_strerror(code).decode()
    if code == L.E_INDEX_OUT_OF_BOUNDS:
        raise IndexOutOfBoundsException(code, msg)
    if code == L.E_INVALID_ARGUMENT:
        raise IllegalArgumentException(code, msg)
    if code == L.E_SINGULAR:
        raise MatrixSingularException(code, msg)
    if code == L.E_UNSUPPORTED:
        raise UnsupportedOperationException(code, msg)
    if code == L.E_IO:
        raise IOException(code, msg)
    if code == L.E_ZERO_DIAGONAL:
        raise ZeroDiagonalException(code, msg, row)
    if code == L.E_MISSING_DIAGONAL:
        raise MissingDiagonalException(code, msg, row)
    if code == L.E_ZERO_PIVOT:
        raise ZeroPivotException(code, msg, row)
    raise MtjgError(code, msg)


def _finalize(obj):
    """Releases a handle when its Python object dies.  At interpreter exit the
    process is ending and the driver reclaims device memory, and handles may
    be finalized in any order, so nothing is released then."""
    if sys.is_finalizing():
        return
    try:
        obj.close()
    except Exception:
        pass


def _ck(code):
    if code != L.MTJG_OK:
        _raise(code)


# ---------------------------------------------------------------------------
# context

class Context:
    """Device + HIP stream + reduction workspace (one per host thread)."""

    def __init__(self, device=0):
        self._lib = load()
        h = C.c_void_p()
        _ck(self._lib.mtjg_ctx_create(device, C.byref(h)))
        self.handle = h
        self.device = device

    def synchronize(self):
        _ck(self._lib.mtjg_ctx_synchronize(self.handle))

    @property
    def stream(self):
        return self._lib.mtjg_ctx_stream(self.handle)

    def close(self):
        if getattr(self, "handle", None):
            self._lib.mtjg_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):
        _finalize(self)


_default = None


def default_context():
    global _default
    if _default is None:
        _default = Context(0)
    return _default


# ---------------------------------------------------------------------------
# DenseVector

class DenseVector:
    """Device-resident DenseVector (M/DenseVector.java)."""

    def __init__(self, data, ctx=None, dtype=np.float64):
        self.ctx = ctx or default_context()
        self._lib = self.ctx._lib
        if isinstance(data, DenseVector):  # copy constructor
            src = data
            data = src.size()
            dtype = src.dtype
        else:
            src = None
        self.dtype = np.dtype(dtype)
        code = L.F64 if self.dtype == np.float64 else L.F32
        h = C.c_void_p()
        if np.isscalar(data):
            n = int(data)
            _ck(self._lib.mtjg_vec_create(self.ctx.handle, n, code, C.byref(h)))
            self.handle = h
            if src is not None:
                self.set(src)
        else:
            arr = np.ascontiguousarray(np.asarray(data, dtype=self.dtype))
            _ck(self._lib.mtjg_vec_create(self.ctx.handle, arr.size, code, C.byref(h)))
            self.handle = h
            _ck(self._lib.mtjg_vec_upload(h, arr.ctypes.data))

    # sizes and transfers
    def size(self):
        return int(self._lib.mtjg_vec_size(self.handle))

    def __len__(self):
        return self.size()

    def device_ptr(self):
        return self._lib.mtjg_vec_device_ptr(self.handle)

    def getData(self):
        """Host copy of the contents (DenseVector.getData, M/DenseVector.java:324)."""
        out = np.empty(self.size(), dtype=self.dtype)
        _ck(self._lib.mtjg_vec_download(self.handle, out.ctypes.data))
        return out

    numpy = getData

    def upload(self, arr):
        arr = np.ascontiguousarray(np.asarray(arr, dtype=self.dtype))
        if arr.size != self.size():
            raise IndexOutOfBoundsException(L.E_INDEX_OUT_OF_BOUNDS, "size mismatch")
        _ck(self._lib.mtjg_vec_upload(self.handle, arr.ctypes.data))
        return self

    def get(self, i):
        return float(self.getData()[i])

    def copy(self):
        return DenseVector(self, self.ctx)

    # MTJ Vector operations (M/DenseVector.java:189-318)
    def zero(self):
        _ck(self._lib.mtjg_vec_zero(self.handle))
        return self

    def scale(self, alpha):
        _ck(self._lib.mtjg_vec_scale(self.handle, float(alpha)))
        return self

    def set(self, a, y=None):
        if y is None:
            _ck(self._lib.mtjg_vec_set(self.handle, a.handle))
        else:
            _ck(self._lib.mtjg_vec_set_scaled(self.handle, float(a), y.handle))
        return self

    def add(self, a, y=None):
        if y is None:
            _ck(self._lib.mtjg_vec_add(self.handle, a.handle))
        else:
            _ck(self._lib.mtjg_vec_add_scaled(self.handle, float(a), y.handle))
        return self

    def dot(self, y):
        out = C.c_double()
        _ck(self._lib.mtjg_vec_dot(self.handle, y.handle, C.byref(out)))
        return out.value

    def norm(self, norm_type=Norm.Two):
        out = C.c_double()
        _ck(self._lib.mtjg_vec_norm(self.handle, norm_type.value, C.byref(out)))
        return out.value

    def close(self):
        if getattr(self, "handle", None):
            self._lib.mtjg_vec_destroy(self.handle)
            self.handle = None

    def __del__(self):
        _finalize(self)


# ---------------------------------------------------------------------------
# CompRowMatrix

class CompRowMatrix:
    """Device-resident CSR matrix (S/CompRowMatrix.java)."""

    def __init__(self, n_rows, n_cols=None, row_pointer=None, column_index=None, data=None,
                 ctx=None):
        self.ctx = ctx or default_context()
        self._lib = self.ctx._lib
        if hasattr(n_rows, "indptr"):  # scipy.sparse CSR
            m = n_rows.tocsr()
            n_rows, n_cols = m.shape
            row_pointer, column_index, data = m.indptr, m.indices, m.data
        rp = np.ascontiguousarray(row_pointer, dtype=np.int32)
        ci = np.ascontiguousarray(column_index, dtype=np.int32)
        va = np.ascontiguousarray(data, dtype=np.float64)
        h = C.c_void_p()
        _ck(self._lib.mtjg_csr_create(self.ctx.handle, int(n_rows), int(n_cols), int(ci.size),
                                      rp.ctypes.data, ci.ctypes.data if ci.size else None,
                                      va.ctypes.data if va.size else None, C.byref(h)))
        self.handle = h

    def shape(self):
        r, c, z = C.c_int32(), C.c_int32(), C.c_int64()
        _ck(self._lib.mtjg_csr_shape(self.handle, C.byref(r), C.byref(c), C.byref(z)))
        return r.value, c.value, z.value

    def format(self):
        """("csr" | "sell" | "csr_xcd", entries streamed by a product incl.
        padding)."""
        f, z = C.c_int32(), C.c_int64()
        _ck(self._lib.mtjg_csr_format(self.handle, C.byref(f), C.byref(z)))
        return {0: "csr", 1: "sell", 2: "csr_xcd"}[f.value], z.value

    def fp32_diagonals(self):
        """Diagonals of the FP32 stencil copy the mixed-precision products
        read (0: they use SELL-64 or CSR)."""
        c = C.c_int32()
        _ck(self._lib.mtjg_csr_fp32_diagonals(self.handle, C.byref(c)))
        return c.value

    def numRows(self):
        return self.shape()[0]

    def numColumns(self):
        return self.shape()[1]

    def isSquare(self):
        r, c, _ = self.shape()
        return r == c

    def arrays(self):
        """(getRowPointers, getColumnIndices, getData) as host copies."""
        r, _, z = self.shape()
        rp = np.empty(r + 1, np.int32)
        ci = np.empty(z, np.int32)
        va = np.empty(z, np.float64)
        _ck(self._lib.mtjg_csr_download(self.handle, rp.ctypes.data, ci.ctypes.data, va.ctypes.data))
        return rp, ci, va

    def getRowPointers(self):
        return self.arrays()[0]

    def getColumnIndices(self):
        return self.arrays()[1]

    def getData(self):
        return self.arrays()[2]

    def setValues(self, data):
        va = np.ascontiguousarray(data, dtype=np.float64)
        _ck(self._lib.mtjg_csr_set_values(self.handle, va.ctypes.data))
        return self

    def get(self, row, column):
        out = C.c_double()
        _ck(self._lib.mtjg_csr_get(self.handle, int(row), int(column), C.byref(out)))
        return out.value

    def set(self, row, column, value):
        _ck(self._lib.mtjg_csr_set(self.handle, int(row), int(column), float(value)))

    # products (S/CompRowMatrix.java:294-388, M/AbstractMatrix.java:144-198)
    def mult(self, *args):
        if len(args) == 2:
            x, y = args
            _ck(self._lib.mtjg_csr_mult(self.handle, x.handle, y.handle))
        else:
            alpha, x, y = args
            _ck(self._lib.mtjg_csr_mult_scaled(self.handle, float(alpha), x.handle, y.handle))
        return y

    def multAdd(self, *args):
        if len(args) == 2:
            alpha, (x, y) = 1.0, args
        else:
            alpha, x, y = args
        _ck(self._lib.mtjg_csr_mult_add(self.handle, float(alpha), x.handle, y.handle))
        return y

    def transMult(self, *args):
        if len(args) == 2:
            x, y = args
            _ck(self._lib.mtjg_csr_trans_mult(self.handle, x.handle, y.handle))
        else:
            alpha, x, y = args
            _ck(self._lib.mtjg_csr_trans_mult_scaled(self.handle, float(alpha), x.handle, y.handle))
        return y

    def transMultAdd(self, *args):
        if len(args) == 2:
            alpha, (x, y) = 1.0, args
        else:
            alpha, x, y = args
        _ck(self._lib.mtjg_csr_trans_mult_add(self.handle, float(alpha), x.handle, y.handle))
        return y

    def row_histogram(self, nbins=24):
        bins = np.zeros(nbins, np.int64)
        _ck(self._lib.mtjg_csr_row_histogram(self.handle, bins.ctypes.data_as(C.POINTER(C.c_int64)),
                                             nbins))
        return bins

    def close(self):
        if getattr(self, "handle", None):
            self._lib.mtjg_csr_destroy(self.handle)
            self.handle = None

    def __del__(self):
        _finalize(self)


# ---------------------------------------------------------------------------
# preconditioners

def _grid_dims(self):
    """(nx, ny, nz) of the grid sweeps the preconditioner runs (ILU, ICC and
    SSOR on a natural-order 5- or 7-point stencil), or None when the pattern
    is not such a stencil (level schedule)."""
    d = (C.c_int32 * 3)()
    _ck(self._lib.mtjg_prec_ilu_grid(self.handle, d))
    return tuple(d) if d[0] else None


class _Preconditioner:
    def setMatrix(self, A):
        bad = C.c_int32(-1)
        code = self._lib.mtjg_prec_set_matrix(self.handle, A.handle, C.byref(bad))
        if code != L.MTJG_OK:
            _raise(code, bad.value)

    def apply(self, b, x):
        _ck(self._lib.mtjg_prec_apply(self.handle, b.handle, x.handle))
        return x

    def transApply(self, b, x):
        _ck(self._lib.mtjg_prec_trans_apply(self.handle, b.handle, x.handle))
        return x

    def close(self):
        if getattr(self, "handle", None):
            self._lib.mtjg_prec_destroy(self.handle)
            self.handle = None

    def __del__(self):
        _finalize(self)


class DiagonalPreconditioner(_Preconditioner):
    """S/DiagonalPreconditioner.java."""

    def __init__(self, n, ctx=None):
        self.ctx = ctx or default_context()
        self._lib = self.ctx._lib
        h = C.c_void_p()
        _ck(self._lib.mtjg_prec_diagonal_create(self.ctx.handle, int(n), C.byref(h)))
        self.handle = h


class FP32CGPreconditioner(_Preconditioner):
    """A Preconditioner whose apply(b, x) is an FP32 Jacobi-CG solve of
    A x = b on the FP32 copy of A (x = ||b|| d, d from CG on b / ||b|| to the
    relative tolerance inner_rtol); with IR it is FP32-inner / FP64 iterative
    refinement through MTJ's own plug-in points (S/IR.java:57-70,
    S/Preconditioner.java:42)."""

    def __init__(self, n, inner_rtol=1e-4, inner_max_iter=1000, ctx=None):
        self.ctx = ctx or default_context()
        self._lib = self.ctx._lib
        h = C.c_void_p()
        _ck(self._lib.mtjg_prec_cg32_create(self.ctx.handle, int(n), float(inner_rtol), int(inner_max_iter),
                                            C.byref(h)))
        self.handle = h

    def inner_iterations(self):
        t = C.c_int64()
        _ck(self._lib.mtjg_prec_inner_iterations(self.handle, C.byref(t)))
        return t.value


class ILU(_Preconditioner):
    """S/ILU.java: ILU(0) in the pattern of the matrix given to setMatrix."""

    def __init__(self, LU, ctx=None):
        self.ctx = ctx or LU.ctx
        self._lib = self.ctx._lib
        h = C.c_void_p()
        _ck(self._lib.mtjg_prec_ilu_create(self.ctx.handle, LU.handle, C.byref(h)))
        self.handle = h

    grid = _grid_dims


class ICC(_Preconditioner):
    """S/ICC.java: incomplete Cholesky without fill-in, in the pattern of the
    matrix given to setMatrix.  transApply is apply."""

    def __init__(self, R, ctx=None):
        self.ctx = ctx or R.ctx
        self._lib = self.ctx._lib
        h = C.c_void_p()
        _ck(self._lib.mtjg_prec_icc_create(self.ctx.handle, R.handle, C.byref(h)))
        self.handle = h

    grid = _grid_dims


class AMG(_Preconditioner):
    """S/AMG.java: smoothed-aggregation algebraic multigrid with SSOR smoothing.

    AMG() uses MTJ's defaults (1, 1.85, 1.85, 1, 1, 1, 1, 40, 2/3);
    AMG(omegaPreF, omegaPreR, omegaPostF, omegaPostR, nu1, nu2, gamma, min,
    omega) runs reverse (symmetric) sweeps; AMG(omegaPre, omegaPost, nu1, nu2,
    gamma, min, omega) runs forward sweeps only.  The hierarchy is built by
    setMatrix; apply starts its cycle from x's incoming values."""

    def __init__(self, *params, ctx=None):
        if not params:
            params = (1.0, 1.85, 1.85, 1.0, 1, 1, 1, 40, 2.0 / 3)
        if len(params) == 7:
            w1, w2, nu1, nu2, gamma, mn, omega = params
            full, reverse = (w1, w1, w2, w2, nu1, nu2, gamma, mn, omega), 0
        elif len(params) == 9:
            full, reverse = tuple(params), 1
        else:
            raise TypeError("AMG takes 0, 7 or 9 parameters")
        self.ctx = ctx or default_context()
        self._lib = self.ctx._lib
        wpf, wpr, wqf, wqr, nu1, nu2, gamma, mn, omega = full
        h = C.c_void_p()
        _ck(self._lib.mtjg_prec_amg_create(self.ctx.handle, float(wpf), float(wpr), float(wqf), float(wqr),
                                           int(nu1), int(nu2), int(gamma), int(mn), float(omega), reverse,
                                           C.byref(h)))
        self.handle = h

    def levels(self):
        """Sizes of the hierarchy built by the last setMatrix, finest first."""
        m = C.c_int32(0)
        _ck(self._lib.mtjg_prec_amg_levels(self.handle, C.byref(m), None, 0))
        sizes = (C.c_int32 * max(m.value, 1))()
        _ck(self._lib.mtjg_prec_amg_levels(self.handle, C.byref(m), sizes, m.value))
        return [int(sizes[k]) for k in range(m.value)]


class SSOR(_Preconditioner):
    """S/SSOR.java: SSOR(F, reverse=True, omegaF=1, omegaR=1).  apply starts
    its sweeps from x's incoming values."""

    def __init__(self, F, reverse=True, omegaF=1.0, omegaR=1.0, ctx=None):
        self.ctx = ctx or F.ctx
        self._lib = self.ctx._lib
        h = C.c_void_p()
        _ck(self._lib.mtjg_prec_ssor_create(self.ctx.handle, F.handle, int(bool(reverse)), float(omegaF),
                                            float(omegaR), C.byref(h)))
        self.handle = h

    def setOmega(self, omegaF, omegaR):
        _ck(self._lib.mtjg_prec_ssor_set_omega(self.handle, float(omegaF), float(omegaR)))

    grid = _grid_dims


# ---------------------------------------------------------------------------
# monitor and solvers

class DefaultIterationMonitor:
    """S/DefaultIterationMonitor.java (defaults 100000, 1e-5, 1e-50, 1e5)."""

    def __init__(self, maxIter=100000, rtol=1e-5, atol=1e-50, dtol=1e5):
        self.maxIter, self.rtol, self.atol, self.dtol = int(maxIter), rtol, atol, dtol
        self._iter = 0
        self._residual = 0.0

    def setMaxIterations(self, m):
        self.maxIter = int(m)

    def setRelativeTolerance(self, r):
        self.rtol = r

    def setAbsoluteTolerance(self, a):
        self.atol = a

    def setDivergenceTolerance(self, d):
        self.dtol = d

    def iterations(self):
        return self._iter

    def residual(self):
        return self._residual

    def _struct(self):
        return L.Monitor(self.maxIter, self.rtol, self.atol, self.dtol)


class _Solver:
    KIND = None

    def __init__(self, template, ctx=None):
        self.ctx = ctx or template.ctx
        self._lib = self.ctx._lib
        h = C.c_void_p()
        _ck(self._lib.mtjg_solver_create(self.ctx.handle, self.KIND, template.handle, C.byref(h)))
        self.handle = h
        self.M = None
        self.iter = DefaultIterationMonitor()
        self._reporter = None
        self.last_result = None

    def setPreconditioner(self, M):
        self.M = M
        _ck(self._lib.mtjg_solver_set_preconditioner(self.handle, M.handle if M else None))

    def getPreconditioner(self):
        return self.M

    def setIterationMonitor(self, it):
        self.iter = it

    def getIterationMonitor(self):
        return self.iter

    def setIterationReporter(self, fn):
        """fn(residual, iteration), called at each iteration (forces a sync)."""
        if fn is None:
            self._reporter = None
            _ck(self._lib.mtjg_solver_set_reporter(self.handle, L.REPORTER(), None))
            return
        cb = L.REPORTER(lambda r, i, _u: fn(r, i))
        self._reporter = cb
        _ck(self._lib.mtjg_solver_set_reporter(self.handle, cb, None))

    def solve(self, A, b, x):
        """IterativeSolver.solve(A, b, x) (S/IterativeSolver.java:43)."""
        m = self.iter._struct()
        _ck(self._lib.mtjg_solver_set_monitor(self.handle, C.byref(m)))
        res = L.Result()
        code = self._lib.mtjg_solver_solve(self.handle, A.handle, b.handle, x.handle, C.byref(res))
        self.iter._iter = res.iterations
        self.iter._residual = res.residual
        self.last_result = res
        if code in (L.NOT_CONVERGED_ITERATIONS, L.NOT_CONVERGED_DIVERGENCE,
                    L.NOT_CONVERGED_BREAKDOWN):
            raise IterativeSolverNotConvergedException(Reason(code), res.iterations, res.residual,
                                                       (self._lib.mtjg_last_error() or b"").decode())
        _ck(code)
        return x

    def close(self):
        if getattr(self, "handle", None):
            self._lib.mtjg_solver_destroy(self.handle)
            self.handle = None

    def __del__(self):
        _finalize(self)


class CG(_Solver):
    """S/CG.java."""
    KIND = L.CG


class BiCGstab(_Solver):
    """S/BiCGstab.java."""
    KIND = L.BICGSTAB


class IR(_Solver):
    """S/IR.java."""
    KIND = L.IR


class BiCG(_Solver):
    """S/BiCG.java (uses transMult and the preconditioner's transApply)."""
    KIND = L.BICG


class CGS(_Solver):
    """S/CGS.java."""
    KIND = L.CGS


class QMR(_Solver):
    """S/QMR.java with M1 = M2 = the preconditioner (setPreconditioner form)."""
    KIND = L.QMR


class GMRES(_Solver):
    """S/GMRES.java (restart 30 by default)."""
    KIND = L.GMRES

    def __init__(self, template, restart=30, ctx=None):
        super().__init__(template, ctx)
        self.setRestart(restart)

    def setRestart(self, restart):
        _ck(self._lib.mtjg_gmres_set_restart(self.handle, int(restart)))
        self.restart = int(restart)


class MixedCG(_Solver):
    """FP32 inner Jacobi-CG under FP64 iterative refinement (no MTJ counterpart;
    the outer loop has the shape of S/IR.java:57-70)."""
    KIND = L.MIXED_CG

    def __init__(self, template, inner_rtol=1e-4, inner_max_iter=1000, ctx=None):
        super().__init__(template, ctx)
        self.setInner(inner_rtol, inner_max_iter)

    def setInner(self, inner_rtol, inner_max_iter):
        _ck(self._lib.mtjg_mixed_set_inner(self.handle, float(inner_rtol), int(inner_max_iter)))

    def inner_iterations(self):
        return int(self._lib.mtjg_solver_inner_iterations(self.handle))
