"""ctypes view of the CPU restatement of MTJ's loops (build/libmtjoracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() as
the checker, and by bench.py's CPU-baseline leg.  The product never loads it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "build", "libmtjoracle.so")

OK, ITERATIONS, DIVERGENCE, BREAKDOWN = 0, 1, 2, 3
ZERO_DIAGONAL, MISSING_DIAGONAL, ZERO_PIVOT, SINGULAR, BAD_ARG, UNSUPPORTED = 10, 11, 12, 13, 20, 22
PREC_IDENTITY, PREC_DIAGONAL, PREC_ILU, PREC_SSOR, PREC_ICC, PREC_AMG, PREC_CG32 = 0, 1, 2, 3, 4, 5, 6


class Monitor(C.Structure):
    _fields_ = [("iter", C.c_int), ("max_iter", C.c_int), ("rtol", C.c_double),
                ("atol", C.c_double), ("dtol", C.c_double), ("init_r", C.c_double),
                ("residual", C.c_double)]


class Prec(C.Structure):
    _fields_ = [("kind", C.c_int), ("n", C.c_int), ("invdiag", C.c_void_p), ("rp", C.c_void_p),
                ("ci", C.c_void_p), ("lu", C.c_void_p), ("diagind", C.c_void_p),
                ("y", C.c_void_p), ("omega_f", C.c_double), ("omega_r", C.c_double),
                ("reverse", C.c_int), ("xx", C.c_void_p), ("amg", C.c_void_p), ("a32", C.c_void_p),
                ("inv32", C.c_void_p), ("rtol", C.c_double), ("max_iter", C.c_int), ("inner_total", C.c_long)]


_PI = C.POINTER(C.c_int32)
_PD = C.POINTER(C.c_double)


class AmgS(C.Structure):
    _fields_ = [("omega_pre_f", C.c_double), ("omega_pre_r", C.c_double), ("omega_post_f", C.c_double),
                ("omega_post_r", C.c_double), ("omega", C.c_double), ("nu1", C.c_int), ("nu2", C.c_int),
                ("gamma", C.c_int), ("min_size", C.c_int), ("reverse", C.c_int), ("m", C.c_int),
                ("n", _PI), ("arp", C.POINTER(_PI)), ("aci", C.POINTER(_PI)), ("ava", C.POINTER(_PD)),
                ("pt", C.POINTER(_PI)), ("icp", C.POINTER(_PI)), ("iri", C.POINTER(_PI)),
                ("iva", C.POINTER(_PD)), ("ssor", C.c_void_p), ("lu", _PD), ("piv", _PI),
                ("singular", C.c_int), ("u", C.POINTER(_PD)), ("f", C.POINTER(_PD)), ("r", C.POINTER(_PD)),
                ("transpose", C.c_int), ("error", C.c_int)]


_lib = None


def build():
    subprocess.check_call(["make", "-s", "-C", HERE])


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        lib = C.CDLL(LIB)
        P = C.c_void_p
        D = C.c_double
        lib.oc_csr_get.restype = D
        for name in ("oc_dv_dot",):
            getattr(lib, name).restype = D
        for name in ("oc_dv_norm1", "oc_dv_norm2", "oc_dv_norm2_robust", "oc_dv_norm_inf", "oc_now"):
            getattr(lib, name).restype = D
        lib.oc_csr_mult.argtypes = [C.c_int, P, P, P, P, P]
        lib.oc_csr_mult_add.argtypes = [C.c_int, P, P, P, D, P, P]
        lib.oc_csr_trans_mult.argtypes = [C.c_int, C.c_int, P, P, P, P, P]
        lib.oc_csr_trans_mult_add.argtypes = [C.c_int, C.c_int, P, P, P, D, P, P]
        lib.oc_csr_get.argtypes = [P, P, P, C.c_int, C.c_int]
        lib.oc_dv_scale.argtypes = [C.c_int, D, P]
        lib.oc_dv_set_scaled.argtypes = [C.c_int, D, P, P]
        lib.oc_dv_add_scaled.argtypes = [C.c_int, D, P, P]
        lib.oc_dv_add.argtypes = [C.c_int, P, P]
        lib.oc_dv_dot.argtypes = [C.c_int, P, P]
        for name in ("oc_dv_norm1", "oc_dv_norm2", "oc_dv_norm2_robust", "oc_dv_norm_inf"):
            getattr(lib, name).argtypes = [C.c_int, P]
        lib.oc_monitor_init.argtypes = [C.POINTER(Monitor), C.c_int, D, D, D]
        lib.oc_monitor_converged.argtypes = [C.POINTER(Monitor), D]
        lib.oc_prec_identity.argtypes = [C.POINTER(Prec), C.c_int]
        lib.oc_prec_diagonal.argtypes = [C.POINTER(Prec), C.c_int, P, P, P, C.POINTER(C.c_int)]
        lib.oc_prec_ilu.argtypes = [C.POINTER(Prec), C.c_int, P, P, P, C.POINTER(C.c_int)]
        lib.oc_prec_icc.argtypes = [C.POINTER(Prec), C.c_int, P, P, P, C.POINTER(C.c_int)]
        lib.oc_amg_params.argtypes = [C.POINTER(AmgS), D, D, D, D, C.c_int, C.c_int, C.c_int, C.c_int, D,
                                      C.c_int]
        lib.oc_amg_setup.argtypes = [C.POINTER(AmgS), C.c_int, P, P, P, C.POINTER(C.c_int),
                                     C.POINTER(C.c_int)]
        lib.oc_amg_apply.argtypes = [C.POINTER(AmgS), P, P, C.c_int]
        lib.oc_amg_free.argtypes = [C.POINTER(AmgS)]
        lib.oc_prec_amg.argtypes = [C.POINTER(Prec), C.POINTER(AmgS)]
        lib.oc_java_hashset_order.argtypes = [P, C.c_int, P]
        lib.oc_dgetf2.argtypes = [C.c_int, P, P]
        lib.oc_dgetrs.argtypes = [C.c_int, C.c_int, P, P, P]
        lib.oc_prec_cg32.argtypes = [C.POINTER(Prec), C.c_int, P, P, P, C.c_double, C.c_int, C.POINTER(C.c_int)]
        lib.oc_prec_ssor.argtypes = [C.POINTER(Prec), C.c_int, P, P, P, C.c_int, C.c_double, C.c_double,
                                     C.POINTER(C.c_int)]
        lib.oc_prec_apply.argtypes = [C.POINTER(Prec), P, P]
        lib.oc_prec_trans_apply.argtypes = [C.POINTER(Prec), P, P]
        lib.oc_prec_free.argtypes = [C.POINTER(Prec)]
        lib.oc_givens.argtypes = [D, D, C.POINTER(D), C.POINTER(D)]
        lib.oc_triu_solve.argtypes = [C.c_int, P, C.c_int, P]
        for name in ("oc_unit_lower_solve", "oc_unit_lower_trans_solve", "oc_upper_solve",
                     "oc_upper_trans_solve"):
            getattr(lib, name).argtypes = [C.c_int, P, P, P, P, P, P]
        sol = [C.c_int, P, P, P, C.POINTER(Prec), P, P, C.POINTER(Monitor)]
        lib.oc_cg.argtypes = sol
        lib.oc_bicgstab.argtypes = sol
        lib.oc_ir.argtypes = sol
        lib.oc_bicg.argtypes = sol
        lib.oc_cgs.argtypes = sol
        lib.oc_qmr.argtypes = sol
        lib.oc_gmres.argtypes = [C.c_int, P, P, P, C.POINTER(Prec), C.c_int, P, P,
                                 C.POINTER(Monitor)]
        _lib = lib
    return _lib


def _p(a):
    return a.ctypes.data if a is not None and a.size else None


class Csr:
    """Host CSR triple kept alive for the oracle calls."""

    def __init__(self, n_rows, n_cols, rp, ci, va):
        self.nr, self.nc = int(n_rows), int(n_cols)
        self.rp = np.ascontiguousarray(rp, np.int32)
        self.ci = np.ascontiguousarray(ci, np.int32)
        self.va = np.ascontiguousarray(va, np.float64)

    def mult(self, x):
        y = np.empty(self.nr)
        load().oc_csr_mult(self.nr, _p(self.rp), _p(self.ci), _p(self.va), _p(np.ascontiguousarray(x, float)), _p(y))
        return y

    def mult_add(self, alpha, x, y):
        y = np.array(y, dtype=float)
        load().oc_csr_mult_add(self.nr, _p(self.rp), _p(self.ci), _p(self.va), alpha,
                               _p(np.ascontiguousarray(x, float)), _p(y))
        return y

    def trans_mult(self, x):
        y = np.empty(self.nc)
        load().oc_csr_trans_mult(self.nr, self.nc, _p(self.rp), _p(self.ci), _p(self.va),
                                 _p(np.ascontiguousarray(x, float)), _p(y))
        return y

    def trans_mult_add(self, alpha, x, y):
        y = np.array(y, dtype=float)
        load().oc_csr_trans_mult_add(self.nr, self.nc, _p(self.rp), _p(self.ci), _p(self.va), alpha,
                                     _p(np.ascontiguousarray(x, float)), _p(y))
        return y

    def get(self, i, j):
        return load().oc_csr_get(_p(self.rp), _p(self.ci), _p(self.va), i, j)


class Precond:
    """Identity / DiagonalPreconditioner / ILU / SSOR / ICC restatement."""

    def __init__(self, kind, A=None, n=None, reverse=True, omega_f=1.0, omega_r=1.0, inner_rtol=1e-4,
                 inner_max_iter=1000):
        self.s = Prec()
        self.bad = C.c_int(-1)
        lib = load()
        if kind == PREC_IDENTITY:
            lib.oc_prec_identity(C.byref(self.s), n if n is not None else A.nr)
            self.code = OK
        elif kind == PREC_DIAGONAL:
            self.code = lib.oc_prec_diagonal(C.byref(self.s), A.nr, _p(A.rp), _p(A.ci), _p(A.va),
                                             C.byref(self.bad))
        elif kind == PREC_SSOR:
            self.code = lib.oc_prec_ssor(C.byref(self.s), A.nr, _p(A.rp), _p(A.ci), _p(A.va), int(reverse),
                                         float(omega_f), float(omega_r), C.byref(self.bad))
        elif kind == PREC_CG32:
            self.code = lib.oc_prec_cg32(C.byref(self.s), A.nr, _p(A.rp), _p(A.ci), _p(A.va), float(inner_rtol),
                                         int(inner_max_iter), C.byref(self.bad))
        elif kind == PREC_ICC:
            self.code = lib.oc_prec_icc(C.byref(self.s), A.nr, _p(A.rp), _p(A.ci), _p(A.va),
                                        C.byref(self.bad))
        else:
            self.code = lib.oc_prec_ilu(C.byref(self.s), A.nr, _p(A.rp), _p(A.ci), _p(A.va),
                                        C.byref(self.bad))
        self.A = A

    def apply(self, b, x0=None):
        """x = M^-1 b; x0 is x's incoming content (read by SSOR), zero if None."""
        x = np.zeros(self.s.n) if x0 is None else np.array(x0, dtype=float)
        load().oc_prec_apply(C.byref(self.s), _p(np.ascontiguousarray(b, float)), _p(x))
        return x

    def trans_apply(self, b, x0=None):
        x = np.zeros(self.s.n) if x0 is None else np.array(x0, dtype=float)
        load().oc_prec_trans_apply(C.byref(self.s), _p(np.ascontiguousarray(b, float)), _p(x))
        return x

    def ilu_values(self):
        nnz = int(self.A.rp[-1])
        buf = (C.c_double * nnz).from_address(self.s.lu)
        return np.frombuffer(buf, dtype=np.float64).copy()

    def ilu_diagind(self):
        buf = (C.c_int32 * self.s.n).from_address(self.s.diagind)
        return np.frombuffer(buf, dtype=np.int32).copy()

    def __del__(self):
        try:
            load().oc_prec_free(C.byref(self.s))
        except Exception:
            pass


class Amg:
    """S/AMG.java restatement.  Amg(A) uses AMG()'s defaults; params are
    (omegaPreF, omegaPreR, omegaPostF, omegaPostR, nu1, nu2, gamma, min,
    omega) with reverse=True (the nine-argument constructor) or
    (omegaPre, omegaPost, nu1, nu2, gamma, min, omega) with reverse=False."""

    def __init__(self, A, *params):
        if not params:
            params = (1.0, 1.85, 1.85, 1.0, 1, 1, 1, 40, 2.0 / 3)
        if len(params) == 7:
            wpre, wpost, nu1, nu2, gamma, mn, omega = params
            full, reverse = (wpre, wpre, wpost, wpost, nu1, nu2, gamma, mn, omega), 0
        else:
            full, reverse = tuple(params), 1
        self.g = AmgS()
        self.bad = C.c_int(-1)
        self.bad_level = C.c_int(-1)
        lib = load()
        wpf, wpr, wqf, wqr, nu1, nu2, gamma, mn, omega = full
        lib.oc_amg_params(C.byref(self.g), float(wpf), float(wpr), float(wqf), float(wqr), int(nu1), int(nu2),
                          int(gamma), int(mn), float(omega), reverse)
        self.A = A
        self.code = lib.oc_amg_setup(C.byref(self.g), A.nr, _p(A.rp), _p(A.ci), _p(A.va), C.byref(self.bad),
                                     C.byref(self.bad_level))

    def levels(self):
        return [int(self.g.n[k]) for k in range(self.g.m)]

    def matrix(self, k):
        """(rowptr, colind, values) of A_k."""
        n = self.g.n[k]
        nnz = self.g.arp[k][n]
        return (np.ctypeslib.as_array(self.g.arp[k], (n + 1,)).copy(),
                np.ctypeslib.as_array(self.g.aci[k], (max(nnz, 1),))[:nnz].copy(),
                np.ctypeslib.as_array(self.g.ava[k], (max(nnz, 1),))[:nnz].copy())

    def interpolation(self, k):
        """(colptr, rowind, values) of I_k (n_k x n_{k+1})."""
        c = self.g.n[k + 1]
        nnz = self.g.icp[k][c]
        return (np.ctypeslib.as_array(self.g.icp[k], (c + 1,)).copy(),
                np.ctypeslib.as_array(self.g.iri[k], (max(nnz, 1),))[:nnz].copy(),
                np.ctypeslib.as_array(self.g.iva[k], (max(nnz, 1),))[:nnz].copy())

    def aggregates(self, k):
        return np.ctypeslib.as_array(self.g.pt[k], (self.g.n[k],)).copy()

    def coarse_lu(self):
        nc = self.g.n[self.g.m - 1]
        lu = np.ctypeslib.as_array(self.g.lu, (nc * nc + 1,))[:nc * nc].reshape(nc, nc, order="F").copy()
        piv = np.ctypeslib.as_array(self.g.piv, (max(nc, 1),))[:nc].copy()
        return lu, piv, bool(self.g.singular)

    def apply(self, b, x0=None, transpose=False):
        x = np.zeros(self.g.n[0]) if x0 is None else np.array(x0, dtype=float)
        rc = load().oc_amg_apply(C.byref(self.g), _p(np.ascontiguousarray(b, float)), _p(x), int(transpose))
        return rc, x

    def precond(self):
        """This AMG as a Precond for solve()."""
        P = Precond.__new__(Precond)
        P.s = Prec()
        P.bad = C.c_int(-1)
        load().oc_prec_amg(C.byref(P.s), C.byref(self.g))
        P.code = OK
        P.A = self.A
        P._owner = self
        return P

    def __del__(self):
        try:
            load().oc_amg_free(C.byref(self.g))
        except Exception:
            pass


def java_hashset_order(vals):
    """Iteration order of a java.util.HashSet<Integer> filled with vals in order."""
    v = np.ascontiguousarray(vals, np.int32)
    out = np.zeros(max(len(v), 1), np.int32)
    rc = load().oc_java_hashset_order(_p(v), len(v), _p(out))
    return rc, out[:len(v)]


def dgetf2(a):
    """Reference LAPACK dgetf2 on a copy of the square matrix a: (info, lu, piv 1-based)."""
    n = a.shape[0]
    lu = np.asfortranarray(np.array(a, dtype=float))
    piv = np.zeros(max(n, 1), np.int32)
    info = load().oc_dgetf2(n, _p(lu), _p(piv))
    return info, lu, piv[:n]


def dgetrs(lu, piv, b, trans=False):
    n = lu.shape[0]
    x = np.array(b, dtype=float)
    load().oc_dgetrs(int(trans), n, _p(np.asfortranarray(lu)), _p(np.ascontiguousarray(piv, np.int32)), _p(x))
    return x


def monitor(max_iter=100000, rtol=1e-5, atol=1e-50, dtol=1e5):
    m = Monitor()
    load().oc_monitor_init(C.byref(m), max_iter, rtol, atol, dtol)
    return m


def solve(method, A, b, x0, prec=None, mon=None, restart=30):
    """Runs one restated solver; returns (code, x, iterations, residual)."""
    lib = load()
    mon = mon or monitor()
    prec = prec or Precond(PREC_IDENTITY, n=A.nr)
    x = np.array(x0, dtype=float)
    b = np.ascontiguousarray(b, float)
    args = (A.nr, _p(A.rp), _p(A.ci), _p(A.va), C.byref(prec.s))
    if method == "cg":
        code = lib.oc_cg(*args, _p(b), _p(x), C.byref(mon))
    elif method == "bicgstab":
        code = lib.oc_bicgstab(*args, _p(b), _p(x), C.byref(mon))
    elif method == "ir":
        code = lib.oc_ir(*args, _p(b), _p(x), C.byref(mon))
    elif method == "gmres":
        code = lib.oc_gmres(*args, restart, _p(b), _p(x), C.byref(mon))
    elif method in ("bicg", "cgs", "qmr"):
        code = getattr(lib, "oc_" + method)(*args, _p(b), _p(x), C.byref(mon))
    else:
        raise ValueError(method)
    return code, x, mon.iter, mon.residual


def dv_dot(x, y):
    return load().oc_dv_dot(len(x), _p(np.ascontiguousarray(x, float)), _p(np.ascontiguousarray(y, float)))


def dv_norm(x, kind):
    lib = load()
    x = np.ascontiguousarray(x, float)
    fn = {"one": lib.oc_dv_norm1, "two": lib.oc_dv_norm2, "robust": lib.oc_dv_norm2_robust,
          "inf": lib.oc_dv_norm_inf}[kind]
    return fn(len(x), _p(x))


def givens(x, y):
    c, s = C.c_double(), C.c_double()
    load().oc_givens(x, y, C.byref(c), C.byref(s))
    return c.value, s.value


def triu_solve(H, k, s):
    """In-place back-substitution of the leading k x k block of column-major H."""
    H = np.asfortranarray(H, dtype=float)
    s = np.array(s, dtype=float)
    code = load().oc_triu_solve(k, H.ctypes.data, H.shape[0], s.ctypes.data)
    return code, s


class _MM(C.Structure):
    _fields_ = [("rows", C.c_int32), ("cols", C.c_int32), ("symmetry", C.c_int32), ("nnz", C.c_long),
                ("rp", C.POINTER(C.c_int32)), ("ci", C.POINTER(C.c_int32)), ("val", C.POINTER(C.c_double))]


IO, UNSUPPORTED, INDEX = 21, 22, 23


def mm_matrix(text):
    """CompRowMatrix(MatrixVectorReader) on `text` (bytes).  Returns
    (code, rows, cols, rp, ci, val, symmetry); arrays are None on error."""
    lib = load()
    m = _MM()
    code = lib.oc_mm_matrix(text, C.c_long(len(text)), C.byref(m))
    if code != OK:
        return code, None, None, None, None, None, None
    try:
        rp = np.ctypeslib.as_array(m.rp, (m.rows + 1,)).copy()
        ci = np.ctypeslib.as_array(m.ci, (max(m.nnz, 1),))[:m.nnz].copy()
        va = np.ctypeslib.as_array(m.val, (max(m.nnz, 1),))[:m.nnz].copy()
        return code, m.rows, m.cols, rp, ci, va, m.symmetry
    finally:
        lib.oc_mm_free(C.byref(m))


def mm_vector(text):
    """DenseVector(MatrixVectorReader) on `text` (bytes): (code, x)."""
    lib = load()
    m = _MM()
    code = lib.oc_mm_vector(text, C.c_long(len(text)), C.byref(m))
    if code != OK:
        return code, None
    try:
        return code, np.ctypeslib.as_array(m.val, (max(m.rows, 1),))[:m.rows].copy()
    finally:
        lib.oc_mm_free(C.byref(m))
