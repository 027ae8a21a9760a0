"""Loader and ctypes prototypes for libmtjgpu.so (the C ABI in include/mtjgpu.h).

The HIP library is the product: there is no CPU fallback.  If the shared
library has not been built, importing this module raises ImportError naming
the build command.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
PKG_ROOT = os.path.dirname(_HERE)
LIB_DIR = os.path.join(PKG_ROOT, "lib")
LIB_PATH = os.path.join(LIB_DIR, "libmtjgpu.so")

MTJG_OK = 0
NOT_CONVERGED_ITERATIONS = 1
NOT_CONVERGED_DIVERGENCE = 2
NOT_CONVERGED_BREAKDOWN = 3
E_INVALID_ARGUMENT = -1
E_INDEX_OUT_OF_BOUNDS = -2
E_ZERO_DIAGONAL = -3
E_MISSING_DIAGONAL = -4
E_ZERO_PIVOT = -5
E_SINGULAR = -6
E_UNSUPPORTED = -7
E_OUT_OF_MEMORY = -8
E_DEVICE = -9
E_NULL = -10
E_IO = -11

F64, F32 = 0, 1
NORM_ONE, NORM_TWO, NORM_TWO_ROBUST, NORM_INFINITY = 0, 1, 2, 3
CG, BICGSTAB, GMRES, IR, MIXED_CG, BICG, CGS, QMR = 0, 1, 2, 3, 4, 5, 6, 7


class Monitor(C.Structure):
    _fields_ = [("max_iter", C.c_int32), ("rtol", C.c_double), ("atol", C.c_double),
                ("dtol", C.c_double)]


class Result(C.Structure):
    _fields_ = [("iterations", C.c_int32), ("residual", C.c_double), ("reason", C.c_int32)]


REPORTER = C.CFUNCTYPE(None, C.c_double, C.c_int32, C.c_void_p)

# (name, restype, argtypes) for every symbol include/mtjgpu.h declares.
P = C.c_void_p
I32 = C.c_int32
I64 = C.c_int64
D = C.c_double
PI32 = C.POINTER(C.c_int32)
PI64 = C.POINTER(C.c_int64)
PD = C.POINTER(C.c_double)
PP = C.POINTER(C.c_void_p)

PROTOTYPES = [
    ("mtjg_strerror", C.c_char_p, [C.c_int]),
    ("mtjg_last_error", C.c_char_p, []),
    ("mtjg_ctx_create", C.c_int, [C.c_int, PP]),
    ("mtjg_ctx_destroy", C.c_int, [P]),
    ("mtjg_ctx_synchronize", C.c_int, [P]),
    ("mtjg_ctx_stream", P, [P]),
    ("mtjg_vec_create", C.c_int, [P, I64, C.c_int, PP]),
    ("mtjg_vec_wrap", C.c_int, [P, I64, C.c_int, P, PP]),
    ("mtjg_vec_destroy", C.c_int, [P]),
    ("mtjg_vec_size", I64, [P]),
    ("mtjg_vec_device_ptr", P, [P]),
    ("mtjg_vec_upload", C.c_int, [P, P]),
    ("mtjg_vec_download", C.c_int, [P, P]),
    ("mtjg_vec_zero", C.c_int, [P]),
    ("mtjg_vec_scale", C.c_int, [P, D]),
    ("mtjg_vec_set", C.c_int, [P, P]),
    ("mtjg_vec_set_scaled", C.c_int, [P, D, P]),
    ("mtjg_vec_add", C.c_int, [P, P]),
    ("mtjg_vec_add_scaled", C.c_int, [P, D, P]),
    ("mtjg_vec_dot", C.c_int, [P, P, PD]),
    ("mtjg_vec_norm", C.c_int, [P, C.c_int, PD]),
    ("mtjg_csr_create", C.c_int, [P, I32, I32, I64, P, P, P, PP]),
    ("mtjg_csr_destroy", C.c_int, [P]),
    ("mtjg_csr_shape", C.c_int, [P, PI32, PI32, PI64]),
    ("mtjg_csr_format", C.c_int, [P, PI32, PI64]),
    ("mtjg_csr_fp32_diagonals", C.c_int, [P, PI32]),
    ("mtjg_csr_download", C.c_int, [P, P, P, P]),
    ("mtjg_csr_set_values", C.c_int, [P, P]),
    ("mtjg_csr_get", C.c_int, [P, I32, I32, PD]),
    ("mtjg_csr_set", C.c_int, [P, I32, I32, D]),
    ("mtjg_csr_mult", C.c_int, [P, P, P]),
    ("mtjg_csr_mult_add", C.c_int, [P, D, P, P]),
    ("mtjg_csr_mult_scaled", C.c_int, [P, D, P, P]),
    ("mtjg_csr_trans_mult", C.c_int, [P, P, P]),
    ("mtjg_csr_trans_mult_add", C.c_int, [P, D, P, P]),
    ("mtjg_csr_trans_mult_scaled", C.c_int, [P, D, P, P]),
    ("mtjg_csr_row_histogram", C.c_int, [P, PI64, C.c_int]),
    ("mtjg_prec_diagonal_create", C.c_int, [P, I32, PP]),
    ("mtjg_prec_ilu_create", C.c_int, [P, P, PP]),
    ("mtjg_prec_ilu_grid", C.c_int, [P, PI32]),
    ("mtjg_prec_cg32_create", C.c_int, [P, I32, D, I32, PP]),
    ("mtjg_prec_inner_iterations", C.c_int, [P, C.POINTER(C.c_int64)]),
    ("mtjg_prec_icc_create", C.c_int, [P, P, PP]),
    ("mtjg_prec_amg_create", C.c_int, [P, D, D, D, D, I32, I32, I32, I32, D, I32, PP]),
    ("mtjg_prec_amg_levels", C.c_int, [P, C.POINTER(C.c_int32), C.POINTER(C.c_int32), I32]),
    ("mtjg_prec_ssor_create", C.c_int, [P, P, I32, D, D, PP]),
    ("mtjg_prec_ssor_set_omega", C.c_int, [P, D, D]),
    ("mtjg_prec_set_matrix", C.c_int, [P, P, PI32]),
    ("mtjg_prec_apply", C.c_int, [P, P, P]),
    ("mtjg_prec_trans_apply", C.c_int, [P, P, P]),
    ("mtjg_prec_destroy", C.c_int, [P]),
    ("mtjg_solver_create", C.c_int, [P, C.c_int, P, PP]),
    ("mtjg_solver_destroy", C.c_int, [P]),
    ("mtjg_solver_set_preconditioner", C.c_int, [P, P]),
    ("mtjg_solver_set_monitor", C.c_int, [P, C.POINTER(Monitor)]),
    ("mtjg_solver_set_reporter", C.c_int, [P, REPORTER, P]),
    ("mtjg_gmres_set_restart", C.c_int, [P, I32]),
    ("mtjg_mixed_set_inner", C.c_int, [P, D, I32]),
    ("mtjg_solver_solve", C.c_int, [P, P, P, P, C.POINTER(Result)]),
    ("mtjg_cg_solve_noalloc", C.c_int, [P, P, P, P, P, C.POINTER(Monitor), C.POINTER(Result)]),
    ("mtjg_batch_create", C.c_int, [P, I32, I32, I64, I32, I32, PP]),
    ("mtjg_batch_destroy", C.c_int, [P]),
    ("mtjg_batch_submit", C.c_int, [P, I32, I64, P, P, P, P, D, I32, D, I32, PI64]),
    ("mtjg_batch_step", C.c_int, [P, PI32, PI32]),
    ("mtjg_batch_result", C.c_int, [P, I64, P, C.POINTER(Result), PI32]),
    ("mtjg_batch_completed", C.c_int, [P, PI64, I32, PI32]),
    ("mtjg_batch_info", C.c_int, [P, PI32, PI32, PI64, PI64]),
    ("mtjg_batch_stats", C.c_int, [P, C.POINTER(C.c_double), PI64, C.POINTER(C.c_double)]),
    ("mtjg_mm_read_matrix", C.c_int, [C.c_char_p, PP]),
    ("mtjg_mm_parse_matrix", C.c_int, [C.c_char_p, I64, PP]),
    ("mtjg_mm_read_vector", C.c_int, [C.c_char_p, PP]),
    ("mtjg_mm_parse_vector", C.c_int, [C.c_char_p, I64, PP]),
    ("mtjg_mm_shape", C.c_int, [P, PI32, PI32, PI64, PI32]),
    ("mtjg_mm_copy_csr", C.c_int, [P, P, P, P]),
    ("mtjg_mm_copy_vector", C.c_int, [P, P]),
    ("mtjg_mm_to_csr", C.c_int, [P, P, PP]),
    ("mtjg_mm_to_vec", C.c_int, [P, P, PP]),
    ("mtjg_mm_destroy", C.c_int, [P]),
]

_lib = None


def load():
    """Loads libmtjgpu.so and binds its prototypes (once)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            "libmtjgpu.so is not built (%s); run `make -C matrix-toolkits-java_amd` "
            "or __graft_entry__.build()" % LIB_PATH)
    lib = C.CDLL(LIB_PATH)
    for name, res, args in PROTOTYPES:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    extra = getattr(lib, "mtjg_solver_inner_iterations", None)
    if extra is not None:
        extra.restype = C.c_int
        extra.argtypes = [P]
    _lib = lib
    return lib
