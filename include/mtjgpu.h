/*
 * mtjgpu.h -- C ABI of the MI355X (gfx950) implementation of MTJ's
 * unstructured-sparse hot path: CompRowMatrix CSR products, DenseVector
 * operations, DiagonalPreconditioner / ILU(0), and the CG / BiCGstab /
 * GMRES / IR solve loops, with all operands resident in device memory.
 *
 * Every entry point returns an int: 0 (MTJG_OK) on success, a positive
 * MTJG_NOT_CONVERGED_* code when a solve ended the way MTJ ends it with an
 * IterativeSolverNotConvergedException, or a negative MTJG_E_* code where MTJ
 * throws a RuntimeException subclass.  No entry point throws across the ABI.
 * mtjg_last_error() returns the calling thread's message for the last
 * failure, worded like MTJ's exception messages.
 *
 * Threading: objects are not thread-safe (MTJ solvers hold scratch state,
 * S/CG.java:40).  Use one mtjg_ctx per host thread; calls on one context
 * are ordered on its HIP stream.  Calls that return host scalars (dot,
 * norm, get, download, solve) synchronise the stream.
 *
 * Reference citations: S/ = src/main/java/no/uib/cipr/matrix/sparse/,
 * M/ = src/main/java/no/uib/cipr/matrix/.
 */
#ifndef MTJGPU_H
#define MTJGPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mtjg_ctx_s *mtjg_ctx;
typedef struct mtjg_vec_s *mtjg_vec;
typedef struct mtjg_csr_s *mtjg_csr;
typedef struct mtjg_prec_s *mtjg_prec;
typedef struct mtjg_solver_s *mtjg_solver;

/* ---- status codes ---------------------------------------------------------- */
enum {
    MTJG_OK = 0,
    /* IterativeSolverNotConvergedException reasons
     * (M/NotConvergedException.java:33-49) */
    MTJG_NOT_CONVERGED_ITERATIONS = 1,
    MTJG_NOT_CONVERGED_DIVERGENCE = 2,
    MTJG_NOT_CONVERGED_BREAKDOWN = 3,
    /* IllegalArgumentException (S/AbstractIterativeSolver.java:70-77,
     * S/CompRowMatrix.java:192-203, S/GMRES.java:114-116) */
    MTJG_E_INVALID_ARGUMENT = -1,
    /* IndexOutOfBoundsException (M/AbstractMatrix.java:169-176,
     * M/AbstractVector.java:174-178, S/CompRowMatrix.java:426-430) */
    MTJG_E_INDEX_OUT_OF_BOUNDS = -2,
    /* RuntimeException("Zero diagonal on row i+1")
     * (S/DiagonalPreconditioner.java:71-72) */
    MTJG_E_ZERO_DIAGONAL = -3,
    /* RuntimeException("Missing diagonal entry on row k+1")
     * (S/ILU.java:134-136) */
    MTJG_E_MISSING_DIAGONAL = -4,
    /* RuntimeException("Zero pivot encountered ...") (S/ILU.java:105-107) */
    MTJG_E_ZERO_PIVOT = -5,
    /* MatrixSingularException from the GMRES back-substitution
     * (M/AbstractTriangDenseMatrix.java:251-252) */
    MTJG_E_SINGULAR = -6,
    /* UnsupportedOperationException / wrong operand kind */
    MTJG_E_UNSUPPORTED = -7,
    MTJG_E_OUT_OF_MEMORY = -8,
    MTJG_E_DEVICE = -9,
    MTJG_E_NULL = -10,
    MTJG_E_IO = -11             /* IOException / EOFException (unreadable file, bad header, early end) */
};

/* Vector element types. */
enum { MTJG_F64 = 0, MTJG_F32 = 1 };

/* Vector.Norm (M/Vector.java:159-180). */
enum { MTJG_NORM_ONE = 0, MTJG_NORM_TWO = 1, MTJG_NORM_TWO_ROBUST = 2,
       MTJG_NORM_INFINITY = 3 };

/* Solver kinds. */
enum {
    MTJG_CG = 0,        /* S/CG.java */
    MTJG_BICGSTAB = 1,  /* S/BiCGstab.java */
    MTJG_GMRES = 2,     /* S/GMRES.java */
    MTJG_IR = 3,        /* S/IR.java */
    MTJG_MIXED_CG = 4,  /* FP32 inner Jacobi-CG under FP64 refinement; no
                           MTJ counterpart (S/IR.java:57-70 is its outer
                           loop's shape) */
    MTJG_BICG = 5,      /* S/BiCG.java (uses transMult and transApply) */
    MTJG_CGS = 6,       /* S/CGS.java */
    MTJG_QMR = 7        /* S/QMR.java with M1 = M2 = the preconditioner
                           (uses transMultAdd and transApply) */
};

const char *mtjg_strerror(int code);
/* Message of the calling thread's last failure ("" if none). */
const char *mtjg_last_error(void);

/* ---- context --------------------------------------------------------------- */
/* Binds a device and creates the context's HIP stream and workspace.  No MTJ
 * counterpart (MTJ has no device); every object below belongs to a context. */
int mtjg_ctx_create(int device, mtjg_ctx *out);
int mtjg_ctx_destroy(mtjg_ctx ctx);
int mtjg_ctx_synchronize(mtjg_ctx ctx);
/* The context's stream (hipStream_t), for callers that interleave work. */
void *mtjg_ctx_stream(mtjg_ctx ctx);

/* ---- DenseVector (M/DenseVector.java) --------------------------------------- */
/* new DenseVector(size): zero-filled (M/DenseVector.java:66-74). */
int mtjg_vec_create(mtjg_ctx ctx, int64_t n, int dtype, mtjg_vec *out);
/* Wraps caller-owned device memory of n elements (not freed on destroy). */
int mtjg_vec_wrap(mtjg_ctx ctx, int64_t n, int dtype, void *device_ptr,
                  mtjg_vec *out);
int mtjg_vec_destroy(mtjg_vec v);
int64_t mtjg_vec_size(mtjg_vec v);
void *mtjg_vec_device_ptr(mtjg_vec v);
/* Host <-> device copies of all n elements (element type of the vector). */
int mtjg_vec_upload(mtjg_vec v, const void *host);
int mtjg_vec_download(mtjg_vec v, void *host);
/* zero() (M/DenseVector.java:189-193). */
int mtjg_vec_zero(mtjg_vec x);
/* scale(alpha): x *= alpha, no shortcut (M/DenseVector.java:195-200). */
int mtjg_vec_scale(mtjg_vec x, double alpha);
/* set(y): x = y (M/DenseVector.java:202-213). */
int mtjg_vec_set(mtjg_vec x, mtjg_vec y);
/* set(alpha, y): x = alpha*y, alpha == 0 zeroes (M/DenseVector.java:215-231). */
int mtjg_vec_set_scaled(mtjg_vec x, double alpha, mtjg_vec y);
/* add(y): x += y (M/DenseVector.java:233-246). */
int mtjg_vec_add(mtjg_vec x, mtjg_vec y);
/* add(alpha, y): x += alpha*y, alpha == 0 is a no-op
 * (M/DenseVector.java:248-264). */
int mtjg_vec_add_scaled(mtjg_vec x, double alpha, mtjg_vec y);
/* dot(y) (M/DenseVector.java:266-279). */
int mtjg_vec_dot(mtjg_vec x, mtjg_vec y, double *out);
/* norm(type) (M/AbstractVector.java:180-190, M/DenseVector.java:281-318). */
int mtjg_vec_norm(mtjg_vec x, int norm_type, double *out);

/* ---- CompRowMatrix (S/CompRowMatrix.java) ----------------------------------- */
/* CompRowMatrix(numRows, numColumns, nz) + set(i,j,v) for every entry
 * (S/CompRowMatrix.java:178-208): copies the CSR to the device.  Columns
 * within a row are sorted (values follow) and must be unique and in
 * [0, n_cols); row pointers must be non-decreasing from 0 to nnz. */
int mtjg_csr_create(mtjg_ctx ctx, int32_t n_rows, int32_t n_cols, int64_t nnz,
                    const int32_t *row_pointer, const int32_t *column_index,
                    const double *data, mtjg_csr *out);
int mtjg_csr_destroy(mtjg_csr A);
int mtjg_csr_shape(mtjg_csr A, int32_t *n_rows, int32_t *n_cols, int64_t *nnz);
/* getRowPointers / getColumnIndices / getData (S/CompRowMatrix.java:255-271),
 * copied to caller memory; any pointer may be NULL. */
int mtjg_csr_download(mtjg_csr A, int32_t *row_pointer, int32_t *column_index,
                      double *data);
/* Device storage chosen for the products (no MTJ counterpart): the row-block
 * CSR kernels (MTJG_FORMAT_CSR); for patterns whose rows in each group of
 * 64 have similar lengths, an added SELL-64 copy (MTJG_FORMAT_SELL; padding
 * at most a quarter of the stored entries); or, for irregular patterns whose
 * columns scatter over an x of 16 to 128 MB, an added copy of the entries
 * ordered by column blocks of 2 to 4 MB of x, each gathered on one XCD, whose
 * products a first pass forms before the row-block kernels sum them
 * (MTJG_FORMAT_CSR_XCD; results bitwise those of MTJG_FORMAT_CSR).  *stored counts the entries the
 * products stream, padding included.  MTJG_SPMV_FORMAT=csr|sell|xcd in the
 * environment at creation forces the choice. */
enum { MTJG_FORMAT_CSR = 0, MTJG_FORMAT_SELL = 1, MTJG_FORMAT_CSR_XCD = 2 };
int mtjg_csr_format(mtjg_csr A, int32_t *format, int64_t *stored);
/* Diagonals of the FP32 stencil copy (no MTJ counterpart): the FP32 products
 * of the mixed-precision solvers (MTJG_MIXED_CG, MTJG_PREC_CG32) read a copy
 * that stores only values, diagonal by diagonal, when every entry lies on one
 * of at most 8 diagonals (column - row offsets) and the pattern got a SELL-64
 * copy; *count is their number, or 0 when the FP32 products use SELL-64 or
 * CSR.  Results are bitwise those of the SELL-64 FP32 product.
 * MTJG_SPMV_STENCIL=0 in the environment at creation turns the copy off. */
int mtjg_csr_fp32_diagonals(mtjg_csr A, int32_t *count);
/* set(Matrix) with an identical pattern (S/CompRowMatrix.java:449-471). */
int mtjg_csr_set_values(mtjg_csr A, const double *data);
/* get(row, column): 0 outside the pattern (S/CompRowMatrix.java:406-417). */
int mtjg_csr_get(mtjg_csr A, int32_t row, int32_t column, double *out);
/* set(row, column, value): IndexOutOfBounds outside the pattern
 * (S/CompRowMatrix.java:390-396, :422-431). */
int mtjg_csr_set(mtjg_csr A, int32_t row, int32_t column, double value);
/* mult(x, y): y = A x (S/CompRowMatrix.java:294-314). */
int mtjg_csr_mult(mtjg_csr A, mtjg_vec x, mtjg_vec y);
/* multAdd(alpha, x, y): y += alpha A x (S/CompRowMatrix.java:329-347). */
int mtjg_csr_mult_add(mtjg_csr A, double alpha, mtjg_vec x, mtjg_vec y);
/* mult(alpha, x, y) = multAdd(alpha, x, y.zero()) (M/AbstractMatrix.java:
 * 148-150). */
int mtjg_csr_mult_scaled(mtjg_csr A, double alpha, mtjg_vec x, mtjg_vec y);
/* transMult(x, y): y = A' x (S/CompRowMatrix.java:349-366). */
int mtjg_csr_trans_mult(mtjg_csr A, mtjg_vec x, mtjg_vec y);
/* transMultAdd(alpha, x, y): y = alpha((y/alpha) + A' x)
 * (S/CompRowMatrix.java:368-388). */
int mtjg_csr_trans_mult_add(mtjg_csr A, double alpha, mtjg_vec x, mtjg_vec y);
/* transMult(alpha, x, y) = transMultAdd(alpha, x, y.zero())
 * (M/AbstractMatrix.java:182-184). */
int mtjg_csr_trans_mult_scaled(mtjg_csr A, double alpha, mtjg_vec x,
                               mtjg_vec y);
/* Row-length histogram in power-of-two bins: bins[b] counts rows with length
 * in [2^(b-1), 2^b) (bins[0] counts empty rows); nbins entries are written. */
int mtjg_csr_row_histogram(mtjg_csr A, int64_t *bins, int nbins);

/* ---- Preconditioners (S/Preconditioner.java) -------------------------------- */
/* new DiagonalPreconditioner(n) (S/DiagonalPreconditioner.java:43-45). */
int mtjg_prec_diagonal_create(mtjg_ctx ctx, int32_t n, mtjg_prec *out);
/* new ILU(CompRowMatrix LU): LU supplies the pattern of the factor
 * (S/ILU.java:54-62). */
int mtjg_prec_ilu_create(mtjg_ctx ctx, mtjg_csr LU, mtjg_prec *out);
/* Grid of an ILU's, ICC's or SSOR's sweeps (no MTJ counterpart): when the
 * pattern given to setMatrix is a 5- or 7-point stencil of an nx x ny (x nz)
 * grid in natural order (row x + nx (y + ny z), off-diagonal columns among
 * its grid neighbours, ny <= 1024), the triangular solves of apply (and of
 * ILU's transApply) and SSOR's two sweeps run as grid sweeps over the
 * anti-diagonals x + y + z, bitwise equal to the level-scheduled ones; dims
 * receives nx, ny, nz, or zeros when the level schedule is used.
 * MTJG_ILU_GRID=0 in the environment at creation turns grid sweeps off. */
int mtjg_prec_ilu_grid(mtjg_prec M, int32_t dims[3]);
/* new SSOR(F, reverse, omegaF, omegaR) (S/SSOR.java:81-93): symmetric
 * successive over-relaxation; F supplies the size (square, else
 * MTJG_E_INVALID_ARGUMENT), and each relaxation must lie in [0, 2]
 * (S/SSOR.java:116-124).  setMatrix(A) takes A's pattern and values and
 * fails with MTJG_E_MISSING_DIAGONAL on a row without a diagonal entry.
 * apply(b, x) starts its sweeps from x's incoming values, as MTJ's does
 * (S/SSOR.java:143-194); transApply is apply (S/SSOR.java:196-199). */
int mtjg_prec_ssor_create(mtjg_ctx ctx, mtjg_csr F, int32_t reverse,
                          double omega_f, double omega_r, mtjg_prec *out);
/* new ICC(CompRowMatrix R) (S/ICC.java:57-63): incomplete Cholesky without
 * fill-in; R supplies the size (square, else MTJG_E_INVALID_ARGUMENT).
 * setMatrix(A) takes A's pattern and values and factors as
 * S/ICC.java:90-136 does, bit for bit (the factorization runs on the host:
 * each row is eliminated against the preceding rows its fill reaches);
 * MTJG_E_MISSING_DIAGONAL, or MTJG_E_ZERO_DIAGONAL when a row's pivot is
 * zero before its square root, with *bad_row.  A negative pivot gives NaN,
 * as in MTJ.  apply(b, x) is R'y = b then R x = y on the device
 * (S/ICC.java:65-71); transApply is apply (S/ICC.java:73-75); b and x may be
 * the same vector. */
int mtjg_prec_icc_create(mtjg_ctx ctx, mtjg_csr R, mtjg_prec *out);
/* new AMG(omegaPreF, omegaPreR, omegaPostF, omegaPostR, nu1, nu2, gamma,
 * min, omega) with reverse = 1 (S/AMG.java:140-156; AMG() is
 * (1, 1.85, 1.85, 1, 1, 1, 1, 40, 2/3), :196-205), or
 * AMG(omegaPre, omegaPost, nu1, nu2, gamma, min, omega) with reverse = 0 and
 * omegaPreF = omegaPreR = omegaPre, omegaPostF = omegaPostR = omegaPost
 * (:179-194): smoothed-aggregation algebraic multigrid with SSOR smoothing.
 * setMatrix(A) builds the hierarchy on the host bit for bit as
 * S/AMG.java:227-289 does and uploads it; A must be square
 * (MTJG_E_INVALID_ARGUMENT); MTJG_E_MISSING_DIAGONAL with *bad_row for a
 * level-0 row without a diagonal entry; MTJG_E_INVALID_ARGUMENT for a
 * relaxation outside [0, 2] once smoothers are built; MTJG_E_UNSUPPORTED when
 * a strong neighbourhood would form a java.util.HashSet tree bin, the
 * hierarchy exceeds 100 levels, or the coarsest level exceeds 4096 rows.
 * apply(b, x) runs one cycle starting from x's incoming values
 * (S/AMG.java:207-214); transApply uses the transposed coarse solve; both fail
 * with MTJG_E_SINGULAR when the coarsest matrix is singular, leaving x as it
 * was.  b and x may be the same vector. */
int mtjg_prec_amg_create(mtjg_ctx ctx, double omega_pre_f, double omega_pre_r,
                         double omega_post_f, double omega_post_r, int32_t nu1,
                         int32_t nu2, int32_t gamma, int32_t min, double omega,
                         int32_t reverse, mtjg_prec *out);
/* A Preconditioner (S/Preconditioner.java:42, :63) whose apply(b, x) is an
 * FP32 Jacobi-CG solve: x = ||b|| d, where d solves A32 d = (float)(b/||b||)
 * from d = 0 by CG (the loop of S/CG.java:58-87 on FP32 vectors, with
 * DiagonalPreconditioner's multiply by the reciprocal diagonal, in FP32) on
 * the FP32 copy of A's values, until ||r|| < inner_rtol ||r_0|| or
 * inner_max_iter iterations; sums are accumulated in double-double.
 * x's incoming values are ignored; b = 0 gives x = 0 and a non-finite b
 * gives NaN.  Passed to IR (S/IR.java:57-70: r = b - A x, M.apply(r, z),
 * x += z) this is FP32-inner / FP64 iterative refinement through MTJ's own
 * plug-in points.  setMatrix(A) builds the FP32 values and inverse diagonal
 * (MTJG_E_ZERO_DIAGONAL with *bad_row) and keeps a reference to A, which must
 * outlive the preconditioner's use.  transApply is apply (A symmetric).
 * apply synchronises the context's stream.  No MTJ counterpart as a class. */
int mtjg_prec_cg32_create(mtjg_ctx ctx, int32_t n, double inner_rtol,
                          int32_t inner_max_iter, mtjg_prec *out);
/* Inner CG iterations of an MTJG_PREC_CG32 preconditioner over all its
 * applies. */
int mtjg_prec_inner_iterations(mtjg_prec M, int64_t *total);
/* The hierarchy built by the last setMatrix: *levels, and the first
 * max_sizes level sizes (finest first) into sizes (may be NULL). */
int mtjg_prec_amg_levels(mtjg_prec M, int32_t *levels, int32_t *sizes, int32_t max_sizes);
/* setOmega(omegaF, omegaR) (S/SSOR.java:116-124). */
int mtjg_prec_ssor_set_omega(mtjg_prec M, double omega_f, double omega_r);
/* setMatrix(A) (S/Preconditioner.java:63; S/DiagonalPreconditioner.java:
 * 64-76; S/ILU.java:80-84).  On MTJG_E_ZERO_DIAGONAL, MTJG_E_MISSING_DIAGONAL
 * or MTJG_E_ZERO_PIVOT, *bad_row (may be NULL) is the 0-based row. */
int mtjg_prec_set_matrix(mtjg_prec M, mtjg_csr A, int32_t *bad_row);
/* apply(b, x): x = M^-1 b (S/Preconditioner.java:42). */
int mtjg_prec_apply(mtjg_prec M, mtjg_vec b, mtjg_vec x);
/* transApply(b, x): x = M^-T b (S/Preconditioner.java:54). */
int mtjg_prec_trans_apply(mtjg_prec M, mtjg_vec b, mtjg_vec x);
int mtjg_prec_destroy(mtjg_prec M);

/* ---- Iterative solvers (S/IterativeSolver.java) ----------------------------- */
/* DefaultIterationMonitor(maxIter, rtol, atol, dtol)
 * (S/DefaultIterationMonitor.java:72-90; defaults 100000, 1e-5, 1e-50, 1e5). */
typedef struct {
    int32_t max_iter;
    double rtol;
    double atol;
    double dtol;
} mtjg_monitor;

/* IterationMonitor.iterations() / residual() and the exception reason
 * (S/IterationMonitor.java:50-55, S/IterativeSolverNotConvergedException.java:
 * 54-89). */
typedef struct {
    int32_t iterations;
    double residual;
    int32_t reason; /* 0 converged, else MTJG_NOT_CONVERGED_* or MTJG_E_* */
} mtjg_result;

/* IterationReporter.monitor(r, iter) (S/IterationReporter.java:28-52).
 * Installing one makes the solver synchronise at every convergence check. */
typedef void (*mtjg_reporter)(double residual, int32_t iteration, void *user);

/* new CG(template) / new BiCGstab(template) / new GMRES(template) /
 * new IR(template) (S/CG.java:51-56, S/BiCGstab.java:53-63,
 * S/GMRES.java:84-104, S/IR.java:51-54): scratch vectors sized like the
 * template.  Identity preconditioner and default monitor
 * (S/AbstractIterativeSolver.java:45-48). */
int mtjg_solver_create(mtjg_ctx ctx, int kind, mtjg_vec template_vec,
                       mtjg_solver *out);
int mtjg_solver_destroy(mtjg_solver S);
/* setPreconditioner(M); NULL restores the identity
 * (S/AbstractIterativeSolver.java:50-52). */
int mtjg_solver_set_preconditioner(mtjg_solver S, mtjg_prec M);
/* setIterationMonitor(new DefaultIterationMonitor(...))
 * (S/AbstractIterativeSolver.java:62-64). */
int mtjg_solver_set_monitor(mtjg_solver S, const mtjg_monitor *m);
int mtjg_solver_set_reporter(mtjg_solver S, mtjg_reporter cb, void *user);
/* GMRES.setRestart(restart) (S/GMRES.java:112-125). */
int mtjg_gmres_set_restart(mtjg_solver S, int32_t restart);
/* Inner solve of MTJG_MIXED_CG: FP32 Jacobi-CG relative to its own initial
 * residual, to the larger of inner_rtol and half the reduction the outer
 * residual still lacks (so the last refinement is not over-solved), at most
 * inner_max_iter iterations per refinement step.  The batch engine's
 * inner_rtol has the same meaning.  No MTJ counterpart. */
int mtjg_mixed_set_inner(mtjg_solver S, double inner_rtol,
                         int32_t inner_max_iter);
/* solve(A, b, x): x holds the initial guess and receives the solution
 * (S/IterativeSolver.java:43).  result may be NULL. */
int mtjg_solver_solve(mtjg_solver S, mtjg_csr A, mtjg_vec b, mtjg_vec x,
                      mtjg_result *result);

/* ---- Allocation-free CG solve ------------------------------------------------ */
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the solve of an MTJG_CG solver with no preconditioner or
 * a DiagonalPreconditioner (S/CG.java:58-87, same kernels, same x, iteration
 * count and residual bit for bit as mtjg_solver_solve), with every scratch
 * vector, the solver state and the reduction workspace taken from a fixed
 * device array of MTJG_NOALLOC_SCRATCH_BYTES per GPU, held until the process
 * ends.  Makes no hipMalloc or hipFree.  Systems of at most
 * MTJG_NOALLOC_MAX_ROWS rows (else MTJG_E_INVALID_ARGUMENT); other
 * preconditioners give MTJG_E_UNSUPPORTED; monitor may be NULL
 * (DefaultIterationMonitor's defaults); no reporter.  Returns like
 * mtjg_solver_solve.  Every call on one GPU shares that GPU's array: calls
 * that overlap on one GPU (from several contexts or threads) race. */
#define MTJG_NOALLOC_MAX_ROWS (1 << 20)
#define MTJG_NOALLOC_SCRATCH_BYTES (41 << 20)
int mtjg_cg_solve_noalloc(mtjg_ctx ctx, mtjg_csr A, mtjg_prec M, mtjg_vec b,
                          mtjg_vec x, const mtjg_monitor *monitor,
                          mtjg_result *result);

/* ---- Batched serving of many systems (no MTJ counterpart) ------------------ */
/* A batch solves many independent SPD systems with mixed-precision CG (FP64
 * residual r = b - A x with the FP64 values, FP32 inner Jacobi-CG on r/||r||,
 * FP64 update of x), all active systems advancing together in the same
 * launches.  Requests queue on the host and are admitted between steps while
 * a slot, work-table room (ceil(n / rows_per_workgroup) workgroups per
 * system) and arena memory are free; finished systems retire after each
 * step.  A system's result bits depend only on the system and
 * rows_per_workgroup, not on the rest of the batch.  The outer loop follows
 * DefaultIterationMonitor (rtol relative to ||b||, atol 1e-50, dtol 1e5,
 * max_outer refinement steps). */
typedef struct mtjg_batch_s *mtjg_batch;

/* Admitted systems (CSR in FP64 and FP32, 3 FP64 and 6 FP32 vectors) live in
 * a device arena.  arena_bytes > 0 fixes it at that size; arena_bytes == 0
 * admits by free HBM: the arena grows in chunks while hipMemGetInfo shows the
 * chunk plus a reserve (4 GiB or 1/16 of HBM, the larger) free, and keeps
 * its chunks until the batch is destroyed.  Every host<->device copy of the
 * engine goes through pinned staging owned by the batch.  Each step runs
 * iters_per_step batch iterations from one captured hipGraph. */
int mtjg_batch_create(mtjg_ctx ctx, int32_t max_slots, int32_t max_workgroups,
                      int64_t arena_bytes, int32_t rows_per_workgroup,
                      int32_t iters_per_step, mtjg_batch *out);
int mtjg_batch_destroy(mtjg_batch B);
/* Queues A x = b (CSR copied on the host; x0 = 0); *id names the request.
 * The CSR structure is checked (row pointers 0..nnz non-decreasing, columns
 * in range, a stored nonzero diagonal in every row), in parallel on the
 * batch's host threads for large systems; the error names the lowest failing
 * row, as a serial check would. */
int mtjg_batch_submit(mtjg_batch B, int32_t n, int64_t nnz,
                      const int32_t *row_pointer, const int32_t *column_index,
                      const double *data, const double *b, double rtol,
                      int32_t max_outer, double inner_rtol, int32_t max_inner,
                      int64_t *id);
/* Queues A x = b (CSR copied on the host; x0 = 0); *id names the request.
 * The CSR structure is checked as the single-system entry points check it
 * (row pointers 0..nnz non-decreasing, columns in range, a stored nonzero
 * diagonal in every row), in parallel on the batch's host threads for large
 * systems; the error names the lowest failing row. *//* Admits what fits, runs one step, retires finished systems.  A system whose
 * step finished it is reported completed by the next call (its solution is
 * downloaded while that step's successor runs); *n_active counts admitted
 * requests not yet reported completed. */
int mtjg_batch_step(mtjg_batch B, int32_t *n_active, int32_t *n_completed);
/* Collects a finished request (x may be NULL); returns its status
 * (MTJG_OK or MTJG_NOT_CONVERGED_*); iterations = refinement steps. */
int mtjg_batch_result(mtjg_batch B, int64_t id, double *x, mtjg_result *result,
                      int32_t *inner_iterations);
/* Ids of requests retired since the last call (at most max_ids; the rest
 * stay reported for the next call). */
int mtjg_batch_completed(mtjg_batch B, int64_t *ids, int32_t max_ids,
                         int32_t *count);
int mtjg_batch_info(mtjg_batch B, int32_t *queued, int32_t *active,
                    int64_t *arena_used, int64_t *arena_size);
/* Device time of the replayed graphs so far (HIP events around each step's
 * graph, counted once its snapshot has been consumed), the steps run, and
 * (host_ms[8], may be NULL) the host time (ms) spent in: submit, admission,
 * waiting for step snapshots, retiring finished systems, retiring's wait for
 * solution downloads, admission's wait for uploads, publishing work tables,
 * launching steps. */
int mtjg_batch_stats(mtjg_batch B, double *busy_ms, int64_t *steps,
                     double *host_ms);

/* ------------------------------------------------------------------------
 * Matrix Market input (host side).  A parsed file is held in an mtjg_mm
 * handle: a CSR matrix exactly as CompRowMatrix(MatrixVectorReader) builds it
 * (S/CompRowMatrix.java:86-166: sorted unique columns per row, last duplicate
 * wins, symmetric / skew-symmetric files mirrored after the direct entries),
 * or a dense vector as DenseVector(MatrixVectorReader) builds it
 * (M/DenseVector.java:52-96).  Header, tokenizer and number rules follow
 * MatrixVectorReader (M/io/MatrixVectorReader.java:68-87, :121-211, :606-657).
 * Errors: MTJG_E_IO (IOException/EOFException), MTJG_E_UNSUPPORTED (pattern,
 * dense or complex matrix; pattern or complex vector), MTJG_E_INVALID_ARGUMENT
 * (number format, MatrixInfo validation, column index outside the matrix),
 * MTJG_E_INDEX_OUT_OF_BOUNDS (row index outside the matrix).
 * ---------------------------------------------------------------------- */

typedef struct mtjg_mm_s *mtjg_mm;

/* MatrixInfo.MatrixSymmetry of a parsed matrix. */
enum { MTJG_MM_GENERAL = 0, MTJG_MM_SYMMETRIC = 1, MTJG_MM_SKEW_SYMMETRIC = 2 };

/* new CompRowMatrix(new MatrixVectorReader(new FileReader(path))). */
int mtjg_mm_read_matrix(const char *path, mtjg_mm *out);
/* The same from text already in memory. */
int mtjg_mm_parse_matrix(const char *text, int64_t length, mtjg_mm *out);
/* new DenseVector(new MatrixVectorReader(new FileReader(path))). */
int mtjg_mm_read_vector(const char *path, mtjg_mm *out);
int mtjg_mm_parse_vector(const char *text, int64_t length, mtjg_mm *out);
/* numRows(), numColumns(), stored entries (vector: entries read), symmetry. */
int mtjg_mm_shape(mtjg_mm m, int32_t *rows, int32_t *cols, int64_t *nnz,
                  int32_t *symmetry);
/* getRowPointers(), getColumnIndices(), getData() into caller arrays of
 * rows+1, nnz and nnz elements (any may be null). */
int mtjg_mm_copy_csr(mtjg_mm m, int32_t *row_pointer, int32_t *column_index,
                     double *data);
/* getData() of a parsed vector into x[rows]. */
int mtjg_mm_copy_vector(mtjg_mm m, double *x);
/* Device CompRowMatrix / DenseVector from a parsed file. */
int mtjg_mm_to_csr(mtjg_ctx ctx, mtjg_mm m, mtjg_csr *out);
int mtjg_mm_to_vec(mtjg_ctx ctx, mtjg_mm m, mtjg_vec *out);
int mtjg_mm_destroy(mtjg_mm m);

#ifdef __cplusplus
}
#endif
#endif
