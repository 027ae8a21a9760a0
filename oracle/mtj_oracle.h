/*
 * mtj_oracle.h -- CPU restatement of MTJ's unstructured-sparse hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP
 * implementation: only the test suite, the smoke check and the CPU-baseline
 * leg of the benchmark load it.  Nothing in the product library links or
 * calls it.
 *
 * Every function restates one MTJ routine in the same operation order, with
 * IEEE binary64 arithmetic and no FMA contraction (built with
 * -O2 -ffp-contract=off -fno-fast-math), so that results equal what a JVM
 * running MTJ computes on the same inputs.  Reference citations use
 *   S/ = src/main/java/no/uib/cipr/matrix/sparse/
 *   M/ = src/main/java/no/uib/cipr/matrix/
 *
 * Parity pinning (what checks each function independently of its own
 * formula) is listed in the test files tests/test_oracle_*.py; the one
 * function whose arithmetic is fixed by a third-party library (the LAPACK
 * dtrtrs back-substitution reached from GMRES) follows reference BLAS dtrsm
 * and is pinned against scipy.linalg.solve_triangular.
 */
#ifndef MTJ_ORACLE_H
#define MTJ_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Outcome codes, mirroring IterativeSolverNotConvergedException.Reason
 * (M/NotConvergedException.java:33-49) and the RuntimeExceptions of the
 * preconditioners. */
enum {
    OC_OK = 0,
    OC_ITERATIONS = 1,      /* Reason.Iterations */
    OC_DIVERGENCE = 2,      /* Reason.Divergence */
    OC_BREAKDOWN = 3,       /* Reason.Breakdown  */
    OC_ZERO_DIAGONAL = 10,  /* S/DiagonalPreconditioner.java:71-72 */
    OC_MISSING_DIAGONAL = 11, /* S/ILU.java:134-136 */
    OC_ZERO_PIVOT = 12,     /* S/ILU.java:105-107 */
    OC_SINGULAR = 13,       /* MatrixSingularException from dtrtrs */
    OC_BAD_ARG = 20,        /* IllegalArgumentException (incl. NumberFormatException) */
    OC_IO = 21,             /* IOException / EOFException of MatrixVectorReader */
    OC_UNSUPPORTED = 22,    /* UnsupportedOperationException */
    OC_INDEX = 23           /* IndexOutOfBoundsException */
};

/* ---- CompRowMatrix (S/CompRowMatrix.java) -------------------------------- */
void oc_csr_mult(int nr, const int32_t *rp, const int32_t *ci, const double *a,
                 const double *x, double *y);
void oc_csr_mult_add(int nr, const int32_t *rp, const int32_t *ci,
                     const double *a, double alpha, const double *x, double *y);
void oc_csr_trans_mult(int nr, int nc, const int32_t *rp, const int32_t *ci,
                       const double *a, const double *x, double *y);
void oc_csr_trans_mult_add(int nr, int nc, const int32_t *rp, const int32_t *ci,
                           const double *a, double alpha, const double *x,
                           double *y);
double oc_csr_get(const int32_t *rp, const int32_t *ci, const double *a, int row,
                  int col);

/* ---- DenseVector (M/DenseVector.java) ------------------------------------ */
void oc_dv_zero(int n, double *x);
void oc_dv_scale(int n, double alpha, double *x);
void oc_dv_set(int n, const double *y, double *x);
void oc_dv_set_scaled(int n, double alpha, const double *y, double *x);
void oc_dv_add(int n, const double *y, double *x);
void oc_dv_add_scaled(int n, double alpha, const double *y, double *x);
double oc_dv_dot(int n, const double *x, const double *y);
double oc_dv_norm1(int n, const double *x);
double oc_dv_norm2(int n, const double *x);
double oc_dv_norm2_robust(int n, const double *x);
double oc_dv_norm_inf(int n, const double *x);

/* ---- DefaultIterationMonitor (S/DefaultIterationMonitor.java) ------------ */
typedef struct {
    int iter;
    int max_iter;
    double rtol, atol, dtol;
    double init_r;
    double residual;
} oc_monitor;

void oc_monitor_init(oc_monitor *m, int max_iter, double rtol, double atol,
                     double dtol);
/* 1 = converged, 0 = continue, OC_DIVERGENCE / OC_ITERATIONS as negative
 * values: -OC_DIVERGENCE, -OC_ITERATIONS. */
int oc_monitor_converged(oc_monitor *m, double r);

/* ---- Preconditioners ------------------------------------------------------ */
enum { OC_PREC_IDENTITY = 0, OC_PREC_DIAGONAL = 1, OC_PREC_ILU = 2, OC_PREC_SSOR = 3, OC_PREC_ICC = 4,
       OC_PREC_AMG = 5, OC_PREC_CG32 = 6 };

struct oc_amg_s;

typedef struct {
    int kind;
    int n;
    double *invdiag;     /* DiagonalPreconditioner */
    int32_t *rp, *ci;    /* ILU: copied pattern */
    double *lu;          /* ILU: factor values */
    int32_t *diagind;    /* ILU: diagonal positions */
    double *y;           /* ILU: temporary */
    double omega_f, omega_r; /* SSOR relaxation of the forward / reverse sweep */
    int reverse;         /* SSOR: run the reverse sweep */
    double *xx;          /* SSOR: half iterate */
    struct oc_amg_s *amg; /* AMG: the hierarchy (owned by the caller) */
    float *a32;          /* CG32: binary32 copy of A's values (pattern in rp, ci) */
    float *inv32;        /* CG32: binary32 reciprocal diagonal */
    double rtol;         /* CG32: inner relative tolerance */
    int max_iter;        /* CG32: inner iteration limit */
    long inner_total;    /* CG32: inner iterations over all applies */
} oc_prec;

void oc_prec_identity(oc_prec *p, int n);
/* Returns OC_OK or OC_ZERO_DIAGONAL (row in *bad_row, 0-based). */
int oc_prec_diagonal(oc_prec *p, int n, const int32_t *rp, const int32_t *ci,
                     const double *a, int *bad_row);
/* Returns OC_OK, OC_MISSING_DIAGONAL or OC_ZERO_PIVOT (row in *bad_row). */
int oc_prec_ilu(oc_prec *p, int n, const int32_t *rp, const int32_t *ci,
                const double *a, int *bad_row);
/* SSOR(F, reverse, omegaF, omegaR) + setMatrix(A) (S/SSOR.java:81-141):
 * OC_BAD_ARG for a relaxation outside [0, 2], OC_MISSING_DIAGONAL (row in
 * *bad_row).  Its apply reads x's incoming values as the first iterate. */
int oc_prec_ssor(oc_prec *p, int n, const int32_t *rp, const int32_t *ci,
                 const double *a, int reverse, double omega_f, double omega_r,
                 int *bad_row);
/* ICC(R) + setMatrix(A) (S/ICC.java:57-139): the factor R (upper part in
 * place, lower part keeps A's values) over A's pattern.  OC_MISSING_DIAGONAL,
 * OC_ZERO_PIVOT (a zero R_ii, S/ICC.java:108-110) or OC_ZERO_DIAGONAL (the
 * pivot of row k is zero before its square root, S/ICC.java:127-130), row in
 * *bad_row. */
int oc_prec_icc(oc_prec *p, int n, const int32_t *rp, const int32_t *ci,
                const double *a, int *bad_row);
/* FP32 inner-CG preconditioner + setMatrix(A): OC_ZERO_DIAGONAL (row in
 * *bad_row).  apply(b, x): x = ||b|| d, d from CG in binary32 on
 * (float)(b / ||b||); see mtj_oracle.c. */
int oc_prec_cg32(oc_prec *p, int n, const int32_t *rp, const int32_t *ci,
                 const double *a, double rtol, int max_iter, int *bad_row);
void oc_prec_apply(const oc_prec *p, const double *b, double *x);
void oc_prec_trans_apply(const oc_prec *p, const double *b, double *x);
void oc_prec_free(oc_prec *p);
/* Accessors for tests: the factored values and diagonal index (ILU). */
const double *oc_prec_ilu_values(const oc_prec *p);
const int32_t *oc_prec_ilu_diagind(const oc_prec *p);

/* ---- AMG (S/AMG.java) ---------------------------------------------------- */
/* Parameters (S/AMG.java:113-205), then the hierarchy built by setMatrix:
 * m levels; levels 0..m-2 carry A_k (CSR), the aggregate of each node
 * (pt_k, -1 for none), the interpolation I_k (n_k x n_{k+1}, compressed
 * columns, rows ascending) and one SSOR relaxation (the pre- and
 * post-smoothers differ only in their relaxations); the coarsest level a
 * column-major dense LU with 1-based pivots. */
typedef struct oc_amg_s {
    double omega_pre_f, omega_pre_r, omega_post_f, omega_post_r, omega;
    int nu1, nu2, gamma, min_size, reverse;
    int m;
    int32_t *n;
    int32_t **arp, **aci;
    double **ava;
    int32_t **pt;
    int32_t **icp, **iri;
    double **iva;
    oc_prec *ssor;
    double *lu;
    int32_t *piv;
    int singular;
    double **u, **f, **r;
    int transpose, error;
} oc_amg;

void oc_amg_params(oc_amg *g, double omega_pre_f, double omega_pre_r, double omega_post_f,
                   double omega_post_r, int nu1, int nu2, int gamma, int min_size, double omega,
                   int reverse);
/* AMG.setMatrix (S/AMG.java:227-289): OC_OK, OC_MISSING_DIAGONAL (row and
 * level in *bad_row, *bad_level), OC_BAD_ARG (a relaxation outside [0, 2],
 * checked when SSOR levels are built), OC_UNSUPPORTED (a strong
 * neighbourhood whose HashSet would hold a tree bin, or 100 levels). */
int oc_amg_setup(oc_amg *g, int n, const int32_t *rp, const int32_t *ci, const double *a,
                 int *bad_row, int *bad_level);
/* apply (transpose = 0) / transApply (1) (S/AMG.java:207-225): OC_OK, or
 * OC_SINGULAR when the coarsest matrix is singular (x left unchanged). */
int oc_amg_apply(oc_amg *g, const double *b, double *x, int transpose);
void oc_amg_free(oc_amg *g);
/* An AMG as a Preconditioner for the restated solvers. */
void oc_prec_amg(oc_prec *p, oc_amg *g);
/* Iteration order of a java.util.HashSet<Integer> filled with vals[0..cnt)
 * in that order: OC_OK, or OC_UNSUPPORTED when a tree bin would form. */
int oc_java_hashset_order(const int32_t *vals, int cnt, int32_t *order);
/* Reference LAPACK dgetf2 (returns info) and dgetrs (trans 0 / 1), n x n,
 * column-major, pivots 1-based. */
int oc_dgetf2(int n, double *a, int32_t *piv);
void oc_dgetrs(int trans, int n, const double *a, const int32_t *piv, double *b);

/* Triangular views over an ILU factor (S/UnitLowerCompRowMatrix.java,
 * S/UpperCompRowMatrix.java). */
void oc_unit_lower_solve(int n, const int32_t *rp, const int32_t *ci,
                         const double *a, const int32_t *diagind,
                         const double *b, double *x);
void oc_unit_lower_trans_solve(int n, const int32_t *rp, const int32_t *ci,
                               const double *a, const int32_t *diagind,
                               const double *b, double *x);
void oc_upper_solve(int n, const int32_t *rp, const int32_t *ci,
                    const double *a, const int32_t *diagind, const double *b,
                    double *x);
void oc_upper_trans_solve(int n, const int32_t *rp, const int32_t *ci,
                          const double *a, const int32_t *diagind,
                          const double *b, double *x);

/* ---- GMRES helpers --------------------------------------------------------- */
/* GivensRotation (M/GivensRotation.java:42-58). */
void oc_givens(double x, double y, double *c, double *s);
/* Upper-triangular solve of the leading k x k block of column-major H (leading
 * dimension ld) in place on s; reference-BLAS dtrsm column order as reached
 * through dtrtrs (M/AbstractTriangDenseMatrix.java:228-257).  Returns OC_OK or
 * OC_SINGULAR. */
int oc_triu_solve(int k, const double *h, int ld, double *s);

/* ---- Solvers ---------------------------------------------------------------- */
/* All solvers take x as the initial guess and return the solution in x.
 * Return OC_OK on convergence, else OC_ITERATIONS / OC_DIVERGENCE /
 * OC_BREAKDOWN / OC_SINGULAR; m->iter and m->residual hold the monitor state
 * at exit, as IterativeSolverNotConvergedException reports them. */
int oc_cg(int n, const int32_t *rp, const int32_t *ci, const double *a,
          const oc_prec *M, const double *b, double *x, oc_monitor *m);
int oc_bicgstab(int n, const int32_t *rp, const int32_t *ci, const double *a,
                const oc_prec *M, const double *b, double *x, oc_monitor *m);
int oc_gmres(int n, const int32_t *rp, const int32_t *ci, const double *a,
             const oc_prec *M, int restart, const double *b, double *x,
             oc_monitor *m);
int oc_ir(int n, const int32_t *rp, const int32_t *ci, const double *a,
          const oc_prec *M, const double *b, double *x, oc_monitor *m);

/* ---- timing helper for the CPU baseline ------------------------------------ */
/* S/BiCG.java:63-102, S/CGS.java:65-105, S/QMR.java:117-213 (M1 = M2 = M). */
int oc_bicg(int n, const int32_t *rp, const int32_t *ci, const double *a,
            const oc_prec *M, const double *b, double *x, oc_monitor *m);
int oc_cgs(int n, const int32_t *rp, const int32_t *ci, const double *a,
           const oc_prec *M, const double *b, double *x, oc_monitor *m);
int oc_qmr(int n, const int32_t *rp, const int32_t *ci, const double *a,
           const oc_prec *M, const double *b, double *x, oc_monitor *m);
double oc_now(void);

/* Matrix Market input (mtj_oracle_mm.c): CompRowMatrix(MatrixVectorReader)
 * (S/CompRowMatrix.java:86-166) and DenseVector(MatrixVectorReader)
 * (M/DenseVector.java:52-96) on text in memory.  For a vector, val holds the
 * dense entries and rp/ci are null. */
typedef struct {
    int32_t rows, cols, symmetry;
    long nnz;
    int32_t *rp, *ci;
    double *val;
} oc_mm;
int oc_mm_matrix(const char *text, long len, oc_mm *out);
int oc_mm_vector(const char *text, long len, oc_mm *out);
void oc_mm_free(oc_mm *m);

#ifdef __cplusplus
}
#endif
#endif
