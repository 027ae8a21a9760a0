/*
 * mtj_oracle.c -- CPU restatement of MTJ's CSR SpMV, DenseVector operations,
 * DefaultIterationMonitor, DiagonalPreconditioner, ILU(0) with its triangular
 * views, GivensRotation, and the CG / BiCGstab / GMRES / IR solve loops.
 *
 * TEST INFRASTRUCTURE ONLY (the checker for the HIP path, and the timed CPU
 * baseline).  Build: -O2 -ffp-contract=off -fno-fast-math, single thread.
 *
 * Citations: S/ = src/main/java/no/uib/cipr/matrix/sparse/,
 *            M/ = src/main/java/no/uib/cipr/matrix/.
 */
#include "mtj_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

/* java.lang.Math.max(double, double): NaN wins, max(-0.0, 0.0) is 0.0. */
static double java_max(double a, double b) {
    if (a != a) return a;
    if (b != b) return b;
    if (a == 0.0 && b == 0.0) return signbit(a) ? b : a;
    return (a >= b) ? a : b;
}

/* ======================= CompRowMatrix ==================================== */

/* S/CompRowMatrix.java:294-314 (DenseVector fast path): y.zero(); per row a
 * sequential dot in storage order; y[i] is set only when dot != 0. */
void oc_csr_mult(int nr, const int32_t *rp, const int32_t *ci, const double *a,
                 const double *x, double *y) {
    for (int i = 0; i < nr; ++i) y[i] = 0.0;
    for (int i = 0; i < nr; ++i) {
        double dot = 0;
        for (int32_t j = rp[i]; j < rp[i + 1]; j++) dot += a[j] * x[ci[j]];
        if (dot != 0) y[i] = dot;
    }
}

/* S/CompRowMatrix.java:329-347: yd[i] += alpha * dot. */
void oc_csr_mult_add(int nr, const int32_t *rp, const int32_t *ci,
                     const double *a, double alpha, const double *x, double *y) {
    for (int i = 0; i < nr; ++i) {
        double dot = 0;
        for (int32_t j = rp[i]; j < rp[i + 1]; ++j) dot += a[j] * x[ci[j]];
        y[i] += alpha * dot;
    }
}

/* S/CompRowMatrix.java:349-366: y.zero(); scatter rows in ascending order. */
void oc_csr_trans_mult(int nr, int nc, const int32_t *rp, const int32_t *ci,
                       const double *a, const double *x, double *y) {
    for (int i = 0; i < nc; ++i) y[i] = 0.0;
    for (int i = 0; i < nr; ++i)
        for (int32_t j = rp[i]; j < rp[i + 1]; ++j) y[ci[j]] += a[j] * x[i];
}

/* S/CompRowMatrix.java:368-388: y.scale(1/alpha); scatter; y.scale(alpha).
 * DenseVector.scale has no alpha==0 shortcut (M/DenseVector.java:195-200). */
void oc_csr_trans_mult_add(int nr, int nc, const int32_t *rp, const int32_t *ci,
                           const double *a, double alpha, const double *x,
                           double *y) {
    oc_dv_scale(nc, 1. / alpha, y);
    for (int i = 0; i < nr; ++i)
        for (int32_t j = rp[i]; j < rp[i + 1]; ++j) y[ci[j]] += a[j] * x[i];
    oc_dv_scale(nc, alpha, y);
}

/* java.util.Arrays.binarySearch(int[], from, to, key) as wrapped by
 * S/Arrays.java:119-121: index if found, else -(insertion point) - 1. */
static int32_t bsearch_range(const int32_t *ci, int32_t key, int32_t lo,
                             int32_t hi_excl) {
    int32_t low = lo, high = hi_excl - 1;
    while (low <= high) {
        int32_t mid = (int32_t)(((uint32_t)low + (uint32_t)high) >> 1);
        int32_t v = ci[mid];
        if (v < key) low = mid + 1;
        else if (v > key) high = mid - 1;
        else return mid;
    }
    return -(low + 1);
}

/* S/CompRowMatrix.java:406-417. */
double oc_csr_get(const int32_t *rp, const int32_t *ci, const double *a, int row,
                  int col) {
    int32_t idx = bsearch_range(ci, col, rp[row], rp[row + 1]);
    return idx >= 0 ? a[idx] : 0;
}

/* ======================= DenseVector ====================================== */

void oc_dv_zero(int n, double *x) { for (int i = 0; i < n; ++i) x[i] = 0; }

/* M/DenseVector.java:195-200 (no shortcut for alpha == 0 or 1). */
void oc_dv_scale(int n, double alpha, double *x) {
    for (int i = 0; i < n; ++i) x[i] *= alpha;
}

/* M/DenseVector.java:202-213. */
void oc_dv_set(int n, const double *y, double *x) {
    memmove(x, y, (size_t)n * sizeof(double));
}

/* M/DenseVector.java:215-231: alpha == 0 zeroes (NaN in y is not copied). */
void oc_dv_set_scaled(int n, double alpha, const double *y, double *x) {
    if (alpha == 0) { oc_dv_zero(n, x); return; }
    for (int i = 0; i < n; ++i) x[i] = alpha * y[i];
}

/* M/DenseVector.java:233-246. */
void oc_dv_add(int n, const double *y, double *x) {
    for (int i = 0; i < n; i++) x[i] += y[i];
}

/* M/DenseVector.java:248-264: alpha == 0 is a no-op. */
void oc_dv_add_scaled(int n, double alpha, const double *y, double *x) {
    if (alpha == 0) return;
    for (int i = 0; i < n; i++) x[i] += alpha * y[i];
}

/* M/DenseVector.java:266-279 (this.data[i] * y[i], sequential). */
double oc_dv_dot(int n, const double *x, const double *y) {
    double dot = 0.;
    for (int i = 0; i < n; ++i) dot += x[i] * y[i];
    return dot;
}

/* M/DenseVector.java:281-287. */
double oc_dv_norm1(int n, const double *x) {
    double sum = 0;
    for (int i = 0; i < n; ++i) sum += fabs(x[i]);
    return sum;
}

/* M/DenseVector.java:289-295. */
double oc_dv_norm2(int n, const double *x) {
    double norm = 0;
    for (int i = 0; i < n; ++i) norm += x[i] * x[i];
    return sqrt(norm);
}

/* M/DenseVector.java:297-310. */
double oc_dv_norm2_robust(int n, const double *x) {
    double scale = 0, ssq = 1;
    for (int i = 0; i < n; ++i)
        if (x[i] != 0) {
            double absxi = fabs(x[i]);
            if (scale < absxi) {
                ssq = 1 + ssq * (scale / absxi) * (scale / absxi);
                scale = absxi;
            } else
                ssq += (absxi / scale) * (absxi / scale);
        }
    return scale * sqrt(ssq);
}

/* M/DenseVector.java:312-318 (Math.max keeps NaN). */
double oc_dv_norm_inf(int n, const double *x) {
    double max = 0;
    for (int i = 0; i < n; ++i) max = java_max(fabs(x[i]), max);
    return max;
}

/* ======================= DefaultIterationMonitor ========================== */

/* S/DefaultIterationMonitor.java:72-78. */
void oc_monitor_init(oc_monitor *m, int max_iter, double rtol, double atol,
                     double dtol) {
    m->iter = 0;
    m->max_iter = max_iter;
    m->rtol = rtol;
    m->atol = atol;
    m->dtol = dtol;
    m->init_r = 0;
    m->residual = 0;
}

/* S/AbstractIterationMonitor.java:81-93 then S/DefaultIterationMonitor.java:
 * 132-156: store the residual; capture initR on the first check; converged,
 * then divergence, then the iteration limit, then NaN. */
int oc_monitor_converged(oc_monitor *m, double r) {
    m->residual = r;
    if (m->iter == 0) m->init_r = r;
    if (r < java_max(m->rtol * m->init_r, m->atol)) return 1;
    if (r > m->dtol * m->init_r) return -OC_DIVERGENCE;
    if (m->iter >= m->max_iter) return -OC_ITERATIONS;
    if (r != r) return -OC_DIVERGENCE;
    return 0;
}

/* ======================= Preconditioners ================================== */

void oc_prec_identity(oc_prec *p, int n) {
    memset(p, 0, sizeof(*p));
    p->kind = OC_PREC_IDENTITY;
    p->n = n;
}

/* S/DiagonalPreconditioner.java:64-76: invdiag[i] = 1 / A.get(i,i); a zero
 * (or structurally missing) diagonal is an error. */
int oc_prec_diagonal(oc_prec *p, int n, const int32_t *rp, const int32_t *ci,
                     const double *a, int *bad_row) {
    memset(p, 0, sizeof(*p));
    p->kind = OC_PREC_DIAGONAL;
    p->n = n;
    p->invdiag = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    for (int i = 0; i < n; ++i) {
        p->invdiag[i] = oc_csr_get(rp, ci, a, i, i);
        if (p->invdiag[i] == 0) {
            if (bad_row) *bad_row = i;
            return OC_ZERO_DIAGONAL;
        }
        p->invdiag[i] = 1 / p->invdiag[i];
    }
    return OC_OK;
}

/* S/ILU.java:80-140: copy A into the factor storage, locate the diagonals,
 * then the IKJ ILU(0) sweep restricted to the pattern.  The inner merge walk
 * is restated literally (S/ILU.java:113-120), including its read of colind[l]
 * when l has reached the end of row k; a read past the last stored entry
 * would raise ArrayIndexOutOfBoundsException in Java and is reported here as
 * OC_BAD_ARG. */
int oc_prec_ilu(oc_prec *p, int n, const int32_t *rp, const int32_t *ci,
                const double *a, int *bad_row) {
    memset(p, 0, sizeof(*p));
    p->kind = OC_PREC_ILU;
    p->n = n;
    int32_t nnz = rp[n];
    p->rp = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n + 1));
    p->ci = (int32_t *)malloc(sizeof(int32_t) * (size_t)(nnz > 0 ? nnz : 1));
    p->lu = (double *)malloc(sizeof(double) * (size_t)(nnz > 0 ? nnz : 1));
    p->diagind = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
    p->y = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    memcpy(p->rp, rp, sizeof(int32_t) * (size_t)(n + 1));
    memcpy(p->ci, ci, sizeof(int32_t) * (size_t)nnz);
    memcpy(p->lu, a, sizeof(double) * (size_t)nnz);
    int32_t *colind = p->ci, *rowptr = p->rp, *diagind = p->diagind;
    double *data = p->lu;

    /* findDiagonalIndices, S/ILU.java:127-140 */
    for (int k = 0; k < n; ++k) {
        diagind[k] = bsearch_range(colind, k, rowptr[k], rowptr[k + 1]);
        if (diagind[k] < 0) {
            if (bad_row) *bad_row = k;
            return OC_MISSING_DIAGONAL;
        }
    }
    /* factor, S/ILU.java:97-121 */
    for (int k = 1; k < n; ++k)
        for (int32_t i = rowptr[k]; i < diagind[k]; ++i) {
            int32_t index = colind[i];
            double LUii = data[diagind[index]];
            if (LUii == 0) {
                if (bad_row) *bad_row = k;
                return OC_ZERO_PIVOT;
            }
            double LUki = (data[i] /= LUii);
            for (int32_t j = diagind[index] + 1, l = rowptr[k] + 1;
                 j < rowptr[index + 1]; ++j) {
                while (l < rowptr[k + 1] && colind[l] < colind[j]) l++;
                if (l >= nnz) {
                    if (bad_row) *bad_row = k;
                    return OC_BAD_ARG;
                }
                if (colind[l] == colind[j]) data[l] -= LUki * data[j];
            }
        }
    return OC_OK;
}

/* S/ICC.java:80-139, restated literally: for each row k, expand row k of R
 * (still A's values) into the dense Rk; for every i < k in turn take
 * Rki = Rk[i] / R_ii and, unless it is zero, subtract Rki * R_ij from Rk at
 * every strictly-upper entry (i, j) of the stored rows; then store
 * Rk[j] / sqrt(Rk[k]) over row k's upper pattern.  Fill produced in Rk is
 * used within the row and dropped on storing.  Work is O(n^2) as in MTJ. */
int oc_prec_icc(oc_prec *p, int n, const int32_t *rp, const int32_t *ci,
                const double *a, int *bad_row) {
    memset(p, 0, sizeof(*p));
    p->kind = OC_PREC_ICC;
    p->n = n;
    int32_t nnz = rp[n];
    p->rp = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n + 1));
    p->ci = (int32_t *)malloc(sizeof(int32_t) * (size_t)(nnz > 0 ? nnz : 1));
    p->lu = (double *)malloc(sizeof(double) * (size_t)(nnz > 0 ? nnz : 1));
    p->diagind = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
    p->y = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    memcpy(p->rp, rp, sizeof(int32_t) * (size_t)(n + 1));
    memcpy(p->ci, ci, sizeof(int32_t) * (size_t)nnz);
    memcpy(p->lu, a, sizeof(double) * (size_t)nnz);
    const int32_t *colind = p->ci, *rowptr = p->rp;
    int32_t *diagind = p->diagind;
    double *data = p->lu;

    /* findDiagonalIndices, S/ICC.java:141-155 */
    for (int k = 0; k < n; ++k) {
        diagind[k] = bsearch_range(colind, k, rowptr[k], rowptr[k + 1]);
        if (diagind[k] < 0) {
            if (bad_row) *bad_row = k;
            return OC_MISSING_DIAGONAL;
        }
    }
    double *Rk = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    int rc = OC_OK;
    for (int k = 0; k < n && rc == OC_OK; ++k) {
        for (int i = 0; i < n; ++i) Rk[i] = 0;
        for (int32_t i = rowptr[k]; i < rowptr[k + 1]; ++i) Rk[colind[i]] = data[i];
        for (int i = 0; i < k; ++i) {
            double Rii = data[diagind[i]];
            if (Rii == 0) {
                if (bad_row) *bad_row = i;
                rc = OC_ZERO_PIVOT;
                break;
            }
            double Rki = Rk[i] / Rii;
            if (Rki == 0) continue;
            for (int32_t j = diagind[i] + 1; j < rowptr[i + 1]; ++j)
                Rk[colind[j]] -= Rki * data[j];
        }
        if (rc != OC_OK) break;
        if (Rk[k] == 0) {
            if (bad_row) *bad_row = k;
            rc = OC_ZERO_DIAGONAL;
            break;
        }
        double sqRkk = sqrt(Rk[k]);
        for (int32_t i = diagind[k]; i < rowptr[k + 1]; ++i) data[i] = Rk[colind[i]] / sqRkk;
    }
    free(Rk);
    return rc;
}

/* S/UnitLowerCompRowMatrix.java:48-67: x_i = b_i - sum_{j<diag} L_ij x_j. */
void oc_unit_lower_solve(int n, const int32_t *rp, const int32_t *ci,
                         const double *a, const int32_t *diagind,
                         const double *b, double *x) {
    for (int i = 0; i < n; ++i) {
        double sum = 0;
        for (int32_t j = rp[i]; j < diagind[i]; ++j) sum += a[j] * x[ci[j]];
        x[i] = b[i] - sum;
    }
}

/* S/UnitLowerCompRowMatrix.java:69-86. */
void oc_unit_lower_trans_solve(int n, const int32_t *rp, const int32_t *ci,
                               const double *a, const int32_t *diagind,
                               const double *b, double *x) {
    oc_dv_set(n, b, x);
    for (int i = n - 1; i >= 0; --i)
        for (int32_t j = rp[i]; j < diagind[i]; ++j) x[ci[j]] -= a[j] * x[i];
}

/* S/UpperCompRowMatrix.java:48-67: x_i = (b_i - sum_{j>diag} U_ij x_j)/U_ii. */
void oc_upper_solve(int n, const int32_t *rp, const int32_t *ci,
                    const double *a, const int32_t *diagind, const double *b,
                    double *x) {
    for (int i = n - 1; i >= 0; --i) {
        double sum = 0;
        for (int32_t j = diagind[i] + 1; j < rp[i + 1]; ++j)
            sum += a[j] * x[ci[j]];
        x[i] = (b[i] - sum) / a[diagind[i]];
    }
}

/* S/UpperCompRowMatrix.java:69-90. */
void oc_upper_trans_solve(int n, const int32_t *rp, const int32_t *ci,
                          const double *a, const int32_t *diagind,
                          const double *b, double *x) {
    oc_dv_set(n, b, x);
    for (int i = 0; i < n; ++i) {
        x[i] /= a[diagind[i]];
        for (int32_t j = diagind[i] + 1; j < rp[i + 1]; ++j)
            x[ci[j]] -= a[j] * x[i];
    }
}

/* Preconditioner.apply: IdentityPreconditioner x.set(b)
 * (S/AbstractIterativeSolver.java:84-86); DiagonalPreconditioner
 * x_i = b_i * invdiag_i (S/DiagonalPreconditioner.java:54-55); ILU L.solve
 * then U.solve through y (S/ILU.java:64-70). */
/* S/SSOR.java:81-141: the relaxations are checked by setOmega (:116-124),
 * setMatrix copies A (F.set(A)) and locates the diagonals (:126-141). */
int oc_prec_ssor(oc_prec *p, int n, const int32_t *rp, const int32_t *ci,
                 const double *a, int reverse, double omega_f, double omega_r,
                 int *bad_row) {
    memset(p, 0, sizeof(*p));
    p->kind = OC_PREC_SSOR;
    p->n = n;
    if (omega_f < 0 || omega_f > 2 || omega_r < 0 || omega_r > 2) return OC_BAD_ARG;
    p->omega_f = omega_f;
    p->omega_r = omega_r;
    p->reverse = reverse;
    int32_t nnz = rp[n];
    p->rp = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n + 1));
    p->ci = (int32_t *)malloc(sizeof(int32_t) * (size_t)(nnz > 0 ? nnz : 1));
    p->lu = (double *)malloc(sizeof(double) * (size_t)(nnz > 0 ? nnz : 1));
    p->diagind = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
    p->xx = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    memcpy(p->rp, rp, sizeof(int32_t) * (size_t)(n + 1));
    memcpy(p->ci, ci, sizeof(int32_t) * (size_t)nnz);
    memcpy(p->lu, a, sizeof(double) * (size_t)nnz);
    for (int k = 0; k < n; ++k) {
        p->diagind[k] = bsearch_range(p->ci, k, p->rp[k], p->rp[k + 1]);
        if (p->diagind[k] < 0) {
            if (bad_row) *bad_row = k;
            return OC_MISSING_DIAGONAL;
        }
    }
    return OC_OK;
}

/* S/SSOR.java:143-194: forward sweep from x's incoming values into xx, then
 * (reverse) the backward sweep from xx into x. */
static void ssor_apply(const oc_prec *p, const double *bd, double *xd) {
    int n = p->n;
    const int32_t *rowptr = p->rp, *colind = p->ci, *diagind = p->diagind;
    const double *data = p->lu;
    double *xx = p->xx;
    memcpy(xx, xd, sizeof(double) * (size_t)n);
    for (int i = 0; i < n; ++i) {
        double sigma = 0;
        for (int j = rowptr[i]; j < diagind[i]; ++j) sigma += data[j] * xx[colind[j]];
        for (int j = diagind[i] + 1; j < rowptr[i + 1]; ++j) sigma += data[j] * xd[colind[j]];
        sigma = (bd[i] - sigma) / data[diagind[i]];
        xx[i] = xd[i] + p->omega_f * (sigma - xd[i]);
    }
    if (!p->reverse) {
        memcpy(xd, xx, sizeof(double) * (size_t)n);
        return;
    }
    for (int i = n - 1; i >= 0; --i) {
        double sigma = 0;
        for (int j = rowptr[i]; j < diagind[i]; ++j) sigma += data[j] * xx[colind[j]];
        for (int j = diagind[i] + 1; j < rowptr[i + 1]; ++j) sigma += data[j] * xd[colind[j]];
        sigma = (bd[i] - sigma) / data[diagind[i]];
        xd[i] = xx[i] + p->omega_r * (sigma - xx[i]);
    }
}

/* ---- FP32 inner-CG preconditioner ------------------------------------------
 * Not an MTJ class: the library's MTJG_PREC_CG32, a Preconditioner
 * (S/Preconditioner.java:42) whose apply is an FP32 solve, restated here so
 * that IR (S/IR.java:57-70) with it can be checked.  setMatrix keeps A's
 * pattern, rounds its values to binary32 and forms the binary32 reciprocal
 * diagonal 1/a_ii (rounded from double), as DiagonalPreconditioner does
 * (S/DiagonalPreconditioner.java:64-76).  apply(b, x):
 *   nb = ||b|| (exactly rounded sum of squares), rin = (float)(b * (1/nb)),
 *   then MTJ's CG loop (S/CG.java:58-87) from d = 0 on binary32 vectors:
 *   z = r * inv32 (S/DiagonalPreconditioner.java:54-55), rho = r.z,
 *   p = z | p = p*beta + z (two roundings), q = A32 p (each row summed in
 *   storage order in binary32, a zero sum stored as +0), alpha = rho / p.q,
 *   d += alpha p and r += (-alpha) q skipped when alpha == 0 (DenseVector's
 *   add(0, y) shortcut), with alpha and beta rounded to binary32 where they
 *   scale a vector; the monitor (atol 0, dtol 1e300) checks ||r|| against
 *   rtol ||rin|| at the top of every iteration; finally x = nb * (double)d.
 *   Dots and norms are accumulated in double-double (TwoSum / TwoProduct) in
 *   index order, so each is the exact sum rounded once (up to a tie), as the
 *   device's compensated reductions are.  b = 0 gives x = 0, a non-finite
 *   ||b|| gives NaN. */
typedef struct { double hi, lo; } oc_dd;

static void dd_acc(oc_dd *a, double x, double y) {
    double p = x * y;
    double e = fma(x, y, -p);
    double s = a->hi + p;
    double bb = s - a->hi;
    double t = (a->hi - (s - bb)) + (p - bb);
    a->hi = s;
    a->lo += t + e;
}

static double dd_val(oc_dd a) { return a.hi + a.lo; }

int oc_prec_cg32(oc_prec *p, int n, const int32_t *rp, const int32_t *ci,
                 const double *a, double rtol, int max_iter, int *bad_row) {
    memset(p, 0, sizeof(*p));
    p->kind = OC_PREC_CG32;
    p->n = n;
    p->rtol = rtol;
    p->max_iter = max_iter;
    int nnz = rp[n];
    p->rp = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n + 1));
    p->ci = (int32_t *)malloc(sizeof(int32_t) * (size_t)(nnz > 0 ? nnz : 1));
    p->a32 = (float *)malloc(sizeof(float) * (size_t)(nnz > 0 ? nnz : 1));
    p->inv32 = (float *)malloc(sizeof(float) * (size_t)(n > 0 ? n : 1));
    memcpy(p->rp, rp, sizeof(int32_t) * (size_t)(n + 1));
    if (nnz) memcpy(p->ci, ci, sizeof(int32_t) * (size_t)nnz);
    for (int k = 0; k < nnz; ++k) p->a32[k] = (float)a[k];
    for (int i = 0; i < n; ++i) {
        double d = 0;
        for (int k = rp[i]; k < rp[i + 1]; ++k)
            if (ci[k] == i) d = a[k];
        if (d == 0) {
            if (bad_row) *bad_row = i;
            return OC_ZERO_DIAGONAL;
        }
        p->inv32[i] = (float)(1 / d);
    }
    return OC_OK;
}

static void cg32_apply(oc_prec *p, const double *b, double *x) {
    int n = p->n;
    oc_dd bb = {0, 0};
    for (int i = 0; i < n; ++i) dd_acc(&bb, b[i], b[i]);
    double nb = sqrt(dd_val(bb));
    if (!(nb > 0) || !isfinite(nb)) {
        for (int i = 0; i < n; ++i) x[i] = nb == 0 ? 0.0 : NAN;
        return;
    }
    size_t sz = sizeof(float) * (size_t)(n > 0 ? n : 1);
    float *r = (float *)malloc(sz), *z = (float *)malloc(sz), *pp = (float *)malloc(sz),
          *q = (float *)malloc(sz), *d = (float *)malloc(sz);
    double sc = 1.0 / nb;
    for (int i = 0; i < n; ++i) {
        r[i] = (float)(b[i] * sc);
        d[i] = 0.0f;
    }
    oc_monitor m;
    oc_monitor_init(&m, p->max_iter, p->rtol, 0.0, 1e300);
    double rho = 0, rho_1 = 0, beta = 0, alpha = 0;
    for (m.iter = 0;; m.iter++) {
        if (m.iter > 0) {
            float fa = (float)alpha;
            if (fa != 0.0f)
                for (int i = 0; i < n; ++i) {
                    d[i] = d[i] + fa * pp[i];
                    r[i] = r[i] + (-fa) * q[i];
                }
        }
        oc_dd rz = {0, 0}, rr = {0, 0};
        for (int i = 0; i < n; ++i) {
            z[i] = r[i] * p->inv32[i];
            dd_acc(&rz, (double)r[i], (double)z[i]);
            dd_acc(&rr, (double)r[i], (double)r[i]);
        }
        if (oc_monitor_converged(&m, sqrt(dd_val(rr))) != 0) break;
        rho = dd_val(rz);
        if (m.iter == 0) {
            for (int i = 0; i < n; ++i) pp[i] = z[i];
        } else {
            beta = rho / rho_1;
            float fb = (float)beta;
            for (int i = 0; i < n; ++i) {
                float t = pp[i] * fb;
                pp[i] = t + z[i];
            }
        }
        oc_dd pq = {0, 0};
        for (int i = 0; i < n; ++i) {
            float s = 0.0f;
            for (int k = p->rp[i]; k < p->rp[i + 1]; ++k) s += p->a32[k] * pp[p->ci[k]];
            q[i] = s != 0.0f ? s : 0.0f;
            dd_acc(&pq, (double)pp[i], (double)q[i]);
        }
        alpha = rho / dd_val(pq);
        rho_1 = rho;
    }
    p->inner_total += m.iter;
    for (int i = 0; i < n; ++i) x[i] = nb * (double)d[i];
    free(r); free(z); free(pp); free(q); free(d);
}

void oc_prec_apply(const oc_prec *p, const double *b, double *x) {
    int n = p->n;
    if (p->kind == OC_PREC_AMG) {
        oc_amg_apply(p->amg, b, x, 0);
    } else if (p->kind == OC_PREC_SSOR) {
        ssor_apply(p, b, x);
    } else if (p->kind == OC_PREC_IDENTITY) {
        oc_dv_set(n, b, x);
    } else if (p->kind == OC_PREC_CG32) {
        cg32_apply((oc_prec *)p, b, x);
    } else if (p->kind == OC_PREC_DIAGONAL) {
        for (int i = 0; i < n; ++i) x[i] = b[i] * p->invdiag[i];
    } else if (p->kind == OC_PREC_ICC) {
        /* R'y = b, then R x = y (S/ICC.java:65-71) */
        oc_upper_trans_solve(n, p->rp, p->ci, p->lu, p->diagind, b, p->y);
        oc_upper_solve(n, p->rp, p->ci, p->lu, p->diagind, p->y, x);
    } else {
        oc_unit_lower_solve(n, p->rp, p->ci, p->lu, p->diagind, b, p->y);
        oc_upper_solve(n, p->rp, p->ci, p->lu, p->diagind, p->y, x);
    }
}

/* transApply: identity and Jacobi are symmetric, SSOR and ICC assume a
 * symmetric matrix and apply themselves (S/SSOR.java:196-199,
 * S/ICC.java:73-75); ILU is U' then L' (S/ILU.java:72-78). */
void oc_prec_trans_apply(const oc_prec *p, const double *b, double *x) {
    int n = p->n;
    if (p->kind == OC_PREC_AMG) {
        oc_amg_apply(p->amg, b, x, 1);
    } else if (p->kind != OC_PREC_ILU) {
        oc_prec_apply(p, b, x);
    } else {
        oc_upper_trans_solve(n, p->rp, p->ci, p->lu, p->diagind, b, p->y);
        oc_unit_lower_trans_solve(n, p->rp, p->ci, p->lu, p->diagind, p->y, x);
    }
}

void oc_prec_free(oc_prec *p) {
    free(p->invdiag);
    free(p->rp);
    free(p->ci);
    free(p->lu);
    free(p->diagind);
    free(p->y);
    free(p->xx);
    free(p->a32);
    free(p->inv32);
    memset(p, 0, sizeof(*p));
}

const double *oc_prec_ilu_values(const oc_prec *p) { return p->lu; }
const int32_t *oc_prec_ilu_diagind(const oc_prec *p) { return p->diagind; }

/* ======================= GMRES helpers ==================================== */

/* M/GivensRotation.java:42-58. */
void oc_givens(double x, double y, double *c, double *s) {
    double roe = fabs(x) > fabs(y) ? x : y;
    double scale = fabs(x) + fabs(y);
    if (scale != 0) {
        double xs = x / scale;
        double ys = y / scale;
        double r = scale * sqrt(xs * xs + ys * ys);
        if (roe < 0) r *= -1;
        *c = x / r;
        *s = y / r;
    } else {
        *c = 1;
        *s = 0;
    }
}

/* M/GivensRotation.java:72-76 (matrix column) and :88-92 (vector). */
static void givens_apply(double c, double s, double *a1, double *a2) {
    double temp = c * *a1 + s * *a2;
    *a2 = -s * *a1 + c * *a2;
    *a1 = temp;
}

/* LAPACK dtrtrs('U','N','N', k, 1, H, ld, s, ...): a zero diagonal gives
 * info > 0 (MatrixSingularException, M/AbstractTriangDenseMatrix.java:
 * 251-252); otherwise reference-BLAS dtrsm, left/upper/no-transpose, one
 * right-hand side: for k = n..1, if b(k) != 0 then b(k) /= A(k,k) and
 * b(i) -= b(k) * A(i,k) for i < k. */
int oc_triu_solve(int k, const double *h, int ld, double *s) {
    for (int i = 0; i < k; ++i)
        if (h[i + (size_t)i * ld] == 0) return OC_SINGULAR;
    for (int kk = k - 1; kk >= 0; --kk) {
        if (s[kk] != 0) {
            s[kk] = s[kk] / h[kk + (size_t)kk * ld];
            for (int i = 0; i < kk; ++i) s[i] = s[i] - s[kk] * h[i + (size_t)kk * ld];
        }
    }
    return OC_OK;
}

/* ======================= Solvers ========================================== */

static double *vec_alloc(int n) {
    return (double *)calloc((size_t)(n > 0 ? n : 1), sizeof(double));
}

/* Converts a monitor verdict into the loop decision: 1 stop (converged),
 * 0 continue, >1 failure code. */
static int check(oc_monitor *m, double r) {
    int c = oc_monitor_converged(m, r);
    if (c == 1) return 1;
    if (c < 0) return 1 + (-c);
    return 0;
}

/* S/CG.java:58-87. */
int oc_cg(int n, const int32_t *rp, const int32_t *ci, const double *a,
          const oc_prec *M, const double *b, double *x, oc_monitor *m) {
    double *p = vec_alloc(n), *z = vec_alloc(n), *q = vec_alloc(n),
           *r = vec_alloc(n);
    double alpha = 0, beta = 0, rho = 0, rho_1 = 0;
    int ret = OC_OK;
    (void)beta;

    oc_dv_set(n, b, r);
    oc_csr_mult_add(n, rp, ci, a, -1, x, r);

    for (m->iter = 0;; m->iter++) {
        int c = check(m, oc_dv_norm2(n, r));
        if (c) { ret = c - 1; break; }
        oc_prec_apply(M, r, z);
        rho = oc_dv_dot(n, r, z);
        if (m->iter == 0)
            oc_dv_set(n, z, p);
        else {
            beta = rho / rho_1;
            oc_dv_scale(n, beta, p);
            oc_dv_add(n, z, p);
        }
        oc_csr_mult(n, rp, ci, a, p, q);
        alpha = rho / oc_dv_dot(n, p, q);
        oc_dv_add_scaled(n, alpha, p, x);
        oc_dv_add_scaled(n, -alpha, q, r);
        rho_1 = rho;
    }
    free(p); free(z); free(q); free(r);
    return ret;
}

/* S/BiCGstab.java:65-116, including the half-step convergence check on s
 * (which re-captures initR while iter == 0) and the two breakdown tests. */
int oc_bicgstab(int n, const int32_t *rp, const int32_t *ci, const double *a,
                const oc_prec *M, const double *b, double *x, oc_monitor *m) {
    double *p = vec_alloc(n), *s = vec_alloc(n), *phat = vec_alloc(n),
           *shat = vec_alloc(n), *t = vec_alloc(n), *v = vec_alloc(n),
           *temp = vec_alloc(n), *r = vec_alloc(n), *rtilde = vec_alloc(n);
    double rho_1 = 1, rho_2 = 1, alpha = 1, beta = 1, omega = 1;
    int ret = OC_OK;

    oc_dv_set(n, b, r);
    oc_csr_mult_add(n, rp, ci, a, -1, x, r);
    oc_dv_set(n, r, rtilde);

    for (m->iter = 0;; m->iter++) {
        int c = check(m, oc_dv_norm2(n, r));
        if (c) { ret = c - 1; break; }
        rho_1 = oc_dv_dot(n, rtilde, r);
        if (rho_1 == 0) { ret = OC_BREAKDOWN; break; }
        if (omega == 0) { ret = OC_BREAKDOWN; break; }
        if (m->iter == 0)
            oc_dv_set(n, r, p);
        else {
            beta = (rho_1 / rho_2) * (alpha / omega);
            oc_dv_set_scaled(n, -omega, v, temp);
            oc_dv_add(n, p, temp);
            oc_dv_set(n, r, p);
            oc_dv_add_scaled(n, beta, temp, p);
        }
        oc_prec_apply(M, p, phat);
        oc_csr_mult(n, rp, ci, a, phat, v);
        alpha = rho_1 / oc_dv_dot(n, rtilde, v);
        oc_dv_set(n, r, s);
        oc_dv_add_scaled(n, -alpha, v, s);

        oc_dv_add_scaled(n, alpha, phat, x);
        c = check(m, oc_dv_norm2(n, s));
        if (c) { ret = c - 1; break; }

        oc_prec_apply(M, s, shat);
        oc_csr_mult(n, rp, ci, a, shat, t);
        omega = oc_dv_dot(n, t, s) / oc_dv_dot(n, t, t);
        oc_dv_add_scaled(n, omega, shat, x);
        oc_dv_set(n, s, r);
        oc_dv_add_scaled(n, -omega, t, r);

        rho_2 = rho_1;
    }
    free(p); free(s); free(phat); free(shat); free(t); free(v); free(temp);
    free(r); free(rtilde);
    return ret;
}

/* S/GMRES.java:127-176: left-preconditioned restarted GMRES with modified
 * Gram-Schmidt, Givens rotations on the column-major (restart+1) x restart
 * Hessenberg matrix, and the dtrtrs back-substitution. */
int oc_gmres(int n, const int32_t *rp, const int32_t *ci, const double *a,
             const oc_prec *M, int restart, const double *b, double *x,
             oc_monitor *m) {
    if (restart <= 0) return OC_BAD_ARG; /* S/GMRES.java:114-116 */
    int ld = restart + 1;
    double *w = vec_alloc(n), *u = vec_alloc(n), *r = vec_alloc(n);
    double *V = (double *)calloc((size_t)(restart + 1) * (size_t)(n > 0 ? n : 1),
                                 sizeof(double));
    double *s = vec_alloc(restart + 1);
    double *H = (double *)calloc((size_t)ld * restart, sizeof(double));
    double *rc = vec_alloc(restart + 1), *rs = vec_alloc(restart + 1);
    int ret = OC_OK;
#define VV(k) (V + (size_t)(k) * (size_t)n)
#define HH(i, j) H[(i) + (size_t)(j) * ld]

    oc_dv_set(n, b, u);
    oc_csr_mult_add(n, rp, ci, a, -1, x, u);
    oc_prec_apply(M, u, r);
    double normr = oc_dv_norm2(n, r);
    oc_prec_apply(M, b, u);

    for (m->iter = 0;; m->iter++) {
        int c = check(m, oc_dv_norm2(n, r));
        if (c) { ret = c - 1; break; }

        oc_dv_set_scaled(n, 1 / normr, r, VV(0));
        oc_dv_zero(restart + 1, s);
        s[0] = normr;
        int i = 0;
        int fail = 0;
        for (; i < restart; i++, m->iter++) {
            c = check(m, fabs(s[i]));
            if (c == 1) break;
            if (c > 1) { fail = c - 1; break; }
            oc_csr_mult(n, rp, ci, a, VV(i), u);
            oc_prec_apply(M, u, w);
            for (int k = 0; k <= i; k++) {
                HH(k, i) = oc_dv_dot(n, w, VV(k));
                oc_dv_add_scaled(n, -HH(k, i), VV(k), w);
            }
            HH(i + 1, i) = oc_dv_norm2(n, w);
            oc_dv_set_scaled(n, 1. / HH(i + 1, i), w, VV(i + 1));
            for (int k = 0; k < i; ++k)
                givens_apply(rc[k], rs[k], &HH(k, i), &HH(k + 1, i));
            oc_givens(HH(i, i), HH(i + 1, i), &rc[i], &rs[i]);
            givens_apply(rc[i], rs[i], &HH(i, i), &HH(i + 1, i));
            givens_apply(rc[i], rs[i], &s[i], &s[i + 1]);
        }
        if (fail) { ret = fail; break; }

        if (oc_triu_solve(i, H, ld, s) != OC_OK) { ret = OC_SINGULAR; break; }
        for (int j = 0; j < i; j++) oc_dv_add_scaled(n, s[j], VV(j), x);

        oc_dv_set(n, b, u);
        oc_csr_mult_add(n, rp, ci, a, -1, x, u);
        oc_prec_apply(M, u, r);
        normr = oc_dv_norm2(n, r);
    }
#undef VV
#undef HH
    free(w); free(u); free(r); free(V); free(s); free(H); free(rc); free(rs);
    return ret;
}

/* S/IR.java:57-70: preconditioned Richardson / iterative refinement. */
int oc_ir(int n, const int32_t *rp, const int32_t *ci, const double *a,
          const oc_prec *M, const double *b, double *x, oc_monitor *m) {
    double *z = vec_alloc(n), *r = vec_alloc(n);
    int ret = OC_OK;
    oc_dv_set(n, b, r);
    oc_csr_mult_add(n, rp, ci, a, -1, x, r);
    for (m->iter = 0;; m->iter++) {
        int c = check(m, oc_dv_norm2(n, r));
        if (c) { ret = c - 1; break; }
        oc_prec_apply(M, r, z);
        oc_dv_add(n, z, x);
        oc_dv_set(n, b, r);
        oc_csr_mult_add(n, rp, ci, a, -1, x, r);
    }
    free(z); free(r);
    return ret;
}

/* S/BiCG.java:63-102: biconjugate gradients; the shadow system uses
 * transApply and transMult. */
int oc_bicg(int n, const int32_t *rp, const int32_t *ci, const double *a,
            const oc_prec *M, const double *b, double *x, oc_monitor *m) {
    double *z = vec_alloc(n), *p = vec_alloc(n), *q = vec_alloc(n), *r = vec_alloc(n),
           *zt = vec_alloc(n), *pt = vec_alloc(n), *qt = vec_alloc(n), *rt = vec_alloc(n);
    double rho_1 = 1, rho_2 = 1, alpha = 1, beta = 1;
    int ret = OC_OK;
    oc_dv_set(n, b, r);
    oc_csr_mult_add(n, rp, ci, a, -1, x, r);
    oc_dv_set(n, r, rt);
    for (m->iter = 0;; m->iter++) {
        int c = check(m, oc_dv_norm2(n, r));
        if (c) { ret = c - 1; break; }
        oc_prec_apply(M, r, z);
        oc_prec_trans_apply(M, rt, zt);
        rho_1 = oc_dv_dot(n, z, rt);
        if (rho_1 == 0.) { ret = OC_BREAKDOWN; break; }
        if (m->iter == 0) {
            oc_dv_set(n, z, p);
            oc_dv_set(n, zt, pt);
        } else {
            beta = rho_1 / rho_2;
            oc_dv_scale(n, beta, p);
            oc_dv_add(n, z, p);
            oc_dv_scale(n, beta, pt);
            oc_dv_add(n, zt, pt);
        }
        oc_csr_mult(n, rp, ci, a, p, q);
        oc_csr_trans_mult(n, n, rp, ci, a, pt, qt);
        alpha = rho_1 / oc_dv_dot(n, pt, q);
        oc_dv_add_scaled(n, alpha, p, x);
        oc_dv_add_scaled(n, -alpha, q, r);
        oc_dv_add_scaled(n, -alpha, qt, rt);
        rho_2 = rho_1;
    }
    (void)beta;
    free(z); free(p); free(q); free(r); free(zt); free(pt); free(qt); free(rt);
    return ret;
}

/* S/CGS.java:65-105: conjugate gradients squared. */
int oc_cgs(int n, const int32_t *rp, const int32_t *ci, const double *a,
           const oc_prec *M, const double *b, double *x, oc_monitor *m) {
    double *p = vec_alloc(n), *q = vec_alloc(n), *u = vec_alloc(n), *phat = vec_alloc(n),
           *vhat = vec_alloc(n), *uhat = vec_alloc(n), *qhat = vec_alloc(n), *sum = vec_alloc(n),
           *r = vec_alloc(n), *rt = vec_alloc(n);
    double rho_1 = 0, rho_2 = 0, alpha = 0, beta = 0;
    int ret = OC_OK;
    oc_dv_set(n, b, r);
    oc_csr_mult_add(n, rp, ci, a, -1, x, r);
    oc_dv_set(n, r, rt);
    for (m->iter = 0;; m->iter++) {
        int c = check(m, oc_dv_norm2(n, r));
        if (c) { ret = c - 1; break; }
        rho_1 = oc_dv_dot(n, rt, r);
        if (rho_1 == 0) { ret = OC_BREAKDOWN; break; }
        if (m->iter == 0) {
            oc_dv_set(n, r, u);
            oc_dv_set(n, u, p);
        } else {
            beta = rho_1 / rho_2;
            oc_dv_set(n, r, u);
            oc_dv_add_scaled(n, beta, q, u);
            oc_dv_set(n, q, sum);
            oc_dv_add_scaled(n, beta, p, sum);
            oc_dv_set(n, u, p);
            oc_dv_add_scaled(n, beta, sum, p);
        }
        oc_prec_apply(M, p, phat);
        oc_csr_mult(n, rp, ci, a, phat, vhat);
        alpha = rho_1 / oc_dv_dot(n, rt, vhat);
        oc_dv_set_scaled(n, -alpha, vhat, q);
        oc_dv_add(n, u, q);
        oc_dv_set(n, u, sum);
        oc_dv_add(n, q, sum);
        oc_prec_apply(M, sum, uhat);
        oc_dv_add_scaled(n, alpha, uhat, x);
        oc_csr_mult(n, rp, ci, a, uhat, qhat);
        oc_dv_add_scaled(n, -alpha, qhat, r);
        rho_2 = rho_1;
    }
    free(p); free(q); free(u); free(phat); free(vhat); free(uhat); free(qhat); free(sum);
    free(r); free(rt);
    return ret;
}

/* S/QMR.java:117-213: quasi-minimal residual with M1 = M2 = M (the
 * setPreconditioner form, S/QMR.java:214-220). */
int oc_qmr(int n, const int32_t *rp, const int32_t *ci, const double *a,
           const oc_prec *M, const double *b, double *x, oc_monitor *m) {
    double *r = vec_alloc(n), *y = vec_alloc(n), *z = vec_alloc(n), *v = vec_alloc(n),
           *w = vec_alloc(n), *vt = vec_alloc(n), *wt = vec_alloc(n), *yt = vec_alloc(n),
           *zt = vec_alloc(n), *p = vec_alloc(n), *q = vec_alloc(n), *pt = vec_alloc(n),
           *d = vec_alloc(n), *s = vec_alloc(n);
    double rho = 0, rho_1 = 0, xi = 0, gamma = 1., gamma_1 = 0, theta = 0, theta_1 = 0,
           eta = -1., delta = 0, ep = 0, beta = 0;
    int ret = OC_OK;
    oc_dv_set(n, b, r);
    oc_csr_mult_add(n, rp, ci, a, -1, x, r);
    oc_dv_set(n, r, vt);
    oc_prec_apply(M, vt, y);
    rho = oc_dv_norm2(n, y);
    oc_dv_set(n, r, wt);
    oc_prec_trans_apply(M, wt, z);
    xi = oc_dv_norm2(n, z);
    for (m->iter = 0;; m->iter++) {
        int c = check(m, oc_dv_norm2(n, r));
        if (c) { ret = c - 1; break; }
        if (rho == 0) { ret = OC_BREAKDOWN; break; }
        if (xi == 0) { ret = OC_BREAKDOWN; break; }
        oc_dv_set_scaled(n, 1 / rho, vt, v);
        oc_dv_scale(n, 1 / rho, y);
        oc_dv_set_scaled(n, 1 / xi, wt, w);
        oc_dv_scale(n, 1 / xi, z);
        delta = oc_dv_dot(n, z, y);
        if (delta == 0) { ret = OC_BREAKDOWN; break; }
        oc_prec_apply(M, y, yt);
        oc_prec_trans_apply(M, z, zt);
        if (m->iter == 0) {
            oc_dv_set(n, yt, p);
            oc_dv_set(n, zt, q);
        } else {
            oc_dv_scale(n, -xi * delta / ep, p);
            oc_dv_add(n, yt, p);
            oc_dv_scale(n, -rho * delta / ep, q);
            oc_dv_add(n, zt, q);
        }
        oc_csr_mult(n, rp, ci, a, p, pt);
        ep = oc_dv_dot(n, q, pt);
        if (ep == 0) { ret = OC_BREAKDOWN; break; }
        beta = ep / delta;
        if (beta == 0) { ret = OC_BREAKDOWN; break; }
        oc_dv_set_scaled(n, -beta, v, vt);
        oc_dv_add(n, pt, vt);
        oc_prec_apply(M, vt, y);
        rho_1 = rho;
        rho = oc_dv_norm2(n, y);
        oc_dv_set_scaled(n, -beta, w, wt);
        oc_csr_trans_mult_add(n, n, rp, ci, a, 1, q, wt);
        oc_prec_trans_apply(M, wt, z);
        xi = oc_dv_norm2(n, z);
        gamma_1 = gamma;
        theta_1 = theta;
        theta = rho / (gamma_1 * beta);
        gamma = 1 / sqrt(1 + theta * theta);
        if (gamma == 0) { ret = OC_BREAKDOWN; break; }
        eta = -eta * rho_1 * gamma * gamma / (beta * gamma_1 * gamma_1);
        if (m->iter == 0) {
            oc_dv_set_scaled(n, eta, p, d);
            oc_dv_set_scaled(n, eta, pt, s);
        } else {
            double val = theta_1 * theta_1 * gamma * gamma;
            oc_dv_scale(n, val, d);
            oc_dv_add_scaled(n, eta, p, d);
            oc_dv_scale(n, val, s);
            oc_dv_add_scaled(n, eta, pt, s);
        }
        oc_dv_add(n, d, x);
        oc_dv_add_scaled(n, -1, s, r);
    }
    free(r); free(y); free(z); free(v); free(w); free(vt); free(wt); free(yt); free(zt);
    free(p); free(q); free(pt); free(d); free(s);
    return ret;
}

double oc_now(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}
