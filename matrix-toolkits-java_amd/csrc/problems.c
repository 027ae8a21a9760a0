/*
 * problems.c -- seeded synthetic workloads for the CSR solve path.
 *
 * Host-only C, shared by the benchmark, the tests and the serving driver so
 * that every consumer builds bit-identical inputs from a seed:
 *   - splitmix64 stream; uniform [-1,1) as (u >> 11) * 2^-53 * 2 - 1;
 *     N(0,1) by Box-Muller (one value per pair of draws);
 *   - 2-D 5-point and 3-D 7-point Poisson (natural ordering, Dirichlet by
 *     omission), constant or log-normal cell coefficients with harmonic-mean
 *     faces;
 *   - 3-D 7-point first-order upwind convection-diffusion;
 *   - power-law random CSR (row lengths P(k) ~ k^-alpha on [1, kmax],
 *     distinct uniform columns, N(0,1) values, dominant diagonal).
 * Every generated row has sorted, unique column indices and a stored
 * diagonal, as MTJ's CompRowMatrix(int,int,int[][]) constructor produces
 * (S/CompRowMatrix.java:178-208).
 */
#include <math.h>
#include <stdio.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

static uint64_t sm64(uint64_t *s) {
    uint64_t z = (*s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

static double u01(uint64_t *s) { return (double)(sm64(s) >> 11) * 0x1.0p-53; }

static double normal(uint64_t *s) {
    double a = 1.0 - u01(s); /* (0, 1] */
    double b = u01(s);
    return sqrt(-2.0 * log(a)) * cos(6.283185307179586 * b);
}

void mtjp_fill_uniform(int64_t n, uint64_t seed, double *out) {
    uint64_t s = seed;
    for (int64_t i = 0; i < n; ++i) out[i] = u01(&s) * 2.0 - 1.0;
}

void mtjp_fill_normal(int64_t n, uint64_t seed, double *out) {
    uint64_t s = seed;
    for (int64_t i = 0; i < n; ++i) out[i] = normal(&s);
}

/* ---- structured grids ---------------------------------------------------- */

/* Number of stored entries of the 2d+1-point stencil on an nx*ny*nz grid
 * (nz = 1 for 2-D). */
int64_t mtjp_stencil_nnz(int dim, int nx, int ny, int nz) {
    int64_t n = (int64_t)nx * ny * (dim == 3 ? nz : 1);
    int64_t nnz = n;
    nnz += 2 * (int64_t)(nx - 1) * ny * (dim == 3 ? nz : 1);
    nnz += 2 * (int64_t)nx * (ny - 1) * (dim == 3 ? nz : 1);
    if (dim == 3) nnz += 2 * (int64_t)nx * ny * (nz - 1);
    return nnz;
}

/* Generic stencil assembly.  kind 0: constant-coefficient Poisson (diagonal
 * 2*dim, off-diagonal -1).  kind 1: log-normal kappa per cell, harmonic-mean
 * face coefficients, boundary faces weighted by the cell's own kappa.
 * kind 2: 3-D convection-diffusion, scaled by h^2, with diffusion eps = 1/Pe
 * and first-order upwinding of the unit velocity (vx, vy, vz) drawn from the
 * seed.  Columns are emitted in ascending order. */
static int64_t stencil(int kind, int dim, int nx, int ny, int nz, double pe,
                       uint64_t seed, int32_t *rp, int32_t *ci, double *val,
                       double *vel_out) {
    if (dim != 3) nz = 1;
    int64_t n = (int64_t)nx * ny * nz;
    int64_t sxy = (int64_t)nx * ny;
    double *kap = NULL;
    if (kind == 1) {
        kap = (double *)malloc(sizeof(double) * (size_t)n);
        uint64_t s = seed;
        for (int64_t i = 0; i < n; ++i) kap[i] = exp(normal(&s));
    }
    double v[3] = {0, 0, 0}, eps = 0, h = 0;
    if (kind == 2) {
        uint64_t s = seed;
        double nrm = 0;
        for (int d = 0; d < 3; ++d) { v[d] = normal(&s); nrm += v[d] * v[d]; }
        nrm = sqrt(nrm);
        for (int d = 0; d < 3; ++d) v[d] /= nrm;
        eps = 1.0 / pe;
        h = 1.0 / (double)(nx + 1);
        if (vel_out) for (int d = 0; d < 3; ++d) vel_out[d] = v[d];
    }
    int64_t pos = 0;
    rp[0] = 0;
    for (int64_t k = 0; k < nz; ++k)
        for (int64_t j = 0; j < ny; ++j)
            for (int64_t i = 0; i < nx; ++i) {
                int64_t row = i + nx * (j + ny * k);
                /* neighbour offsets in ascending column order */
                int64_t off[7];
                int has[7];
                int dir[7], sgn[7];
                int m = 0;
                if (dim == 3) { off[m] = -sxy; has[m] = k > 0; dir[m] = 2; sgn[m] = -1; m++; }
                off[m] = -nx; has[m] = j > 0; dir[m] = 1; sgn[m] = -1; m++;
                off[m] = -1; has[m] = i > 0; dir[m] = 0; sgn[m] = -1; m++;
                off[m] = 0; has[m] = 1; dir[m] = -1; sgn[m] = 0; m++;
                off[m] = 1; has[m] = i < nx - 1; dir[m] = 0; sgn[m] = 1; m++;
                off[m] = nx; has[m] = j < ny - 1; dir[m] = 1; sgn[m] = 1; m++;
                if (dim == 3) { off[m] = sxy; has[m] = k < nz - 1; dir[m] = 2; sgn[m] = 1; m++; }
                double diag = 0;
                int64_t dpos = -1;
                for (int t = 0; t < m; ++t) {
                    if (off[t] == 0) {
                        ci[pos] = (int32_t)row;
                        dpos = pos++;
                        continue;
                    }
                    double face;
                    if (kind == 0) {
                        face = 1.0;
                    } else if (kind == 1) {
                        double kp = kap[row];
                        if (has[t]) {
                            double kq = kap[row + off[t]];
                            face = 2.0 * kp * kq / (kp + kq);
                        } else {
                            face = kp;
                        }
                    } else {
                        /* diffusion eps plus the upwind share of |v_d| * h */
                        double vd = v[dir[t]];
                        double up = (sgn[t] < 0) ? (vd > 0 ? vd : 0) : (vd < 0 ? -vd : 0);
                        face = eps + h * up;
                    }
                    if (kind == 2) {
                        /* diagonal gets eps per face plus |v_d| h once per axis */
                        diag += eps + (sgn[t] < 0 ? h * fabs(v[dir[t]]) : 0.0);
                    } else {
                        diag += face;
                    }
                    if (has[t]) {
                        ci[pos] = (int32_t)(row + off[t]);
                        val[pos] = -face;
                        pos++;
                    }
                }
                val[dpos] = (kind == 0) ? (double)(2 * dim) : diag;
                rp[row + 1] = (int32_t)pos;
            }
    free(kap);
    return pos;
}

int64_t mtjp_poisson(int dim, int nx, int ny, int nz, int32_t *rp, int32_t *ci,
                     double *val) {
    return stencil(0, dim, nx, ny, nz, 0, 0, rp, ci, val, NULL);
}

int64_t mtjp_varcoef_poisson(int dim, int nx, int ny, int nz, uint64_t seed,
                             int32_t *rp, int32_t *ci, double *val) {
    return stencil(1, dim, nx, ny, nz, 0, seed, rp, ci, val, NULL);
}

int64_t mtjp_convdiff3d(int nx, int ny, int nz, double peclet, uint64_t seed,
                        int32_t *rp, int32_t *ci, double *val, double *vel) {
    return stencil(2, 3, nx, ny, nz, peclet, seed, rp, ci, val, vel);
}

/* ---- power-law random CSR ------------------------------------------------ */

/* Off-diagonal row lengths L_i ~ P(k) proportional to k^-alpha on [1, kmax],
 * by inverse CDF.  Returns the total number of stored entries including one
 * diagonal per row (lengths are capped at n - 1). */
int64_t mtjp_powerlaw_lengths(int32_t n, double alpha, int32_t kmax,
                              uint64_t seed, int32_t *len) {
    double *cdf = (double *)malloc(sizeof(double) * (size_t)kmax);
    double acc = 0;
    for (int32_t k = 1; k <= kmax; ++k) { acc += pow((double)k, -alpha); cdf[k - 1] = acc; }
    for (int32_t k = 0; k < kmax; ++k) cdf[k] /= acc;
    uint64_t s = seed;
    int64_t total = 0;
    for (int32_t i = 0; i < n; ++i) {
        double u = u01(&s);
        int32_t lo = 0, hi = kmax - 1;
        while (lo < hi) {
            int32_t mid = (lo + hi) / 2;
            if (cdf[mid] < u) lo = mid + 1; else hi = mid;
        }
        int32_t L = lo + 1;
        if (L > n - 1) L = n - 1;
        len[i] = L;
        total += L + 1;
    }
    free(cdf);
    return total;
}

static int cmp_i32(const void *a, const void *b) {
    int32_t x = *(const int32_t *)a, y = *(const int32_t *)b;
    return (x > y) - (x < y);
}

/* Fills the CSR of the power-law matrix: distinct uniform off-diagonal
 * columns, N(0,1) values, diagonal 1 + sum |a_ij| (strict dominance). */
void mtjp_powerlaw_fill(int32_t n, const int32_t *len, uint64_t seed,
                        int32_t *rp, int32_t *ci, double *val) {
    uint64_t s = seed ^ 0xA5A5A5A5DEADBEEFULL;
    rp[0] = 0;
    for (int32_t i = 0; i < n; ++i) {
        int32_t L = len[i];
        int32_t base = rp[i];
        int32_t *c = ci + base;
        int32_t have = 0;
        /* draw, sort, dedupe, redraw the shortfall */
        while (have < L) {
            while (have < L) {
                int32_t col = (int32_t)(sm64(&s) % (uint64_t)(n - 1));
                if (col >= i) col++; /* skip the diagonal */
                c[have++] = col;
            }
            qsort(c, (size_t)have, sizeof(int32_t), cmp_i32);
            int32_t u = 0;
            for (int32_t t = 0; t < have; ++t)
                if (t == 0 || c[t] != c[t - 1]) c[u++] = c[t];
            have = u;
        }
        /* insert the diagonal in order */
        int32_t pos = L;
        while (pos > 0 && c[pos - 1] > i) { c[pos] = c[pos - 1]; pos--; }
        c[pos] = i;
        double sum = 0;
        for (int32_t t = 0; t <= L; ++t) {
            if (t == pos) continue;
            double v = normal(&s);
            val[base + t] = v;
            sum += fabs(v);
        }
        val[base + pos] = 1.0 + sum;
        rp[i + 1] = base + L + 1;
    }
}

/* Symmetric positive definite companion of a CSR matrix B with a stored
 * diagonal: S = B_off + B_off^T (B_off = B without its diagonal), then every
 * diagonal entry set to 1 + sum_j |s_ij| summed in column order, so S is
 * strictly diagonally dominant with a positive diagonal.  Rows of S hold the
 * union of row i and column i of B_off, sorted, plus the diagonal.  With
 * sci == NULL only srp is filled; returns nnz(S), or -1 if out of memory. */
int64_t mtjp_symmetrize(int32_t n, const int32_t *rp, const int32_t *ci, const double *val,
                        int32_t *srp, int32_t *sci, double *sval) {
    int64_t nnz = rp[n];
    int32_t *tp = (int32_t *)calloc((size_t)n + 1, sizeof(int32_t));
    int32_t *tr = (int32_t *)malloc(sizeof(int32_t) * (size_t)(nnz > 0 ? nnz : 1));
    double *tv = (double *)malloc(sizeof(double) * (size_t)(nnz > 0 ? nnz : 1));
    if (!tp || !tr || !tv) { free(tp); free(tr); free(tv); return -1; }
    for (int32_t i = 0; i < n; ++i)
        for (int32_t k = rp[i]; k < rp[i + 1]; ++k)
            if (ci[k] != i) tp[ci[k] + 1]++;
    for (int32_t j = 0; j < n; ++j) tp[j + 1] += tp[j];
    int32_t *fill = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
    if (!fill) { free(tp); free(tr); free(tv); return -1; }
    memcpy(fill, tp, sizeof(int32_t) * (size_t)n);
    /* rows visited in order, so each column of the transpose is sorted */
    for (int32_t i = 0; i < n; ++i)
        for (int32_t k = rp[i]; k < rp[i + 1]; ++k)
            if (ci[k] != i) { int32_t p = fill[ci[k]]++; tr[p] = i; tv[p] = val[k]; }
    free(fill);
    int64_t pos = 0;
    srp[0] = 0;
    for (int32_t i = 0; i < n; ++i) {
        int32_t a = rp[i], ae = rp[i + 1], b = tp[i], be = tp[i + 1];
        int64_t diag = -1;
        double sum = 0;
        for (;;) {
            while (a < ae && ci[a] == i) a++;
            int32_t ca = a < ae ? ci[a] : INT32_MAX, cb = b < be ? tr[b] : INT32_MAX;
            int32_t c = ca < cb ? ca : cb;
            if (diag < 0 && i < c) {
                diag = pos;
                if (sci) sci[pos] = i;
                pos++;
            }
            if (c == INT32_MAX) break;
            double v = 0;
            if (ca == c && cb == c) v = val[a++] + tv[b++];
            else if (ca == c) v = val[a++];
            else v = tv[b++];
            if (sci) { sci[pos] = c; sval[pos] = v; sum += fabs(v); }
            pos++;
        }
        if (sci) sval[diag] = 1.0 + sum;
        srp[i + 1] = (int32_t)pos;
    }
    free(tp); free(tr); free(tv);
    return pos;
}

/* Writes a CSR matrix as a Matrix Market coordinate file in the layout of
 * MTJ's MatrixVectorWriter (M/io/MatrixVectorWriter.java:89-116, :249-263):
 * MatrixInfo header, "%10d %10d %19d" size line, "%10d %10d % .12e" entries,
 * 1-based indices.  With lower_only, only entries on or below the diagonal
 * are written and the header says "symmetric".  Returns 0 on success. */
int mtjp_write_mm(const char *path, int32_t nr, int32_t nc, const int32_t *rp,
                  const int32_t *ci, const double *val, int lower_only) {
    FILE *f = fopen(path, "w");
    if (!f) return -1;
    static char buf[1 << 20];
    setvbuf(f, buf, _IOFBF, sizeof buf);
    int64_t count = 0;
    for (int32_t i = 0; i < nr; ++i)
        for (int32_t k = rp[i]; k < rp[i + 1]; ++k)
            if (!lower_only || ci[k] <= i) count++;
    fprintf(f, "%%%%MatrixMarket matrix coordinate real %s\n", lower_only ? "symmetric" : "general");
    fprintf(f, "%10d %10d %19lld\n", nr, nc, (long long)count);
    for (int32_t i = 0; i < nr; ++i)
        for (int32_t k = rp[i]; k < rp[i + 1]; ++k)
            if (!lower_only || ci[k] <= i) fprintf(f, "%10d %10d % .12e\n", i + 1, ci[k] + 1, val[k]);
    return fclose(f) == 0 ? 0 : -1;
}
