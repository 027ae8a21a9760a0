/* CPU restatement of MTJ's algebraic multigrid preconditioner (S/AMG.java)
 * and of the reference-LAPACK routines behind its coarsest-level DenseLU
 * (M/DenseLU.java:80-94, :176-205).
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the device AMG; the product never
 * links it.  Every loop follows the cited MTJ source in MTJ's order with MTJ's
 * dense work arrays (O(n * c) for the prolongation and O(c * nnz) for the
 * Galerkin product, as in MTJ), compiled without FMA contraction.
 *
 * java.util.HashSet iteration order.  The enlargement step
 * (S/AMG.java:542-590) breaks ties between aggregates by the order in which it
 * walks a node's strong neighbourhood, a HashSet<Integer>.  OpenJDK's
 * HashMap (the published algorithm of java.util.HashMap, JDK 8 and later)
 * puts an Integer v in bucket (h ^ (h >>> 16)) & (cap - 1) with h = v; the
 * table is created with 16 buckets on the first insertion and doubles when
 * the size exceeds 3/4 of the buckets, or when an insertion finds 8 entries
 * in its bucket while the table has fewer than 64 buckets (treeifyBin);
 * resizing keeps the insertion order inside a bucket, and iteration walks
 * the buckets in index order.  A bucket reaching 9 entries in a table of 64
 * or more buckets becomes a red-black tree whose iteration order is not
 * restated here: such a neighbourhood is reported as OC_UNSUPPORTED.
 *
 * LAPACK.  DenseLU.factor calls dgetrf and solve / transSolve call dgetrs
 * (netlib-java 1.1.2, pom.xml:87-92, not under the reference tree).  For
 * n <= 64 reference dgetrf is the unblocked dgetf2, restated here as the
 * published LAPACK 3.2+ algorithm (IDAMAX pivot, row swap, reciprocal scaling
 * when |pivot| >= the safe minimum, DGER rank-1 update), and dgetrs as DLASWP
 * plus the reference-BLAS DTRSM loops.  For larger coarse matrices LAPACK's
 * blocked dgetrf rounds differently; the restatement stays unblocked. */
#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "mtj_oracle.h"

#define AMG_MAX_LEVELS 100

static int32_t bsearch_col(const int32_t *ci, int32_t key, int32_t lo, int32_t hi) {
    hi -= 1;
    while (lo <= hi) {
        int32_t mid = (lo + hi) >> 1;
        if (ci[mid] < key) lo = mid + 1;
        else if (ci[mid] > key) hi = mid - 1;
        else return mid;
    }
    return -1;
}

/* ---- java.util.HashSet<Integer> ------------------------------------------ */

static uint32_t jhash(int32_t v) {
    uint32_t h = (uint32_t)v;
    return h ^ (h >> 16);
}

int oc_java_hashset_order(const int32_t *vals, int cnt, int32_t *order) {
    uint32_t cap = 16;
    int size = 0;
    for (int t = 0; t < cnt; ++t) {
        int same = 0;
        for (int s = 0; s < t; ++s)
            if (((jhash(vals[s]) ^ jhash(vals[t])) & (cap - 1)) == 0) same++;
        if (same >= 8) { /* HashMap.putVal -> treeifyBin */
            if (cap < 64) cap *= 2;
            else return OC_UNSUPPORTED;
        }
        if (++size > (int)(cap / 4 * 3)) cap *= 2;
    }
    /* buckets in index order, insertion order inside a bucket */
    int k = 0;
    for (uint32_t b = 0; b < cap && k < cnt; ++b)
        for (int t = 0; t < cnt; ++t)
            if ((jhash(vals[t]) & (cap - 1)) == b) order[k++] = vals[t];
    return OC_OK;
}

typedef struct {
    int32_t *v; /* members in HashSet iteration order */
    int cnt;
} jset;

static int jset_contains(const jset *s, int32_t x) {
    for (int t = 0; t < s->cnt; ++t)
        if (s->v[t] == x) return 1;
    return 0;
}

/* ---- LAPACK ---------------------------------------------------------------- */

int oc_dgetf2(int n, double *a, int32_t *piv) {
    int info = 0;
    const double sfmin = DBL_MIN;
    for (int j = 0; j < n; ++j) {
        int jp = j; /* IDAMAX: first index of the largest |a| */
        double dmax = fabs(a[j + (size_t)j * n]);
        for (int i = j + 1; i < n; ++i)
            if (fabs(a[i + (size_t)j * n]) > dmax) {
                jp = i;
                dmax = fabs(a[i + (size_t)j * n]);
            }
        piv[j] = jp + 1;
        if (a[jp + (size_t)j * n] != 0) {
            if (jp != j)
                for (int c = 0; c < n; ++c) {
                    double t = a[j + (size_t)c * n];
                    a[j + (size_t)c * n] = a[jp + (size_t)c * n];
                    a[jp + (size_t)c * n] = t;
                }
            if (j < n - 1) {
                if (fabs(a[j + (size_t)j * n]) >= sfmin) {
                    double r = 1.0 / a[j + (size_t)j * n];
                    for (int i = j + 1; i < n; ++i) a[i + (size_t)j * n] = r * a[i + (size_t)j * n];
                } else {
                    for (int i = j + 1; i < n; ++i)
                        a[i + (size_t)j * n] = a[i + (size_t)j * n] / a[j + (size_t)j * n];
                }
            }
        } else if (info == 0) {
            info = j + 1;
        }
        if (j < n - 1) /* DGER(alpha = -1): A22 += x * (-y') */
            for (int c = j + 1; c < n; ++c) {
                double y = a[j + (size_t)c * n];
                if (y != 0) {
                    double temp = -1.0 * y;
                    for (int i = j + 1; i < n; ++i)
                        a[i + (size_t)c * n] = a[i + (size_t)c * n] + a[i + (size_t)j * n] * temp;
                }
            }
    }
    return info;
}

void oc_dgetrs(int trans, int n, const double *a, const int32_t *piv, double *b) {
    if (!trans) {
        for (int i = 0; i < n; ++i) { /* DLASWP forward */
            int ip = piv[i] - 1;
            if (ip != i) {
                double t = b[i];
                b[i] = b[ip];
                b[ip] = t;
            }
        }
        for (int k = 0; k < n; ++k) /* DTRSM Left Lower NoTrans Unit */
            if (b[k] != 0)
                for (int i = k + 1; i < n; ++i) b[i] = b[i] - b[k] * a[i + (size_t)k * n];
        for (int k = n - 1; k >= 0; --k) /* DTRSM Left Upper NoTrans NonUnit */
            if (b[k] != 0) {
                b[k] = b[k] / a[k + (size_t)k * n];
                for (int i = 0; i < k; ++i) b[i] = b[i] - b[k] * a[i + (size_t)k * n];
            }
    } else {
        for (int i = 0; i < n; ++i) { /* DTRSM Left Upper Trans NonUnit */
            double temp = 1.0 * b[i];
            for (int k = 0; k < i; ++k) temp = temp - a[k + (size_t)i * n] * b[k];
            b[i] = temp / a[i + (size_t)i * n];
        }
        for (int i = n - 1; i >= 0; --i) { /* DTRSM Left Lower Trans Unit */
            double temp = 1.0 * b[i];
            for (int k = i + 1; k < n; ++k) temp = temp - a[k + (size_t)i * n] * b[k];
            b[i] = temp;
        }
        for (int i = n - 1; i >= 0; --i) { /* DLASWP backward */
            int ip = piv[i] - 1;
            if (ip != i) {
                double t = b[i];
                b[i] = b[ip];
                b[ip] = t;
            }
        }
    }
}

/* ---- Aggregator (S/AMG.java:379-629) ------------------------------------- */

/* findNodeNeighborhood (S/AMG.java:481-505): j is strongly coupled to i when
 * |a_ij| >= eps sqrt(a_ii a_jj); members in HashSet order. */
static int neighbourhoods(int n, const int32_t *rp, const int32_t *ci, const double *a,
                          const int32_t *diagind, double eps, jset *N) {
    int32_t maxlen = 0;
    for (int i = 0; i < n; ++i)
        if (rp[i + 1] - rp[i] > maxlen) maxlen = rp[i + 1] - rp[i];
    int32_t *tmp = (int32_t *)malloc(sizeof(int32_t) * (size_t)(maxlen > 0 ? maxlen : 1));
    int rc = OC_OK;
    for (int i = 0; i < n && rc == OC_OK; ++i) {
        double aii = a[diagind[i]];
        int cnt = 0;
        for (int32_t j = rp[i]; j < rp[i + 1]; ++j) {
            double aij = a[j];
            double ajj = a[diagind[ci[j]]];
            if (fabs(aij) >= eps * sqrt(aii * ajj)) tmp[cnt++] = ci[j];
        }
        N[i].cnt = cnt;
        N[i].v = (int32_t *)malloc(sizeof(int32_t) * (size_t)(cnt > 0 ? cnt : 1));
        rc = oc_java_hashset_order(tmp, cnt, N[i].v);
    }
    free(tmp);
    return rc;
}

/* createInitialR, createInitialAggregates, enlargeAggregates,
 * createFinalAggregates (S/AMG.java:510-629); the aggregates are recorded as
 * the aggregate index of each node (-1: none), their count returned. */
static int aggregate(int n, const int32_t *rp, const int32_t *ci, const double *a, const jset *N,
                     int32_t *pt) {
    char *R = (char *)malloc((size_t)(n > 0 ? n : 1));
    for (int i = 0; i < n; ++i) {
        R[i] = 0;
        for (int32_t j = rp[i]; j < rp[i + 1]; ++j)
            if (ci[j] != i && a[j] != 0) {
                R[i] = 1;
                break;
            }
        pt[i] = -1;
    }
    int c = 0;
    for (int i = 0; i < n; ++i) {
        if (!R[i]) continue;
        int fr = 1;
        for (int t = 0; t < N[i].cnt; ++t) fr &= R[N[i].v[t]];
        if (fr) {
            for (int t = 0; t < N[i].cnt; ++t) {
                pt[N[i].v[t]] = c;
                R[N[i].v[t]] = 0;
            }
            c++;
        }
    }
    int32_t *belong = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
    memcpy(belong, pt, sizeof(int32_t) * (size_t)n);
    int *intersect = (int *)malloc(sizeof(int) * (size_t)(c > 0 ? c : 1));
    for (int i = 0; i < n; ++i) {
        if (!R[i]) continue;
        for (int k = 0; k < c; ++k) intersect[k] = 0;
        int largest = 0, maxv = 0;
        for (int t = 0; t < N[i].cnt; ++t) {
            int k = belong[N[i].v[t]];
            if (k < 0) continue;
            intersect[k]++;
            if (intersect[k] > maxv) {
                largest = k;
                maxv = intersect[largest];
            }
        }
        if (maxv > 0) {
            R[i] = 0;
            pt[i] = largest;
        }
    }
    for (int i = 0; i < n; ++i) {
        if (!R[i]) continue;
        int any = 0;
        for (int t = 0; t < N[i].cnt; ++t) {
            int32_t j = N[i].v[t];
            if (R[j]) {
                R[j] = 0;
                pt[j] = c;
                any = 1;
            }
        }
        if (any) c++;
    }
    free(R);
    free(belong);
    free(intersect);
    return c;
}

/* ---- Interpolator (S/AMG.java:635-928) ----------------------------------- */

typedef struct {
    int32_t *cp, *ri; /* compressed columns */
    double *va;
} ccs;

/* Column-sorted CSR (or row-sorted CCS) from triplets emitted with the minor
 * index ascending within each major index. */
static void triplets_to_compressed(int nmajor, int64_t cnt, const int32_t *maj, const int32_t *mnr,
                                   const double *val, int32_t **ptr, int32_t **idx, double **va) {
    *ptr = (int32_t *)calloc((size_t)nmajor + 1, sizeof(int32_t));
    *idx = (int32_t *)malloc(sizeof(int32_t) * (size_t)(cnt > 0 ? cnt : 1));
    *va = (double *)malloc(sizeof(double) * (size_t)(cnt > 0 ? cnt : 1));
    for (int64_t t = 0; t < cnt; ++t) (*ptr)[maj[t] + 1]++;
    for (int i = 0; i < nmajor; ++i) (*ptr)[i + 1] += (*ptr)[i];
    int32_t *pos = (int32_t *)malloc(sizeof(int32_t) * (size_t)(nmajor > 0 ? nmajor : 1));
    memcpy(pos, *ptr, sizeof(int32_t) * (size_t)nmajor);
    for (int64_t t = 0; t < cnt; ++t) {
        int32_t p = pos[maj[t]]++;
        (*idx)[p] = mnr[t];
        (*va)[p] = val[t];
    }
    free(pos);
}

typedef struct {
    int64_t cnt, cap;
    int32_t *maj, *mnr;
    double *val;
} tlist;

static void tl_push(tlist *L, int32_t maj, int32_t mnr, double v) {
    if (L->cnt == L->cap) {
        L->cap = L->cap ? 2 * L->cap : 1024;
        L->maj = (int32_t *)realloc(L->maj, sizeof(int32_t) * (size_t)L->cap);
        L->mnr = (int32_t *)realloc(L->mnr, sizeof(int32_t) * (size_t)L->cap);
        L->val = (double *)realloc(L->val, sizeof(double) * (size_t)L->cap);
    }
    L->maj[L->cnt] = maj;
    L->mnr[L->cnt] = mnr;
    L->val[L->cnt] = v;
    L->cnt++;
}

static void tl_free(tlist *L) {
    free(L->maj);
    free(L->mnr);
    free(L->val);
}

/* createSmoothedProlongation (S/AMG.java:823-885) with its dense row of c
 * entries, then createInterpolationMatrix(P, n) (:751-781): column j holds
 * the rows i with a nonzero entry, ascending. */
static void smoothed_prolongation(int n, const int32_t *rp, const int32_t *ci, const double *a,
                                  const int32_t *diagind, const jset *N, const int32_t *pt, int c,
                                  double omega, ccs *I) {
    double *dot = (double *)malloc(sizeof(double) * (size_t)(c > 0 ? c : 1));
    tlist L = {0, 0, NULL, NULL, NULL};
    for (int i = 0; i < n; ++i) {
        if (pt[i] == -1) continue;
        for (int j = 0; j < c; ++j) dot[j] = 0;
        double weak = 0;
        for (int32_t e = rp[i]; e < rp[i + 1]; ++e) {
            if (pt[ci[e]] == -1) continue;
            double aij = a[e];
            if (aij != 0 && !jset_contains(&N[i], ci[e])) {
                weak += aij;
                continue;
            }
            dot[pt[ci[e]]] += aij;
        }
        dot[pt[i]] -= weak;
        double scale = -omega / a[diagind[i]];
        for (int j = 0; j < c; ++j) dot[j] *= scale;
        dot[pt[i]] += 1;
        for (int j = 0; j < c; ++j)
            if (dot[j] != 0) tl_push(&L, j, i, dot[j]);
    }
    triplets_to_compressed(c, L.cnt, L.maj, L.mnr, L.val, &I->cp, &I->ri, &I->va);
    tl_free(&L);
    free(dot);
}

/* createInterpolationMatrix(pt, c) (S/AMG.java:786-794): ones at (i, pt_i). */
static void tentative_interpolation(int n, const int32_t *pt, int c, ccs *I) {
    tlist L = {0, 0, NULL, NULL, NULL};
    for (int i = 0; i < n; ++i)
        if (pt[i] != -1) tl_push(&L, pt[i], i, 1.0);
    triplets_to_compressed(c, L.cnt, L.maj, L.mnr, L.val, &I->cp, &I->ri, &I->va);
    tl_free(&L);
}

/* createGalerkinSlow (S/AMG.java:891-927): for each coarse column k, A times
 * column k of I by CompRowMatrix.mult, then I' times that by
 * CompColMatrix.transMult (S/CompColMatrix.java:302-320); nonzeros stored. */
static void galerkin_slow(int n, const int32_t *rp, const int32_t *ci, const double *a, const ccs *I,
                          int c, int32_t **arp, int32_t **aci, double **ava) {
    double *iCol = (double *)calloc((size_t)(n > 0 ? n : 1), sizeof(double));
    double *aiCol = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    double *itai = (double *)malloc(sizeof(double) * (size_t)(c > 0 ? c : 1));
    tlist L = {0, 0, NULL, NULL, NULL};
    for (int k = 0; k < c; ++k) {
        for (int i = 0; i < n; ++i) iCol[i] = 0;
        for (int32_t e = I->cp[k]; e < I->cp[k + 1]; ++e) iCol[I->ri[e]] = I->va[e];
        oc_csr_mult(n, rp, ci, a, iCol, aiCol);
        for (int i = 0; i < c; ++i) {
            double dot = 0;
            for (int32_t e = I->cp[i]; e < I->cp[i + 1]; ++e) dot += I->va[e] * aiCol[I->ri[e]];
            itai[i] = dot;
        }
        for (int i = 0; i < c; ++i)
            if (itai[i] != 0) tl_push(&L, i, k, itai[i]);
    }
    triplets_to_compressed(c, L.cnt, L.maj, L.mnr, L.val, arp, aci, ava);
    tl_free(&L);
    free(iCol);
    free(aiCol);
    free(itai);
}

/* createGalerkinFast (S/AMG.java:730-746): FlexCompRowMatrix.add of a_ij at
 * (pt_i, pt_j), i then j ascending; every added position is stored, also when
 * its sum is zero. */
static void galerkin_fast(int n, const int32_t *rp, const int32_t *ci, const double *a,
                          const int32_t *pt, int c, int32_t **arp, int32_t **aci, double **ava) {
    /* fine rows of each aggregate, ascending */
    int32_t *fp = (int32_t *)calloc((size_t)c + 1, sizeof(int32_t));
    for (int i = 0; i < n; ++i)
        if (pt[i] != -1) fp[pt[i] + 1]++;
    for (int p = 0; p < c; ++p) fp[p + 1] += fp[p];
    int32_t *fr = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
    int32_t *pos = (int32_t *)malloc(sizeof(int32_t) * (size_t)(c > 0 ? c : 1));
    memcpy(pos, fp, sizeof(int32_t) * (size_t)c);
    for (int i = 0; i < n; ++i)
        if (pt[i] != -1) fr[pos[pt[i]]++] = i;
    double *acc = (double *)calloc((size_t)(c > 0 ? c : 1), sizeof(double));
    char *used = (char *)calloc((size_t)(c > 0 ? c : 1), 1);
    tlist L = {0, 0, NULL, NULL, NULL};
    for (int p = 0; p < c; ++p) {
        for (int32_t t = fp[p]; t < fp[p + 1]; ++t) {
            int i = fr[t];
            for (int32_t e = rp[i]; e < rp[i + 1]; ++e) {
                int32_t q = pt[ci[e]];
                if (q == -1) continue;
                acc[q] += a[e];
                used[q] = 1;
            }
        }
        for (int q = 0; q < c; ++q)
            if (used[q]) {
                tl_push(&L, p, q, acc[q]);
                acc[q] = 0;
                used[q] = 0;
            }
    }
    triplets_to_compressed(c, L.cnt, L.maj, L.mnr, L.val, arp, aci, ava);
    tl_free(&L);
    free(fp);
    free(fr);
    free(pos);
    free(acc);
    free(used);
}

/* ---- AMG ----------------------------------------------------------------- */

void oc_amg_params(oc_amg *g, double omega_pre_f, double omega_pre_r, double omega_post_f,
                   double omega_post_r, int nu1, int nu2, int gamma, int min_size, double omega,
                   int reverse) {
    memset(g, 0, sizeof(*g));
    g->omega_pre_f = omega_pre_f;
    g->omega_pre_r = omega_pre_r;
    g->omega_post_f = omega_post_f;
    g->omega_post_r = omega_post_r;
    g->nu1 = nu1;
    g->nu2 = nu2;
    g->gamma = gamma;
    g->min_size = min_size;
    g->omega = omega;
    g->reverse = reverse;
}

static void push_level(oc_amg *g, int32_t n, int32_t *rp, int32_t *ci, double *va) {
    g->n = (int32_t *)realloc(g->n, sizeof(int32_t) * (size_t)(g->m + 1));
    g->arp = (int32_t **)realloc(g->arp, sizeof(int32_t *) * (size_t)(g->m + 1));
    g->aci = (int32_t **)realloc(g->aci, sizeof(int32_t *) * (size_t)(g->m + 1));
    g->ava = (double **)realloc(g->ava, sizeof(double *) * (size_t)(g->m + 1));
    g->pt = (int32_t **)realloc(g->pt, sizeof(int32_t *) * (size_t)(g->m + 1));
    g->icp = (int32_t **)realloc(g->icp, sizeof(int32_t *) * (size_t)(g->m + 1));
    g->iri = (int32_t **)realloc(g->iri, sizeof(int32_t *) * (size_t)(g->m + 1));
    g->iva = (double **)realloc(g->iva, sizeof(double *) * (size_t)(g->m + 1));
    g->n[g->m] = n;
    g->arp[g->m] = rp;
    g->aci[g->m] = ci;
    g->ava[g->m] = va;
    g->pt[g->m] = NULL;
    g->icp[g->m] = NULL;
    g->iri[g->m] = NULL;
    g->iva[g->m] = NULL;
    g->m++;
}

/* AMG.setMatrix (S/AMG.java:227-289). */
int oc_amg_setup(oc_amg *g, int n, const int32_t *rp, const int32_t *ci, const double *a,
                 int *bad_row, int *bad_level) {
    int32_t nnz = rp[n];
    int32_t *rp0 = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n + 1));
    int32_t *ci0 = (int32_t *)malloc(sizeof(int32_t) * (size_t)(nnz > 0 ? nnz : 1));
    double *va0 = (double *)malloc(sizeof(double) * (size_t)(nnz > 0 ? nnz : 1));
    memcpy(rp0, rp, sizeof(int32_t) * (size_t)(n + 1));
    memcpy(ci0, ci, sizeof(int32_t) * (size_t)nnz);
    memcpy(va0, a, sizeof(double) * (size_t)nnz);
    push_level(g, n, rp0, ci0, va0);
    for (int k = 0; g->n[k] > g->min_size; ++k) {
        if (k + 1 >= AMG_MAX_LEVELS) return OC_UNSUPPORTED;
        const int nk = g->n[k];
        const int32_t *arp = g->arp[k], *aci = g->aci[k];
        const double *ava = g->ava[k];
        double eps = 0.08 * pow(0.5, k);
        int32_t *diagind = (int32_t *)malloc(sizeof(int32_t) * (size_t)(nk > 0 ? nk : 1));
        for (int i = 0; i < nk; ++i) {
            diagind[i] = bsearch_col(aci, i, arp[i], arp[i + 1]);
            if (diagind[i] < 0) {
                if (bad_row) *bad_row = i;
                if (bad_level) *bad_level = k;
                free(diagind);
                return OC_MISSING_DIAGONAL;
            }
        }
        jset *N = (jset *)calloc((size_t)(nk > 0 ? nk : 1), sizeof(jset));
        int rc = neighbourhoods(nk, arp, aci, ava, diagind, eps, N);
        int32_t *pt = (int32_t *)malloc(sizeof(int32_t) * (size_t)(nk > 0 ? nk : 1));
        int c = rc == OC_OK ? aggregate(nk, arp, aci, ava, N, pt) : 0;
        int stop = rc != OC_OK || c == 0;
        if (!stop) {
            ccs I;
            int32_t *crp, *cci;
            double *cva;
            if (g->omega != 0) {
                smoothed_prolongation(nk, arp, aci, ava, diagind, N, pt, c, g->omega, &I);
                galerkin_slow(nk, arp, aci, ava, &I, c, &crp, &cci, &cva);
            } else {
                galerkin_fast(nk, arp, aci, ava, pt, c, &crp, &cci, &cva);
                tentative_interpolation(nk, pt, c, &I);
            }
            g->pt[k] = pt;
            g->icp[k] = I.cp;
            g->iri[k] = I.ri;
            g->iva[k] = I.va;
            push_level(g, c, crp, cci, cva);
        } else {
            free(pt);
        }
        for (int i = 0; i < nk; ++i) free(N[i].v);
        free(N);
        free(diagind);
        if (rc != OC_OK) return rc;
        if (stop) break;
    }
    /* coarsest level: DenseMatrix copy and DenseLU (S/AMG.java:263-266) */
    const int m = g->m, nc = g->n[m - 1];
    g->lu = (double *)calloc((size_t)nc * (size_t)nc + 1, sizeof(double));
    g->piv = (int32_t *)malloc(sizeof(int32_t) * (size_t)(nc > 0 ? nc : 1));
    for (int i = 0; i < nc; ++i)
        for (int32_t e = g->arp[m - 1][i]; e < g->arp[m - 1][i + 1]; ++e)
            g->lu[i + (size_t)g->aci[m - 1][e] * nc] = g->ava[m - 1][e];
    g->singular = oc_dgetf2(nc, g->lu, g->piv) > 0;
    /* vectors and the SSOR relaxations (S/AMG.java:268-288) */
    g->u = (double **)calloc((size_t)m, sizeof(double *));
    g->f = (double **)calloc((size_t)m, sizeof(double *));
    g->r = (double **)calloc((size_t)m, sizeof(double *));
    for (int k = 0; k < m; ++k) {
        g->u[k] = (double *)calloc((size_t)(g->n[k] > 0 ? g->n[k] : 1), sizeof(double));
        g->f[k] = (double *)calloc((size_t)(g->n[k] > 0 ? g->n[k] : 1), sizeof(double));
        g->r[k] = (double *)calloc((size_t)(g->n[k] > 0 ? g->n[k] : 1), sizeof(double));
    }
    g->ssor = (oc_prec *)calloc((size_t)(m > 1 ? m - 1 : 1), sizeof(oc_prec));
    for (int k = 0; k < m - 1; ++k) {
        int bad = -1;
        int rc = oc_prec_ssor(&g->ssor[k], g->n[k], g->arp[k], g->aci[k], g->ava[k], g->reverse,
                              g->omega_pre_f, g->omega_pre_r, &bad);
        if (rc == OC_OK && (g->omega_post_f < 0 || g->omega_post_f > 2 || g->omega_post_r < 0 ||
                            g->omega_post_r > 2))
            rc = OC_BAD_ARG;
        if (rc != OC_OK) {
            if (bad_row) *bad_row = bad;
            if (bad_level) *bad_level = k;
            return rc;
        }
    }
    return OC_OK;
}

static void relax(oc_amg *g, int k, int times, double wf, double wr) {
    for (int i = 0; i < times; ++i) {
        g->ssor[k].omega_f = wf;
        g->ssor[k].omega_r = wr;
        oc_prec_apply(&g->ssor[k], g->f[k], g->u[k]); /* transApply is apply */
    }
}

/* cycle (S/AMG.java:297-328) and directSolve (:333-342). */
static int cycle(oc_amg *g, int k) {
    if (k == g->m - 1) {
        const int nc = g->n[k];
        oc_dv_set(nc, g->f[k], g->u[k]);
        if (g->singular) return OC_SINGULAR;
        oc_dgetrs(g->transpose, nc, g->lu, g->piv, g->u[k]);
        return OC_OK;
    }
    const int nk = g->n[k], nc = g->n[k + 1];
    relax(g, k, g->nu1, g->omega_pre_f, g->omega_pre_r);
    for (int i = 0; i < nc; ++i) g->u[k + 1][i] = 0;
    oc_dv_set(nk, g->f[k], g->r[k]);
    oc_csr_mult_add(nk, g->arp[k], g->aci[k], g->ava[k], -1, g->u[k], g->r[k]);
    /* I.transMult(r, f_{k+1}) (S/CompColMatrix.java:302-320) */
    for (int c = 0; c < nc; ++c) {
        double dot = 0;
        for (int32_t e = g->icp[k][c]; e < g->icp[k][c + 1]; ++e) dot += g->iva[k][e] * g->r[k][g->iri[k][e]];
        g->f[k + 1][c] = dot;
    }
    for (int i = 0; i < g->gamma; ++i) {
        int rc = cycle(g, k + 1);
        if (rc != OC_OK) return rc;
    }
    /* I.multAdd(1, u_{k+1}, u_k) (S/CompColMatrix.java:280-300) */
    for (int i = 0; i < nk; ++i) g->u[k][i] *= 1.0 / 1.0;
    for (int c = 0; c < nc; ++c)
        for (int32_t e = g->icp[k][c]; e < g->icp[k][c + 1]; ++e)
            g->u[k][g->iri[k][e]] += g->iva[k][e] * g->u[k + 1][c];
    for (int i = 0; i < nk; ++i) g->u[k][i] *= 1.0;
    relax(g, k, g->nu2, g->omega_post_f, g->omega_post_r);
    return OC_OK;
}

/* apply / transApply (S/AMG.java:207-225): u_0 = x, f_0 = b, cycle, x = u_0. */
int oc_amg_apply(oc_amg *g, const double *b, double *x, int transpose) {
    oc_dv_set(g->n[0], x, g->u[0]);
    oc_dv_set(g->n[0], b, g->f[0]);
    g->transpose = transpose;
    int rc = cycle(g, 0);
    g->error = rc;
    if (rc == OC_OK) oc_dv_set(g->n[0], g->u[0], x);
    return rc;
}

void oc_amg_free(oc_amg *g) {
    for (int k = 0; k < g->m; ++k) {
        free(g->arp[k]);
        free(g->aci[k]);
        free(g->ava[k]);
        free(g->pt[k]);
        free(g->icp[k]);
        free(g->iri[k]);
        free(g->iva[k]);
        if (g->u) free(g->u[k]);
        if (g->f) free(g->f[k]);
        if (g->r) free(g->r[k]);
    }
    if (g->ssor)
        for (int k = 0; k < g->m - 1; ++k) oc_prec_free(&g->ssor[k]);
    free(g->ssor);
    free(g->n);
    free(g->arp);
    free(g->aci);
    free(g->ava);
    free(g->pt);
    free(g->icp);
    free(g->iri);
    free(g->iva);
    free(g->u);
    free(g->f);
    free(g->r);
    free(g->lu);
    free(g->piv);
    memset(g, 0, sizeof(*g));
}

void oc_prec_amg(oc_prec *p, oc_amg *g) {
    memset(p, 0, sizeof(*p));
    p->kind = OC_PREC_AMG;
    p->n = g->n[0];
    p->amg = g;
}
