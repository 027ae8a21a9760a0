// Host-side setup of MTJ's algebraic multigrid preconditioner
// (S/AMG.java:227-289, smoothed aggregation after Vanek, Mandel and Brezina):
// the aggregates, the interpolation operators, the Galerkin operators and the
// LU factors of the coarsest matrix, bit for bit as MTJ forms them.  Plain
// C++ (no HIP types): the device library uploads the result and runs the
// cycles; the CPU tests compile this header directly.
//
// MTJ's setup is sequential and partly quadratic (a dense row of c entries
// per fine node for the prolongation, one dense product A * I(:, k) per
// coarse node for the Galerkin operator).  Here the same scalar operations
// run in the same order on sparse work sets: every term MTJ adds that is
// omitted here is a product with an exact zero, which cannot change a sum
// that starts at +0.0 when all entries are finite.  When a level holds a
// non-finite entry (0 * Inf is NaN) the Galerkin product falls back to MTJ's
// dense loops, and a prolongation row whose damping factor is not finite
// (a zero diagonal) is formed densely as MTJ does.
//
// java.util.HashSet iteration order decides ties in the aggregate
// enlargement (S/AMG.java:560-580).  OpenJDK's HashMap puts an Integer v in
// bucket (h ^ (h >>> 16)) & (cap - 1), h = v; the table starts with 16
// buckets on the first insertion and doubles when the size exceeds 3/4 of the
// buckets or when an insertion finds 8 entries in its bucket while the table
// has fewer than 64 buckets; inside a bucket the insertion order survives
// resizing, and iteration walks the buckets in index order.  A bucket that
// reaches 9 entries in a table of 64 or more buckets becomes a tree bin with
// a different order: setup reports it as unsupported.
#pragma once
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <utility>
#include <vector>

namespace mtjg {
namespace amg {

struct Params {
  double omega_pre_f = 1, omega_pre_r = 1.85, omega_post_f = 1.85, omega_post_r = 1;
  int32_t nu1 = 1, nu2 = 1, gamma = 1, min = 40;
  double omega = 2.0 / 3;
  bool reverse = true;
};

struct Csr {  // square, columns ascending in each row
  int32_t n = 0;
  std::vector<int32_t> rp{0}, ci;
  std::vector<double> va;
};

struct Ccs {  // n_rows x n_cols, rows ascending in each column
  int32_t nr = 0, nc = 0;
  std::vector<int32_t> cp{0}, ri;
  std::vector<double> va;
};

struct Hierarchy {
  std::vector<Csr> A;                  // A[0..m-1]; A[m-1] is the coarsest
  std::vector<Ccs> I;                  // I[k]: n_k x n_{k+1}
  std::vector<std::vector<int32_t>> pt;  // aggregate of each node of level k (-1: none)
  std::vector<double> lu;              // coarsest matrix, column-major LU
  std::vector<int32_t> piv;            // 1-based LAPACK pivots
  bool singular = false;
  int levels() const { return (int)A.size(); }
};

enum Status : int { OK = 0, MISSING_DIAGONAL = 1, UNSUPPORTED = 2, BAD_RELAXATION = 3 };

constexpr int MAX_LEVELS = 100;
// Largest coarsest level the device LU solve holds (its right-hand side
// lives in 32 KB of LDS).
constexpr int COARSE_MAX = 4096;

inline uint32_t java_hash(int32_t v) {
  uint32_t h = (uint32_t)v;
  return h ^ (h >> 16);
}

// Iteration order of a java.util.HashSet<Integer> filled with vals[0..cnt)
// in that order.  Returns false when a tree bin would form.
inline bool java_hashset_order(const int32_t *vals, int cnt, int32_t *order, std::vector<int32_t> &counts) {
  uint32_t cap = 16;
  counts.assign(cap, 0);
  int size = 0;
  for (int t = 0; t < cnt; ++t) {
    if (counts[java_hash(vals[t]) & (cap - 1)] >= 8) {  // HashMap.putVal -> treeifyBin
      if (cap >= 64) return false;
      cap *= 2;
      counts.assign(cap, 0);
      for (int s = 0; s < t; ++s) counts[java_hash(vals[s]) & (cap - 1)]++;
    }
    counts[java_hash(vals[t]) & (cap - 1)]++;
    if (++size > (int)(cap / 4 * 3)) {
      cap *= 2;
      counts.assign(cap, 0);
      for (int s = 0; s <= t; ++s) counts[java_hash(vals[s]) & (cap - 1)]++;
    }
  }
  // stable counting sort by bucket
  std::vector<int32_t> start(cap + 1, 0);
  for (int t = 0; t < cnt; ++t) start[(java_hash(vals[t]) & (cap - 1)) + 1]++;
  for (uint32_t b = 0; b < cap; ++b) start[b + 1] += start[b];
  for (int t = 0; t < cnt; ++t) order[start[java_hash(vals[t]) & (cap - 1)]++] = vals[t];
  return true;
}

// Reference LAPACK dgetf2 (IDAMAX pivoting, row swaps, reciprocal scaling
// when |pivot| >= the safe minimum, DGER update), n x n column-major;
// returns info.
inline int dgetf2(int n, double *a, int32_t *piv) {
  int info = 0;
  const double sfmin = 2.2250738585072014e-308;
  auto at = [&](int i, int j) -> double & { return a[(size_t)i + (size_t)j * (size_t)n]; };
  for (int j = 0; j < n; ++j) {
    int jp = j;
    double dmax = std::fabs(at(j, j));
    for (int i = j + 1; i < n; ++i)
      if (std::fabs(at(i, j)) > dmax) {
        jp = i;
        dmax = std::fabs(at(i, j));
      }
    piv[j] = jp + 1;
    if (at(jp, j) != 0) {
      if (jp != j)
        for (int c = 0; c < n; ++c) std::swap(at(j, c), at(jp, c));
      if (j < n - 1) {
        if (std::fabs(at(j, j)) >= sfmin) {
          const double r = 1.0 / at(j, j);
          for (int i = j + 1; i < n; ++i) at(i, j) = r * at(i, j);
        } else {
          for (int i = j + 1; i < n; ++i) at(i, j) = at(i, j) / at(j, j);
        }
      }
    } else if (info == 0) {
      info = j + 1;
    }
    if (j < n - 1)
      for (int c = j + 1; c < n; ++c) {
        const double y = at(j, c);
        if (y != 0) {
          const double temp = -1.0 * y;
          for (int i = j + 1; i < n; ++i) at(i, c) = at(i, c) + at(i, j) * temp;
        }
      }
  }
  return info;
}

namespace detail {

inline int32_t find_col(const Csr &A, int32_t row, int32_t col) {
  int32_t lo = A.rp[(size_t)row], hi = A.rp[(size_t)row + 1] - 1;
  while (lo <= hi) {
    int32_t mid = (lo + hi) >> 1;
    int32_t c = A.ci[(size_t)mid];
    if (c < col) lo = mid + 1;
    else if (c > col) hi = mid - 1;
    else return mid;
  }
  return -1;
}

// Strongly coupled neighbourhoods (S/AMG.java:481-505) in HashSet order,
// stored as one CSR-like list.
inline bool neighbourhoods(const Csr &A, const std::vector<int32_t> &diag, double eps, std::vector<int32_t> &np,
                           std::vector<int32_t> &nv) {
  const int32_t n = A.n;
  np.assign((size_t)n + 1, 0);
  nv.clear();
  nv.reserve(A.ci.size());
  std::vector<int32_t> tmp, counts;
  for (int32_t i = 0; i < n; ++i) {
    const double aii = A.va[(size_t)diag[(size_t)i]];
    tmp.clear();
    for (int32_t j = A.rp[(size_t)i]; j < A.rp[(size_t)i + 1]; ++j) {
      const double aij = A.va[(size_t)j];
      const double ajj = A.va[(size_t)diag[(size_t)A.ci[(size_t)j]]];
      if (std::fabs(aij) >= eps * std::sqrt(aii * ajj)) tmp.push_back(A.ci[(size_t)j]);
    }
    const size_t base = nv.size();
    nv.resize(base + tmp.size());
    if (!java_hashset_order(tmp.data(), (int)tmp.size(), nv.data() + base, counts)) return false;
    np[(size_t)i + 1] = (int32_t)nv.size();
  }
  return true;
}

// createInitialR / createInitialAggregates / enlargeAggregates /
// createFinalAggregates (S/AMG.java:510-629); returns the aggregate count.
inline int32_t aggregate(const Csr &A, const std::vector<int32_t> &np, const std::vector<int32_t> &nv,
                         std::vector<int32_t> &pt) {
  const int32_t n = A.n;
  std::vector<uint8_t> R((size_t)n, 0);
  pt.assign((size_t)n, -1);
  for (int32_t i = 0; i < n; ++i)
    for (int32_t j = A.rp[(size_t)i]; j < A.rp[(size_t)i + 1]; ++j)
      if (A.ci[(size_t)j] != i && A.va[(size_t)j] != 0) {
        R[(size_t)i] = 1;
        break;
      }
  int32_t c = 0;
  for (int32_t i = 0; i < n; ++i) {
    if (!R[(size_t)i]) continue;
    bool fr = true;
    for (int32_t t = np[(size_t)i]; t < np[(size_t)i + 1]; ++t) fr = fr && R[(size_t)nv[(size_t)t]];
    if (!fr) continue;
    for (int32_t t = np[(size_t)i]; t < np[(size_t)i + 1]; ++t) {
      pt[(size_t)nv[(size_t)t]] = c;
      R[(size_t)nv[(size_t)t]] = 0;
    }
    c++;
  }
  const std::vector<int32_t> belong(pt);
  std::vector<int32_t> intersect((size_t)c, 0), touched;
  for (int32_t i = 0; i < n; ++i) {
    if (!R[(size_t)i]) continue;
    int32_t largest = 0, maxv = 0;
    touched.clear();
    for (int32_t t = np[(size_t)i]; t < np[(size_t)i + 1]; ++t) {
      const int32_t k = belong[(size_t)nv[(size_t)t]];
      if (k < 0) continue;
      if (intersect[(size_t)k]++ == 0) touched.push_back(k);
      if (intersect[(size_t)k] > maxv) {
        largest = k;
        maxv = intersect[(size_t)k];
      }
    }
    for (int32_t k : touched) intersect[(size_t)k] = 0;
    if (maxv > 0) {
      R[(size_t)i] = 0;
      pt[(size_t)i] = largest;
    }
  }
  for (int32_t i = 0; i < n; ++i) {
    if (!R[(size_t)i]) continue;
    bool any = false;
    for (int32_t t = np[(size_t)i]; t < np[(size_t)i + 1]; ++t) {
      const int32_t j = nv[(size_t)t];
      if (R[(size_t)j]) {
        R[(size_t)j] = 0;
        pt[(size_t)j] = c;
        any = true;
      }
    }
    if (any) c++;
  }
  return c;
}

// Compressed storage from (major, minor, value) triplets emitted with the
// minor index ascending inside each major index.
struct Triplets {
  std::vector<int32_t> maj, mnr;
  std::vector<double> val;
  void push(int32_t a, int32_t b, double v) {
    maj.push_back(a);
    mnr.push_back(b);
    val.push_back(v);
  }
  void compress(int32_t nmajor, std::vector<int32_t> &ptr, std::vector<int32_t> &idx, std::vector<double> &va) const {
    ptr.assign((size_t)nmajor + 1, 0);
    for (int32_t a : maj) ptr[(size_t)a + 1]++;
    for (int32_t i = 0; i < nmajor; ++i) ptr[(size_t)i + 1] += ptr[(size_t)i];
    std::vector<int32_t> pos(ptr.begin(), ptr.end() - 1);
    idx.resize(maj.size());
    va.resize(maj.size());
    for (size_t t = 0; t < maj.size(); ++t) {
      const int32_t p = pos[(size_t)maj[t]]++;
      idx[(size_t)p] = mnr[t];
      va[(size_t)p] = val[t];
    }
  }
};

// createSmoothedProlongation + createInterpolationMatrix(P, n)
// (S/AMG.java:751-781, :823-885).
inline void smoothed_prolongation(const Csr &A, const std::vector<int32_t> &diag, const std::vector<int32_t> &np,
                                  const std::vector<int32_t> &nv, const std::vector<int32_t> &pt, int32_t c,
                                  double omega, Ccs &I) {
  const int32_t n = A.n;
  std::vector<double> dot((size_t)c, 0.0);
  std::vector<uint8_t> used((size_t)c, 0), strong((size_t)n, 0);
  std::vector<int32_t> touched;
  Triplets T;
  for (int32_t i = 0; i < n; ++i) {
    const int32_t pi = pt[(size_t)i];
    if (pi == -1) continue;
    for (int32_t t = np[(size_t)i]; t < np[(size_t)i + 1]; ++t) strong[(size_t)nv[(size_t)t]] = 1;
    touched.clear();
    auto touch = [&](int32_t q) {
      if (!used[(size_t)q]) {
        used[(size_t)q] = 1;
        dot[(size_t)q] = 0.0;
        touched.push_back(q);
      }
    };
    double weak = 0;
    for (int32_t e = A.rp[(size_t)i]; e < A.rp[(size_t)i + 1]; ++e) {
      const int32_t col = A.ci[(size_t)e];
      if (pt[(size_t)col] == -1) continue;
      const double aij = A.va[(size_t)e];
      if (aij != 0 && !strong[(size_t)col]) {
        weak += aij;
        continue;
      }
      touch(pt[(size_t)col]);
      dot[(size_t)pt[(size_t)col]] += aij;
    }
    touch(pi);
    dot[(size_t)pi] -= weak;
    const double scale = -omega / A.va[(size_t)diag[(size_t)i]];
    if (std::isfinite(scale)) {
      for (int32_t q : touched) dot[(size_t)q] *= scale;
      dot[(size_t)pi] += 1;
      for (int32_t q : touched)
        if (dot[(size_t)q] != 0) T.push(q, i, dot[(size_t)q]);
    } else {
      // every untouched entry is 0 * scale = NaN, and stored, as in MTJ
      for (int32_t q = 0; q < c; ++q) {
        double v = used[(size_t)q] ? dot[(size_t)q] : 0.0;
        v *= scale;
        if (q == pi) v += 1;
        if (v != 0) T.push(q, i, v);
      }
    }
    for (int32_t q : touched) used[(size_t)q] = 0;
    for (int32_t t = np[(size_t)i]; t < np[(size_t)i + 1]; ++t) strong[(size_t)nv[(size_t)t]] = 0;
  }
  I.nr = n;
  I.nc = c;
  T.compress(c, I.cp, I.ri, I.va);
}

// createInterpolationMatrix(pt, c) (S/AMG.java:786-794).
inline void tentative_interpolation(int32_t n, const std::vector<int32_t> &pt, int32_t c, Ccs &I) {
  Triplets T;
  for (int32_t i = 0; i < n; ++i)
    if (pt[(size_t)i] != -1) T.push(pt[(size_t)i], i, 1.0);
  I.nr = n;
  I.nc = c;
  T.compress(c, I.cp, I.ri, I.va);
}

inline bool all_finite(const std::vector<double> &v) {
  for (double x : v)
    if (!std::isfinite(x)) return false;
  return true;
}

// createGalerkinSlow (S/AMG.java:891-927): column k of Ac is I' (A I(:, k)),
// A I(:, k) by CompRowMatrix.mult (stored where nonzero) and I' times it by
// CompColMatrix.transMult; nonzeros stored.
inline void galerkin_slow(const Csr &A, const Ccs &I, Csr &Ac) {
  const int32_t n = A.n, c = I.nc;
  Triplets T;
  if (all_finite(A.va) && all_finite(I.va)) {
    // columns of A, and rows of I (fine row -> coarse columns)
    std::vector<int32_t> acp((size_t)n + 1, 0), ari(A.ci.size());
    for (int32_t j : A.ci) acp[(size_t)j + 1]++;
    for (int32_t j = 0; j < n; ++j) acp[(size_t)j + 1] += acp[(size_t)j];
    {
      std::vector<int32_t> pos(acp.begin(), acp.end() - 1);
      for (int32_t r = 0; r < n; ++r)
        for (int32_t e = A.rp[(size_t)r]; e < A.rp[(size_t)r + 1]; ++e) ari[(size_t)pos[(size_t)A.ci[(size_t)e]]++] = r;
    }
    std::vector<int32_t> irp((size_t)n + 1, 0), ici(I.ri.size());
    for (int32_t r : I.ri) irp[(size_t)r + 1]++;
    for (int32_t r = 0; r < n; ++r) irp[(size_t)r + 1] += irp[(size_t)r];
    {
      std::vector<int32_t> pos(irp.begin(), irp.end() - 1);
      for (int32_t q = 0; q < c; ++q)
        for (int32_t e = I.cp[(size_t)q]; e < I.cp[(size_t)q + 1]; ++e) ici[(size_t)pos[(size_t)I.ri[(size_t)e]]++] = q;
    }
    std::vector<double> iv((size_t)n, 0.0), aiv((size_t)n, 0.0);
    std::vector<int32_t> in_col((size_t)n, -1), row_seen((size_t)n, -1), in_ai((size_t)n, -1), coarse_seen((size_t)c, -1);
    std::vector<int32_t> rows, cand;
    for (int32_t k = 0; k < c; ++k) {
      for (int32_t e = I.cp[(size_t)k]; e < I.cp[(size_t)k + 1]; ++e) {
        iv[(size_t)I.ri[(size_t)e]] = I.va[(size_t)e];
        in_col[(size_t)I.ri[(size_t)e]] = k;
      }
      rows.clear();
      for (int32_t e = I.cp[(size_t)k]; e < I.cp[(size_t)k + 1]; ++e) {
        const int32_t j = I.ri[(size_t)e];
        for (int32_t t = acp[(size_t)j]; t < acp[(size_t)j + 1]; ++t) {
          const int32_t r = ari[(size_t)t];
          if (row_seen[(size_t)r] != k) {
            row_seen[(size_t)r] = k;
            rows.push_back(r);
          }
        }
      }
      cand.clear();
      for (int32_t r : rows) {
        double dot = 0;
        for (int32_t e = A.rp[(size_t)r]; e < A.rp[(size_t)r + 1]; ++e) {
          const int32_t j = A.ci[(size_t)e];
          if (in_col[(size_t)j] == k) dot += A.va[(size_t)e] * iv[(size_t)j];
        }
        if (dot != 0) {
          aiv[(size_t)r] = dot;
          in_ai[(size_t)r] = k;
          for (int32_t t = irp[(size_t)r]; t < irp[(size_t)r + 1]; ++t) {
            const int32_t q = ici[(size_t)t];
            if (coarse_seen[(size_t)q] != k) {
              coarse_seen[(size_t)q] = k;
              cand.push_back(q);
            }
          }
        }
      }
      std::sort(cand.begin(), cand.end());
      for (int32_t q : cand) {
        double dot = 0;
        for (int32_t e = I.cp[(size_t)q]; e < I.cp[(size_t)q + 1]; ++e) {
          const int32_t r = I.ri[(size_t)e];
          if (in_ai[(size_t)r] == k) dot += I.va[(size_t)e] * aiv[(size_t)r];
        }
        if (dot != 0) T.push(q, k, dot);
      }
    }
  } else {
    std::vector<double> icol((size_t)n), aicol((size_t)n);
    for (int32_t k = 0; k < c; ++k) {
      std::fill(icol.begin(), icol.end(), 0.0);
      for (int32_t e = I.cp[(size_t)k]; e < I.cp[(size_t)k + 1]; ++e) icol[(size_t)I.ri[(size_t)e]] = I.va[(size_t)e];
      for (int32_t r = 0; r < n; ++r) {  // CompRowMatrix.mult
        double dot = 0;
        for (int32_t e = A.rp[(size_t)r]; e < A.rp[(size_t)r + 1]; ++e) dot += A.va[(size_t)e] * icol[(size_t)A.ci[(size_t)e]];
        aicol[(size_t)r] = dot != 0 ? dot : 0.0;
      }
      for (int32_t q = 0; q < c; ++q) {  // CompColMatrix.transMult
        double dot = 0;
        for (int32_t e = I.cp[(size_t)q]; e < I.cp[(size_t)q + 1]; ++e) dot += I.va[(size_t)e] * aicol[(size_t)I.ri[(size_t)e]];
        if (dot != 0) T.push(q, k, dot);
      }
    }
  }
  Ac.n = c;
  T.compress(c, Ac.rp, Ac.ci, Ac.va);
}

// createGalerkinFast (S/AMG.java:730-746): a_ij added at (pt_i, pt_j), i then
// j ascending; every position added is stored, also when its sum is zero.
inline void galerkin_fast(const Csr &A, const std::vector<int32_t> &pt, int32_t c, Csr &Ac) {
  const int32_t n = A.n;
  std::vector<int32_t> fp((size_t)c + 1, 0), fr((size_t)n);
  for (int32_t i = 0; i < n; ++i)
    if (pt[(size_t)i] != -1) fp[(size_t)pt[(size_t)i] + 1]++;
  for (int32_t p = 0; p < c; ++p) fp[(size_t)p + 1] += fp[(size_t)p];
  {
    std::vector<int32_t> pos(fp.begin(), fp.end() - 1);
    for (int32_t i = 0; i < n; ++i)
      if (pt[(size_t)i] != -1) fr[(size_t)pos[(size_t)pt[(size_t)i]]++] = i;
  }
  std::vector<double> acc((size_t)c, 0.0);
  std::vector<uint8_t> used((size_t)c, 0);
  std::vector<int32_t> touched;
  Triplets T;
  for (int32_t p = 0; p < c; ++p) {
    touched.clear();
    for (int32_t t = fp[(size_t)p]; t < fp[(size_t)p + 1]; ++t) {
      const int32_t i = fr[(size_t)t];
      for (int32_t e = A.rp[(size_t)i]; e < A.rp[(size_t)i + 1]; ++e) {
        const int32_t q = pt[(size_t)A.ci[(size_t)e]];
        if (q == -1) continue;
        if (!used[(size_t)q]) {
          used[(size_t)q] = 1;
          acc[(size_t)q] = 0.0;
          touched.push_back(q);
        }
        acc[(size_t)q] += A.va[(size_t)e];
      }
    }
    std::sort(touched.begin(), touched.end());
    for (int32_t q : touched) {
      T.push(p, q, acc[(size_t)q]);
      used[(size_t)q] = 0;
    }
  }
  Ac.n = c;
  T.compress(c, Ac.rp, Ac.ci, Ac.va);
}

}  // namespace detail

// AMG.setMatrix (S/AMG.java:227-289) for a square CSR matrix.  On
// MISSING_DIAGONAL, *bad_row / *bad_level name the row and level; on
// BAD_RELAXATION a relaxation lies outside [0, 2] (checked, as MTJ's SSOR
// constructor does, only when a level gets smoothers).
inline Status build(const Csr &A0, const Params &p, Hierarchy &H, int32_t *bad_row, int32_t *bad_level) {
  H = Hierarchy();
  H.A.push_back(A0);
  for (int k = 0; H.A[(size_t)k].n > p.min; ++k) {
    if (k + 1 >= MAX_LEVELS) return UNSUPPORTED;
    const Csr &A = H.A[(size_t)k];
    const double eps = 0.08 * std::pow(0.5, k);
    std::vector<int32_t> diag((size_t)A.n);
    for (int32_t i = 0; i < A.n; ++i) {
      diag[(size_t)i] = detail::find_col(A, i, i);
      if (diag[(size_t)i] < 0) {
        *bad_row = i;
        *bad_level = k;
        return MISSING_DIAGONAL;
      }
    }
    std::vector<int32_t> np, nv, pt;
    if (!detail::neighbourhoods(A, diag, eps, np, nv)) return UNSUPPORTED;
    const int32_t c = detail::aggregate(A, np, nv, pt);
    if (c == 0) break;
    Ccs I;
    Csr Ac;
    if (p.omega != 0) {
      detail::smoothed_prolongation(A, diag, np, nv, pt, c, p.omega, I);
      detail::galerkin_slow(A, I, Ac);
    } else {
      detail::galerkin_fast(A, pt, c, Ac);
      detail::tentative_interpolation(A.n, pt, c, I);
    }
    H.pt.push_back(std::move(pt));
    H.I.push_back(std::move(I));
    H.A.push_back(std::move(Ac));
  }
  const Csr &C = H.A.back();
  const size_t nc = (size_t)C.n;
  H.lu.assign(nc * nc, 0.0);
  for (int32_t i = 0; i < C.n; ++i)
    for (int32_t e = C.rp[(size_t)i]; e < C.rp[(size_t)i + 1]; ++e)
      H.lu[(size_t)i + (size_t)C.ci[(size_t)e] * nc] = C.va[(size_t)e];
  H.piv.assign(nc, 0);
  H.singular = dgetf2(C.n, H.lu.data(), H.piv.data()) > 0;
  if (H.levels() > 1) {
    auto bad = [](double w) { return w < 0 || w > 2; };
    if (bad(p.omega_pre_f) || bad(p.omega_pre_r) || bad(p.omega_post_f) || bad(p.omega_post_r)) {
      *bad_level = 0;
      return BAD_RELAXATION;
    }
  }
  return OK;
}

}  // namespace amg
}  // namespace mtjg
