// Host-side ICC factorization (S/ICC.java:90-136), shared by the device
// library's ICC.setMatrix.  Plain C++: no HIP types.
#pragma once
#include <cmath>
#include <cstdint>
#include <vector>

namespace mtjg {

// ICC.factor (S/ICC.java:90-136) with MTJ's arithmetic in MTJ's order.  MTJ
// scans every i < k; only the rows whose term can be nonzero are visited
// here: those reached from row k's strictly-lower entries through the stored
// strictly-upper entries of earlier rows (elsewhere Rk[i] is exactly zero,
// and 0 / R_ii is zero unless R_ii is NaN), plus every earlier row whose
// pivot is NaN.  Pending rows are bits in a bitmap scanned upward from the
// lowest one; a row only ever adds rows above itself, so they are taken in
// ascending order and each Rk entry receives its subtractions in MTJ's
// order.  Entries right of the diagonal that row k does not store are never
// read, so they are not formed.  Returns 0, 1 (zero pivot R_ii, row *bad) or
// 2 (Rk[k] == 0 before the square root, row *bad).
inline int icc_factor_host(int32_t n, const std::vector<int32_t> &rp, const std::vector<int32_t> &ci,
                           const std::vector<int32_t> &diag, std::vector<double> &data, int32_t *bad) {
  std::vector<double> Rk((size_t)n, 0.0);
  std::vector<uint8_t> seen((size_t)n, 0);
  std::vector<uint64_t> pending((size_t)n / 64 + 1, 0);
  std::vector<int32_t> touched, nan_rows;
  for (int32_t k = 0; k < n; ++k) {
    touched.clear();
    int32_t lo = k;
    auto queue_row = [&](int32_t c) {
      pending[(size_t)c >> 6] |= 1ull << (c & 63);
      lo = c < lo ? c : lo;
    };
    for (int32_t e = rp[k]; e < rp[k + 1]; ++e) {
      const int32_t c = ci[(size_t)e];
      Rk[(size_t)c] = data[(size_t)e];
      seen[(size_t)c] = 1;
      touched.push_back(c);
      if (c < k) queue_row(c);
    }
    for (int32_t r : nan_rows)
      if (!seen[(size_t)r]) {
        seen[(size_t)r] = 1;
        touched.push_back(r);
        queue_row(r);
      }
    for (int32_t w = lo >> 6; k > 0 && w <= (k - 1) >> 6;) {
      const uint64_t word = pending[(size_t)w];
      if (!word) {
        ++w;
        continue;
      }
      const int b = __builtin_ctzll(word);
      pending[(size_t)w] = word & (word - 1);
      const int32_t i = (w << 6) + b;
      const double Rii = data[(size_t)diag[(size_t)i]];
      if (Rii == 0) {
        *bad = i;
        return 1;
      }
      const double Rki = Rk[(size_t)i] / Rii;
      if (Rki == 0) continue;
      for (int32_t j = diag[(size_t)i] + 1; j < rp[(size_t)i + 1]; ++j) {
        const int32_t c = ci[(size_t)j];
        if (!seen[(size_t)c]) {
          if (c > k) continue;  // fill right of row k's pattern: never read
          seen[(size_t)c] = 1;  // fill left of the diagonal: an elimination row
          touched.push_back(c);
          queue_row(c);
        }
        Rk[(size_t)c] -= Rki * data[(size_t)j];
      }
    }
    if (Rk[(size_t)k] == 0) {
      *bad = k;
      return 2;
    }
    const double sqRkk = std::sqrt(Rk[(size_t)k]);
    for (int32_t e = diag[(size_t)k]; e < rp[(size_t)k + 1]; ++e)
      data[(size_t)e] = Rk[(size_t)ci[(size_t)e]] / sqRkk;
    if (std::isnan(data[(size_t)diag[(size_t)k]])) nan_rows.push_back(k);
    for (int32_t c : touched) {
      Rk[(size_t)c] = 0;
      seen[(size_t)c] = 0;
    }
  }
  return 0;
}

}  // namespace mtjg
