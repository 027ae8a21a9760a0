// Sequential sparse row sums with their loads issued in batches: the
// triangular solves, SSOR sweeps and multigrid transfers must add a row's
// products one by one in storage order (MTJ's loops), but nothing stops the
// loads: a batch of ROWDOT_K column indices and values is loaded, then the
// x values they address, then the products are added in order.  The result is
// bitwise that of the plain loop; a row costs about two memory latencies per
// batch instead of two per entry.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace mtjg {

constexpr int ROWDOT_K = 8;

// s (+ or -)= a[j] * x[ci[j]] for the cnt entries j = j0, j0 + 1, ...
// (DESC: j0, j0 - 1, ...), in that order.
template <bool SUB, bool DESC = false>
__device__ __forceinline__ double rowdot(double s, int32_t j0, int32_t cnt, const int32_t *__restrict__ ci,
                                         const double *__restrict__ a, const double *x) {
  for (int32_t u0 = 0; u0 < cnt; u0 += ROWDOT_K) {
    int32_t c[ROWDOT_K];
    double v[ROWDOT_K], xv[ROWDOT_K];
#pragma unroll
    for (int u = 0; u < ROWDOT_K; ++u)
      if (u0 + u < cnt) {
        const int32_t j = DESC ? j0 - (u0 + u) : j0 + (u0 + u);
        c[u] = ci[j];
        v[u] = a[j];
      }
#pragma unroll
    for (int u = 0; u < ROWDOT_K; ++u)
      if (u0 + u < cnt) xv[u] = x[c[u]];
#pragma unroll
    for (int u = 0; u < ROWDOT_K; ++u)
      if (u0 + u < cnt) {
        if (SUB) s -= v[u] * xv[u];
        else s += v[u] * xv[u];
      }
  }
  return s;
}

}  // namespace mtjg
