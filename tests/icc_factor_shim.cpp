// Test shim: exposes the library's host ICC factorization (csrc/icc_factor.hpp)
// through a C entry point so the CPU tests can compare it with the oracle.
#include <cstring>

#include "icc_factor.hpp"

extern "C" int icc_factor_shim(int n, const int32_t *rp, const int32_t *ci, const int32_t *diag, double *val,
                               int32_t *bad) {
  std::vector<int32_t> r(rp, rp + n + 1), c(ci, ci + rp[n]), d(diag, diag + n);
  std::vector<double> v(val, val + rp[n]);
  int rc = mtjg::icc_factor_host(n, r, c, d, v, bad);
  std::memcpy(val, v.data(), sizeof(double) * v.size());
  return rc;
}
