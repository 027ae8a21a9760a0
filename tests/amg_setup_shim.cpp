// Test shim: exposes the library's host AMG setup (csrc/amg_setup.hpp)
// through C entry points so the CPU tests can compare the hierarchy it
// builds with the oracle's, array for array.
#include <cstring>

#include "amg_setup.hpp"

using mtjg::amg::Hierarchy;

extern "C" void *amg_shim_build(int n, const int32_t *rp, const int32_t *ci, const double *va, double wpf,
                                double wpr, double wqf, double wqr, int nu1, int nu2, int gamma, int mn,
                                double omega, int reverse, int *status, int *bad_row, int *bad_level) {
  mtjg::amg::Csr A;
  A.n = n;
  A.rp.assign(rp, rp + n + 1);
  A.ci.assign(ci, ci + rp[n]);
  A.va.assign(va, va + rp[n]);
  mtjg::amg::Params p;
  p.omega_pre_f = wpf;
  p.omega_pre_r = wpr;
  p.omega_post_f = wqf;
  p.omega_post_r = wqr;
  p.nu1 = nu1;
  p.nu2 = nu2;
  p.gamma = gamma;
  p.min = mn;
  p.omega = omega;
  p.reverse = reverse != 0;
  auto *H = new Hierarchy();
  int32_t br = -1, bl = -1;
  *status = mtjg::amg::build(A, p, *H, &br, &bl);
  *bad_row = br;
  *bad_level = bl;
  return H;
}

extern "C" int amg_shim_levels(void *h) { return ((Hierarchy *)h)->levels(); }
extern "C" int amg_shim_n(void *h, int k) { return ((Hierarchy *)h)->A[(size_t)k].n; }
extern "C" int amg_shim_nnz(void *h, int k) { return (int)((Hierarchy *)h)->A[(size_t)k].ci.size(); }
extern "C" int amg_shim_inz(void *h, int k) { return (int)((Hierarchy *)h)->I[(size_t)k].ri.size(); }
extern "C" int amg_shim_singular(void *h) { return ((Hierarchy *)h)->singular; }

extern "C" void amg_shim_level(void *h, int k, int32_t *rp, int32_t *ci, double *va) {
  const auto &A = ((Hierarchy *)h)->A[(size_t)k];
  std::memcpy(rp, A.rp.data(), sizeof(int32_t) * A.rp.size());
  std::memcpy(ci, A.ci.data(), sizeof(int32_t) * A.ci.size());
  std::memcpy(va, A.va.data(), sizeof(double) * A.va.size());
}

extern "C" void amg_shim_interp(void *h, int k, int32_t *cp, int32_t *ri, double *va, int32_t *pt) {
  const auto *H = (Hierarchy *)h;
  const auto &I = H->I[(size_t)k];
  std::memcpy(cp, I.cp.data(), sizeof(int32_t) * I.cp.size());
  std::memcpy(ri, I.ri.data(), sizeof(int32_t) * I.ri.size());
  std::memcpy(va, I.va.data(), sizeof(double) * I.va.size());
  std::memcpy(pt, H->pt[(size_t)k].data(), sizeof(int32_t) * H->pt[(size_t)k].size());
}

extern "C" void amg_shim_lu(void *h, double *lu, int32_t *piv) {
  const auto *H = (Hierarchy *)h;
  std::memcpy(lu, H->lu.data(), sizeof(double) * H->lu.size());
  std::memcpy(piv, H->piv.data(), sizeof(int32_t) * H->piv.size());
}

extern "C" void amg_shim_free(void *h) { delete (Hierarchy *)h; }
