// Algebraic multigrid preconditioner on the device (S/AMG.java).
//
// setMatrix builds the hierarchy on the host (amg_setup.hpp: aggregation,
// smoothed interpolation, Galerkin operators and the LU factors of the
// coarsest matrix, bit for bit as MTJ forms them) and uploads it once: per
// level the matrix as a device CompRowMatrix (its residual uses the library's
// SELL-64 product when the level has that copy, else a one-thread-per-row
// sequential kernel; both sum each row in MTJ's order), an SSOR relaxation in
// the level's pattern, the
// interpolation I both as CSR of I (prolongation, one row per fine node) and
// as CSR of I' (restriction, one row per coarse node, which is MTJ's
// compressed-column storage of I), and the vectors u, f, r.  The coarsest
// level keeps its column-major LU factors and pivots.
//
// apply / transApply run MTJ's recursive cycle (S/AMG.java:297-342) from the
// host as stream-ordered launches, in MTJ's operation order:
//   pre-smoothing (SSOR sweeps, nu1 times), u_{k+1} = 0,
//   r_k = f_k - A_k u_k (CompRowMatrix.multAdd),
//   f_{k+1} = I' r_k (CompColMatrix.transMult: one sequential dot per coarse
//   row), gamma recursive cycles,
//   u_k += I u_{k+1} (CompColMatrix.multAdd: the scatter of column j adds
//   into row i in ascending j, restated as one sequential sum per fine row),
//   post-smoothing (nu2 times),
// and at the coarsest level LAPACK dgetrs (row interchanges plus the
// reference-BLAS DTRSM loops) in one workgroup with the right-hand side in
// LDS.  transApply differs only in solving with the transposed LU (SSOR's
// transApply is its apply).
#include <algorithm>
#include <initializer_list>

#include "amg_setup.hpp"
#include "mtjg_internal.hpp"
#include "rowdot.hpp"

namespace mtjg {

struct AmgLevel {
  int32_t n = 0;
  mtjg_csr A = nullptr;
  mtjg_prec ssor = nullptr;
  int32_t *r_ptr = nullptr, *r_idx = nullptr;  // I' (restriction)
  double *r_val = nullptr;
  int32_t *p_ptr = nullptr, *p_idx = nullptr;  // I (prolongation)
  double *p_val = nullptr;
  double *u = nullptr, *f = nullptr, *r = nullptr;
};

struct AmgDevice {
  mtjg_ctx ctx = nullptr;
  amg::Params p;
  std::vector<AmgLevel> L;
  int32_t nc = 0;
  double *lu = nullptr;
  int32_t *piv = nullptr;
  bool singular = false;
};

// f_c[i] = sum over the entries of coarse row i of I' (ascending fine row)
// of I * r (S/CompColMatrix.java:302-320).
__global__ __launch_bounds__(BS) void k_amg_restrict(int32_t nc, const int32_t *__restrict__ ptr,
                                                     const int32_t *__restrict__ idx,
                                                     const double *__restrict__ val,
                                                     const double *__restrict__ r, double *__restrict__ fc) {
  const int32_t i = blockIdx.x * BS + threadIdx.x;
  if (i >= nc) return;
  fc[i] = rowdot<false>(0.0, ptr[i], ptr[i + 1] - ptr[i], idx, val, r);
}

// u[i] += I(i, j) * u_c[j] for the entries of fine row i in ascending j
// (S/CompColMatrix.java:280-300 with alpha = 1).
__global__ __launch_bounds__(BS) void k_amg_prolong(int32_t nf, const int32_t *__restrict__ ptr,
                                                    const int32_t *__restrict__ idx,
                                                    const double *__restrict__ val,
                                                    const double *__restrict__ uc, double *__restrict__ u) {
  const int32_t i = blockIdx.x * BS + threadIdx.x;
  if (i >= nf) return;
  u[i] = rowdot<false>(u[i], ptr[i], ptr[i + 1] - ptr[i], idx, val, uc);
}

// r = f + (-1) (A u) with each row's dot summed sequentially in storage
// order (CompRowMatrix.multAdd, S/CompRowMatrix.java:329-347): the residual
// of levels whose matrix the library's products would reduce with lane
// groups (short rows in small row-blocks).
__global__ __launch_bounds__(BS) void k_amg_resid(int32_t n, const int32_t *__restrict__ rp,
                                                  const int32_t *__restrict__ ci, const double *__restrict__ va,
                                                  const double *__restrict__ u, const double *__restrict__ f,
                                                  double *__restrict__ r) {
  const int32_t i = blockIdx.x * BS + threadIdx.x;
  if (i >= n) return;
  const double dot = rowdot<false>(0.0, rp[i], rp[i + 1] - rp[i], ci, va, u);
  r[i] = f[i] + -1.0 * dot;
}

__global__ __launch_bounds__(BS) void k_amg_zero(int32_t n, double *x) {
  for (int32_t i = blockIdx.x * BS + threadIdx.x; i < n; i += gridDim.x * BS) x[i] = 0.0;
}

// directSolve (S/AMG.java:333-342): u = f, then dgetrs with the LU factors
// of the coarsest matrix (column-major n x n, 1-based pivots).  One
// workgroup; the right-hand side lives in LDS.  The column sweeps of DTRSM
// (NoTrans) update every later entry in parallel after each pivot step;
// the transposed upper solve is reorganised the same way, each entry still
// receiving its subtractions in ascending k; the transposed unit-lower solve
// keeps DTRSM's row-wise dot products (descending rows, ascending k) in one
// thread.
__global__ __launch_bounds__(BS) void k_amg_coarse(int32_t n, const double *__restrict__ lu,
                                                   const int32_t *__restrict__ piv, const double *__restrict__ f,
                                                   double *__restrict__ u, int transpose) {
  __shared__ double x[amg::COARSE_MAX];
  const int tid = threadIdx.x;
  for (int i = tid; i < n; i += BS) x[i] = f[i];
  __syncthreads();
  auto A = [&](int i, int j) { return lu[(size_t)i + (size_t)j * (size_t)n]; };
  if (!transpose) {
    if (tid == 0)
      for (int i = 0; i < n; ++i) {  // DLASWP, forward
        const int ip = piv[i] - 1;
        if (ip != i) {
          const double t = x[i];
          x[i] = x[ip];
          x[ip] = t;
        }
      }
    __syncthreads();
    for (int k = 0; k < n; ++k) {  // DTRSM Left Lower NoTrans Unit
      const double bk = x[k];
      if (bk != 0)
        for (int i = k + 1 + tid; i < n; i += BS) x[i] = x[i] - bk * A(i, k);
      __syncthreads();
    }
    for (int k = n - 1; k >= 0; --k) {  // DTRSM Left Upper NoTrans NonUnit
      double bk = x[k];
      __syncthreads();
      if (bk != 0) {
        bk = bk / A(k, k);
        for (int i = tid; i < k; i += BS) x[i] = x[i] - bk * A(i, k);
        if (tid == 0) x[k] = bk;
      }
      __syncthreads();
    }
  } else {
    for (int i = 0; i < n; ++i) {  // DTRSM Left Upper Trans NonUnit
      if (tid == 0) x[i] = x[i] / A(i, i);
      __syncthreads();
      const double xi = x[i];
      for (int j = i + 1 + tid; j < n; j += BS) x[j] = x[j] - A(i, j) * xi;
      __syncthreads();
    }
    if (tid == 0) {
      for (int i = n - 1; i >= 0; --i) {  // DTRSM Left Lower Trans Unit
        double temp = x[i];
        for (int k = i + 1; k < n; ++k) temp = temp - A(k, i) * x[k];
        x[i] = temp;
      }
      for (int i = n - 1; i >= 0; --i) {  // DLASWP, backward
        const int ip = piv[i] - 1;
        if (ip != i) {
          const double t = x[i];
          x[i] = x[ip];
          x[ip] = t;
        }
      }
    }
    __syncthreads();
  }
  for (int i = tid; i < n; i += BS) u[i] = x[i];
}

static void ck(int rc) {
  if (rc != MTJG_OK) fail(rc, mtjg_last_error());
}

template <typename T>
static T *upload(const std::vector<T> &v, hipStream_t s) {
  T *d = nullptr;
  HIPCHK(hipMalloc(&d, sizeof(T) * std::max<size_t>(v.size(), 1)));
  if (!v.empty()) HIPCHK(hipMemcpyAsync(d, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice, s));
  return d;
}

static double *dvec(int32_t n) {
  double *d = nullptr;
  HIPCHK(hipMalloc(&d, sizeof(double) * (size_t)std::max(n, 1)));
  return d;
}

void amg_destroy_device(AmgDevice *H) {
  if (!H) return;
  (void)hipStreamSynchronize(H->ctx->stream);
  for (AmgLevel &L : H->L) {
    if (L.ssor) mtjg_prec_destroy(L.ssor);
    if (L.A) mtjg_csr_destroy(L.A);
    for (void *p : {(void *)L.r_ptr, (void *)L.r_idx, (void *)L.r_val, (void *)L.p_ptr, (void *)L.p_idx,
                    (void *)L.p_val, (void *)L.u, (void *)L.f, (void *)L.r})
      (void)hipFree(p);
  }
  (void)hipFree(H->lu);
  (void)hipFree(H->piv);
  delete H;
}

// AMG.setMatrix (S/AMG.java:227-289).
AmgDevice *amg_create_device(mtjg_ctx ctx, mtjg_csr A, const amg::Params &p, int32_t *bad_row) {
  hipStream_t s = ctx->stream;
  amg::Csr A0;
  A0.n = A->nr;
  A0.rp = A->h_rp;
  A0.ci.resize((size_t)A->nnz);
  A0.va.resize((size_t)A->nnz);
  if (A->nnz) {
    HIPCHK(hipMemcpyAsync(A0.ci.data(), A->ci, sizeof(int32_t) * (size_t)A->nnz, hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(A0.va.data(), A->val, sizeof(double) * (size_t)A->nnz, hipMemcpyDeviceToHost, s));
  }
  HIPCHK(hipStreamSynchronize(s));
  amg::Hierarchy Hh;
  int32_t brow = -1, blev = -1;
  const amg::Status st = amg::build(A0, p, Hh, &brow, &blev);
  if (st == amg::MISSING_DIAGONAL) {
    if (bad_row && blev == 0) *bad_row = brow;
    fail(MTJG_E_MISSING_DIAGONAL, "Matrix is missing a diagonal entry on row " + std::to_string(brow + 1) +
                                      (blev > 0 ? " of AMG level " + std::to_string(blev) : std::string()));
  }
  if (st == amg::UNSUPPORTED)
    fail(MTJG_E_UNSUPPORTED, "AMG setup: a strong neighbourhood would hash into a tree bin, or more than " +
                                 std::to_string(amg::MAX_LEVELS) + " levels");
  if (st == amg::BAD_RELAXATION) fail(MTJG_E_INVALID_ARGUMENT, "omega must be between 0 and 2");
  const int m = Hh.levels();
  if (Hh.A[(size_t)m - 1].n > amg::COARSE_MAX)
    fail(MTJG_E_UNSUPPORTED, "coarsest AMG level has " + std::to_string(Hh.A[(size_t)m - 1].n) +
                                 " rows; the device LU solve holds at most " + std::to_string(amg::COARSE_MAX));
  auto *H = new AmgDevice();
  H->ctx = ctx;
  H->p = p;
  H->L.resize((size_t)m);
  try {
    for (int k = 0; k < m; ++k) {
      AmgLevel &L = H->L[(size_t)k];
      const amg::Csr &Ak = Hh.A[(size_t)k];
      L.n = Ak.n;
      L.u = dvec(L.n);
      L.f = dvec(L.n);
      if (k == m - 1) break;
      L.r = dvec(L.n);
      ck(mtjg_csr_create(ctx, Ak.n, Ak.n, (int64_t)Ak.ci.size(), Ak.rp.data(), Ak.ci.data(), Ak.va.data(), &L.A));
      ck(mtjg_prec_ssor_create(ctx, L.A, p.reverse ? 1 : 0, p.omega_pre_f, p.omega_pre_r, &L.ssor));
      ck(mtjg_prec_set_matrix(L.ssor, L.A, nullptr));
      const amg::Ccs &I = Hh.I[(size_t)k];
      L.r_ptr = upload(I.cp, s);
      L.r_idx = upload(I.ri, s);
      L.r_val = upload(I.va, s);
      // CSR of I: fine rows, coarse columns ascending
      std::vector<int32_t> pp((size_t)I.nr + 1, 0), pi(I.ri.size());
      std::vector<double> pv(I.ri.size());
      for (int32_t r : I.ri) pp[(size_t)r + 1]++;
      for (int32_t r = 0; r < I.nr; ++r) pp[(size_t)r + 1] += pp[(size_t)r];
      std::vector<int32_t> pos(pp.begin(), pp.end() - 1);
      for (int32_t c = 0; c < I.nc; ++c)
        for (int32_t e = I.cp[(size_t)c]; e < I.cp[(size_t)c + 1]; ++e) {
          const int32_t q = pos[(size_t)I.ri[(size_t)e]]++;
          pi[(size_t)q] = c;
          pv[(size_t)q] = I.va[(size_t)e];
        }
      L.p_ptr = upload(pp, s);
      L.p_idx = upload(pi, s);
      L.p_val = upload(pv, s);
      HIPCHK(hipStreamSynchronize(s));  // host staging vectors go out of scope
    }
    H->nc = Hh.A[(size_t)m - 1].n;
    H->lu = upload(Hh.lu, s);
    H->piv = upload(Hh.piv, s);
    H->singular = Hh.singular;
    HIPCHK(hipStreamSynchronize(s));
  } catch (...) {
    amg_destroy_device(H);
    throw;
  }
  return H;
}

static void relax(AmgDevice *H, AmgLevel &L, int times, double wf, double wr, hipStream_t s) {
  if (times <= 0) return;
  ck(mtjg_prec_ssor_set_omega(L.ssor, wf, wr));
  for (int i = 0; i < times; ++i) prec_apply_dev(L.ssor, L.n, L.f, L.u, s, nullptr);
  (void)H;
}

static void cycle(AmgDevice *H, int k, bool transpose, hipStream_t s) {
  const int m = (int)H->L.size();
  AmgLevel &L = H->L[(size_t)k];
  if (k == m - 1) {
    if (H->nc > 0) k_amg_coarse<<<1, BS, 0, s>>>(H->nc, H->lu, H->piv, L.f, L.u, transpose ? 1 : 0);
    HIPCHK(hipGetLastError());
    return;
  }
  AmgLevel &C = H->L[(size_t)k + 1];
  relax(H, L, H->p.nu1, H->p.omega_pre_f, H->p.omega_pre_r, s);
  if (C.n > 0) k_amg_zero<<<grid_for(C.n), BS, 0, s>>>(C.n, C.u);
  if (L.A->sell) {  // SELL-64 rows are summed sequentially by one lane each
    SpmvArgs ra;
    ra.mode = SPMV_MULTADD;
    ra.alpha = -1;
    ra.y0 = L.f;
    spmv_f64(L.A, L.u, L.r, ra, s);
  } else if (L.n > 0) {
    k_amg_resid<<<(L.n + BS - 1) / BS, BS, 0, s>>>(L.n, L.A->rp, L.A->ci, L.A->val, L.u, L.f, L.r);
  }
  if (C.n > 0) k_amg_restrict<<<(C.n + BS - 1) / BS, BS, 0, s>>>(C.n, L.r_ptr, L.r_idx, L.r_val, L.r, C.f);
  for (int i = 0; i < H->p.gamma; ++i) cycle(H, k + 1, transpose, s);
  if (L.n > 0) k_amg_prolong<<<(L.n + BS - 1) / BS, BS, 0, s>>>(L.n, L.p_ptr, L.p_idx, L.p_val, C.u, L.u);
  relax(H, L, H->p.nu2, H->p.omega_post_f, H->p.omega_post_r, s);
  HIPCHK(hipGetLastError());
}

// apply / transApply (S/AMG.java:207-225): u_0 = x, f_0 = b, cycle,
// x = u_0.  A singular coarsest matrix raises MatrixSingularException before
// any work, leaving x as it was (MTJ throws from the coarse solve, before
// x.set(u_0)).
void amg_apply_device(AmgDevice *H, const double *b, double *x, bool transpose, hipStream_t s) {
  if (H->singular) fail(MTJG_E_SINGULAR, "MatrixSingularException: singular coarsest AMG matrix");
  AmgLevel &L0 = H->L[0];
  const size_t bytes = sizeof(double) * (size_t)L0.n;
  if (L0.n > 0) {
    HIPCHK(hipMemcpyAsync(L0.u, x, bytes, hipMemcpyDeviceToDevice, s));
    HIPCHK(hipMemcpyAsync(L0.f, b, bytes, hipMemcpyDeviceToDevice, s));
  }
  cycle(H, 0, transpose, s);
  if (L0.n > 0) HIPCHK(hipMemcpyAsync(x, L0.u, bytes, hipMemcpyDeviceToDevice, s));
}

void amg_check_device(AmgDevice *H) {
  for (AmgLevel &L : H->L)
    if (L.ssor) prec_check(L.ssor);
}

int32_t amg_level_sizes(const AmgDevice *H, int32_t *sizes, int32_t cap) {
  const int32_t m = (int32_t)H->L.size();
  for (int32_t k = 0; k < m && k < cap; ++k) sizes[k] = H->L[(size_t)k].n;
  return m;
}

}  // namespace mtjg
