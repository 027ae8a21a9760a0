// Iterative solvers (S/CG.java, S/BiCGstab.java, S/GMRES.java, S/IR.java,
// S/BiCG.java, S/CGS.java, S/QMR.java) with every operand device resident;
// CG, BiCGstab, GMRES and IR drive fused HIP kernels.
//
// Each solver keeps its scalars (rho, alpha, omega, beta, the Hessenberg
// matrix, the iteration-monitor state) in a device struct.  Reductions end in
// the last workgroup of their kernel, which evaluates the next scalar and the
// DefaultIterationMonitor predicate on the device and sets a status word.
// Every kernel of an iteration returns at once when the status is no longer
// "running", so the host can queue several iterations ahead and only reads
// the status word between batches ("chunks") of iterations.
//
// Vector updates apply the same per-element operation as MTJ (no FMA:
// compiled with -ffp-contract=off) and therefore equal MTJ bitwise given the
// same scalars; dot products and norms differ only in summation order.
#include <algorithm>
#include <initializer_list>
#include <cstddef>
#include <cmath>
#include <cstring>
#include <limits>
#include <mutex>

#include "mtjg_internal.hpp"
#include "reduce.hpp"

using namespace mtjg;

struct mtjg_solver_s {
  mtjg_ctx ctx = nullptr;
  int kind = MTJG_CG;
  int64_t n = 0;
  mtjg_prec M = nullptr;
  mtjg_monitor mon{100000, 1e-5, 1e-50, 1e5};
  mtjg_reporter rep = nullptr;
  void *rep_user = nullptr;
  int32_t restart = 30;
  double inner_rtol = 1e-4;
  int32_t inner_max = 1000;
  // device state
  KState *st = nullptr;      // outer solver state
  KState *st_in = nullptr;   // inner (FP32) solve of MTJG_MIXED_CG
  KState *h_st = nullptr;    // pinned staging [2]
  int32_t *h_slot = nullptr; // pinned status slots [2]
  GridRed red{};             // two-level reduction workspace (grown on demand)
  hipEvent_t ev[2] = {nullptr, nullptr};
  std::vector<double *> v64;  // scratch FP64 vectors
  std::vector<float *> v32;   // scratch FP32 vectors (mixed)
  // GMRES
  double *V = nullptr;   // (restart+1) x n basis, vector-major
  double *H = nullptr;   // (restart+1) x restart, column-major
  double *gs = nullptr;  // s, rotation c and s: 3 * (restart+1)
  // cooperative Gram-Schmidt (k_gmres_mgs_all): published-step counter,
  // poll-budget flag, poll budget; mgs_coop: 0 undecided, 1 usable, -1 not
  uint32_t *mgs_bar = nullptr;
  int32_t *mgs_err = nullptr;
  int64_t mgs_budget = 0;
  int mgs_coop = 0;
  int32_t inner_total = 0;
};

namespace {

// Vector kernels split the vector into tiles of BS * VPT elements; a thread
// owns elements tile + threadIdx.x + u * BS (unit stride across the wave, all
// loads of a thread independent).  The grid holds at most VGRID_MAX
// workgroups (eight per CU, all co-resident) that walk the tiles in
// grid-stride order, so a reducing kernel takes at most VGRID_MAX tickets.
constexpr int VPT = 4;
constexpr int VGRID_MAX = 2048;
inline int vgrid(int64_t n) {
  return (int)std::min<int64_t>(VGRID_MAX, std::max<int64_t>(1, (n + BS * VPT - 1) / (BS * VPT)));
}
#define TILES(base)                                                                          \
  for (int64_t base = (int64_t)blockIdx.x * (BS * VPT) + threadIdx.x; base - threadIdx.x < n; \
       base += (int64_t)gridDim.x * (BS * VPT))
#define VLOOP(i)                                                                                  \
  TILES(i##_b)                                                                                    \
  _Pragma("unroll") for (int64_t i##_u = 0, i = i##_b; i##_u < VPT; ++i##_u, i += BS) if (i < n)

// ---------------------------------------------------------------------------
// CG kernels

enum PrecK : int { PK_IDENTITY = 0, PK_JACOBI = 1, PK_EXTERNAL = 2 };

// Residual step of CG.  With UPDATE it first applies x += alpha p and
// r += (-alpha) q (S/CG.java:80-81) and advances the iteration counter
// (iter.next()); then z = M r for identity / Jacobi
// (S/DiagonalPreconditioner.java:54-55), and the sums r.z and r.r.  The last
// workgroup evaluates the monitor on ||r|| (S/CG.java:66) and, while
// running, rho = r.z (S/CG.java:68) and beta = rho / rho_1 (S/CG.java:73).
// Vector loads and stores of the CG streaming kernels.  With NT they carry
// the non-temporal hint: for vectors too large for the 256 MB Infinity Cache
// to pass from one kernel to the next (3D Poisson 256^3: 0.326 -> 0.311 s per
// solve) - for smaller ones it keeps the reuse and is not used (192^3:
// 0.171 s plain, 0.175 s with the hint).
template <bool NT, typename T>
__device__ __forceinline__ T vld(const T *p) {
  if (NT) return __builtin_nontemporal_load(p);
  return *p;
}
template <bool NT, typename T>
__device__ __forceinline__ void vst(T *p, T v) {
  if (NT) __builtin_nontemporal_store(v, p);
  else *p = v;
}

// 16-byte accesses of the vector kernels' VEC form: a tile's elements are
// taken in rounds of BS * VW (VW elements of 16 bytes per thread), so every
// access instruction of a wavefront covers 1 KB contiguously; with 4-byte
// elements the scalar form moves only 256 B per instruction (FP32 CG
// residual step 4.0 TB/s).  Groups that run past n are done element by
// element.
template <typename T>
struct V16;
template <>
struct V16<float> {
  typedef float type __attribute__((ext_vector_type(4)));
};
template <>
struct V16<double> {
  typedef double type __attribute__((ext_vector_type(2)));
};
template <bool NT, typename T>
__device__ __forceinline__ void vld16(const T *p, T *o) {
  using VT = typename V16<T>::type;
  constexpr int VW = 16 / sizeof(T);
  const VT v = NT ? __builtin_nontemporal_load((const VT *)p) : *(const VT *)p;
#pragma unroll
  for (int k = 0; k < VW; ++k) o[k] = v[k];
}
template <bool NT, typename T>
__device__ __forceinline__ void vst16(T *p, const T *o) {
  using VT = typename V16<T>::type;
  constexpr int VW = 16 / sizeof(T);
  VT v;
#pragma unroll
  for (int k = 0; k < VW; ++k) v[k] = o[k];
  if (NT) __builtin_nontemporal_store(v, (VT *)p);
  else *(VT *)p = v;
}
// Element index of slot u (0 .. VPT-1) of this thread in the tile at `tile`.
template <typename T>
__device__ __forceinline__ int64_t vec_index(int64_t tile, int u) {
  constexpr int VW = 16 / sizeof(T);
  return tile + (int64_t)(u / VW) * (BS * VW) + (int64_t)threadIdx.x * VW + u % VW;
}
// This thread's VPT elements of the tile at `tile` for the VEC form: loads
// of any number of vectors, all issued before any value is used.  A tile
// that lies wholly below n (the test is uniform over the workgroup) takes
// straight-line 16-byte loads with no branch between them; the last tile is
// read element by element.  Each operand is a (source, registers) pair.
struct VIo {
  const double *p;
  double *v;
};
__device__ __forceinline__ bool vtile_full(int64_t n, int64_t tile) { return tile + BS * VPT <= n; }
template <bool NT = false, typename... Io>
__device__ __forceinline__ void vload_tile(int64_t n, int64_t tile, Io... io) {
  constexpr int VW = 2;
  if (vtile_full(n, tile)) {
#pragma unroll
    for (int g = 0; g < VPT; g += VW) {
      const int64_t i = vec_index<double>(tile, g);
      (vld16<NT>(io.p + i, io.v + g), ...);
    }
  } else {
#pragma unroll
    for (int u = 0; u < VPT; ++u) {
      const int64_t i = vec_index<double>(tile, u);
      if (i < n) ((io.v[u] = io.p[i]), ...);
    }
  }
}
__device__ __forceinline__ void vstore_tile(int64_t n, int64_t tile, double *a, const double *av) {
  constexpr int VW = 2;
  if (vtile_full(n, tile)) {
#pragma unroll
    for (int g = 0; g < VPT; g += VW) vst16<false>(a + vec_index<double>(tile, g), av + g);
  } else {
#pragma unroll
    for (int u = 0; u < VPT; ++u) {
      const int64_t i = vec_index<double>(tile, u);
      if (i < n) a[i] = av[u];
    }
  }
}
// Whether slot u of the tile at `tile` is an element (< n).
__device__ __forceinline__ bool vslot_ok(int64_t n, int64_t tile, int u) {
  return vtile_full(n, tile) || vec_index<double>(tile, u) < n;
}
#define VEC_TILES(tile) \
  for (int64_t tile = (int64_t)blockIdx.x * (BS * VPT); tile < n; tile += (int64_t)gridDim.x * (BS * VPT))

// Whether every pointer is 16-byte aligned (the VEC form's precondition).
inline bool aligned16(std::initializer_list<const void *> ps) {
  for (const void *p : ps)
    if (p && ((uintptr_t)p & 15)) return false;
  return true;
}

template <typename T, int PK, bool UPDATE, bool NT = false, bool VEC = false>
__global__ __launch_bounds__(BS) void k_cg_resid(int64_t n, T *x, const T *p, T *r, const T *q,
                                                 const T *invdiag, T *z, KState *st, GridRed red) {
  if (st->status != ST_RUNNING) return;
  const T alpha = (T)st->alpha;
  dd v[2] = {{0.0, 0.0}, {0.0, 0.0}};
  if (VEC) {
    constexpr int VW = 16 / sizeof(T);
    static_assert(VPT % VW == 0, "a thread's slots are whole 16-byte groups");
    for (int64_t tile = (int64_t)blockIdx.x * (BS * VPT); tile < n; tile += (int64_t)gridDim.x * (BS * VPT)) {
      T ri[VPT], xi[VPT], pi[VPT], qi[VPT], di[VPT];
#pragma unroll
      for (int g = 0; g < VPT; g += VW) {
        const int64_t i = vec_index<T>(tile, g);
        if (i + VW <= n) {
          vld16<NT>(&r[i], &ri[g]);
          if (UPDATE) {
            vld16<NT>(&x[i], &xi[g]);
            vld16<NT>(&p[i], &pi[g]);
            vld16<NT>(&q[i], &qi[g]);
          }
          if (PK == PK_JACOBI) vld16<NT>(&invdiag[i], &di[g]);
        } else {
#pragma unroll
          for (int k = 0; k < VW; ++k)
            if (i + k < n) {
              ri[g + k] = r[i + k];
              if (UPDATE) {
                xi[g + k] = x[i + k];
                pi[g + k] = p[i + k];
                qi[g + k] = q[i + k];
              }
              if (PK == PK_JACOBI) di[g + k] = invdiag[i + k];
            }
        }
      }
      const bool upd = UPDATE && alpha != T(0);
#pragma unroll
      for (int u = 0; u < VPT; ++u) {
        if (upd) {
          xi[u] = xi[u] + alpha * pi[u];
          ri[u] = ri[u] + (-alpha) * qi[u];
        }
        if (PK == PK_JACOBI) di[u] = ri[u] * di[u];  // z
      }
#pragma unroll
      for (int g = 0; g < VPT; g += VW) {
        const int64_t i = vec_index<T>(tile, g);
        if (i + VW <= n) {
          if (upd) {
            vst16<NT>(&x[i], &xi[g]);
            vst16<NT>(&r[i], &ri[g]);
          }
          if (PK == PK_JACOBI) vst16<NT>(&z[i], &di[g]);
        } else {
#pragma unroll
          for (int k = 0; k < VW; ++k)
            if (i + k < n) {
              if (upd) {
                x[i + k] = xi[g + k];
                r[i + k] = ri[g + k];
              }
              if (PK == PK_JACOBI) z[i + k] = di[g + k];
            }
        }
#pragma unroll
        for (int k = 0; k < VW; ++k)
          if (i + k < n) {
            const int u = g + k;
            if (PK == PK_JACOBI) dd_fma(v[0], (double)ri[u], (double)di[u]);
            else if (PK == PK_IDENTITY) dd_fma(v[0], (double)ri[u], (double)ri[u]);
            dd_fma(v[1], (double)ri[u], (double)ri[u]);
          }
      }
    }
  } else TILES(base) {
    T ri[VPT];
#pragma unroll
    for (int u = 0; u < VPT; ++u) {
      const int64_t i = base + u * BS;
      if (i < n) {
        ri[u] = vld<NT>(&r[i]);
        if (UPDATE && alpha != T(0)) {
          vst<NT>(&x[i], vld<NT>(&x[i]) + alpha * vld<NT>(&p[i]));
          ri[u] = ri[u] + (-alpha) * vld<NT>(&q[i]);
          vst<NT>(&r[i], ri[u]);
        }
      }
    }
#pragma unroll
    for (int u = 0; u < VPT; ++u) {
      const int64_t i = base + u * BS;
      if (i < n) {
        if (PK == PK_JACOBI) {
          T zi = ri[u] * vld<NT>(&invdiag[i]);
          vst<NT>(&z[i], zi);
          dd_fma(v[0], (double)ri[u], (double)zi);
        } else if (PK == PK_IDENTITY) {
          dd_fma(v[0], (double)ri[u], (double)ri[u]);
        }
        dd_fma(v[1], (double)ri[u], (double)ri[u]);
      }
    }
  }
  double tot[2];
  if (grid_sum_dd<2>(v, red, tot) && threadIdx.x == 0) {
    if (UPDATE) st->iter++;
    int s = monitor_check(st, sqrt(tot[1]));
    if (s != ST_RUNNING) {
      st->status = s;
      return;
    }
    if (PK != PK_EXTERNAL) {
      st->rho = tot[0];
      if (st->iter > 0) st->beta = st->rho / st->rho_1;
    }
  }
}

// rho = r.z for a preconditioner applied by separate launches (ILU).
template <typename T>
__global__ __launch_bounds__(BS) void k_cg_rho(int64_t n, const T *r, const T *z, KState *st, GridRed red) {
  if (st->status != ST_RUNNING) return;
  dd v[1] = {{0.0, 0.0}};
  VLOOP(i) dd_fma(v[0], (double)r[i], (double)z[i]);
  double tot[1];
  if (grid_sum_dd<1>(v, red, tot) && threadIdx.x == 0) {
    st->rho = tot[0];
    if (st->iter > 0) st->beta = st->rho / st->rho_1;
  }
}

// p = z on the first iteration, else p.scale(beta).add(z) (S/CG.java:70-75):
// two roundings, p*beta then + z.
template <typename T, bool NT = false, bool VEC = false>
__global__ __launch_bounds__(BS) void k_cg_p(int64_t n, T *p, const T *z, const KState *st) {
  if (st->status != ST_RUNNING) return;
  const bool first = st->iter == 0;
  const T beta = (T)st->beta;
  if (VEC) {
    constexpr int VW = 16 / sizeof(T);
    for (int64_t tile = (int64_t)blockIdx.x * (BS * VPT); tile < n; tile += (int64_t)gridDim.x * (BS * VPT)) {
#pragma unroll
      for (int g = 0; g < VPT; g += VW) {
        const int64_t i = vec_index<T>(tile, g);
        T pv[VW], zv[VW];
        if (i + VW <= n) {
          vld16<NT>(&z[i], zv);
          if (!first) vld16<NT>(&p[i], pv);
#pragma unroll
          for (int k = 0; k < VW; ++k) pv[k] = first ? zv[k] : pv[k] * beta + zv[k];
          vst16<NT>(&p[i], pv);
        } else {
          for (int k = 0; k < VW; ++k)
            if (i + k < n) p[i + k] = first ? z[i + k] : p[i + k] * beta + z[i + k];
        }
      }
    }
    return;
  }
  VLOOP(i) {
    if (first) {
      vst<NT>(&p[i], vld<NT>(&z[i]));
    } else {
      T pb = vld<NT>(&p[i]) * beta;
      vst<NT>(&p[i], pb + vld<NT>(&z[i]));
    }
  }
}

// Launches of the CG residual step and direction update, in the VEC form
// when every operand is 16-byte aligned (always for the library's own
// vectors; a caller's vector view may not be).
template <typename T, int PK, bool UPDATE>
void launch_cg_resid(bool nt, bool vec, int g, hipStream_t s, int64_t n, T *x, const T *p, T *r, const T *q,
                     const T *invd, T *z, KState *st, const GridRed &red) {
  if (nt) {
    if (vec) k_cg_resid<T, PK, UPDATE, true, true><<<g, BS, 0, s>>>(n, x, p, r, q, invd, z, st, red);
    else k_cg_resid<T, PK, UPDATE, true, false><<<g, BS, 0, s>>>(n, x, p, r, q, invd, z, st, red);
  } else {
    if (vec) k_cg_resid<T, PK, UPDATE, false, true><<<g, BS, 0, s>>>(n, x, p, r, q, invd, z, st, red);
    else k_cg_resid<T, PK, UPDATE, false, false><<<g, BS, 0, s>>>(n, x, p, r, q, invd, z, st, red);
  }
}
template <typename T>
void launch_cg_p(bool nt, bool vec, int g, hipStream_t s, int64_t n, T *p, const T *z, const KState *st) {
  if (nt) {
    if (vec) k_cg_p<T, true, true><<<g, BS, 0, s>>>(n, p, z, st);
    else k_cg_p<T, true, false><<<g, BS, 0, s>>>(n, p, z, st);
  } else {
    if (vec) k_cg_p<T, false, true><<<g, BS, 0, s>>>(n, p, z, st);
    else k_cg_p<T, false, false><<<g, BS, 0, s>>>(n, p, z, st);
  }
}

// ---------------------------------------------------------------------------
// BiCGstab kernels (S/BiCGstab.java:74-113)

// Top of an iteration.  With UPDATE: x += omega shat, r = s + (-omega) t
// (S/BiCGstab.java:109-110), rho_2 = rho_1 and iter.next(); then the sums
// r.r and rtilde.r.  The last workgroup checks the monitor on ||r||, sets
// rho_1 = rtilde.r, raises breakdown when rho_1 == 0 or omega == 0
// (:77-83) and forms beta (:88).
template <bool UPDATE, bool VEC = false>
__global__ __launch_bounds__(BS) void k_bicg_top(int64_t n, double *x, const double *shat, double *r,
                                                 const double *s, const double *t,
                                                 const double *rtilde, KState *st, GridRed red) {
  if (st->status != ST_RUNNING) return;
  const double omega = st->omega;
  dd v[2] = {{0.0, 0.0}, {0.0, 0.0}};
  if (VEC) {
    VEC_TILES(tile) {
      double xv[VPT], hv[VPT], rv[VPT], tv[VPT], qv[VPT];
      if (UPDATE) vload_tile(n, tile, VIo{x, xv}, VIo{shat, hv}, VIo{s, rv}, VIo{t, tv});
      else vload_tile(n, tile, VIo{r, rv});
      vload_tile(n, tile, VIo{rtilde, qv});
      if (UPDATE && omega != 0) {
#pragma unroll
        for (int u = 0; u < VPT; ++u) {
          xv[u] = xv[u] + omega * hv[u];
          rv[u] = rv[u] + (-omega) * tv[u];
        }
        vstore_tile(n, tile, x, xv);
      }
      if (UPDATE) vstore_tile(n, tile, r, rv);
#pragma unroll
      for (int u = 0; u < VPT; ++u)
        if (vslot_ok(n, tile, u)) {
          dd_fma(v[0], rv[u], rv[u]);
          dd_fma(v[1], qv[u], rv[u]);
        }
    }
  } else VLOOP(i) {
    double ri;
    if (UPDATE) {
      if (omega != 0) x[i] = x[i] + omega * shat[i];
      ri = s[i];
      if (omega != 0) ri = ri + (-omega) * t[i];
      r[i] = ri;
    } else {
      ri = r[i];
    }
    dd_fma(v[0], ri, ri);
    dd_fma(v[1], rtilde[i], ri);
  }
  double tot[2];
  if (grid_sum_dd<2>(v, red, tot) && threadIdx.x == 0) {
    if (UPDATE) {
      st->rho_2 = st->rho_1;
      st->iter++;
    }
    int stt = monitor_check(st, sqrt(tot[0]));
    if (stt != ST_RUNNING) {
      st->status = stt;
      return;
    }
    st->rho_1 = tot[1];
    if (st->rho_1 == 0 || st->omega == 0) {
      st->status = ST_BREAKDOWN;
      return;
    }
    if (st->iter > 0) st->beta = (st->rho_1 / st->rho_2) * (st->alpha / st->omega);
  }
}

// p = r (first) or p = r + beta * ((-omega) v + p) (S/BiCGstab.java:85-95).
template <bool VEC = false>
__global__ __launch_bounds__(BS) void k_bicg_p(int64_t n, double *p, const double *r, const double *v,
                                               const KState *st) {
  if (st->status != ST_RUNNING) return;
  const bool first = st->iter == 0;
  const double beta = st->beta, omega = st->omega;
  if (VEC) {
    VEC_TILES(tile) {
      double pv[VPT], rv[VPT], vv[VPT];
      if (first) {
        vload_tile(n, tile, VIo{r, rv});
        vstore_tile(n, tile, p, rv);
        continue;
      }
      vload_tile(n, tile, VIo{r, rv}, VIo{v, vv}, VIo{p, pv});
#pragma unroll
      for (int u = 0; u < VPT; ++u) {
        double temp = (-omega) * vv[u];
        temp = temp + pv[u];
        double pi = rv[u];
        if (beta != 0) pi = pi + beta * temp;
        pv[u] = pi;
      }
      vstore_tile(n, tile, p, pv);
    }
    return;
  }
  VLOOP(i) {
    if (first) {
      p[i] = r[i];
    } else {
      double temp = (-omega) * v[i];
      temp = temp + p[i];
      double pi = r[i];
      if (beta != 0) pi = pi + beta * temp;
      p[i] = pi;
    }
  }
}

// s = r + (-alpha) v; x += alpha phat; then the half-step check on ||s||
// (S/BiCGstab.java:100-104).
template <bool VEC = false>
__global__ __launch_bounds__(BS) void k_bicg_s(int64_t n, double *x, const double *phat, const double *r,
                                               const double *v, double *s, KState *st, GridRed red) {
  if (st->status != ST_RUNNING) return;
  const double alpha = st->alpha;
  dd acc[1] = {{0.0, 0.0}};
  if (VEC) {
    VEC_TILES(tile) {
      double rv[VPT], vv[VPT], xv[VPT], hv[VPT];
      vload_tile(n, tile, VIo{r, rv}, VIo{v, vv}, VIo{x, xv}, VIo{phat, hv});
      if (alpha != 0) {
#pragma unroll
        for (int u = 0; u < VPT; ++u) {
          rv[u] = rv[u] + (-alpha) * vv[u];
          xv[u] = xv[u] + alpha * hv[u];
        }
        vstore_tile(n, tile, x, xv);
      }
      vstore_tile(n, tile, s, rv);
#pragma unroll
      for (int u = 0; u < VPT; ++u)
        if (vslot_ok(n, tile, u)) dd_fma(acc[0], rv[u], rv[u]);
    }
  } else VLOOP(i) {
    double si = r[i];
    if (alpha != 0) {
      si = si + (-alpha) * v[i];
      x[i] = x[i] + alpha * phat[i];
    }
    s[i] = si;
    dd_fma(acc[0], si, si);
  }
  double tot[1];
  if (grid_sum_dd<1>(acc, red, tot) && threadIdx.x == 0) {
    int stt = monitor_check(st, sqrt(tot[0]));
    if (stt != ST_RUNNING) st->status = stt;
  }
}

__global__ void k_bicg_omega(KState *st) {
  if (st->status != ST_RUNNING) return;
  st->omega = st->s[0] / st->s[1];  // S/BiCGstab.java:108
}

// ---------------------------------------------------------------------------
// GMRES kernels (S/GMRES.java:127-176)

struct GmresDev {
  double *H;   // column-major (m+1) x m
  double *gs;  // s[m+1], c[m+1], sn[m+1]
  int m;
};

// Start of an outer cycle.  ||r|| (r = M^-1 (b - A x)); unless first,
// iter.next() of the outer loop; the outer check; then s = normr e1 and
// the inner check at i = 0 (S/GMRES.java:137-144).
__global__ __launch_bounds__(BS) void k_gmres_start(int64_t n, const double *r, KState *st,
                                                    GmresDev g, int first, GridRed red) {
  if (st->status != ST_RUNNING) return;
  dd v[1] = {{0.0, 0.0}};
  VLOOP(i)
    dd_fma(v[0], r[i], r[i]);
  double tot[1];
  if (grid_sum_dd<1>(v, red, tot) && threadIdx.x == 0) {
    double normr = sqrt(tot[0]);
    if (!first) st->iter++;
    int stt = monitor_check(st, normr);
    if (stt != ST_RUNNING) {
      st->status = stt;
      return;
    }
    st->normr = normr;
    for (int k = 0; k <= g.m; ++k) g.gs[k] = 0.0;
    g.gs[0] = normr;
    st->inner_i = 0;
    st->cycle_done = 0;
    stt = monitor_check(st, fabs(g.gs[0]));
    if (stt == ST_CONVERGED) st->cycle_done = 1;
    else if (stt != ST_RUNNING) st->status = stt;
  }
}

// v0 = (1/normr) r, with DenseVector.set(alpha, y)'s alpha == 0 zeroing.
template <bool VEC = false>
__global__ __launch_bounds__(BS) void k_gmres_v0(int64_t n, const double *r, double *v0, const KState *st) {
  if (st->status != ST_RUNNING) return;
  const double a = 1 / st->normr;
  if (VEC) {
    VEC_TILES(tile) {
      double rv[VPT];
      vload_tile(n, tile, VIo{r, rv});
#pragma unroll
      for (int u = 0; u < VPT; ++u) rv[u] = (a == 0) ? 0.0 : a * rv[u];
      vstore_tile(n, tile, v0, rv);
    }
    return;
  }
  VLOOP(i)
    v0[i] = (a == 0) ? 0.0 : a * r[i];
}

// Modified Gram-Schmidt step k of Arnoldi column i: w += (-H(k-1,i)) v_{k-1}
// (S/GMRES.java:151) for k > 0, then H(k,i) = w.v_k (:150).
template <bool VEC = false>
__global__ __launch_bounds__(BS) void k_gmres_mgs(int64_t n, double *w, const double *vprev,
                                                  const double *vk, int k, int i, KState *st,
                                                  GmresDev g, GridRed red) {
  if (st->status != ST_RUNNING) return;
  const int ld = g.m + 1;
  const double h = (k > 0) ? -g.H[(k - 1) + (size_t)i * ld] : 0.0;
  dd v[1] = {{0.0, 0.0}};
  if (VEC) {
    VEC_TILES(tile) {
      double wv[VPT], pv[VPT], kv[VPT];
      if (h != 0) {
        vload_tile(n, tile, VIo{w, wv}, VIo{vprev, pv}, VIo{vk, kv});
#pragma unroll
        for (int u = 0; u < VPT; ++u) wv[u] = wv[u] + h * pv[u];
        vstore_tile(n, tile, w, wv);
      } else {
        vload_tile(n, tile, VIo{w, wv}, VIo{vk, kv});
      }
#pragma unroll
      for (int u = 0; u < VPT; ++u)
        if (vslot_ok(n, tile, u)) dd_fma(v[0], wv[u], kv[u]);
    }
  } else VLOOP(e) {
    double we = w[e];
    if (h != 0) {
      we = we + h * vprev[e];
      w[e] = we;
    }
    dd_fma(v[0], we, vk[e]);
  }
  double tot[1];
  if (grid_sum_dd<1>(v, red, tot) && threadIdx.x == 0) g.H[k + (size_t)i * ld] = tot[0];
}

// The i + 1 Gram-Schmidt steps of Arnoldi step i (k_gmres_mgs for
// k = 0 .. i) in one cooperative launch: each workgroup keeps its NT tiles
// of w and of the previous basis vector in registers across the steps, so a
// step reads one basis vector instead of w, two basis vectors and a write of
// w.  Every workgroup walks the same tiles in the same order as k_gmres_mgs
// with the same grid, and the sums take the same reduction tree, so H and w
// are bitwise those of the per-step launches.
//
// Step barrier: *bar counts published steps.  Thread 0 of each workgroup
// reads it at the start (no step of this launch can be published before
// every workgroup has taken its first ticket, so all read the same base).
// The workgroup that completes step k's sum publishes H(k, i) (agent-scope
// store, waited for), then base + k + 1 in *bar; the others wait for that
// number before reading H(k, i), polling at most `budget` times (<= 0 admits
// no wait).  A workgroup that runs out sets *err and the status ST_DEVICE
// and leaves; the others then run out at their next wait, so the grid always
// drains (the reduction's tickets are left part-taken and the host resets
// them).  The reduction's counters are reset by their last users before the
// step is published, so the next step's sum starts clean.  The last step
// needs no barrier (nothing here reads H(i, i)).  All workgroups must be
// resident at once: the host launches this only cooperatively.
// 8 wavefronts per SIMD (at most 64 VGPRs): the 2048-workgroup grid of
// 256-thread workgroups is then resident at once on 256 CUs.
template <int NT>
__global__ __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(8, 8))) void k_gmres_mgs_all(int64_t n, double *w, const double *V, int i, KState *st,
                                                      GmresDev g, GridRed red, uint32_t *bar, int64_t budget,
                                                      int32_t *err) {
  __shared__ int quit;
  if (st->status != ST_RUNNING) return;
  const int ld = g.m + 1;
  const int64_t t0 = (int64_t)blockIdx.x * (BS * VPT), stride = (int64_t)gridDim.x * (BS * VPT);
  uint32_t base = 0;
  if (threadIdx.x == 0) base = __hip_atomic_load(bar, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  double wv[NT][VPT], pv[NT][VPT];
#pragma unroll
  for (int j = 0; j < NT; ++j)
    if (t0 + j * stride < n) vload_tile(n, t0 + j * stride, VIo{w, wv[j]});
  for (int k = 0; k <= i; ++k) {
    const double h = (k > 0) ? -load_agent(&g.H[(k - 1) + (size_t)i * ld]) : 0.0;
    const double *vk = V + (size_t)k * (size_t)n;
    dd v[1] = {{0.0, 0.0}};
#pragma unroll
    for (int j = 0; j < NT; ++j) {
      const int64_t tile = t0 + j * stride;
      if (tile < n) {
        double kv[VPT];
        vload_tile(n, tile, VIo{vk, kv});
        if (h != 0) {
#pragma unroll
          for (int u = 0; u < VPT; ++u) wv[j][u] = wv[j][u] + h * pv[j][u];
        }
#pragma unroll
        for (int u = 0; u < VPT; ++u) {
          if (vslot_ok(n, tile, u)) dd_fma(v[0], wv[j][u], kv[u]);
          pv[j][u] = kv[u];
        }
      }
    }
    double tot[1];
    const bool fin = grid_sum_dd<1>(v, red, tot);
    if (k == i) {
      if (fin && threadIdx.x == 0) g.H[k + (size_t)i * ld] = tot[0];
      break;
    }
    if (threadIdx.x == 0) {
      const uint32_t step = base + (uint32_t)k + 1u;
      quit = 0;
      if (fin) {
        store_agent(&g.H[k + (size_t)i * ld], tot[0]);
        stores_complete();
        __hip_atomic_store(bar, step, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      } else {
        int64_t left = budget;
        while ((int32_t)(__hip_atomic_load(bar, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) - step) < 0) {
          if (--left < 0) {
            atomicExch(err, 1);
            st->status = ST_DEVICE;
            quit = 1;
            break;
          }
          // ~1 us between polls: with 2048 pollers, polling without a
          // sleep delays the publishing store (GMRES(30) + ILU(0) on 3-D
          // convection-diffusion 128^3: 0.122 s against 0.105 s; a sleep
          // of 1 gives 0.107 s, of 4 0.1055 s)
          __builtin_amdgcn_s_sleep(16);
        }
      }
    }
    __syncthreads();
    if (quit) return;
  }
#pragma unroll
  for (int j = 0; j < NT; ++j)
    if (t0 + j * stride < n) vstore_tile(n, t0 + j * stride, w, wv[j]);
}

// GivensRotation constructor and apply (M/GivensRotation.java:42-92).
__device__ void givens(double x, double y, double &c, double &s) {
  double roe = fabs(x) > fabs(y) ? x : y;
  double scale = fabs(x) + fabs(y);
  if (scale != 0) {
    double xs = x / scale;
    double ys = y / scale;
    double r = scale * sqrt(xs * xs + ys * ys);
    if (roe < 0) r *= -1;
    c = x / r;
    s = y / r;
  } else {
    c = 1;
    s = 0;
  }
}
__device__ void givens_apply(double c, double s, double &a1, double &a2) {
  double temp = c * a1 + s * a2;
  a2 = -s * a1 + c * a2;
  a1 = temp;
}

// Last MGS subtraction w += (-H(i,i)) v_i, then H(i+1,i) = ||w||
// (S/GMRES.java:153); the last workgroup applies the previous rotations to
// column i, forms rotation i, rotates H and s (:157-162), advances i and the
// iteration counter, and runs the next inner check (:144).
template <bool VEC = false>
__global__ __launch_bounds__(BS) void k_gmres_arnoldi_end(int64_t n, double *w, const double *vi, int i,
                                                          KState *st, GmresDev g, GridRed red) {
  if (st->status != ST_RUNNING) return;
  const int ld = g.m + 1;
  const double h = -g.H[i + (size_t)i * ld];
  dd v[1] = {{0.0, 0.0}};
  if (VEC) {
    VEC_TILES(tile) {
      double wv[VPT], pv[VPT];
      if (h != 0) {
        vload_tile(n, tile, VIo{w, wv}, VIo{vi, pv});
#pragma unroll
        for (int u = 0; u < VPT; ++u) wv[u] = wv[u] + h * pv[u];
        vstore_tile(n, tile, w, wv);
      } else {
        vload_tile(n, tile, VIo{w, wv});
      }
#pragma unroll
      for (int u = 0; u < VPT; ++u)
        if (vslot_ok(n, tile, u)) dd_fma(v[0], wv[u], wv[u]);
    }
  } else VLOOP(e) {
    double we = w[e];
    if (h != 0) {
      we = we + h * vi[e];
      w[e] = we;
    }
    dd_fma(v[0], we, we);
  }
  double tot[1];
  if (grid_sum_dd<1>(v, red, tot) && threadIdx.x == 0) {
    double *Hc = g.H + (size_t)i * ld;
    double *s = g.gs, *rc = g.gs + ld, *rs = g.gs + 2 * ld;
    Hc[i + 1] = sqrt(tot[0]);
    st->s[2] = Hc[i + 1];  // ||w||, used to scale v_{i+1} (:154)
    for (int k = 0; k < i; ++k) givens_apply(rc[k], rs[k], Hc[k], Hc[k + 1]);
    givens(Hc[i], Hc[i + 1], rc[i], rs[i]);
    givens_apply(rc[i], rs[i], Hc[i], Hc[i + 1]);
    givens_apply(rc[i], rs[i], s[i], s[i + 1]);
    st->inner_i = i + 1;
    st->iter++;
    if (i + 1 < g.m) {
      int stt = monitor_check(st, fabs(s[i + 1]));
      if (stt == ST_CONVERGED) st->cycle_done = 1;
      else if (stt != ST_RUNNING) st->status = stt;
    } else {
      st->cycle_done = 1;
    }
  }
}

// v_{i+1} = (1/H(i+1,i)) w (S/GMRES.java:154).
template <bool VEC = false>
__global__ __launch_bounds__(BS) void k_gmres_scale(int64_t n, const double *w, double *vnext, const KState *st) {
  if (st->status != ST_RUNNING) return;
  const double a = 1. / st->s[2];
  if (VEC) {
    VEC_TILES(tile) {
      double wv[VPT];
      vload_tile(n, tile, VIo{w, wv});
#pragma unroll
      for (int u = 0; u < VPT; ++u) wv[u] = (a == 0) ? 0.0 : a * wv[u];
      vstore_tile(n, tile, vnext, wv);
    }
    return;
  }
  VLOOP(e)
    vnext[e] = (a == 0) ? 0.0 : a * w[e];
}

// new UpperTriangDenseMatrix(H, i, false).solve(s, s) (S/GMRES.java:166)
// through dtrtrs: singular on an exactly zero diagonal
// (M/AbstractTriangDenseMatrix.java:247-252), else reference-BLAS dtrsm
// column-order back-substitution.
__global__ void k_gmres_backsolve(KState *st, GmresDev g) {
  if (st->status != ST_RUNNING) return;
  const int k = st->inner_i, ld = g.m + 1;
  double *s = g.gs;
  for (int i = 0; i < k; ++i)
    if (g.H[i + (size_t)i * ld] == 0) {
      st->status = ST_SINGULAR;
      return;
    }
  for (int kk = k - 1; kk >= 0; --kk) {
    if (s[kk] != 0) {
      s[kk] = s[kk] / g.H[kk + (size_t)kk * ld];
      for (int i = 0; i < kk; ++i) s[i] = s[i] - s[kk] * g.H[i + (size_t)kk * ld];
    }
  }
}

// x.add(s_j, v_j) for j < i, in order (S/GMRES.java:167-168).
__global__ __launch_bounds__(BS) void k_gmres_xupdate(int64_t n, double *x, const double *V, const KState *st,
                                                      GmresDev g) {
  if (st->status != ST_RUNNING) return;
  const int k = st->inner_i;
  const double *s = g.gs;
  VLOOP(e) {
    double xe = x[e];
    for (int j = 0; j < k; ++j) {
      const double sj = s[j];
      if (sj != 0) xe = xe + sj * V[(size_t)j * n + e];
    }
    x[e] = xe;
  }
}

// ---------------------------------------------------------------------------
// IR (S/IR.java:57-70) and shared helpers

// ||r|| and the monitor check; with advance, iter.next() first.
__global__ __launch_bounds__(BS) void k_norm_check(int64_t n, const double *r, KState *st, int advance,
                                                   GridRed red) {
  if (st->status != ST_RUNNING) return;
  dd v[1] = {{0.0, 0.0}};
  VLOOP(i)
    dd_fma(v[0], r[i], r[i]);
  double tot[1];
  if (grid_sum_dd<1>(v, red, tot) && threadIdx.x == 0) {
    if (advance) st->iter++;
    int stt = monitor_check(st, sqrt(tot[0]));
    if (stt != ST_RUNNING) st->status = stt;
    st->normr = sqrt(tot[0]);
  }
}

__global__ __launch_bounds__(BS) void k_add_if_running(int64_t n, double *x, const double *z, const KState *st) {
  if (st->status != ST_RUNNING) return;
  VLOOP(i)
    x[i] = x[i] + z[i];
}

// Mixed precision: r32 = (float)(r / ||r||); x += ||r|| * (double)d32.
__global__ __launch_bounds__(BS) void k_demote_scaled(int64_t n, const double *r, float *r32, const KState *st) {
  if (st->status != ST_RUNNING) return;
  const double sc = 1.0 / st->normr;
  VLOOP(i)
    r32[i] = (float)(r[i] * sc);
}

__global__ __launch_bounds__(BS) void k_promote_add(int64_t n, double *x, const float *d32, const KState *st) {
  if (st->status != ST_RUNNING) return;
  const double sc = st->normr;
  VLOOP(i)
    x[i] = x[i] + sc * (double)d32[i];
}

__global__ __launch_bounds__(BS) void k_invdiag32(int64_t n, const int32_t *rp, const int32_t *ci,
                                                  const double *val, float *inv) {
  VLOOP(i) {
    double d = 0;
    for (int32_t j = rp[i]; j < rp[i + 1]; ++j)
      if (ci[j] == (int32_t)i) d = val[j];
    inv[i] = (float)(1 / d);
  }
}

__global__ void k_copy_f32_zero(int64_t n, const float *a, float *b, float *z) {
  VLOOP(i) {
    b[i] = a[i];
    z[i] = 0.0f;
  }
}

// ---------------------------------------------------------------------------
// host driver helpers

void state_init(mtjg_solver S, KState *dst, int32_t max_iter, double rtol, double atol, double dtol) {
  KState &h = S->h_st[dst == S->st ? 0 : 1];
  memset(&h, 0, sizeof(KState));
  h.max_iter = max_iter;
  h.rtol = rtol;
  h.atol = atol;
  h.dtol = dtol;
  h.alpha = 0;
  h.rho_1 = 0;
  h.status = ST_RUNNING;
  HIPCHK(hipMemcpyAsync(dst, &h, sizeof(KState), hipMemcpyHostToDevice, S->ctx->stream));
}

void state_read(mtjg_solver S, KState *src, KState &out) {
  KState *h = &S->h_st[src == S->st ? 0 : 1];
  HIPCHK(hipMemcpyAsync(h, src, sizeof(KState), hipMemcpyDeviceToHost, S->ctx->stream));
  HIPCHK(hipStreamSynchronize(S->ctx->stream));
  out = *h;
}

// Queues status reads behind each chunk of iterations, keeping up to two
// chunks in flight, until the status word leaves "running".
template <typename Launch>
void run_chunks(mtjg_solver S, KState *st, int chunk, Launch &&launch_iteration) {
  hipStream_t s = S->ctx->stream;
  if (S->rep) chunk = 1;
  auto enqueue = [&](int slot) {
    for (int c = 0; c < chunk; ++c) launch_iteration();
    HIPCHK(hipMemcpyAsync(&S->h_slot[slot], &st->status, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    HIPCHK(hipEventRecord(S->ev[slot], s));
  };
  if (S->rep) {
    for (;;) {
      enqueue(0);
      HIPCHK(hipEventSynchronize(S->ev[0]));
      KState k;
      state_read(S, st, k);
      S->rep(k.residual, k.iter, S->rep_user);
      if (k.status != ST_RUNNING) return;
    }
  }
  enqueue(0);
  enqueue(1);
  for (int k = 0;; ++k) {
    int slot = k & 1;
    HIPCHK(hipEventSynchronize(S->ev[slot]));
    if (S->h_slot[slot] != ST_RUNNING) break;
    enqueue(slot);
  }
  HIPCHK(hipStreamSynchronize(s));
}

int status_to_code(int st) {
  switch (st) {
    case ST_CONVERGED: return MTJG_OK;
    case ST_ITERATIONS: return MTJG_NOT_CONVERGED_ITERATIONS;
    case ST_DIVERGENCE: return MTJG_NOT_CONVERGED_DIVERGENCE;
    case ST_BREAKDOWN: return MTJG_NOT_CONVERGED_BREAKDOWN;
    case ST_SINGULAR: return MTJG_E_SINGULAR;
    default: return MTJG_E_DEVICE;
  }
}

double *solver_vec(mtjg_solver S, size_t k) {
  while (S->v64.size() <= k) {
    double *p = nullptr;
    HIPCHK(hipMalloc(&p, sizeof(double) * (size_t)std::max<int64_t>(S->n, 1)));
    HIPCHK(hipMemsetAsync(p, 0, sizeof(double) * (size_t)std::max<int64_t>(S->n, 1), S->ctx->stream));
    S->v64.push_back(p);
  }
  return S->v64[k];
}

float *scratch32(mtjg_solver S, size_t k) {
  while (S->v32.size() <= k) {
    float *p = nullptr;
    HIPCHK(hipMalloc(&p, sizeof(float) * (size_t)std::max<int64_t>(S->n, 1)));
    S->v32.push_back(p);
  }
  return S->v32[k];
}

// Iterations queued per status read: aims at a few hundred microseconds of
// device work per chunk, fewer for large systems.
void free_red(GridRed &r) {
  (void)hipFree(r.partials);
  (void)hipFree(r.gpart);
  (void)hipFree(r.gticket);
  r = GridRed{};
}

// Sizes the reduction workspace for the largest grid a solve of A launches:
// the vector kernels' vgrid(n) and the fused product's one workgroup per
// row-block.  Counters start at zero and are left at zero by every launch.
// Every reducing kernel of a solve runs at most VGRID_MAX (vector kernels)
// workgroups; the fused product's need comes from spmv_dot_red_cap.
void ensure_red(mtjg_solver S, mtjg_csr A) {
  const int need = (int)std::max<int64_t>(VGRID_MAX, spmv_dot_red_cap(A));
  if (need <= S->red.cap) return;
  HIPCHK(hipStreamSynchronize(S->ctx->stream));
  free_red(S->red);
  GridRed r;
  r.cap = need;
  r.gcap = (need + RED_GROUP - 1) / RED_GROUP;
  HIPCHK(hipMalloc(&r.partials, sizeof(double) * RED_MAX_K * (size_t)r.cap));
  HIPCHK(hipMalloc(&r.gpart, sizeof(double) * RED_MAX_K * (size_t)r.gcap));
  HIPCHK(hipMalloc(&r.gticket, sizeof(uint32_t) * red_ticket_words(r.gcap)));
  HIPCHK(hipMemsetAsync(r.gticket, 0, sizeof(uint32_t) * red_ticket_words(r.gcap), S->ctx->stream));
  r.fticket = r.gticket + (size_t)r.gcap * TICKET_STRIDE;
  S->red = r;
}

int chunk_for(mtjg_csr A) {
  double bytes = 12.0 * (double)A->nnz + 108.0 * (double)A->nr;
  int c = (int)(2.0e8 / std::max(bytes, 1.0));
  return std::max(1, std::min(c, 16));
}

int prec_k(mtjg_prec M) {
  int k = prec_kind(M);
  if (k == PREC_IDENTITY) return PK_IDENTITY;
  if (k == PREC_DIAGONAL) return PK_JACOBI;
  return PK_EXTERNAL;
}

// ---------------------------------------------------------------------------
// CG (S/CG.java:58-87)

// Workspace of one CG solve: the scratch vectors (S/CG.java:51-56), the
// device solver state and the reduction workspace.  p2 is the second
// direction buffer of the fused product (equal to p when it is not used).
struct CgWork {
  double *p, *z, *q, *r, *p2;
  KState *st;
  GridRed red;
};

// S/CG.java:64-86 on the device, from an initialised state: r = b - A x,
// the first residual step, then iterations handed to `run(chunk, iteration)`,
// which queues them until the status word leaves running.
template <typename Run>
void cg_iterate(hipStream_t s, int64_t n, mtjg_csr A, mtjg_prec M, double *b, double *x, const CgWork &w,
                Run &&run) {
  double *p = w.p, *z = w.z, *q = w.q, *r = w.r;
  const int pk = prec_k(M);
  const double *invd = prec_invdiag(M);
  if (pk == PK_IDENTITY) z = r;  // identity: z = r bitwise, no copy needed
  // r = b - A x (S/CG.java:64)
  SpmvArgs ra;
  ra.mode = SPMV_MULTADD;
  ra.alpha = -1;
  ra.y0 = b;
  spmv_f64(A, x, r, ra, s);
  const int g = vgrid(n);
  const bool nt = n * (int64_t)sizeof(double) >= NT_VECTOR_BYTES;
  const bool vec = aligned16({x, w.p, w.p2, r, q, invd, z});
  KState *st = w.st;
  const GridRed &red = w.red;
  auto resid = [&](bool update) {
    if (pk == PK_IDENTITY) {
      if (update) launch_cg_resid<double, PK_IDENTITY, true>(false, vec, g, s, n, x, p, r, q, invd, z, st, red);
      else launch_cg_resid<double, PK_IDENTITY, false>(false, vec, g, s, n, x, p, r, q, invd, z, st, red);
    } else if (pk == PK_JACOBI) {
      if (update) launch_cg_resid<double, PK_JACOBI, true>(nt, vec, g, s, n, x, p, r, q, invd, z, st, red);
      else launch_cg_resid<double, PK_JACOBI, false>(nt, vec, g, s, n, x, p, r, q, invd, z, st, red);
    } else {
      if (update) launch_cg_resid<double, PK_EXTERNAL, true>(false, vec, g, s, n, x, p, r, q, invd, z, st, red);
      else launch_cg_resid<double, PK_EXTERNAL, false>(false, vec, g, s, n, x, p, r, q, invd, z, st, red);
      prec_apply_dev(M, n, r, z, s, st);
      k_cg_rho<double><<<g, BS, 0, s>>>(n, r, z, st, red);
    }
  };
  resid(false);
  HIPCHK(hipGetLastError());
  SpmvArgs qa;
  qa.mode = SPMV_DOT;
  qa.w = p;
  qa.st = st;
  qa.fin = FIN_CG_ALPHA;
  qa.red = red;
  // For SELL-64 matrices whose fused product runs one persistent grid the
  // update p = z + beta p rides in the product (spmv_cg_fused), writing p'
  // to the other of two buffers each iteration.
  const bool fused = spmv_cg_fusable(A) && w.p2 != w.p;
  double *pbuf[2] = {p, fused ? w.p2 : p};
  int cur = 0;
  run(pk == PK_EXTERNAL ? 1 : chunk_for(A), [&]() {
    if (fused && spmv_cg_fused(A, pbuf[cur], z, pbuf[cur ^ 1], q, qa, s)) {
      cur ^= 1;
      p = pbuf[cur];
    } else {
      launch_cg_p<double>(nt, vec, g, s, n, p, z, st);
      spmv_f64(A, p, q, qa, s);
    }
    resid(true);
    HIPCHK(hipGetLastError());
  });
}

void solve_cg(mtjg_solver S, mtjg_csr A, double *b, double *x) {
  CgWork w;
  w.p = solver_vec(S, 0);
  w.z = solver_vec(S, 1);
  w.q = solver_vec(S, 2);
  w.r = solver_vec(S, 3);
  w.p2 = spmv_cg_fusable(A) ? solver_vec(S, 4) : w.p;
  state_init(S, S->st, S->mon.max_iter, S->mon.rtol, S->mon.atol, S->mon.dtol);
  ensure_red(S, A);
  w.st = S->st;
  w.red = S->red;
  cg_iterate(S->ctx->stream, S->n, A, S->M, b, x, w,
             [&](int chunk, auto &&iteration) { run_chunks(S, S->st, chunk, iteration); });
}

// ---------------------------------------------------------------------------
// BiCGstab (S/BiCGstab.java:65-116)

void solve_bicgstab(mtjg_solver S, mtjg_csr A, double *b, double *x) {
  hipStream_t s = S->ctx->stream;
  const int64_t n = S->n;
  double *p = solver_vec(S, 0), *sv = solver_vec(S, 1), *phat = solver_vec(S, 2), *shat = solver_vec(S, 3),
         *t = solver_vec(S, 4), *v = solver_vec(S, 5), *r = solver_vec(S, 6), *rtilde = solver_vec(S, 7);
  const int pk = prec_k(S->M);
  if (pk == PK_IDENTITY) {
    phat = p;  // M.apply = x.set(b): the copy is exact
    shat = sv;
  }
  KState *st = S->st;
  state_init(S, st, S->mon.max_iter, S->mon.rtol, S->mon.atol, S->mon.dtol);
  // omega = alpha = rho_1 = rho_2 = 1 (S/BiCGstab.java:69)
  {
    KState &h = S->h_st[0];
    h.rho_1 = h.rho_2 = h.alpha = h.beta = h.omega = 1;
    HIPCHK(hipMemcpyAsync(st, &h, sizeof(KState), hipMemcpyHostToDevice, s));
  }
  SpmvArgs ra;
  ra.mode = SPMV_MULTADD;
  ra.alpha = -1;
  ra.y0 = b;
  spmv_f64(A, x, r, ra, s);
  if (n) HIPCHK(hipMemcpyAsync(rtilde, r, sizeof(double) * (size_t)n, hipMemcpyDeviceToDevice, s));
  const int g = vgrid(n);
  ensure_red(S, A);
  const bool vec = aligned16({x, shat, r, sv, t, rtilde, p, phat, v});
  if (vec) k_bicg_top<false, true><<<g, BS, 0, s>>>(n, x, shat, r, sv, t, rtilde, st, S->red);
  else k_bicg_top<false><<<g, BS, 0, s>>>(n, x, shat, r, sv, t, rtilde, st, S->red);
  HIPCHK(hipGetLastError());
  SpmvArgs va;
  va.mode = SPMV_DOT;
  va.w = rtilde;
  va.st = st;
  va.fin = FIN_BICG_ALPHA;
  va.red = S->red;
  SpmvArgs ta = va;
  ta.w = sv;
  ta.fin = FIN_BICG_OMEGA;
  run_chunks(S, st, pk == PK_EXTERNAL ? 1 : chunk_for(A), [&]() {
    if (vec) k_bicg_p<true><<<g, BS, 0, s>>>(n, p, r, v, st);
    else k_bicg_p<<<g, BS, 0, s>>>(n, p, r, v, st);
    if (pk != PK_IDENTITY) prec_apply_dev(S->M, n, p, phat, s, st);
    spmv_f64(A, phat, v, va, s);
    if (vec) k_bicg_s<true><<<g, BS, 0, s>>>(n, x, phat, r, v, sv, st, S->red);
    else k_bicg_s<<<g, BS, 0, s>>>(n, x, phat, r, v, sv, st, S->red);
    if (pk != PK_IDENTITY) prec_apply_dev(S->M, n, sv, shat, s, st);
    spmv_f64(A, shat, t, ta, s);
    k_bicg_omega<<<1, 1, 0, s>>>(st);
    if (vec) k_bicg_top<true, true><<<g, BS, 0, s>>>(n, x, shat, r, sv, t, rtilde, st, S->red);
    else k_bicg_top<true><<<g, BS, 0, s>>>(n, x, shat, r, sv, t, rtilde, st, S->red);
    HIPCHK(hipGetLastError());
  });
}

// ---------------------------------------------------------------------------
// GMRES (S/GMRES.java:127-176); the host reads the state after each inner
// step to know whether the cycle has ended.

void gmres_alloc(mtjg_solver S) {
  if (S->V) return;
  const int m = S->restart;
  HIPCHK(hipMalloc(&S->V, sizeof(double) * (size_t)(m + 1) * (size_t)std::max<int64_t>(S->n, 1)));
  HIPCHK(hipMemsetAsync(S->V, 0, sizeof(double) * (size_t)(m + 1) * (size_t)std::max<int64_t>(S->n, 1),
                        S->ctx->stream));
  HIPCHK(hipMalloc(&S->H, sizeof(double) * (size_t)(m + 1) * (size_t)m));
  HIPCHK(hipMemsetAsync(S->H, 0, sizeof(double) * (size_t)(m + 1) * (size_t)m, S->ctx->stream));
  HIPCHK(hipMalloc(&S->gs, sizeof(double) * 3 * (size_t)(m + 1)));
  HIPCHK(hipMemsetAsync(S->gs, 0, sizeof(double) * 3 * (size_t)(m + 1), S->ctx->stream));
}

void gmres_free(mtjg_solver S) {
  (void)hipFree(S->V);
  (void)hipFree(S->H);
  (void)hipFree(S->gs);
  (void)hipFree(S->mgs_bar);
  (void)hipFree(S->mgs_err);
  S->V = S->H = S->gs = nullptr;
  S->mgs_bar = nullptr;
  S->mgs_err = nullptr;
  S->mgs_coop = 0;
}

// k_gmres_mgs_all for nt tiles per workgroup (1 or 2), else null (more
// tiles would not fit the 64 registers the resident grid allows).
static const void *mgs_all_kernel(int64_t nt) {
  switch (nt) {
    case 1: return (const void *)k_gmres_mgs_all<1>;
    case 2: return (const void *)k_gmres_mgs_all<2>;
    default: return nullptr;
  }
}

// Whether kernel kf (a k_gmres_mgs_all) can run on grid g: cooperative
// launches supported and g workgroups resident at once.  Decided once per
// solver (its size is fixed); allocates the step counter and flag then.
// MTJG_GMRES_COOP=0 keeps the per-step launches; MTJG_GMRES_POLL_BUDGET sets
// the barrier's poll budget (default 1 << 22).
static bool mgs_coop_ok(mtjg_solver S, const void *kf, int g) {
  if (S->mgs_coop == 0) {
    S->mgs_coop = -1;
    if (const char *e = getenv("MTJG_GMRES_COOP"))
      if (atoi(e) == 0) return false;
    if (!kf) return false;
    int dev = 0, coop = 0, cus = 0, per_cu = 0;
    HIPCHK(hipGetDevice(&dev));
    HIPCHK(hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, dev));
    HIPCHK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kf, BS, 0));
    if (coop && (int64_t)cus * per_cu >= g) {
      S->mgs_budget = 1LL << 22;
      if (const char *e = getenv("MTJG_GMRES_POLL_BUDGET")) S->mgs_budget = atoll(e);
      HIPCHK(hipMalloc(&S->mgs_bar, sizeof(uint32_t)));
      HIPCHK(hipMalloc(&S->mgs_err, sizeof(int32_t)));
      HIPCHK(hipMemsetAsync(S->mgs_bar, 0, sizeof(uint32_t), S->ctx->stream));
      HIPCHK(hipMemsetAsync(S->mgs_err, 0, sizeof(int32_t), S->ctx->stream));
      S->mgs_coop = 1;
    }
  }
  return S->mgs_coop > 0;
}

// Raises if a cooperative Gram-Schmidt barrier ran out of its poll budget
// during the last solve (synchronises the stream).  The workgroups that left
// early abandoned the reduction's tickets part-taken: they are cleared here.
static void gmres_check(mtjg_solver S) {
  if (!S->mgs_err) return;
  int32_t e = 0;
  HIPCHK(hipMemcpyAsync(&e, S->mgs_err, sizeof(int32_t), hipMemcpyDeviceToHost, S->ctx->stream));
  HIPCHK(hipStreamSynchronize(S->ctx->stream));
  if (!e) return;
  HIPCHK(hipMemsetAsync(S->mgs_err, 0, sizeof(int32_t), S->ctx->stream));
  HIPCHK(hipMemsetAsync(S->red.gticket, 0, sizeof(uint32_t) * red_ticket_words(S->red.gcap), S->ctx->stream));
  HIPCHK(hipStreamSynchronize(S->ctx->stream));
  fail(MTJG_E_DEVICE, "GMRES Gram-Schmidt grid barrier exceeded its poll budget");
}

void solve_gmres(mtjg_solver S, mtjg_csr A, double *b, double *x) {
  hipStream_t s = S->ctx->stream;
  const int64_t n = S->n;
  gmres_alloc(S);
  double *w = solver_vec(S, 0), *u = solver_vec(S, 1), *r = solver_vec(S, 2);
  KState *st = S->st;
  state_init(S, st, S->mon.max_iter, S->mon.rtol, S->mon.atol, S->mon.dtol);
  GmresDev gd{S->H, S->gs, S->restart};
  const int g = vgrid(n);
  ensure_red(S, A);
  auto Vk = [&](int k) { return S->V + (size_t)k * (size_t)n; };
  // the basis vectors are 16-byte aligned when n is even (the basis is one allocation)
  const bool vec = aligned16({w, u, r, S->V}) && n % 2 == 0;
  // the whole Gram-Schmidt loop per cooperative launch when every workgroup
  // holds one tile (the grid's tile order is then the per-step kernels')
  const void *mgs_kf = mgs_all_kernel((n + (int64_t)g * (BS * VPT) - 1) / ((int64_t)g * (BS * VPT)));
  const bool coop = vec && mgs_coop_ok(S, mgs_kf, g);
  SpmvArgs ra;
  ra.mode = SPMV_MULTADD;
  ra.alpha = -1;
  ra.y0 = b;
  SpmvArgs ma;
  ma.mode = SPMV_MULT;
  auto residual = [&]() {
    // u = b - A x; r = M^-1 u (S/GMRES.java:131-132, :170-171)
    spmv_f64(A, x, u, ra, s);
    prec_apply_dev(S->M, n, u, r, s, nullptr);
  };
  residual();
  k_gmres_start<<<g, BS, 0, s>>>(n, r, st, gd, 1, S->red);
  HIPCHK(hipGetLastError());
  KState k;
  for (;;) {
    state_read(S, st, k);
    if (k.status != ST_RUNNING) break;
    if (vec) k_gmres_v0<true><<<g, BS, 0, s>>>(n, r, Vk(0), st);
    else k_gmres_v0<<<g, BS, 0, s>>>(n, r, Vk(0), st);
    for (int i = 0; i < S->restart && !k.cycle_done; ++i) {
      spmv_f64(A, Vk(i), u, ma, s);
      prec_apply_dev(S->M, n, u, w, s, nullptr);
      if (coop) {
        int64_t nn = n;
        const double *V0 = S->V;
        void *args[] = {(void *)&nn,     (void *)&w,          (void *)&V0,         (void *)&i,
                        (void *)&st,     (void *)&gd,         (void *)&S->red,     (void *)&S->mgs_bar,
                        (void *)&S->mgs_budget, (void *)&S->mgs_err};
        HIPCHK(hipLaunchCooperativeKernel(mgs_kf, dim3(g), dim3(BS), args, 0, s));
      } else {
        for (int kk = 0; kk <= i; ++kk)
          if (vec)
            k_gmres_mgs<true><<<g, BS, 0, s>>>(n, w, kk > 0 ? Vk(kk - 1) : nullptr, Vk(kk), kk, i, st, gd,
                                                S->red);
          else
            k_gmres_mgs<<<g, BS, 0, s>>>(n, w, kk > 0 ? Vk(kk - 1) : nullptr, Vk(kk), kk, i, st, gd, S->red);
      }
      if (vec) {
        k_gmres_arnoldi_end<true><<<g, BS, 0, s>>>(n, w, Vk(i), i, st, gd, S->red);
        k_gmres_scale<true><<<g, BS, 0, s>>>(n, w, Vk(i + 1), st);
      } else {
        k_gmres_arnoldi_end<<<g, BS, 0, s>>>(n, w, Vk(i), i, st, gd, S->red);
        k_gmres_scale<<<g, BS, 0, s>>>(n, w, Vk(i + 1), st);
      }
      HIPCHK(hipGetLastError());
      state_read(S, st, k);
      if (k.status != ST_RUNNING) break;
    }
    if (k.status != ST_RUNNING) break;
    k_gmres_backsolve<<<1, 1, 0, s>>>(st, gd);
    k_gmres_xupdate<<<g, BS, 0, s>>>(n, x, S->V, st, gd);
    residual();
    k_gmres_start<<<g, BS, 0, s>>>(n, r, st, gd, 0, S->red);
    HIPCHK(hipGetLastError());
  }
}

// ---------------------------------------------------------------------------
// IR (S/IR.java:57-70)

void solve_ir(mtjg_solver S, mtjg_csr A, double *b, double *x) {
  hipStream_t s = S->ctx->stream;
  const int64_t n = S->n;
  double *z = solver_vec(S, 0), *r = solver_vec(S, 1);
  KState *st = S->st;
  state_init(S, st, S->mon.max_iter, S->mon.rtol, S->mon.atol, S->mon.dtol);
  SpmvArgs ra;
  ra.mode = SPMV_MULTADD;
  ra.alpha = -1;
  ra.y0 = b;
  const int g = vgrid(n);
  ensure_red(S, A);
  spmv_f64(A, x, r, ra, s);
  k_norm_check<<<g, BS, 0, s>>>(n, r, st, 0, S->red);
  HIPCHK(hipGetLastError());
  run_chunks(S, st, 1, [&]() {
    prec_apply_dev(S->M, n, r, z, s, st);
    k_add_if_running<<<g, BS, 0, s>>>(n, x, z, st);
    spmv_f64(A, x, r, ra, s);
    k_norm_check<<<g, BS, 0, s>>>(n, r, st, 1, S->red);
    HIPCHK(hipGetLastError());
  });
}

// ---------------------------------------------------------------------------
// Mixed precision: FP64 refinement x += M32^-1 (b - A x) where M32^-1 is an
// FP32 Jacobi-CG solve on the FP32 copy of A's values, run on the residual
// scaled to unit norm.  The FP64 residual uses the FP64 values, so the outer
// tolerance is met for A itself.  Each inner solve stops at the larger of the
// set inner tolerance and half the reduction the outer residual still lacks,
// so the last correction is not solved further than the outer test needs.

void solve_mixed(mtjg_solver S, mtjg_csr A, double *b, double *x) {
  hipStream_t s = S->ctx->stream;
  const int64_t n = S->n;
  double *r = solver_vec(S, 0);
  float *r32 = scratch32(S, 0), *d32 = scratch32(S, 1), *p32 = scratch32(S, 2), *q32 = scratch32(S, 3),
        *z32 = scratch32(S, 4), *rin = scratch32(S, 5), *inv32 = scratch32(S, 6);
  csr_ensure_fp32(A);
  const int g = vgrid(n);
  const bool vec32 = aligned16({d32, p32, rin, q32, inv32, z32});
  ensure_red(S, A);
  k_invdiag32<<<g, BS, 0, s>>>(n, A->rp, A->ci, A->val, inv32);
  KState *st = S->st, *si = S->st_in;
  state_init(S, st, S->mon.max_iter, S->mon.rtol, S->mon.atol, S->mon.dtol);
  SpmvArgs ra;
  ra.mode = SPMV_MULTADD;
  ra.alpha = -1;
  ra.y0 = b;
  spmv_f64(A, x, r, ra, s);
  k_norm_check<<<g, BS, 0, s>>>(n, r, st, 0, S->red);
  HIPCHK(hipGetLastError());
  SpmvArgs qa;
  qa.mode = SPMV_DOT;
  qa.w = p32;
  qa.st = si;
  qa.fin = FIN_CG_ALPHA;
  qa.red = S->red;
  S->inner_total = 0;
  const int chunk = chunk_for(A);
  for (;;) {
    KState k;
    state_read(S, st, k);
    if (S->rep) S->rep(k.residual, k.iter, S->rep_user);
    if (k.status != ST_RUNNING) break;
    // inner FP32 Jacobi-CG on A32 d = r/||r||, from d = 0
    k_demote_scaled<<<g, BS, 0, s>>>(n, r, r32, st);
    k_copy_f32_zero<<<g, BS, 0, s>>>(n, r32, rin, d32);
    const double target = std::max(k.rtol * k.init_r, k.atol);
    double itol = S->inner_rtol;
    if (k.residual > 0 && std::isfinite(k.residual)) itol = std::max(itol, std::min(0.5, 0.5 * target / k.residual));
    state_init(S, si, S->inner_max, itol, 0.0, 1e300);
    launch_cg_resid<float, PK_JACOBI, false>(false, vec32, g, s, n, d32, p32, rin, q32, inv32, z32, si, S->red);
    HIPCHK(hipGetLastError());
    run_chunks(S, si, chunk, [&]() {
      launch_cg_p<float>(false, vec32, g, s, n, p32, z32, si);
      spmv_f32(A, p32, q32, qa, s);
      launch_cg_resid<float, PK_JACOBI, true>(false, vec32, g, s, n, d32, p32, rin, q32, inv32, z32, si, S->red);
      HIPCHK(hipGetLastError());
    });
    KState ki;
    state_read(S, si, ki);
    S->inner_total += ki.iter;
    // x += ||r|| d; r = b - A x; outer check
    k_promote_add<<<g, BS, 0, s>>>(n, x, d32, st);
    spmv_f64(A, x, r, ra, s);
    k_norm_check<<<g, BS, 0, s>>>(n, r, st, 1, S->red);
    HIPCHK(hipGetLastError());
  }
}

// ---------------------------------------------------------------------------
// The FP32 inner-CG preconditioner (MTJG_PREC_CG32): apply(b, x) is the inner
// solve of solve_mixed as a Preconditioner (S/Preconditioner.java:42), so
// that MTJ's own IR (S/IR.java:57-70: r = b - A x, M.apply(r, z), x += z)
// performs the FP64 refinement.  x = ||b|| d, where d solves A32 d =
// (float)(b / ||b||) by Jacobi-CG from d = 0 on the FP32 copy of A's values
// with an FP32 inverse diagonal, to the fixed relative tolerance rtol (at
// most max_iter iterations).  The workspace is an internal solver object.

// x = ||b|| * (double)d32 (the inner solution scaled back, FP64).
__global__ __launch_bounds__(BS) void k_promote_set(int64_t n, double *x, const float *d32, const KState *st) {
  const double sc = st->normr;
  VLOOP(i)
    x[i] = sc * (double)d32[i];
}

// FP32 inverse diagonal; the first zero or structurally missing diagonal is
// recorded in *bad (S/DiagonalPreconditioner.java:69-75).
__global__ __launch_bounds__(BS) void k_invdiag32_check(int64_t n, const int32_t *rp, const int32_t *ci,
                                                        const double *val, float *inv, int32_t *bad) {
  VLOOP(i) {
    double d = 0;
    for (int32_t j = rp[i]; j < rp[i + 1]; ++j)
      if (ci[j] == (int32_t)i) d = val[j];
    if (d == 0) atomicMin(bad, (int32_t)i);
    inv[i] = (float)(1 / d);
  }
}

}  // namespace

namespace mtjg {

mtjg_solver cg32_create(mtjg_ctx ctx, int32_t n, double rtol, int32_t max_iter) {
  mtjg_vec_s t;
  t.ctx = ctx;
  t.n = n;
  mtjg_solver S = nullptr;
  const int rc = mtjg_solver_create(ctx, MTJG_MIXED_CG, &t, &S);
  if (rc != MTJG_OK) fail(rc, mtjg_last_error());
  S->inner_rtol = rtol;
  S->inner_max = max_iter;
  return S;
}

void cg32_set_matrix(mtjg_solver S, mtjg_csr A, int32_t *bad_row) {
  hipStream_t s = S->ctx->stream;
  const int64_t n = S->n;
  csr_ensure_fp32(A);
  ensure_red(S, A);
  int32_t *bad = (int32_t *)solver_vec(S, 1);  // n >= 1 doubles of room
  const int32_t none = INT32_MAX;
  HIPCHK(hipMemcpyAsync(bad, &none, sizeof(int32_t), hipMemcpyHostToDevice, s));
  k_invdiag32_check<<<vgrid(n), BS, 0, s>>>(n, A->rp, A->ci, A->val, scratch32(S, 6), bad);
  HIPCHK(hipGetLastError());
  int32_t b = none;
  HIPCHK(hipMemcpyAsync(&b, bad, sizeof(int32_t), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  if (b != none) {
    if (bad_row) *bad_row = b;
    fail(MTJG_E_ZERO_DIAGONAL, "Zero diagonal on row " + std::to_string(b + 1));
  }
}

void cg32_apply(mtjg_solver S, mtjg_csr A, const double *b, double *x, KState *outer) {
  hipStream_t s = S->ctx->stream;
  const int64_t n = S->n;
  if (outer) {
    // the outer solve's kernels skip work once it has finished; so does this
    KState *h = &S->h_st[1];
    HIPCHK(hipMemcpyAsync(&h->status, &outer->status, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    if (h->status != ST_RUNNING) return;
  }
  float *r32 = scratch32(S, 0), *d32 = scratch32(S, 1), *p32 = scratch32(S, 2), *q32 = scratch32(S, 3),
        *z32 = scratch32(S, 4), *rin = scratch32(S, 5), *inv32 = scratch32(S, 6);
  const int g = vgrid(n);
  const bool vec32 = aligned16({d32, p32, rin, q32, inv32, z32});
  KState *st = S->st, *si = S->st_in;
  // ||b|| (st->normr), with a monitor that cannot stop: rtol = atol = 0
  state_init(S, st, 1, 0.0, 0.0, 1e300);
  k_norm_check<<<g, BS, 0, s>>>(n, b, st, 0, S->red);
  HIPCHK(hipGetLastError());
  KState k;
  state_read(S, st, k);
  if (!(k.normr > 0) || !std::isfinite(k.normr)) {
    // b = 0 gives x = 0; a non-finite b gives NaN, which the outer monitor
    // reports as divergence
    launch_fill(s, n, x, k.normr == 0 ? 0.0 : std::numeric_limits<double>::quiet_NaN());
    return;
  }
  SpmvArgs qa;
  qa.mode = SPMV_DOT;
  qa.w = p32;
  qa.st = si;
  qa.fin = FIN_CG_ALPHA;
  qa.red = S->red;
  k_demote_scaled<<<g, BS, 0, s>>>(n, b, r32, st);
  k_copy_f32_zero<<<g, BS, 0, s>>>(n, r32, rin, d32);
  state_init(S, si, S->inner_max, S->inner_rtol, 0.0, 1e300);
  launch_cg_resid<float, PK_JACOBI, false>(false, vec32, g, s, n, d32, p32, rin, q32, inv32, z32, si, S->red);
  HIPCHK(hipGetLastError());
  run_chunks(S, si, chunk_for(A), [&]() {
    launch_cg_p<float>(false, vec32, g, s, n, p32, z32, si);
    spmv_f32(A, p32, q32, qa, s);
    launch_cg_resid<float, PK_JACOBI, true>(false, vec32, g, s, n, d32, p32, rin, q32, inv32, z32, si, S->red);
    HIPCHK(hipGetLastError());
  });
  KState ki;
  state_read(S, si, ki);
  S->inner_total += ki.iter;
  k_promote_set<<<g, BS, 0, s>>>(n, x, d32, st);
  HIPCHK(hipGetLastError());
}

int64_t cg32_iterations(mtjg_solver S) { return S ? S->inner_total : 0; }

void cg32_destroy(mtjg_solver S) { (void)mtjg_solver_destroy(S); }

}  // namespace mtjg

namespace {

// ---------------------------------------------------------------------------
// Transpose-using solvers: BiCG (S/BiCG.java:63-102), CGS (S/CGS.java:65-105)
// and QMR (S/QMR.java:117-213).  These drive the library's own vector,
// product and preconditioner entry points from the host, one device call per
// MTJ statement in MTJ's order, and read each dot and norm back (as MTJ's
// loops branch on them); the DefaultIterationMonitor predicate runs on the
// host.

// A device vector handle over solver scratch memory (not owned).
struct DV {
  mtjg_vec_s v;
  DV(mtjg_solver S, double *p) {
    v.ctx = S->ctx;
    v.n = S->n;
    v.dtype = MTJG_F64;
    v.ptr = p;
    v.owned = false;
  }
  mtjg_vec h() { return &v; }
};

void ck(int rc) {
  if (rc != MTJG_OK) fail(rc, mtjg_last_error());
}

double dot(DV &x, DV &y) {
  double r = 0;
  ck(mtjg_vec_dot(x.h(), y.h(), &r));
  return r;
}

double norm2(DV &x) {
  double r = 0;
  ck(mtjg_vec_norm(x.h(), MTJG_NORM_TWO, &r));
  return r;
}

// M.apply / M.transApply; the identity preconditioner copies (x.set(b)).
void papply(mtjg_solver S, DV &b, DV &x) {
  if (S->M) ck(mtjg_prec_apply(S->M, b.h(), x.h()));
  else ck(mtjg_vec_set(x.h(), b.h()));
}

void ptrans(mtjg_solver S, DV &b, DV &x) {
  if (S->M) ck(mtjg_prec_trans_apply(S->M, b.h(), x.h()));
  else ck(mtjg_vec_set(x.h(), b.h()));
}

// DefaultIterationMonitor on the host (S/DefaultIterationMonitor.java:
// 132-156, S/AbstractIterationMonitor.java:81-93).
struct HostMonitor {
  mtjg_solver S;
  KState k{};
  explicit HostMonitor(mtjg_solver s) : S(s) {
    k.max_iter = s->mon.max_iter;
    k.rtol = s->mon.rtol;
    k.atol = s->mon.atol;
    k.dtol = s->mon.dtol;
    k.status = ST_RUNNING;
  }
  // iter.converged(r): true when the loop must stop
  bool converged(double r) {
    k.residual = r;
    if (k.iter == 0) k.init_r = r;
    if (r < java_max(k.rtol * k.init_r, k.atol)) k.status = ST_CONVERGED;
    else if (r > k.dtol * k.init_r) k.status = ST_DIVERGENCE;
    else if (k.iter >= k.max_iter) k.status = ST_ITERATIONS;
    else if (r != r) k.status = ST_DIVERGENCE;
    if (S->rep) S->rep(r, k.iter, S->rep_user);
    return k.status != ST_RUNNING;
  }
  void breakdown() { k.status = ST_BREAKDOWN; }
  // publishes the outcome where mtjg_solver_solve reads it
  void finish() {
    HIPCHK(hipMemcpyAsync(S->st, &k, sizeof(KState), hipMemcpyHostToDevice, S->ctx->stream));
    HIPCHK(hipStreamSynchronize(S->ctx->stream));
  }
};

void solve_bicg(mtjg_solver S, mtjg_csr A, mtjg_vec b, mtjg_vec x) {
  DV z(S, solver_vec(S, 0)), p(S, solver_vec(S, 1)), q(S, solver_vec(S, 2)), r(S, solver_vec(S, 3)),
      zt(S, solver_vec(S, 4)), pt(S, solver_vec(S, 5)), qt(S, solver_vec(S, 6)), rt(S, solver_vec(S, 7));
  HostMonitor mon(S);
  double rho_1 = 1, rho_2 = 1, alpha = 1, beta = 1;
  ck(mtjg_vec_set(r.h(), b));
  ck(mtjg_csr_mult_add(A, -1, x, r.h()));
  ck(mtjg_vec_set(rt.h(), r.h()));
  for (mon.k.iter = 0; !mon.converged(norm2(r)); mon.k.iter++) {
    papply(S, r, z);
    ptrans(S, rt, zt);
    rho_1 = dot(z, rt);
    if (rho_1 == 0.) {
      mon.breakdown();
      break;
    }
    if (mon.k.iter == 0) {
      ck(mtjg_vec_set(p.h(), z.h()));
      ck(mtjg_vec_set(pt.h(), zt.h()));
    } else {
      beta = rho_1 / rho_2;
      ck(mtjg_vec_scale(p.h(), beta));
      ck(mtjg_vec_add(p.h(), z.h()));
      ck(mtjg_vec_scale(pt.h(), beta));
      ck(mtjg_vec_add(pt.h(), zt.h()));
    }
    ck(mtjg_csr_mult(A, p.h(), q.h()));
    ck(mtjg_csr_trans_mult(A, pt.h(), qt.h()));
    alpha = rho_1 / dot(pt, q);
    ck(mtjg_vec_add_scaled(x, alpha, p.h()));
    ck(mtjg_vec_add_scaled(r.h(), -alpha, q.h()));
    ck(mtjg_vec_add_scaled(rt.h(), -alpha, qt.h()));
    rho_2 = rho_1;
  }
  (void)beta;
  mon.finish();
}

void solve_cgs(mtjg_solver S, mtjg_csr A, mtjg_vec b, mtjg_vec x) {
  DV p(S, solver_vec(S, 0)), q(S, solver_vec(S, 1)), u(S, solver_vec(S, 2)), phat(S, solver_vec(S, 3)),
      qhat(S, solver_vec(S, 4)), vhat(S, solver_vec(S, 5)), uhat(S, solver_vec(S, 6)), sum(S, solver_vec(S, 7)),
      r(S, solver_vec(S, 8)), rt(S, solver_vec(S, 9));
  HostMonitor mon(S);
  double rho_1 = 0, rho_2 = 0, alpha = 0, beta = 0;
  ck(mtjg_vec_set(r.h(), b));
  ck(mtjg_csr_mult_add(A, -1, x, r.h()));
  ck(mtjg_vec_set(rt.h(), r.h()));
  for (mon.k.iter = 0; !mon.converged(norm2(r)); mon.k.iter++) {
    rho_1 = dot(rt, r);
    if (rho_1 == 0) {
      mon.breakdown();
      break;
    }
    if (mon.k.iter == 0) {
      ck(mtjg_vec_set(u.h(), r.h()));
      ck(mtjg_vec_set(p.h(), u.h()));
    } else {
      beta = rho_1 / rho_2;
      ck(mtjg_vec_set(u.h(), r.h()));
      ck(mtjg_vec_add_scaled(u.h(), beta, q.h()));
      ck(mtjg_vec_set(sum.h(), q.h()));
      ck(mtjg_vec_add_scaled(sum.h(), beta, p.h()));
      ck(mtjg_vec_set(p.h(), u.h()));
      ck(mtjg_vec_add_scaled(p.h(), beta, sum.h()));
    }
    papply(S, p, phat);
    ck(mtjg_csr_mult(A, phat.h(), vhat.h()));
    alpha = rho_1 / dot(rt, vhat);
    ck(mtjg_vec_set_scaled(q.h(), -alpha, vhat.h()));
    ck(mtjg_vec_add(q.h(), u.h()));
    ck(mtjg_vec_set(sum.h(), u.h()));
    ck(mtjg_vec_add(sum.h(), q.h()));
    papply(S, sum, uhat);
    ck(mtjg_vec_add_scaled(x, alpha, uhat.h()));
    ck(mtjg_csr_mult(A, uhat.h(), qhat.h()));
    ck(mtjg_vec_add_scaled(r.h(), -alpha, qhat.h()));
    rho_2 = rho_1;
  }
  mon.finish();
}

void solve_qmr(mtjg_solver S, mtjg_csr A, mtjg_vec b, mtjg_vec x) {
  DV r(S, solver_vec(S, 0)), y(S, solver_vec(S, 1)), z(S, solver_vec(S, 2)), v(S, solver_vec(S, 3)), w(S, solver_vec(S, 4)),
      vt(S, solver_vec(S, 5)), wt(S, solver_vec(S, 6)), yt(S, solver_vec(S, 7)), zt(S, solver_vec(S, 8)),
      p(S, solver_vec(S, 9)), q(S, solver_vec(S, 10)), pt(S, solver_vec(S, 11)), d(S, solver_vec(S, 12)),
      s(S, solver_vec(S, 13));
  HostMonitor mon(S);
  double rho = 0, rho_1 = 0, xi = 0, gamma = 1., gamma_1 = 0, theta = 0, theta_1 = 0, eta = -1., delta = 0, ep = 0,
         beta = 0;
  ck(mtjg_vec_set(r.h(), b));
  ck(mtjg_csr_mult_add(A, -1, x, r.h()));
  ck(mtjg_vec_set(vt.h(), r.h()));
  papply(S, vt, y);
  rho = norm2(y);
  ck(mtjg_vec_set(wt.h(), r.h()));
  ptrans(S, wt, z);
  xi = norm2(z);
  for (mon.k.iter = 0; !mon.converged(norm2(r)); mon.k.iter++) {
    if (rho == 0 || xi == 0) {
      mon.breakdown();
      break;
    }
    ck(mtjg_vec_set_scaled(v.h(), 1 / rho, vt.h()));
    ck(mtjg_vec_scale(y.h(), 1 / rho));
    ck(mtjg_vec_set_scaled(w.h(), 1 / xi, wt.h()));
    ck(mtjg_vec_scale(z.h(), 1 / xi));
    delta = dot(z, y);
    if (delta == 0) {
      mon.breakdown();
      break;
    }
    papply(S, y, yt);
    ptrans(S, z, zt);
    if (mon.k.iter == 0) {
      ck(mtjg_vec_set(p.h(), yt.h()));
      ck(mtjg_vec_set(q.h(), zt.h()));
    } else {
      ck(mtjg_vec_scale(p.h(), -xi * delta / ep));
      ck(mtjg_vec_add(p.h(), yt.h()));
      ck(mtjg_vec_scale(q.h(), -rho * delta / ep));
      ck(mtjg_vec_add(q.h(), zt.h()));
    }
    ck(mtjg_csr_mult(A, p.h(), pt.h()));
    ep = dot(q, pt);
    if (ep == 0) {
      mon.breakdown();
      break;
    }
    beta = ep / delta;
    if (beta == 0) {
      mon.breakdown();
      break;
    }
    ck(mtjg_vec_set_scaled(vt.h(), -beta, v.h()));
    ck(mtjg_vec_add(vt.h(), pt.h()));
    papply(S, vt, y);
    rho_1 = rho;
    rho = norm2(y);
    // A.transMultAdd(q, w_tld.set(-beta, w)): transMultAdd with alpha 1
    ck(mtjg_vec_set_scaled(wt.h(), -beta, w.h()));
    ck(mtjg_csr_trans_mult_add(A, 1.0, q.h(), wt.h()));
    ptrans(S, wt, z);
    xi = norm2(z);
    gamma_1 = gamma;
    theta_1 = theta;
    theta = rho / (gamma_1 * beta);
    gamma = 1 / std::sqrt(1 + theta * theta);
    if (gamma == 0) {
      mon.breakdown();
      break;
    }
    eta = -eta * rho_1 * gamma * gamma / (beta * gamma_1 * gamma_1);
    if (mon.k.iter == 0) {
      ck(mtjg_vec_set_scaled(d.h(), eta, p.h()));
      ck(mtjg_vec_set_scaled(s.h(), eta, pt.h()));
    } else {
      const double val = theta_1 * theta_1 * gamma * gamma;
      ck(mtjg_vec_scale(d.h(), val));
      ck(mtjg_vec_add_scaled(d.h(), eta, p.h()));
      ck(mtjg_vec_scale(s.h(), val));
      ck(mtjg_vec_add_scaled(s.h(), eta, pt.h()));
    }
    ck(mtjg_vec_add(x, d.h()));
    ck(mtjg_vec_add_scaled(r.h(), -1, s.h()));
  }
  mon.finish();
}

}  // namespace

// ===========================================================================
// The allocation-free CG solve (mtjg_cg_solve_noalloc).
//
// Every device buffer of the solve comes from `scratch`, one fixed array per
// GPU that the code object carries, so the call makes no hipMalloc or
// hipFree: the scratch vectors p z q r and the fused product's second
// direction buffer (MTJG_NOALLOC_MAX_ROWS doubles each), the solver state and
// the two-level reduction workspace.  In host code the array's name is not a
// device address: its address on each device is looked up once, with that
// device current, by hipGetSymbolAddress and kept per device; only that
// pointer is handed to kernels.  What must start at zero (the state, the
// reduction tickets) is cleared by a kernel at every call.  Every call on a
// GPU shares that GPU's array, so calls that overlap on one GPU race.

namespace {

constexpr int64_t NOALLOC_VEC = MTJG_NOALLOC_MAX_ROWS;  // doubles per vector
constexpr int NOALLOC_RED_CAP = 16384;                   // reduction workspace columns
constexpr int NOALLOC_RED_GCAP = (NOALLOC_RED_CAP + RED_GROUP - 1) / RED_GROUP;
constexpr size_t NOALLOC_OFF_STATE = 5 * sizeof(double) * (size_t)NOALLOC_VEC;
constexpr size_t NOALLOC_OFF_PART = NOALLOC_OFF_STATE + 256;
constexpr size_t NOALLOC_OFF_GPART = NOALLOC_OFF_PART + sizeof(double) * RED_MAX_K * (size_t)NOALLOC_RED_CAP;
constexpr size_t NOALLOC_OFF_TICKET = NOALLOC_OFF_GPART + sizeof(double) * RED_MAX_K * (size_t)NOALLOC_RED_GCAP;
constexpr size_t NOALLOC_BYTES = NOALLOC_OFF_TICKET + sizeof(uint32_t) * (NOALLOC_RED_GCAP + 1) * TICKET_STRIDE;
static_assert(sizeof(KState) <= 256, "solver state exceeds its slot");
static_assert(NOALLOC_BYTES <= (size_t)MTJG_NOALLOC_SCRATCH_BYTES, "scratch layout exceeds the array");

}  // namespace

static __device__ __align__(256) unsigned char scratch[MTJG_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex g_noalloc_mu;
unsigned char *g_noalloc_ptr[NOALLOC_MAX_DEVICES] = {};

// The device address of `scratch` on `device`, looked up on the first call
// for that device (with it current) and kept.
unsigned char *noalloc_scratch(int device) {
  if (device < 0 || device >= NOALLOC_MAX_DEVICES) fail(MTJG_E_UNSUPPORTED, "device index out of range");
  std::lock_guard<std::mutex> lock(g_noalloc_mu);
  if (!g_noalloc_ptr[device]) {
    int cur = -1;
    HIPCHK(hipGetDevice(&cur));
    if (cur != device) HIPCHK(hipSetDevice(device));
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (cur != device) HIPCHK(hipSetDevice(cur));
    if (e != hipSuccess || !ptr) {
      (void)hipGetLastError();
      fail(MTJG_E_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    }
    g_noalloc_ptr[device] = (unsigned char *)ptr;
  }
  return g_noalloc_ptr[device];
}

// The solver state of a new solve (DefaultIterationMonitor fields, status
// running, every other field zero) and zero reduction tickets; one
// workgroup.
__global__ void k_noalloc_init(KState *st, int32_t max_iter, double rtol, double atol, double dtol,
                               uint32_t *tickets, int nticket) {
  uint32_t *w = reinterpret_cast<uint32_t *>(st);
  for (int i = threadIdx.x; i < (int)(sizeof(KState) / sizeof(uint32_t)); i += blockDim.x) w[i] = 0u;
  for (int i = threadIdx.x; i < nticket; i += blockDim.x) tickets[i] = 0u;
  __syncthreads();
  if (threadIdx.x == 0) {
    st->max_iter = max_iter;
    st->rtol = rtol;
    st->atol = atol;
    st->dtol = dtol;
    st->status = ST_RUNNING;
  }
}

}  // namespace

// ===========================================================================
// C ABI

extern "C" int mtjg_solver_create(mtjg_ctx ctx, int kind, mtjg_vec template_vec, mtjg_solver *out) {
  MTJG_API_BEGIN
  if (!ctx || !template_vec || !out) fail(MTJG_E_NULL, "null argument");
  if (kind < MTJG_CG || kind > MTJG_QMR) fail(MTJG_E_INVALID_ARGUMENT, "unknown solver kind");
  auto *S = new mtjg_solver_s();
  S->ctx = ctx;
  S->kind = kind;
  S->n = template_vec->n;
  HIPCHK(hipMalloc(&S->st, sizeof(KState)));
  HIPCHK(hipMalloc(&S->st_in, sizeof(KState)));
  HIPCHK(hipMemsetAsync(S->st, 0, sizeof(KState), ctx->stream));
  HIPCHK(hipMemsetAsync(S->st_in, 0, sizeof(KState), ctx->stream));
  HIPCHK(hipHostMalloc(&S->h_st, sizeof(KState) * 2, hipHostMallocDefault));
  HIPCHK(hipHostMalloc(&S->h_slot, sizeof(int32_t) * 2, hipHostMallocDefault));
  HIPCHK(hipEventCreateWithFlags(&S->ev[0], hipEventDisableTiming));
  HIPCHK(hipEventCreateWithFlags(&S->ev[1], hipEventDisableTiming));
  // scratch vectors come from template.copy() (S/CG.java:51-56)
  int nvec = kind == MTJG_CG         ? 4
             : kind == MTJG_BICGSTAB ? 8
             : kind == MTJG_GMRES    ? 3
             : kind == MTJG_IR       ? 2
             : kind == MTJG_BICG     ? 8
             : kind == MTJG_CGS      ? 10
             : kind == MTJG_QMR      ? 14
                                     : 1;
  for (int k = 0; k < nvec; ++k) solver_vec(S, (size_t)k);
  HIPCHK(hipStreamSynchronize(ctx->stream));
  *out = S;
  MTJG_API_END
}

extern "C" int mtjg_solver_destroy(mtjg_solver S) {
  MTJG_API_BEGIN
  if (!S) return MTJG_OK;
  HIPCHK(hipStreamSynchronize(S->ctx->stream));
  for (double *p : S->v64) (void)hipFree(p);
  for (float *p : S->v32) (void)hipFree(p);
  gmres_free(S);
  (void)hipFree(S->st);
  (void)hipFree(S->st_in);
  free_red(S->red);
  (void)hipHostFree(S->h_st);
  (void)hipHostFree(S->h_slot);
  (void)hipEventDestroy(S->ev[0]);
  (void)hipEventDestroy(S->ev[1]);
  delete S;
  MTJG_API_END
}

extern "C" int mtjg_solver_set_preconditioner(mtjg_solver S, mtjg_prec M) {
  MTJG_API_BEGIN
  if (!S) fail(MTJG_E_NULL, "null solver");
  S->M = M;
  MTJG_API_END
}

extern "C" int mtjg_solver_set_monitor(mtjg_solver S, const mtjg_monitor *m) {
  MTJG_API_BEGIN
  if (!S || !m) fail(MTJG_E_NULL, "null argument");
  S->mon = *m;
  MTJG_API_END
}

extern "C" int mtjg_solver_set_reporter(mtjg_solver S, mtjg_reporter cb, void *user) {
  MTJG_API_BEGIN
  if (!S) fail(MTJG_E_NULL, "null solver");
  S->rep = cb;
  S->rep_user = user;
  MTJG_API_END
}

extern "C" int mtjg_gmres_set_restart(mtjg_solver S, int32_t restart) {
  MTJG_API_BEGIN
  if (!S) fail(MTJG_E_NULL, "null solver");
  if (restart <= 0) fail(MTJG_E_INVALID_ARGUMENT, "restart must be a positive integer");
  HIPCHK(hipStreamSynchronize(S->ctx->stream));
  gmres_free(S);
  S->restart = restart;
  MTJG_API_END
}

extern "C" int mtjg_mixed_set_inner(mtjg_solver S, double inner_rtol, int32_t inner_max_iter) {
  MTJG_API_BEGIN
  if (!S) fail(MTJG_E_NULL, "null solver");
  if (!(inner_rtol > 0) || inner_max_iter <= 0) fail(MTJG_E_INVALID_ARGUMENT, "bad inner settings");
  S->inner_rtol = inner_rtol;
  S->inner_max = inner_max_iter;
  MTJG_API_END
}

extern "C" int mtjg_solver_solve(mtjg_solver S, mtjg_csr A, mtjg_vec b, mtjg_vec x, mtjg_result *result) {
  MTJG_API_BEGIN
  if (!S || !A || !b || !x) fail(MTJG_E_NULL, "null argument");
  // checkSizes (S/AbstractIterativeSolver.java:70-77)
  if (A->nr != A->nc) fail(MTJG_E_INVALID_ARGUMENT, "!A.isSquare()");
  if (b->n != A->nr) fail(MTJG_E_INVALID_ARGUMENT, "b.size() != A.numRows()");
  if (b->n != x->n) fail(MTJG_E_INVALID_ARGUMENT, "b.size() != x.size()");
  if (b->n != S->n) fail(MTJG_E_INDEX_OUT_OF_BOUNDS, "solver scratch size differs from the system size");
  if (b->dtype != MTJG_F64 || x->dtype != MTJG_F64) fail(MTJG_E_UNSUPPORTED, "solvers take FP64 vectors");
  if (b->ptr == x->ptr && b->n > 0) fail(MTJG_E_INVALID_ARGUMENT, "b and x must not alias");
  if (S->M && prec_n(S->M) != S->n) fail(MTJG_E_INDEX_OUT_OF_BOUNDS, "preconditioner size differs");
  if (A->nr == 0) fail(MTJG_E_INVALID_ARGUMENT, "empty system");
  HIPCHK(hipSetDevice(S->ctx->device));
  double *bp = (double *)b->ptr, *xp = (double *)x->ptr;
  switch (S->kind) {
    case MTJG_CG: solve_cg(S, A, bp, xp); break;
    case MTJG_BICGSTAB: solve_bicgstab(S, A, bp, xp); break;
    case MTJG_GMRES:
      solve_gmres(S, A, bp, xp);
      gmres_check(S);
      break;
    case MTJG_IR: solve_ir(S, A, bp, xp); break;
    case MTJG_MIXED_CG: solve_mixed(S, A, bp, xp); break;
    case MTJG_BICG: solve_bicg(S, A, b, x); break;
    case MTJG_CGS: solve_cgs(S, A, b, x); break;
    case MTJG_QMR: solve_qmr(S, A, b, x); break;
  }
  prec_check(S->M);
  KState k;
  state_read(S, S->st, k);
  int code = status_to_code(k.status);
  if (result) {
    result->iterations = k.iter;
    result->residual = k.residual;
    result->reason = code;
  }
  if (code != MTJG_OK) {
    const char *why = code == MTJG_NOT_CONVERGED_ITERATIONS   ? "Iterations"
                      : code == MTJG_NOT_CONVERGED_DIVERGENCE ? "Divergence"
                      : code == MTJG_NOT_CONVERGED_BREAKDOWN  ? "Breakdown"
                                                              : "MatrixSingular";
    set_last_error(std::string("Solver did not converge: ") + why + ". Residual=" + std::to_string(k.residual) +
                   ", iterations=" + std::to_string(k.iter));
  }
  return code;
  MTJG_API_END
}

extern "C" int mtjg_solver_inner_iterations(mtjg_solver S) { return S ? S->inner_total : -1; }

extern "C" int mtjg_cg_solve_noalloc(mtjg_ctx ctx, mtjg_csr A, mtjg_prec M, mtjg_vec b, mtjg_vec x,
                                     const mtjg_monitor *monitor, mtjg_result *result) {
  MTJG_API_BEGIN
  if (!ctx || !A || !b || !x) fail(MTJG_E_NULL, "null argument");
  // checkSizes (S/AbstractIterativeSolver.java:70-77)
  if (A->nr != A->nc) fail(MTJG_E_INVALID_ARGUMENT, "!A.isSquare()");
  if (b->n != A->nr) fail(MTJG_E_INVALID_ARGUMENT, "b.size() != A.numRows()");
  if (b->n != x->n) fail(MTJG_E_INVALID_ARGUMENT, "b.size() != x.size()");
  if (b->dtype != MTJG_F64 || x->dtype != MTJG_F64) fail(MTJG_E_UNSUPPORTED, "solvers take FP64 vectors");
  if (b->ptr == x->ptr && b->n > 0) fail(MTJG_E_INVALID_ARGUMENT, "b and x must not alias");
  if (A->nr == 0) fail(MTJG_E_INVALID_ARGUMENT, "empty system");
  const int pk = prec_kind(M);
  if (pk != PREC_IDENTITY && pk != PREC_DIAGONAL)
    fail(MTJG_E_UNSUPPORTED, "the allocation-free solve takes no preconditioner or a DiagonalPreconditioner");
  if (M && prec_n(M) != A->nr) fail(MTJG_E_INDEX_OUT_OF_BOUNDS, "preconditioner size differs");
  const int64_t n = A->nr;
  if (n > MTJG_NOALLOC_MAX_ROWS)
    fail(MTJG_E_INVALID_ARGUMENT, "system of " + std::to_string(n) + " rows exceeds the allocation-free solve's " +
                                      std::to_string((long long)MTJG_NOALLOC_MAX_ROWS) + "-row scratch");
  const int64_t need = std::max<int64_t>(VGRID_MAX, spmv_dot_red_cap(A));
  if (need > NOALLOC_RED_CAP) fail(MTJG_E_INVALID_ARGUMENT, "reduction grid exceeds the allocation-free scratch");
  const mtjg_monitor mon = monitor ? *monitor : mtjg_monitor{100000, 1e-5, 1e-50, 1e5};
  HIPCHK(hipSetDevice(ctx->device));
  unsigned char *base = noalloc_scratch(ctx->device);
  hipStream_t s = ctx->stream;
  CgWork w;
  double *v = reinterpret_cast<double *>(base);
  w.p = v;
  w.z = v + NOALLOC_VEC;
  w.q = v + 2 * NOALLOC_VEC;
  w.r = v + 3 * NOALLOC_VEC;
  w.p2 = v + 4 * NOALLOC_VEC;
  w.st = reinterpret_cast<KState *>(base + NOALLOC_OFF_STATE);
  w.red.partials = reinterpret_cast<double *>(base + NOALLOC_OFF_PART);
  w.red.gpart = reinterpret_cast<double *>(base + NOALLOC_OFF_GPART);
  w.red.gticket = reinterpret_cast<uint32_t *>(base + NOALLOC_OFF_TICKET);
  w.red.fticket = w.red.gticket + (size_t)NOALLOC_RED_GCAP * TICKET_STRIDE;
  w.red.cap = NOALLOC_RED_CAP;
  w.red.gcap = NOALLOC_RED_GCAP;
  k_noalloc_init<<<1, BS, 0, s>>>(w.st, mon.max_iter, mon.rtol, mon.atol, mon.dtol, w.red.gticket,
                                  (int)red_ticket_words(NOALLOC_RED_GCAP));
  HIPCHK(hipGetLastError());
  // The status word (with the residual and iteration count beside it) lands
  // in the context's pinned scalars, allocated when the context was made.
  const size_t lo = offsetof(KState, residual), hi = offsetof(KState, flags) + sizeof(int32_t);
  static_assert(offsetof(KState, flags) + sizeof(int32_t) - offsetof(KState, residual) <= 8 * sizeof(double),
                "status tail exceeds the context's pinned scalars");
  KState tail;
  auto read_tail = [&]() {
    HIPCHK(hipMemcpyAsync(ctx->h_out, reinterpret_cast<char *>(w.st) + lo, hi - lo, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    memcpy(reinterpret_cast<char *>(&tail) + lo, ctx->h_out, hi - lo);
  };
  cg_iterate(s, n, A, M, (double *)b->ptr, (double *)x->ptr, w, [&](int chunk, auto &&iteration) {
    for (;;) {
      for (int c = 0; c < chunk; ++c) iteration();
      read_tail();
      if (tail.status != ST_RUNNING) break;
    }
  });
  read_tail();
  const int code = status_to_code(tail.status);
  if (result) {
    result->iterations = tail.iter;
    result->residual = tail.residual;
    result->reason = code;
  }
  if (code != MTJG_OK)
    set_last_error(std::string("Solver did not converge. Residual=") + std::to_string(tail.residual) +
                   ", iterations=" + std::to_string(tail.iter));
  return code;
  MTJG_API_END
}
