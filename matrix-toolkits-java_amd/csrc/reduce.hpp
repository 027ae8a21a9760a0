// Deterministic reductions for 256-thread workgroups (four wave64s).
//
// Order is fixed by construction: each thread accumulates its elements in a
// fixed sequence, a wave combines lanes by an xor butterfly (a + b == b + a in
// IEEE arithmetic, so every lane of the butterfly holds the same bits), the
// four wave results are added in a fixed order, and the last workgroup of a
// grid adds the per-workgroup partials in a fixed order.  Results therefore
// depend only on the grid shape, never on scheduling, and repeated runs are
// bitwise identical.
#pragma once

#include <hip/hip_runtime.h>

#include "mtjg_internal.hpp"

namespace mtjg {

// The value x holds in lane (lane ^ OFF) of the wavefront, OFF one of 1, 2,
// 4, 8, 16, 32: the partner __shfl_xor(x, OFF) reads.  Inside each row of 16
// lanes DPP moves it as a VALU operand permutation instead of a trip through
// the LDS crossbar (ds_bpermute_b32, at LDS latency): quad_perm [1,0,3,2] for
// 1, [2,3,0,1] for 2, the half-row mirror (lane ^ 7) then quad_perm
// [3,2,1,0] (lane ^ 3) for 4, the row mirror (lane ^ 15) then the half-row
// mirror for 8.  16 is a ds_swizzle (xor 0x10 within each 32 lanes, no
// address operand) and 32 stays a ds_bpermute.  (The gfx950 permlane swaps
// gave wrong partners in reductions as this compiler schedules them, though
// right ones in an isolated lane test, so they are not used.)  Every lane of
// the partner pairs used must be active.
template <int OFF>
__device__ __forceinline__ uint32_t lane_xor_u32(uint32_t x) {
  static_assert(OFF == 1 || OFF == 2 || OFF == 4 || OFF == 8 || OFF == 16 || OFF == 32, "lane offset");
  if constexpr (OFF == 1) {
    return __builtin_amdgcn_mov_dpp(x, 0xB1, 0xF, 0xF, false);
  } else if constexpr (OFF == 2) {
    return __builtin_amdgcn_mov_dpp(x, 0x4E, 0xF, 0xF, false);
  } else if constexpr (OFF == 4) {
    return __builtin_amdgcn_mov_dpp(__builtin_amdgcn_mov_dpp(x, 0x141, 0xF, 0xF, false), 0x1B, 0xF, 0xF, false);
  } else if constexpr (OFF == 8) {
    return __builtin_amdgcn_mov_dpp(__builtin_amdgcn_mov_dpp(x, 0x140, 0xF, 0xF, false), 0x141, 0xF, 0xF, false);
  } else if constexpr (OFF == 16) {
    return __builtin_amdgcn_ds_swizzle(x, 0x401F);
  } else {
    return __shfl_xor(x, 32, 64);
  }
}

template <int OFF>
__device__ __forceinline__ double lane_xor(double v) {
  const uint64_t b = (uint64_t)__double_as_longlong(v);
  const uint64_t lo = lane_xor_u32<OFF>((uint32_t)b), hi = lane_xor_u32<OFF>((uint32_t)(b >> 32));
  return __longlong_as_double((long long)((hi << 32) | lo));
}

template <int OFF>
__device__ __forceinline__ float lane_xor(float v) {
  return __uint_as_float(lane_xor_u32<OFF>(__float_as_uint(v)));
}

// lane_xor for an offset known only at run time (uniform over the wavefront).
__device__ __forceinline__ float lane_xor_rt(float v, int off) {
  switch (off) {
    case 1: return lane_xor<1>(v);
    case 2: return lane_xor<2>(v);
    case 4: return lane_xor<4>(v);
    case 8: return lane_xor<8>(v);
    case 16: return lane_xor<16>(v);
    default: return lane_xor<32>(v);
  }
}
__device__ __forceinline__ double lane_xor_rt(double v, int off) {
  switch (off) {
    case 1: return lane_xor<1>(v);
    case 2: return lane_xor<2>(v);
    case 4: return lane_xor<4>(v);
    case 8: return lane_xor<8>(v);
    case 16: return lane_xor<16>(v);
    default: return lane_xor<32>(v);
  }
}

// One step of the xor butterfly: v += v of lane (lane ^ OFF).
template <int OFF, int K>
__device__ __forceinline__ void xor_step(double (&v)[K]) {
#pragma unroll
  for (int k = 0; k < K; ++k) v[k] += lane_xor<OFF>(v[k]);
}

template <int K>
__device__ __forceinline__ void wave_sum(double (&v)[K]) {
  xor_step<32>(v);
  xor_step<16>(v);
  xor_step<8>(v);
  xor_step<4>(v);
  xor_step<2>(v);
  xor_step<1>(v);
}

// Sum over the workgroup; the total is returned in every thread.
template <int K>
__device__ __forceinline__ void block_sum(double (&v)[K]) {
  __shared__ double sh[K][BS / 64];
  wave_sum<K>(v);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < K; ++k) sh[k][wid] = v[k];
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < K; ++k) v[k] = (sh[k][0] + sh[k][1]) + (sh[k][2] + sh[k][3]);
  __syncthreads();
}

__device__ __forceinline__ void store_agent(double *p, double v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ double load_agent(const double *p) {
  return __hip_atomic_load(const_cast<double *>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Publication protocol of the grid reductions.  Partials are written with
// agent-scope atomic stores (global_store ... sc1: written through to the
// device-wide coherence point) and read with agent-scope atomic loads (sc1:
// served from that point, never from a stale line of another XCD's L2).  The
// only ordering needed is that a workgroup's partial stores have completed
// before its ticket increment, which s_waitcnt vmcnt(0) provides; a
// release/acquire fence pair would add an L2 write-back and invalidate per
// workgroup.  Readers issue their loads after the __syncthreads that follows
// the ticket, so they cannot run ahead of it.
__device__ __forceinline__ void stores_complete() { asm volatile("s_waitcnt vmcnt(0)" ::: "memory"); }
__device__ __forceinline__ uint32_t take_ticket(uint32_t *t) {
  return __hip_atomic_fetch_add(t, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Two-level grid-wide sum for grids of any size.  Workgroup b publishes its K
// partials to partials[k * cap + b] and takes a ticket on its group of
// RED_GROUP consecutive workgroups; the last workgroup of a group adds the
// group's partials in index order into gpart[k * gcap + g] and takes a
// ticket on the final counter; the last group adds the group sums in index
// order.  Every sum has a fixed order, so results are bitwise reproducible;
// both counters are reset by their last user.  The workspace is a GridRed.
template <int K>
__device__ bool grid_sum_h(double (&v)[K], const GridRed &r) {
  __shared__ int last;
  block_sum<K>(v);
  const int b = blockIdx.x, g = b / RED_GROUP;
  const int ngroups = (gridDim.x + RED_GROUP - 1) / RED_GROUP;
  const int gsize = min(RED_GROUP, (int)gridDim.x - g * RED_GROUP);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < K; ++k) store_agent(&r.partials[k * r.cap + b], v[k]);
    stores_complete();
    uint32_t t = take_ticket(&r.gticket[g * TICKET_STRIDE]);
    last = (t == (uint32_t)gsize - 1);
  }
  __syncthreads();
  if (!last) return false;
  double acc[K];
#pragma unroll
  for (int k = 0; k < K; ++k)
    acc[k] = (threadIdx.x < gsize) ? load_agent(&r.partials[k * r.cap + g * RED_GROUP + threadIdx.x]) : 0.0;
  block_sum<K>(acc);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < K; ++k) store_agent(&r.gpart[k * r.gcap + g], acc[k]);
    __hip_atomic_store(&r.gticket[g * TICKET_STRIDE], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    stores_complete();
    uint32_t t = take_ticket(r.fticket);
    last = (t == (uint32_t)ngroups - 1);
  }
  __syncthreads();
  if (!last) return false;
#pragma unroll
  for (int k = 0; k < K; ++k) acc[k] = 0.0;
  for (int i = threadIdx.x; i < ngroups; i += BS) {
#pragma unroll
    for (int k = 0; k < K; ++k) acc[k] += load_agent(&r.gpart[k * r.gcap + i]);
  }
  block_sum<K>(acc);
#pragma unroll
  for (int k = 0; k < K; ++k) v[k] = acc[k];
  if (threadIdx.x == 0) __hip_atomic_store(r.fticket, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  return true;
}

// Two-level grid-wide combination of per-workgroup values by an operation
// that need not be a sum (maximum, scaled sum of squares), in a fixed order.
// On entry thread 0 holds its workgroup's K values in v; the protocol is
// grid_sum_h's, except that a finishing workgroup has thread 0 combine the
// values left to right, comb(acc, next) with acc starting at the first.
// Returns true in the last workgroup, with the grid's values in thread 0's v.
template <int K, typename Comb>
__device__ bool grid_combine_h(double (&v)[K], const GridRed &r, Comb comb) {
  static_assert(K <= RED_MAX_K, "too many values");
  __shared__ int last;
  __shared__ double sh[K][RED_GROUP];
  const int b = blockIdx.x, g = b / RED_GROUP;
  const int ngroups = (gridDim.x + RED_GROUP - 1) / RED_GROUP;
  const int gsize = min(RED_GROUP, (int)gridDim.x - g * RED_GROUP);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < K; ++k) store_agent(&r.partials[k * r.cap + b], v[k]);
    stores_complete();
    uint32_t t = take_ticket(&r.gticket[g * TICKET_STRIDE]);
    last = (t == (uint32_t)gsize - 1);
  }
  __syncthreads();
  if (!last) return false;
  auto fold = [&](int cnt) {
    if (threadIdx.x == 0) {
#pragma unroll
      for (int k = 0; k < K; ++k) v[k] = sh[k][0];
      for (int j = 1; j < cnt; ++j) {
        double nx[K];
#pragma unroll
        for (int k = 0; k < K; ++k) nx[k] = sh[k][j];
        comb(v, nx);
      }
    }
  };
  if ((int)threadIdx.x < gsize) {
#pragma unroll
    for (int k = 0; k < K; ++k) sh[k][threadIdx.x] = load_agent(&r.partials[k * r.cap + g * RED_GROUP + threadIdx.x]);
  }
  __syncthreads();
  fold(gsize);
  if (ngroups == 1) {
    if (threadIdx.x == 0)
      __hip_atomic_store(&r.gticket[g * TICKET_STRIDE], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return true;
  }
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < K; ++k) store_agent(&r.gpart[k * r.gcap + g], v[k]);
    __hip_atomic_store(&r.gticket[g * TICKET_STRIDE], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    stores_complete();
    uint32_t t = take_ticket(r.fticket);
    last = (t == (uint32_t)ngroups - 1);
  }
  __syncthreads();
  if (!last) return false;
  // ngroups <= RED_GROUP: the workspace holds at most RED_GROUP^2 workgroups
  if ((int)threadIdx.x < ngroups) {
#pragma unroll
    for (int k = 0; k < K; ++k) sh[k][threadIdx.x] = load_agent(&r.gpart[k * r.gcap + threadIdx.x]);
  }
  __syncthreads();
  fold(ngroups);
  if (threadIdx.x == 0) __hip_atomic_store(r.fticket, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  return true;
}

// ---- compensated sums ------------------------------------------------------
//
// The solvers' dot products and norms are accumulated as unevaluated sums
// hi + lo (Ogita, Rump and Oishi's Dot2: error-free TwoProduct via fma and
// TwoSum; the library is built with -ffp-contract=off, so these
// transformations are not rewritten).  The result is as accurate as a dot
// product evaluated in twice the working precision and then rounded, i.e.
// essentially the correctly rounded value, so iteration counts follow
// exact-arithmetic behaviour instead of one particular summation order.
// Combination order stays fixed, so results remain bitwise reproducible.
struct dd {
  double hi, lo;
};

__device__ __forceinline__ dd dd_add(dd a, dd b) {
  const double s = a.hi + b.hi;
  const double bb = s - a.hi;
  const double e = (a.hi - (s - bb)) + (b.hi - bb);
  return {s, e + (a.lo + b.lo)};
}

// acc += a * b with the product's rounding error kept.
__device__ __forceinline__ void dd_fma(dd &acc, double a, double b) {
  const double p = a * b;
  const double pe = fma(a, b, -p);
  const double s = acc.hi + p;
  const double bb = s - acc.hi;
  const double e = (acc.hi - (s - bb)) + (p - bb);
  acc.hi = s;
  acc.lo += e + pe;
}

template <int OFF, int K>
__device__ __forceinline__ void xor_step_dd(dd (&v)[K]) {
#pragma unroll
  for (int k = 0; k < K; ++k) v[k] = dd_add(v[k], dd{lane_xor<OFF>(v[k].hi), lane_xor<OFF>(v[k].lo)});
}

template <int K>
__device__ __forceinline__ void block_sum_dd(dd (&v)[K]) {
  __shared__ double sh[2 * K][BS / 64];
  xor_step_dd<32>(v);
  xor_step_dd<16>(v);
  xor_step_dd<8>(v);
  xor_step_dd<4>(v);
  xor_step_dd<2>(v);
  xor_step_dd<1>(v);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < K; ++k) {
      sh[2 * k][wid] = v[k].hi;
      sh[2 * k + 1][wid] = v[k].lo;
    }
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < K; ++k) {
    dd a = dd_add(dd{sh[2 * k][0], sh[2 * k + 1][0]}, dd{sh[2 * k][1], sh[2 * k + 1][1]});
    dd b = dd_add(dd{sh[2 * k][2], sh[2 * k + 1][2]}, dd{sh[2 * k][3], sh[2 * k + 1][3]});
    v[k] = dd_add(a, b);
  }
  __syncthreads();
}

// Two-level grid sum of compensated partials (protocol of grid_sum_h; each
// sum occupies two partial slots, so K <= RED_MAX_K / 2).  Returns true in
// the last workgroup with out[k] = hi + lo of the grid total.
template <int K>
__device__ bool grid_sum_dd(dd (&v)[K], const GridRed &r, double (&out)[K]) {
  static_assert(2 * K <= RED_MAX_K, "too many sums");
  __shared__ int last;
  block_sum_dd<K>(v);
  const int b = blockIdx.x, g = b / RED_GROUP;
  const int ngroups = (gridDim.x + RED_GROUP - 1) / RED_GROUP;
  const int gsize = min(RED_GROUP, (int)gridDim.x - g * RED_GROUP);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < K; ++k) {
      store_agent(&r.partials[(2 * k) * r.cap + b], v[k].hi);
      store_agent(&r.partials[(2 * k + 1) * r.cap + b], v[k].lo);
    }
    stores_complete();
    uint32_t t = take_ticket(&r.gticket[g * TICKET_STRIDE]);
    last = (t == (uint32_t)gsize - 1);
  }
  __syncthreads();
  if (!last) return false;
  dd acc[K];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    acc[k] = dd{0.0, 0.0};
    if ((int)threadIdx.x < gsize) {
      acc[k].hi = load_agent(&r.partials[(2 * k) * r.cap + g * RED_GROUP + threadIdx.x]);
      acc[k].lo = load_agent(&r.partials[(2 * k + 1) * r.cap + g * RED_GROUP + threadIdx.x]);
    }
  }
  block_sum_dd<K>(acc);
  if (ngroups == 1) {
    // A single group needs no second level: its last workgroup finishes, with
    // the group sum folded once into an exact zero, which is what the second
    // level's first step does (it only turns a negative zero positive), so
    // the result is bitwise the two-level one, one atomic round trip sooner.
    if (threadIdx.x == 0)
      __hip_atomic_store(&r.gticket[g * TICKET_STRIDE], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const dd z = dd_add(dd{0.0, 0.0}, acc[k]);
      out[k] = z.hi + z.lo;
    }
    return true;
  }
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < K; ++k) {
      store_agent(&r.gpart[(2 * k) * r.gcap + g], acc[k].hi);
      store_agent(&r.gpart[(2 * k + 1) * r.gcap + g], acc[k].lo);
    }
    __hip_atomic_store(&r.gticket[g * TICKET_STRIDE], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    stores_complete();
    uint32_t t = take_ticket(r.fticket);
    last = (t == (uint32_t)ngroups - 1);
  }
  __syncthreads();
  if (!last) return false;
#pragma unroll
  for (int k = 0; k < K; ++k) acc[k] = dd{0.0, 0.0};
  for (int i = threadIdx.x; i < ngroups; i += BS) {
#pragma unroll
    for (int k = 0; k < K; ++k)
      acc[k] = dd_add(acc[k], dd{load_agent(&r.gpart[(2 * k) * r.gcap + i]),
                                 load_agent(&r.gpart[(2 * k + 1) * r.gcap + i])});
  }
  block_sum_dd<K>(acc);
#pragma unroll
  for (int k = 0; k < K; ++k) out[k] = acc[k].hi + acc[k].lo;
  if (threadIdx.x == 0) __hip_atomic_store(r.fticket, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  return true;
}

// java.lang.Math.max(double, double).
__host__ __device__ __forceinline__ double java_max(double a, double b) {
  if (a != a) return a;
  if (b != b) return b;
  if (a == 0.0 && b == 0.0) return __builtin_signbit(a) ? b : a;
  return (a >= b) ? a : b;
}

// DefaultIterationMonitor.convergedI (S/DefaultIterationMonitor.java:132-156)
// applied to residual r at the state's current iteration, preceded by
// AbstractIterationMonitor.converged storing the residual
// (S/AbstractIterationMonitor.java:81-86).  Returns the new status.
__device__ __forceinline__ int monitor_check(KState *st, double r) {
  st->residual = r;
  if (st->iter == 0) st->init_r = r;
  if (r < java_max(st->rtol * st->init_r, st->atol)) return ST_CONVERGED;
  if (r > st->dtol * st->init_r) return ST_DIVERGENCE;
  if (st->iter >= st->max_iter) return ST_ITERATIONS;
  if (r != r) return ST_DIVERGENCE;
  return ST_RUNNING;
}

}  // namespace mtjg
