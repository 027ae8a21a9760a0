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

template <int K>
__device__ __forceinline__ void wave_sum(double (&v)[K]) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
#pragma unroll
    for (int k = 0; k < K; ++k) v[k] += __shfl_xor(v[k], off, 64);
  }
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

// Grid-wide sum.  Every workgroup publishes its K partial sums to
// partials[k * RED_MAX_BLOCKS + blockIdx.x] and takes a ticket; the
// workgroup that takes the last ticket adds all partials in index order and
// returns true with the grid totals in v (every thread).  Others return
// false.  The ticket is reset by the last workgroup, so the counter is ready
// for the next launch without a memset.
template <int K>
__device__ bool grid_sum(double (&v)[K], double *partials, uint32_t *ticket) {
  __shared__ int last;
  block_sum<K>(v);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < K; ++k) store_agent(&partials[k * RED_MAX_BLOCKS + blockIdx.x], v[k]);
    __atomic_thread_fence(__ATOMIC_RELEASE);
    uint32_t t = __hip_atomic_fetch_add(ticket, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    last = (t == gridDim.x - 1);
  }
  __syncthreads();
  if (!last) return false;
  __atomic_thread_fence(__ATOMIC_ACQUIRE);
  double acc[K];
#pragma unroll
  for (int k = 0; k < K; ++k) acc[k] = 0.0;
  for (int i = threadIdx.x; i < (int)gridDim.x; i += BS) {
#pragma unroll
    for (int k = 0; k < K; ++k) acc[k] += load_agent(&partials[k * RED_MAX_BLOCKS + i]);
  }
  block_sum<K>(acc);
#pragma unroll
  for (int k = 0; k < K; ++k) v[k] = acc[k];
  if (threadIdx.x == 0)
    __hip_atomic_store(ticket, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  return true;
}

// java.lang.Math.max(double, double).
__device__ __forceinline__ double java_max(double a, double b) {
  if (a != a) return a;
  if (b != b) return b;
  if (a == 0.0 && b == 0.0) return signbit(a) ? b : a;
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
