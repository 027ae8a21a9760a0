// Batched, continuously admitted mixed-precision CG for many independent SPD
// systems (the serving path).  No MTJ counterpart: each system is solved the
// way solve_mixed solves one (FP64 residual r = b - A x with FP64 values, FP32
// inner Jacobi-CG on r / ||r|| to a relative tolerance, the larger of the
// request's and half the reduction the outer residual still lacks, FP64
// update x += ||r|| d), but every active system advances in the same launches.
//
// Work layout: system s is cut into nb_s row tiles of rows_per_wg rows (a
// multiple of 256, so of the 64-row slices of the system's SELL-64 copy,
// spmv_rowblock.hpp); a device work table maps each workgroup of a launch to
// (slot, tile, the slot's partial-sum offset), and the workgroup's four
// wavefronts take the tile's slices in turn.  One batch iteration is three
// launches over the whole table:
//   B1  phase INNER: q = A32 p and p.q per slot -> alpha
//       phase OUTER: r = b - A64 x (MTJ's multAdd order) and ||r||, then the
//                    DefaultIterationMonitor check of the outer loop
//   B2  phase START: rr = (float)(r / ||r||), d = 0, z = invdiag rr, rr.z
//       phase INNER: d += alpha p, rr -= alpha q, z = invdiag rr, rr.z, ||rr||,
//                    inner convergence test -> beta or finish
//   B3  p = z (first inner step) | p = z + beta p | x += ||r|| d (finish)
// Per-slot sums are fixed-order compensated reductions whose last workgroup
// (per-slot ticket) runs the slot's scalar epilogue, so a slot's result bits
// depend only on its own system and tile size, never on what else is in the
// batch.
//
// Host pipeline: each mtjg_batch_step launches a hipGraph of `iters` batch
// iterations (captured once per grid size) followed by a snapshot of the
// slot table into pinned memory, and only then retires the systems the
// previous step's snapshot shows finished and admits queued requests, so the
// host work overlaps the running graph.  A finished slot is inert (its
// kernels return at once), so the running graph may still list it; new
// systems are uploaded on a second stream into free arena ranges, and their
// slot records and the new work table are copied on the main stream after
// the running graph, before the next one.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "mtjg_internal.hpp"
#include "reduce.hpp"
#include "spmv_rowblock.hpp"

namespace mtjg {
namespace {

enum : int32_t { PH_OUTER = 0, PH_START = 1, PH_INNER = 2 };
enum : int32_t { A3_NONE = 0, A3_PINIT = 1, A3_PUPDATE = 2, A3_PROMOTE = 3 };

struct BSlot {
  // system and work vectors (device pointers into the arena)
  const int32_t *rp;
  const int32_t *ci;
  const double *v64;
  const float *inv;
  const double *b;
  const int32_t *sptr;     // SELL-64 copy: slice offsets, columns, values
  const int32_t *scol;
  const double *sv64;
  const float *sv32;       // FP32 values, or null when the system has a stencil copy:
  const float *st32;       // the FP32 values diagonal by diagonal (st_row),
  const uint8_t *stmask;   // a mask byte per row,
  int32_t st_m;            // the number of diagonals (0: no stencil copy)
  int32_t st_off[ST_MAX_OFF];  // and their column - row offsets
  double *x, *r;
  float *rr, *d, *p, *q, *z;
  int64_t rid;             // request id (host bookkeeping)
  int32_t n, nb, rows_per_wg;
  // control
  int32_t active, status, phase, action3;
  int32_t outer, inner, inner_total, max_outer, max_inner;
  double rtol, atol, dtol, init_r, rnorm;
  double rho, alpha, beta, inner_init, inner_rtol;
};

struct BWork {
  int32_t slot, tile, part_off, pad;
};

struct BDev {
  BSlot *slots;
  const BWork *work;
  const int32_t *nwork;
  double *partials;  // 4 arrays of `cap`: hi/lo of two sums
  int cap;
  // per-slot tickets, slot s's at s * TICKET_STRIDE: each on a memory line
  // of its own, apart from the slot record every workgroup of the slot reads
  uint32_t *ticket;
};

__device__ __forceinline__ bool slot_live(const BSlot &S) { return S.active && S.status == ST_RUNNING; }

// Per-slot two-sum reduction; returns true in the slot's last workgroup with
// the totals in out[].
template <int K>
__device__ bool slot_sum(dd (&v)[K], BSlot &S, const BWork &w, const BDev &d, double (&out)[K]) {
  __shared__ int last;
  block_sum_dd<K>(v);
  const int base = w.part_off;
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < K; ++k) {
      store_agent(&d.partials[(2 * k) * d.cap + base + w.tile], v[k].hi);
      store_agent(&d.partials[(2 * k + 1) * d.cap + base + w.tile], v[k].lo);
    }
    stores_complete();
    uint32_t t = take_ticket(&d.ticket[(size_t)w.slot * TICKET_STRIDE]);
    last = (t == (uint32_t)S.nb - 1);
  }
  __syncthreads();
  if (!last) return false;
  dd acc[K];
#pragma unroll
  for (int k = 0; k < K; ++k) acc[k] = dd{0.0, 0.0};
  for (int j = threadIdx.x; j < S.nb; j += BS) {
#pragma unroll
    for (int k = 0; k < K; ++k)
      acc[k] = dd_add(acc[k], dd{load_agent(&d.partials[(2 * k) * d.cap + base + j]),
                                 load_agent(&d.partials[(2 * k + 1) * d.cap + base + j])});
  }
  block_sum_dd<K>(acc);
#pragma unroll
  for (int k = 0; k < K; ++k) out[k] = acc[k].hi + acc[k].lo;
  if (threadIdx.x == 0)
    __hip_atomic_store(&d.ticket[(size_t)w.slot * TICKET_STRIDE], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  return true;
}

// Rows of tile `tile`: thread t owns rows r0 + t + k*BS.
#define TILE_ROWS(i, S, tile)                                                      \
  for (int32_t i = (tile) * (S).rows_per_wg + threadIdx.x,                         \
               i##_end = min((S).n, ((tile) + 1) * (S).rows_per_wg);               \
       i < i##_end; i += BS)

// Rows of tile `tile` in groups of four: thread t owns rows g .. g+3 with
// g = r0 + 4t + 4k*BS, so each FP32 access of a wavefront is one 1 KB
// contiguous 16-byte-per-lane load or store (4-byte accesses moved 256 B per
// instruction).  Every vector of a slot starts 256-byte aligned in the arena
// and tiles are multiples of 256 rows, so a whole group is 16-byte aligned;
// the group that runs past the system's last row is done row by row.
typedef float f4v __attribute__((ext_vector_type(4)));
typedef double d2v __attribute__((ext_vector_type(2)));
#define TILE_GROUPS(g, full, S, tile)                                                              \
  for (int32_t g = (tile) * (S).rows_per_wg + 4 * (int32_t)threadIdx.x,                           \
               g##_end = min((S).n, ((tile) + 1) * (S).rows_per_wg);                              \
       g < g##_end; g += 4 * BS)                                                                   \
    for (bool full = g + 4 <= g##_end, g##_once = true; g##_once; g##_once = false)
__device__ __forceinline__ f4v ld4f(const float *p, int32_t g, int32_t end, bool full) {
  if (full) return *(const f4v *)(p + g);
  f4v v = {0.f, 0.f, 0.f, 0.f};
  for (int k = 0; k < 4; ++k)
    if (g + k < end) v[k] = p[g + k];
  return v;
}
__device__ __forceinline__ void st4f(float *p, int32_t g, int32_t end, bool full, f4v v) {
  if (full) {
    *(f4v *)(p + g) = v;
    return;
  }
  for (int k = 0; k < 4; ++k)
    if (g + k < end) p[g + k] = v[k];
}

// The three phases of a batch iteration, each for one work-table entry `wi`
// (a tile of one system), run by a whole workgroup.
__device__ __forceinline__ void b1_work(const BDev &d, int wi) {
  const BWork w = d.work[wi];
  BSlot &S = d.slots[w.slot];
  if (!slot_live(S)) return;
  const int phase = S.phase;
  if (phase == PH_START) return;
  const int per_tile = S.rows_per_wg / SELL_C;
  const int nslice = (S.n + SELL_C - 1) / SELL_C;
  const int sl_end = min(nslice, (w.tile + 1) * per_tile);
  const int wave = threadIdx.x / SELL_C, lane = threadIdx.x % SELL_C;
  dd acc0{0.0, 0.0}, acc1{0.0, 0.0};
  if (phase == PH_INNER) {
    // q = A32 p with p.q (FP32 values and vectors): from the stencil copy
    // when the system has one, else the SELL-64 rows; either adds a row's
    // products in column order, so both give the same bits
    if (S.st_m > 0) {
      for (int sl = w.tile * per_tile + wave; sl < sl_end; sl += BS / SELL_C)
        st_row<SPMV_DOT>(sl * SELL_C + lane, S.n, S.st_m, S.st_off, S.st32, S.stmask, S.p, S.q, S.p, acc0, acc1);
    } else {
      for (int sl = w.tile * per_tile + wave; sl < sl_end; sl += BS / SELL_C)
        sell_row<float, float, SPMV_DOT>(sl, lane, S.n, S.sptr, S.scol, S.sv32, S.p, S.q, nullptr, 1.0, 1.0, S.p,
                                         acc0, acc1);
    }
  } else {
    // r = b + (-1) * (A64 x): MTJ's multAdd order, with r.r
    for (int sl = w.tile * per_tile + wave; sl < sl_end; sl += BS / SELL_C)
      sell_row<double, double, SPMV_RESID>(sl, lane, S.n, S.sptr, S.scol, S.sv64, S.x, S.r, S.b, -1.0, 1.0,
                                           nullptr, acc0, acc1);
    acc0 = acc1;
  }
  dd v[1] = {acc0};
  double tot[1];
  if (slot_sum<1>(v, S, w, d, tot) && threadIdx.x == 0) {
    if (phase == PH_INNER) {
      S.alpha = S.rho / tot[0];
    } else {
      const double rn = sqrt(tot[0]);
      S.rnorm = rn;
      S.action3 = A3_NONE;
      if (S.outer == 0) S.init_r = rn;
      if (rn < java_max(S.rtol * S.init_r, S.atol)) S.status = ST_CONVERGED;
      else if (rn > S.dtol * S.init_r) S.status = ST_DIVERGENCE;
      else if (S.outer >= S.max_outer) S.status = ST_ITERATIONS;
      else if (rn != rn) S.status = ST_DIVERGENCE;
      else S.phase = PH_START;
    }
  }
}

__device__ __forceinline__ void b2_work(const BDev &d, int wi) {
  const BWork w = d.work[wi];
  BSlot &S = d.slots[w.slot];
  if (!slot_live(S)) return;
  const int phase = S.phase;
  if (phase == PH_OUTER) return;
  dd v[2] = {{0.0, 0.0}, {0.0, 0.0}};
  if (phase == PH_START) {
    const double sc = 1.0 / S.rnorm;
    TILE_GROUPS(g, full, S, w.tile) {
      const int32_t end = g_end;
      double r4[4];
      if (full) {
        const d2v a0 = *(const d2v *)(S.r + g), a1 = *(const d2v *)(S.r + g + 2);
        r4[0] = a0[0], r4[1] = a0[1], r4[2] = a1[0], r4[3] = a1[1];
      } else {
        for (int k = 0; k < 4; ++k) r4[k] = g + k < end ? S.r[g + k] : 0.0;
      }
      const f4v inv = ld4f(S.inv, g, end, full);
      f4v rr, z;
      for (int k = 0; k < 4; ++k) {
        rr[k] = (float)(r4[k] * sc);
        z[k] = rr[k] * inv[k];
      }
      st4f(S.rr, g, end, full, rr);
      st4f(S.d, g, end, full, f4v{0.f, 0.f, 0.f, 0.f});
      st4f(S.z, g, end, full, z);
      for (int k = 0; k < 4; ++k)
        if (full || g + k < end) {
          dd_fma(v[0], (double)rr[k], (double)z[k]);
          dd_fma(v[1], (double)rr[k], (double)rr[k]);
        }
    }
  } else {
    const float a = (float)S.alpha;
    TILE_GROUPS(g, full, S, w.tile) {
      const int32_t end = g_end;
      f4v dv = ld4f(S.d, g, end, full), rr = ld4f(S.rr, g, end, full);
      const f4v pv = ld4f(S.p, g, end, full), qv = ld4f(S.q, g, end, full), inv = ld4f(S.inv, g, end, full);
      f4v z;
      for (int k = 0; k < 4; ++k) {
        dv[k] = dv[k] + a * pv[k];
        rr[k] = rr[k] + (-a) * qv[k];
        z[k] = rr[k] * inv[k];
      }
      st4f(S.d, g, end, full, dv);
      st4f(S.rr, g, end, full, rr);
      st4f(S.z, g, end, full, z);
      for (int k = 0; k < 4; ++k)
        if (full || g + k < end) {
          dd_fma(v[0], (double)rr[k], (double)z[k]);
          dd_fma(v[1], (double)rr[k], (double)rr[k]);
        }
    }
  }
  double tot[2];
  if (slot_sum<2>(v, S, w, d, tot) && threadIdx.x == 0) {
    const double rn = sqrt(tot[1]);
    if (phase == PH_START) {
      S.rho = tot[0];
      S.inner = 0;
      S.inner_init = rn;
      S.phase = PH_INNER;
      S.action3 = A3_PINIT;
    } else {
      S.inner++;
      S.inner_total++;
      // the inner solve stops at the larger of the request's inner tolerance
      // and half the reduction the outer residual still lacks
      const double itol =
          java_max(S.inner_rtol, fmin(0.5, 0.5 * java_max(S.rtol * S.init_r, S.atol) / S.rnorm));
      if (rn < itol * S.inner_init || S.inner >= S.max_inner || rn != rn) {
        S.action3 = A3_PROMOTE;
        S.phase = PH_OUTER;
      } else {
        S.beta = tot[0] / S.rho;
        S.rho = tot[0];
        S.action3 = A3_PUPDATE;
      }
    }
  }
}

__device__ __forceinline__ void b3_work(const BDev &d, int wi) {
  const BWork w = d.work[wi];
  BSlot &S = d.slots[w.slot];
  if (!slot_live(S)) return;
  const int act = S.action3;
  if (act == A3_PINIT) {
    TILE_GROUPS(g, full, S, w.tile) st4f(S.p, g, g_end, full, ld4f(S.z, g, g_end, full));
  } else if (act == A3_PUPDATE) {
    const float b = (float)S.beta;
    TILE_GROUPS(g, full, S, w.tile) {
      f4v pv = ld4f(S.p, g, g_end, full);
      const f4v zv = ld4f(S.z, g, g_end, full);
      for (int k = 0; k < 4; ++k) {
        const float pb = pv[k] * b;
        pv[k] = pb + zv[k];
      }
      st4f(S.p, g, g_end, full, pv);
    }
  } else if (act == A3_PROMOTE) {
    const double sc = S.rnorm;
    TILE_ROWS(i, S, w.tile) S.x[i] = S.x[i] + sc * (double)S.d[i];
  }
  // the outer counter advances once per refinement step, in the slot's
  // first tile only (single writer; read next by B1's epilogue)
  if (act == A3_PROMOTE && w.tile == 0 && threadIdx.x == 0) S.outer++;
}

// One launch per phase (the captured graph of a step): workgroup w takes
// work entry w.
__global__ __launch_bounds__(BS) void k_batch_b1(BDev d) {
  if ((int)blockIdx.x < *d.nwork) b1_work(d, blockIdx.x);
}
__global__ __launch_bounds__(BS) void k_batch_b2(BDev d) {
  if ((int)blockIdx.x < *d.nwork) b2_work(d, blockIdx.x);
}
__global__ __launch_bounds__(BS) void k_batch_b3(BDev d) {
  if ((int)blockIdx.x < *d.nwork) b3_work(d, blockIdx.x);
}

// Grid barrier of the persistent step (all workgroups co-resident: the
// kernel is launched cooperatively).  Workgroup leaders count arrivals; the
// last resets the count and advances the generation the others wait on.  The
// agent-scope fences on both sides write this workgroup's stores back to the
// device-wide coherence point and invalidate its stale cached lines, as a
// kernel boundary does.  A wait is bounded: past the budget (about a second)
// the workgroup sets `err` and goes on, and once `err` is set no barrier
// waits, so a broken barrier ends the kernel with an error instead of a
// hang.
struct GridBar {
  uint32_t count, gen;
  int32_t err, pad;
};
constexpr int BAR_SPIN_BUDGET = 1 << 20;

__device__ __forceinline__ void grid_sync(GridBar *b) {
  __syncthreads();
  if (threadIdx.x == 0) {
    __threadfence();
    const uint32_t g = __hip_atomic_load(&b->gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (__hip_atomic_fetch_add(&b->count, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == gridDim.x - 1) {
      __hip_atomic_store(&b->count, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __threadfence();
      __hip_atomic_store(&b->gen, g + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    } else {
      int spins = 0;
      while (__hip_atomic_load(&b->gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == g) {
        if (++spins > BAR_SPIN_BUDGET ||
            __hip_atomic_load(&b->err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) {
          __hip_atomic_store(&b->err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          break;
        }
        __builtin_amdgcn_s_sleep(1);
      }
    }
    __threadfence();
  }
  __syncthreads();
}

// A whole step in one cooperative launch: `iters` batch iterations, each the
// three phases over the work table (grid-stride), separated by grid barriers.
__global__ __launch_bounds__(BS) void k_batch_persist(BDev d, int iters, GridBar *bar) {
  const int nw = *d.nwork;
  for (int it = 0; it < iters; ++it) {
    for (int wi = blockIdx.x; wi < nw; wi += gridDim.x) b1_work(d, wi);
    grid_sync(bar);
    for (int wi = blockIdx.x; wi < nw; wi += gridDim.x) b2_work(d, wi);
    grid_sync(bar);
    for (int wi = blockIdx.x; wi < nw; wi += gridDim.x) b3_work(d, wi);
    if (it + 1 < iters) grid_sync(bar);
  }
}

// Diagonals of a system's FP32 stencil copy (m = 0: none).
struct StDiag {
  int32_t m;
  int32_t off[ST_MAX_OFF];
};

// Admission: the FP32 inverse diagonal and the row's lane of the SELL-64
// copy (columns, FP64 values, and FP32 values unless the system has a
// stencil copy; padding was preset to -1 / 0), or the row's FP32 values on
// its diagonals and its mask byte.
__global__ void k_batch_setup(int32_t n, const int32_t *rp, const int32_t *ci, const double *v,
                              const int32_t *sptr, float *inv, int32_t *scol, double *sv64, float *sv32,
                              StDiag sd, float *st32, uint8_t *stmask) {
  for (int32_t i = blockIdx.x * BS + threadIdx.x; i < n; i += gridDim.x * BS) {
    const int32_t off = sptr[i / SELL_C] + i % SELL_C;
    double dg = 0;
    float out[ST_MAX_OFF];
    uint32_t mk = 0;
#pragma unroll
    for (int k = 0; k < ST_MAX_OFF; ++k) out[k] = 0.0f;
    for (int32_t k = 0, j = rp[i]; j < rp[i + 1]; ++k, ++j) {
      if (ci[j] == i) dg = v[j];
      scol[off + k * SELL_C] = ci[j];
      sv64[off + k * SELL_C] = v[j];
      if (sd.m == 0) {
        sv32[off + k * SELL_C] = (float)v[j];
      } else {
        const int32_t dd_ = ci[j] - i;
#pragma unroll
        for (int t = 0; t < ST_MAX_OFF; ++t)
          if (t < sd.m && sd.off[t] == dd_) {
            out[t] = (float)v[j];
            mk |= 1u << t;
          }
      }
    }
    if (sd.m > 0) {
#pragma unroll
      for (int t = 0; t < ST_MAX_OFF; ++t)
        if (t < sd.m) st32[(size_t)t * n + i] = out[t];
      stmask[i] = (uint8_t)mk;
    }
    inv[i] = (float)(1.0 / dg);
  }
}

__global__ void k_zero_f64(int64_t n, double *x) {
  for (int64_t i = blockIdx.x * (int64_t)BS + threadIdx.x; i < n; i += (int64_t)gridDim.x * BS) x[i] = 0.0;
}

// ---- host side ---------------------------------------------------------

inline size_t al256(size_t b) { return (b + 255) & ~(size_t)255; }

// Device memory of the admitted systems: chunks of device memory, each with a
// first-fit free list in 256-byte granules.  With a fixed size, one chunk of
// that size is allocated at creation and never grows.  Without one, a chunk is
// added when no chunk has room, provided hipMemGetInfo still shows the new
// chunk plus a reserve (kept for the rest of the process) free, so admission
// is bounded by free HBM.  Chunks are freed only when the batch is destroyed
// (hipFree would synchronise the device in the middle of a stream).
struct Arena {
  struct Chunk {
    char *base;
    size_t size;
    std::map<size_t, size_t> free_;  // offset -> length
  };
  std::vector<Chunk> chunks;
  bool fixed = false;
  size_t size = 0, used = 0;
  size_t total = 0;  // the device's memory (hipMemGetInfo at creation)
  static constexpr size_t CHUNK = (size_t)4 << 30;

  void add_chunk(char *base, size_t bytes) {
    Chunk c{base, bytes, {}};
    c.free_[0] = bytes;
    chunks.push_back(std::move(c));
    size += bytes;
  }
  void init_fixed(size_t bytes) {
    char *base = nullptr;
    HIPCHK(hipMalloc(&base, bytes));
    add_chunk(base, bytes);
    fixed = true;
  }
  // Free HBM the growing arena leaves to the rest of the process.
  static size_t reserve(size_t total) { return std::max((size_t)4 << 30, total / 16); }
  // Largest system footprint the arena can ever hold.
  size_t capacity() const {
    if (fixed) return size;
    const size_t res = reserve(total);
    return total > res ? total - res : 0;
  }
  // Device address of `bytes` free bytes (its chunk in *chunk), or nullptr.
  char *alloc(size_t bytes, int *chunk) {
    bytes = al256(bytes);
    for (size_t c = 0; c < chunks.size(); ++c) {
      auto &fl = chunks[c].free_;
      for (auto it = fl.begin(); it != fl.end(); ++it) {
        if (it->second < bytes) continue;
        const size_t off = it->first, len = it->second;
        fl.erase(it);
        if (len > bytes) fl[off + bytes] = len - bytes;
        used += bytes;
        *chunk = (int)c;
        return chunks[c].base + off;
      }
    }
    if (fixed) return nullptr;
    size_t fr = 0, tot = 0;
    HIPCHK(hipMemGetInfo(&fr, &tot));
    const size_t want = std::max(CHUNK, (bytes + ((size_t)256 << 20) - 1) & ~(((size_t)256 << 20) - 1));
    if (fr < want + reserve(tot)) return nullptr;
    char *base = nullptr;
    if (hipMalloc(&base, want) != hipSuccess) {
      (void)hipGetLastError();
      return nullptr;
    }
    add_chunk(base, want);
    return alloc(bytes, chunk);
  }
  void release(int chunk, char *p, size_t bytes) {
    bytes = al256(bytes);
    Chunk &C = chunks.at((size_t)chunk);
    const size_t off = (size_t)(p - C.base);
    if (p < C.base || off + bytes > C.size) fail(MTJG_E_DEVICE, "batch arena: release outside its chunk");
    used -= bytes;
    auto &fl = C.free_;
    auto it = fl.emplace(off, bytes).first;
    auto nx = std::next(it);
    if (nx != fl.end() && it->first + it->second == nx->first) {
      it->second += nx->second;
      fl.erase(nx);
    }
    if (it != fl.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        fl.erase(it);
      }
    }
  }
  void free_all() {
    for (auto &c : chunks) (void)hipFree(c.base);
    chunks.clear();
    size = used = 0;
  }
};

// Pinned host staging owned by the batch.  Every copy between a system's
// host arrays and the device goes through it, so the runtime never stages a
// pageable buffer and every transfer's host range has a known owner and size.
// Ranges are handed out in 256-byte steps; the owner resets (or grows) it
// only after the stream that used its ranges has synchronised.
struct Pinned {
  char *base = nullptr;
  size_t cap = 0, used = 0;
  void grow(size_t bytes) {  // only with no copy in flight
    if (bytes <= cap) return;
    const size_t nc = std::max(bytes, 2 * cap);
    if (base) HIPCHK(hipHostFree(base));
    base = nullptr;
    cap = used = 0;
    HIPCHK(hipHostMalloc((void **)&base, nc, hipHostMallocDefault));
    cap = nc;
  }
  char *take(size_t bytes) {
    const size_t b = al256(bytes);
    if (used + b > cap) return nullptr;
    char *p = base + used;
    used += b;
    return p;
  }
  bool holds(const void *p, size_t bytes) const {
    const char *q = (const char *)p;
    return q >= base && bytes <= cap && (size_t)(q - base) <= cap - bytes;
  }
  void release() {
    if (base) (void)hipHostFree(base);
    base = nullptr;
    cap = used = 0;
  }
};

// Device range [lo, lo + len) that a copy may touch.
struct DevRange {
  const char *lo;
  size_t len;
  bool holds(const void *p, size_t bytes) const {
    const char *q = (const char *)p;
    return q >= lo && bytes <= len && (size_t)(q - lo) <= len - bytes;
  }
};

// hipMemcpyAsync between a pinned staging range and a device range, both
// checked against their owners first.
void copy_checked(void *dst, const void *src, size_t bytes, hipMemcpyKind kind, const Pinned &host,
                  const DevRange &dev, hipStream_t s) {
  const void *h = kind == hipMemcpyHostToDevice ? src : dst;
  const void *d = kind == hipMemcpyHostToDevice ? dst : src;
  if (!host.holds(h, bytes) || !dev.holds(d, bytes))
    fail(MTJG_E_DEVICE, "batch: a copy falls outside its staging buffer or device range");
  HIPCHK(hipMemcpyAsync(dst, src, bytes, kind, s));
}

// Host memory of queued requests and finished solutions: blocks recycled
// through a pool, so that a stream of systems keeps reusing memory it has
// already touched instead of taking fresh pages (a page fault per 4 KiB) for
// every request.  Blocks are taken and returned by the API calls only (not by
// the workers).
class HostPool;
struct HostBlock {
  HostPool *pool = nullptr;
  char *p = nullptr;
  size_t cap = 0;
  HostBlock() = default;
  HostBlock(HostPool *pl, char *q, size_t c) : pool(pl), p(q), cap(c) {}
  HostBlock(HostBlock &&o) noexcept : pool(o.pool), p(o.p), cap(o.cap) { o.p = nullptr; }
  HostBlock &operator=(HostBlock &&o) noexcept {
    if (this != &o) {
      reset();
      pool = o.pool;
      p = o.p;
      cap = o.cap;
      o.p = nullptr;
    }
    return *this;
  }
  HostBlock(const HostBlock &) = delete;
  HostBlock &operator=(const HostBlock &) = delete;
  ~HostBlock() { reset(); }
  inline void reset();
};

class HostPool {
 public:
  // spare blocks kept for reuse, at most this many bytes
  static constexpr size_t KEEP = (size_t)4 << 30;
  ~HostPool() {
    for (auto &kv : free_) std::free(kv.second);
  }
  // A block of at least `bytes` bytes: a spare one at most twice that size,
  // else a new one (sizes rounded to a quarter of their power of two).
  HostBlock get(size_t bytes) {
    bytes = std::max<size_t>(bytes, 4096);
    auto it = free_.lower_bound(bytes);
    if (it != free_.end() && it->first <= 2 * bytes) {
      HostBlock b(this, it->second, it->first);
      spare_ -= it->first;
      free_.erase(it);
      return b;
    }
    size_t q = (size_t)1 << (63 - __builtin_clzll(bytes));
    q = std::max<size_t>(q / 4, 4096);
    const size_t cap = (bytes + q - 1) / q * q;
    char *p = (char *)std::malloc(cap);
    if (!p) fail(MTJG_E_OUT_OF_MEMORY, "batch: host memory for a request");
    return HostBlock(this, p, cap);
  }
  void put(char *p, size_t cap) {
    if (spare_ + cap > KEEP) {
      std::free(p);
      return;
    }
    free_.emplace(cap, p);
    spare_ += cap;
  }

 private:
  std::multimap<size_t, char *> free_;
  size_t spare_ = 0;
};

inline void HostBlock::reset() {
  if (p) pool->put(p, cap);
  p = nullptr;
}

// A few worker threads for the host side of large requests: checking and
// copying a system's arrays at submit, packing them into pinned staging at
// admission, copying solutions out at retirement.  run(parts, f) calls
// f(0 .. parts - 1) on the workers and the calling thread and returns when
// all have finished; f must not throw.
class Workers {
 public:
  explicit Workers(int n) {
    for (int i = 0; i < n; ++i) th_.emplace_back([this] { loop(); });
  }
  ~Workers() {
    {
      std::lock_guard<std::mutex> l(mu_);
      stop_ = true;
    }
    cv_.notify_all();
    for (auto &t : th_) t.join();
  }
  int size() const { return (int)th_.size() + 1; }
  void run(int parts, const std::function<void(int)> &f) {
    if (parts <= 1 || th_.empty()) {
      for (int p = 0; p < parts; ++p) f(p);
      return;
    }
    {
      std::lock_guard<std::mutex> l(mu_);
      job_ = &f;
      parts_ = parts;
      next_ = 0;
      left_ = parts;
      ++gen_;
    }
    cv_.notify_all();
    work();
    std::unique_lock<std::mutex> l(mu_);
    done_.wait(l, [this] { return left_ == 0; });
    job_ = nullptr;
  }

 private:
  void work() {
    for (;;) {
      int p;
      const std::function<void(int)> *f;
      {
        std::lock_guard<std::mutex> l(mu_);
        if (!job_ || next_ >= parts_) return;
        p = next_++;
        f = job_;
      }
      (*f)(p);
      {
        std::lock_guard<std::mutex> l(mu_);
        if (--left_ == 0) done_.notify_all();
      }
    }
  }
  void loop() {
    uint64_t seen = 0;
    std::unique_lock<std::mutex> l(mu_);
    for (;;) {
      cv_.wait(l, [&] { return stop_ || gen_ != seen; });
      if (stop_) return;
      seen = gen_;
      l.unlock();
      work();
      l.lock();
    }
  }
  std::vector<std::thread> th_;
  std::mutex mu_;
  std::condition_variable cv_, done_;
  const std::function<void(int)> *job_ = nullptr;
  int parts_ = 0, next_ = 0, left_ = 0;
  uint64_t gen_ = 0;
  bool stop_ = false;
};

// memcpy split over the workers when large (parts of at least 2 MiB).
void par_copy(Workers &W, void *dst, const void *src, size_t bytes) {
  constexpr size_t PART = (size_t)2 << 20;
  const int parts = (int)std::min<size_t>((size_t)W.size(), bytes / PART);
  if (parts <= 1) {
    if (bytes) memcpy(dst, src, bytes);
    return;
  }
  const size_t step = ((bytes + parts - 1) / parts + 63) & ~(size_t)63;
  W.run(parts, [&](int p) {
    const size_t lo = std::min(bytes, (size_t)p * step), hi = std::min(bytes, lo + step);
    if (hi > lo) memcpy((char *)dst + lo, (const char *)src + lo, hi - lo);
  });
}

struct Request {
  int64_t id;
  int32_t n;
  int64_t nnz;
  HostBlock mem;                // rp | ci | val | b, each 64-byte aligned
  int32_t *rp, *ci;
  double *val, *b;
  std::vector<int32_t> sptr;    // SELL-64 slice offsets
  StDiag st{};                  // diagonals of the FP32 stencil copy (st.m = 0: none)
  int32_t nb = 0;               // tiles
  double rtol;
  int32_t max_outer, max_inner;
  double inner_rtol;
};

struct Done {
  int32_t status, outer, inner;
  double residual;
  int32_t n = 0;
  bool ready = false;           // x holds the solution (its download drained)
  HostBlock x;                  // n doubles
};

// A solution download in flight: pinned staging source, and the arena range
// it reads (freed once drained).
struct Download {
  int64_t id;
  const char *h;
  int32_t n;
  int chunk;
  char *dev;
  size_t bytes;
};

struct HostSlot {
  int64_t id = -1;  // request id, -1 when free
  char *dev = nullptr;  // the system's arena range
  size_t bytes = 0;
  int chunk = -1;
  int32_t n = 0, nb = 0;
  double *x = nullptr;
};

inline size_t al64(size_t b) { return (b + 63) & ~(size_t)63; }

// Host bytes of a request's arrays, laid out rp | ci | val | b.
size_t req_host_bytes(size_t n, size_t nnz) { return al64(4 * (n + 1)) + al64(4 * nnz) + al64(8 * nnz) + 8 * n; }

void place_arrays(Request &R) {
  const size_t n = (size_t)R.n, nnz = (size_t)R.nnz;
  char *p = R.mem.p;
  R.rp = (int32_t *)p;
  p += al64(4 * (n + 1));
  R.ci = (int32_t *)p;
  p += al64(4 * nnz);
  R.val = (double *)p;
  p += al64(8 * nnz);
  R.b = (double *)p;
}

// Threads for the host work of large requests: the calling thread plus up to
// seven workers (fewer on a host with fewer hardware threads).
int host_workers() {
  const unsigned hw = std::thread::hardware_concurrency();
  return (int)std::max(0u, std::min(hw, 8u) - (hw > 0 ? 1u : 0u));
}

// Adds the column - row offset d to the ascending set off[0..m) of at most
// ST_MAX_OFF; false when it does not fit.
bool diag_insert(int32_t *off, int32_t &m, int32_t d) {
  int k = 0;
  while (k < m && off[k] < d) ++k;
  if (k < m && off[k] == d) return true;
  if (m == ST_MAX_OFF) return false;
  for (int t = m; t > k; --t) off[t] = off[t - 1];
  off[k] = d;
  ++m;
  return true;
}

// Checks a system's CSR structure and copies it into R (whose block is
// sized for it), in parallel parts of rows on the workers.  Each row is
// checked in the order a serial pass would: its pointers (0 <= rp[i] <=
// rp[i+1] <= nnz), each entry's column, the diagonal's value, the diagonal's
// presence.  A part stops at its first failing row and copies nothing; the
// error reported is the lowest failing row's, as a serial pass would report.
void check_and_copy(Workers &W, Request &R, const int32_t *rp, const int32_t *ci, const double *val,
                    const double *b) {
  const int32_t n = R.n;
  const int64_t nnz = R.nnz;
  enum { OK_ = 0, BAD_PTR, BAD_COL, ZERO_DIAG, NO_DIAG };
  struct Fail {
    int32_t row = INT32_MAX;
    int kind = OK_;
  };
  const int parts = nnz < (1 << 16) ? 1 : std::min(4 * W.size(), std::max(1, n / 256));
  const int32_t per = (int32_t)(((int64_t)n + parts - 1) / parts);
  std::vector<Fail> fails((size_t)parts);
  std::vector<StDiag> diags((size_t)parts);
  W.run(parts, [&](int p) {
    const int32_t r0 = (int32_t)std::min<int64_t>(n, (int64_t)p * per);
    const int32_t r1 = (int32_t)std::min<int64_t>(n, (int64_t)r0 + per);
    // the part's diagonals (column - row offsets) for the FP32 stencil
    // copy: a row whose offsets repeat the previous row's adds nothing
    StDiag &sd = diags[(size_t)p];
    sd.m = 0;
    int32_t last[ST_MAX_OFF];
    int32_t lastn = -1;
    for (int32_t i = r0; i < r1; ++i) {
      const int32_t a = rp[i], e = rp[i + 1];
      if (a < 0 || e < a || e > nnz) {
        fails[p] = Fail{i, BAD_PTR};
        return;
      }
      bool diag = false;
      for (int32_t k = a; k < e; ++k) {
        const int32_t c = ci[k];
        if (c < 0 || c >= n) {
          fails[p] = Fail{i, BAD_COL};
          return;
        }
        if (c == i) {
          if (val[k] == 0.0) {
            fails[p] = Fail{i, ZERO_DIAG};
            return;
          }
          diag = true;
        }
      }
      if (!diag) {
        fails[p] = Fail{i, NO_DIAG};
        return;
      }
      if (sd.m >= 0) {
        bool same = e - a == lastn;
        for (int32_t k = a; same && k < e; ++k) same = ci[k] - i == last[k - a];
        if (!same) {
          if (e - a > ST_MAX_OFF) {
            sd.m = -1;
          } else {
            for (int32_t k = a; k < e && sd.m >= 0; ++k) {
              last[k - a] = ci[k] - i;
              if (!diag_insert(sd.off, sd.m, ci[k] - i)) sd.m = -1;
            }
            lastn = e - a;
          }
        }
      }
    }
    // the part's rows hold the entries rp[r0] .. rp[r1] (checked above)
    if (r1 > r0) {
      const size_t k0 = (size_t)rp[r0], k1 = (size_t)rp[r1];
      memcpy(R.ci + k0, ci + k0, 4 * (k1 - k0));
      memcpy(R.val + k0, val + k0, 8 * (k1 - k0));
    }
    memcpy(R.rp + r0, rp + r0, 4 * (size_t)(r1 - r0));
    memcpy(R.b + r0, b + r0, 8 * (size_t)(r1 - r0));
  });
  for (const Fail &f : fails) {
    if (f.kind == OK_) continue;
    const std::string row = std::to_string(f.row);
    if (f.kind == BAD_PTR) fail(MTJG_E_INVALID_ARGUMENT, "row pointers must be non-decreasing");
    if (f.kind == BAD_COL) fail(MTJG_E_INVALID_ARGUMENT, "column index out of range in row " + row);
    if (f.kind == ZERO_DIAG) fail(MTJG_E_ZERO_DIAGONAL, "zero diagonal in row " + row);
    fail(MTJG_E_MISSING_DIAGONAL, "missing diagonal in row " + row);
  }
  R.rp[n] = rp[n];
  // the union of the parts' diagonals (MTJG_SPMV_STENCIL=0 at submission
  // keeps the SELL-64 FP32 values)
  R.st.m = 0;
  const char *env = getenv("MTJG_SPMV_STENCIL");
  bool fits = !(env && atoi(env) == 0);
  for (size_t p = 0; fits && p < diags.size(); ++p) {
    fits = diags[p].m >= 0;
    for (int k = 0; fits && k < diags[p].m; ++k) fits = diag_insert(R.st.off, R.st.m, diags[p].off[k]);
  }
  if (!fits) R.st.m = 0;
}

// Bytes of the FP32 values: the stencil copy (values per diagonal and a mask
// byte per row) or the SELL-64 FP32 values.
size_t req_fp32_bytes(const Request &R) {
  const size_t n = (size_t)R.n;
  if (R.st.m > 0) return al256(4 * (size_t)R.st.m * n) + al256(n);
  return al256(4 * (size_t)R.sptr.back());
}

size_t req_bytes(const Request &R) {
  const size_t n = (size_t)R.n, nnz = (size_t)R.nnz, sell = (size_t)R.sptr.back();
  return al256(4 * (n + 1)) + al256(4 * nnz) + al256(8 * nnz) + al256(4 * n) + 3 * al256(8 * n) +
         5 * al256(4 * n) + al256(4 * R.sptr.size()) + al256(4 * sell) + al256(8 * sell) + req_fp32_bytes(R);
}

// Slice offsets of the system's SELL-64 copy (each slice as wide as its
// longest row) and its tile count for tiles of `rows` rows.
void build_sell(Workers &W, Request &R, int rows) {
  const int32_t n = R.n;
  const int32_t nslice = (n + SELL_C - 1) / SELL_C;
  R.sptr.assign((size_t)nslice + 1, 0);
  // slice widths in parallel parts, then the prefix sum
  const int parts = std::max(1, std::min(W.size(), nslice / 4096));
  const int32_t per = (nslice + parts - 1) / parts;
  W.run(parts, [&](int p) {
    for (int32_t sl = p * per; sl < std::min(nslice, (p + 1) * per); ++sl) {
      int32_t w = 0;
      for (int32_t r = sl * SELL_C; r < std::min(n, (sl + 1) * SELL_C); ++r) w = std::max(w, R.rp[r + 1] - R.rp[r]);
      R.sptr[(size_t)sl + 1] = w;
    }
  });
  int64_t total = 0;
  for (int32_t sl = 0; sl < nslice; ++sl) {
    total += (int64_t)R.sptr[(size_t)sl + 1] * SELL_C;
    if (total > INT32_MAX) fail(MTJG_E_UNSUPPORTED, "system too large for the batch engine");
    R.sptr[(size_t)sl + 1] = (int32_t)total;
  }
  R.nb = (n + rows - 1) / rows;
}

}  // namespace
}  // namespace mtjg

using namespace mtjg;

struct mtjg_batch_s {
  mtjg_ctx ctx = nullptr;
  hipStream_t copy = nullptr;  // uploads of admitted systems
  hipStream_t dl = nullptr;    // solution downloads
  bool uploads_in_flight = false;
  int max_slots = 0, max_work = 0, iters = 32;
  int rows_per_wg = 1024;
  Arena arena;
  Pinned up, down;              // staging of system uploads and solution downloads
  std::unique_ptr<Workers> workers;
  HostPool pool;                // host blocks of queued requests and solutions (outlives both)
  BSlot *d_slots = nullptr;     // device slot table
  uint32_t *d_ticket = nullptr;  // per-slot tickets (TICKET_STRIDE apart), zero between phases
  BWork *d_work = nullptr;
  int32_t *d_nwork = nullptr;
  double *d_partials = nullptr;
  // persistent step (MTJG_BATCH_PERSISTENT=1): one cooperative launch per
  // step with grid barriers instead of the captured graph of 3 x iters
  // launches.  Bitwise the same results, but on the 8,192-request stream the
  // device time was 12.7 s against 4.2 s for the graph (0.94 against 0.26 ms
  // per batch iteration): the co-resident grid walks the work table several
  // times per phase and every barrier writes back and invalidates each
  // workgroup's L2 view, so the graph stays the default.
  bool persistent = false;
  int max_coresident = 0;       // workgroups of k_batch_persist resident at once
  GridBar *d_bar = nullptr;
  int32_t *h_bar_err[2] = {nullptr, nullptr};
  // pinned staging, double-buffered: work table (+ count), new slot records,
  // slot-table snapshots
  BWork *h_work[2] = {nullptr, nullptr};
  BSlot *h_new[2] = {nullptr, nullptr};
  BSlot *h_snap[2] = {nullptr, nullptr};
  hipEvent_t ev_stage[2] = {nullptr, nullptr};  // staging buffer i consumed
  hipEvent_t ev_snap[2] = {nullptr, nullptr};   // snapshot i written
  hipEvent_t ev_g0[2] = {nullptr, nullptr}, ev_g1[2] = {nullptr, nullptr};  // around step i's graph
  double busy_ms = 0;                           // device time of the graphs whose snapshots were consumed
  // host time (ms) in: submit, admission, waiting for snapshots, retiring,
  // retiring's wait for solution downloads, admission's wait for uploads,
  // publishing tables, launching steps
  double host_ms[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  hipEvent_t ev_copy = nullptr;                 // uploads on the copy stream done
  int stage = 0;
  int snap_pending = -1, snap_hi = 0;          // buffer of the unprocessed snapshot
  int table_work = 0;                           // entries in the device work table
  int64_t table_rows = 0;
  std::vector<HostSlot> slots;
  std::deque<Request> queue;
  std::map<int64_t, Done> done;
  std::vector<Download> downloads;  // started by the last retirement, not yet drained
  std::vector<int64_t> finished;  // ids retired since the last mtjg_batch_completed
  int64_t next_id = 1;
  std::map<int, hipGraphExec_t> graphs;  // per grid size
  int64_t steps = 0;
};

namespace {

double now_ms() {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

BDev bdev(mtjg_batch B) {
  return BDev{B->d_slots, B->d_work, B->d_nwork, B->d_partials, B->max_work, B->d_ticket};
}

// Admitted requests not yet reported finished: those in a slot and those
// whose solution download is still to be drained.
int active_slots(mtjg_batch B) {
  int a = 0;
  for (const auto &h : B->slots) a += h.id >= 0;
  return a + (int)B->downloads.size();
}

int used_work(mtjg_batch B) {
  int u = 0;
  for (const auto &h : B->slots)
    if (h.id >= 0) u += h.nb;
  return u;
}

hipGraphExec_t graph_for(mtjg_batch B, int grid) {
  auto it = B->graphs.find(grid);
  if (it != B->graphs.end()) return it->second;
  hipStream_t st = B->ctx->stream;
  hipGraph_t g;
  BDev d = bdev(B);
  HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
  for (int k = 0; k < B->iters; ++k) {
    k_batch_b1<<<grid, BS, 0, st>>>(d);
    k_batch_b2<<<grid, BS, 0, st>>>(d);
    k_batch_b3<<<grid, BS, 0, st>>>(d);
  }
  HIPCHK(hipStreamEndCapture(st, &g));
  hipGraphExec_t ex;
  HIPCHK(hipGraphInstantiate(&ex, g, nullptr, nullptr, 0));
  HIPCHK(hipGraphDestroy(g));
  HIPCHK(hipGraphUpload(ex, st));
  B->graphs[grid] = ex;
  return ex;
}

// Finishes the solution downloads the previous retirement started: waits for
// the download stream (by now the copies are normally complete), moves each
// solution into its Done record, reports the request finished and frees its
// arena range (kept until here because the copy read it).  Returns the count.
int drain_downloads(mtjg_batch B) {
  auto &dls = B->downloads;
  if (dls.empty()) return 0;
  const double t = now_ms();
  HIPCHK(hipStreamSynchronize(B->dl));
  B->host_ms[4] += now_ms() - t;
  size_t total = 0;
  for (const Download &d : dls) total += 8 * (size_t)d.n;
  auto out = [&](const Download &d) {
    auto it = B->done.find(d.id);
    if (it == B->done.end()) return;  // cannot happen: Done records outlive their downloads
    memcpy(it->second.x.p, d.h, 8 * (size_t)d.n);
  };
  if (total < ((size_t)4 << 20)) {
    for (const Download &d : dls) out(d);
  } else {
    B->workers->run((int)dls.size(), [&](int k) { out(dls[(size_t)k]); });
  }
  for (const Download &d : dls) {
    B->done[d.id].ready = true;
    B->finished.push_back(d.id);
    B->arena.release(d.chunk, d.dev, d.bytes);
  }
  const int c = (int)dls.size();
  dls.clear();
  B->down.used = 0;
  return c;
}

// Retires: finishes the previous retirement's downloads, then, for each
// system the pending snapshot shows finished, records its outcome, starts the
// download of its solution into the pinned download staging (download
// stream) and frees its slot.  The host has waited for the snapshot's event
// before it reads the snapshot, so the kernels that finished these systems
// are complete and the copies, enqueued after that wait, read their final
// solutions; no later kernel touches a finished system's memory (every batch
// kernel returns at once for a slot that is not running), and its arena range
// is freed only when the download has been drained.  The downloads run
// while the step launched last runs and are drained by the next retirement,
// so the host never waits for them in the step that starts them.  Returns
// the number of requests reported finished.
int retire(mtjg_batch B) {
  int n = drain_downloads(B);
  if (B->snap_pending < 0) return n;
  const int i = B->snap_pending;
  B->snap_pending = -1;
  const double tw = now_ms();
  HIPCHK(hipEventSynchronize(B->ev_snap[i]));
  const double tr = now_ms();
  B->host_ms[2] += tr - tw;
  float ms = 0;
  HIPCHK(hipEventElapsedTime(&ms, B->ev_g0[i], B->ev_g1[i]));
  B->busy_ms += ms;
  if (B->persistent && *B->h_bar_err[i])
    fail(MTJG_E_DEVICE, "batch: a grid barrier of the persistent step timed out");
  const BSlot *snap = B->h_snap[i];
  Pinned &dn = B->down;
  if (B->snap_hi > B->max_slots) fail(MTJG_E_DEVICE, "batch: snapshot larger than the slot table");
  for (int s = 0; s < B->snap_hi; ++s) {
    HostSlot &H = B->slots[s];
    const BSlot &S = snap[s];
    if (H.id < 0 || S.rid != H.id || S.status == ST_RUNNING) continue;
    const size_t bytes = 8 * (size_t)H.n;
    char *h = dn.take(bytes);
    if (!h) {
      n += drain_downloads(B);  // frees the staging (waits for the copies just started)
      dn.grow(al256(bytes));
      h = dn.take(bytes);
    }
    Done &D = B->done[H.id];
    D.status = S.status;
    D.outer = S.outer;
    D.inner = S.inner_total;
    D.residual = S.rnorm;
    D.n = H.n;
    D.ready = false;
    D.x = B->pool.get(bytes);
    copy_checked(h, H.x, bytes, hipMemcpyDeviceToHost, dn, DevRange{H.dev, H.bytes}, B->dl);
    B->downloads.push_back(Download{H.id, h, H.n, H.chunk, H.dev, H.bytes});
    H = HostSlot{};
  }
  B->host_ms[3] += now_ms() - tr;
  return n;
}

// Admits queued requests while a slot, work-table room and arena memory are
// free (first come, first served; a request that does not fit blocks the
// queue, so large requests are not starved).  Each system's arrays are
// packed into the batch's pinned upload staging and copied from there on the
// copy stream; the new slot records are staged in h_new[stage].
int admit(mtjg_batch B, std::vector<int> &new_slots) {
  hipStream_t cs = B->copy;
  Pinned &up = B->up;
  if (B->uploads_in_flight) {
    // the previous admission's uploads (normally complete by now) free the staging
    const double t = now_ms();
    HIPCHK(hipStreamSynchronize(cs));
    B->host_ms[5] += now_ms() - t;
    B->uploads_in_flight = false;
  }
  up.used = 0;
  int admitted = 0;
  bool in_flight = false;
  int used = used_work(B);
  while (!B->queue.empty()) {
    Request &R = B->queue.front();
    int slot = -1;
    for (int s = 0; s < B->max_slots; ++s)
      if (B->slots[s].id < 0) {
        slot = s;
        break;
      }
    if (slot < 0) break;
    const int nb = R.nb;
    if (used + nb > B->max_work) break;
    const size_t n = (size_t)R.n, nnz = (size_t)R.nnz, nsp = R.sptr.size(), sell = (size_t)R.sptr.back();
    if (!R.mem.p || R.mem.cap < req_host_bytes(n, nnz))
      fail(MTJG_E_DEVICE, "batch: request arrays do not match its sizes");
    const size_t bytes = req_bytes(R);
    int chunk = -1;
    char *base = B->arena.alloc(bytes, &chunk);
    if (!base) break;
    const DevRange dev{base, bytes};
    char *p = base;
    auto take = [&](size_t b) {
      char *q = p;
      p += al256(b);
      return q;
    };
    int32_t *rp = (int32_t *)take(4 * (n + 1));
    int32_t *ci = (int32_t *)take(4 * nnz);
    double *v64 = (double *)take(8 * nnz);
    float *inv = (float *)take(4 * n);
    double *b = (double *)take(8 * n);
    double *x = (double *)take(8 * n);
    double *r = (double *)take(8 * n);
    float *rr = (float *)take(4 * n);
    float *dv = (float *)take(4 * n);
    float *pv = (float *)take(4 * n);
    float *qv = (float *)take(4 * n);
    float *zv = (float *)take(4 * n);
    int32_t *sptr = (int32_t *)take(4 * nsp);
    int32_t *scol = (int32_t *)take(4 * sell);
    double *sv64 = (double *)take(8 * sell);
    const bool stencil = R.st.m > 0;
    float *sv32 = stencil ? nullptr : (float *)take(4 * sell);
    float *st32 = stencil ? (float *)take(4 * (size_t)R.st.m * n) : nullptr;
    uint8_t *stmask = stencil ? (uint8_t *)take(n) : nullptr;
    if ((size_t)(p - base) != al256(bytes)) fail(MTJG_E_DEVICE, "batch: system layout does not match its size");
    // host side: the five uploaded arrays, packed into the pinned staging
    const size_t hb = al256(4 * (n + 1)) + al256(4 * nnz) + al256(8 * nnz) + al256(8 * n) + al256(4 * nsp);
    if (up.used + hb > up.cap) {
      if (in_flight) HIPCHK(hipStreamSynchronize(cs));
      in_flight = false;
      up.used = 0;
      up.grow(hb);
    }
    auto stage = [&](void *dst, const void *src, size_t bytes_) {
      char *h = up.take(bytes_);
      if (!h) fail(MTJG_E_DEVICE, "batch: upload staging overflow");
      par_copy(*B->workers, h, src, bytes_);
      copy_checked(dst, h, bytes_, hipMemcpyHostToDevice, up, dev, cs);
    };
    stage(rp, R.rp, 4 * (n + 1));
    stage(ci, R.ci, 4 * nnz);
    stage(v64, R.val, 8 * nnz);
    stage(b, R.b, 8 * n);
    stage(sptr, R.sptr.data(), 4 * nsp);
    in_flight = true;
    if (sell) {
      HIPCHK(hipMemsetAsync(scol, 0xFF, 4 * sell, cs));  // padding column -1
      HIPCHK(hipMemsetAsync(sv64, 0, 8 * sell, cs));
      if (sv32) HIPCHK(hipMemsetAsync(sv32, 0, 4 * sell, cs));
    }
    k_zero_f64<<<grid_for(R.n), BS, 0, cs>>>(R.n, x);
    k_batch_setup<<<grid_for(R.n), BS, 0, cs>>>(R.n, rp, ci, v64, sptr, inv, scol, sv64, sv32, R.st, st32, stmask);
    HIPCHK(hipGetLastError());
    BSlot S{};
    S.rp = rp;
    S.ci = ci;
    S.v64 = v64;
    S.sptr = sptr;
    S.scol = scol;
    S.sv64 = sv64;
    S.sv32 = sv32;
    S.st32 = st32;
    S.stmask = stmask;
    S.st_m = R.st.m;
    for (int k = 0; k < ST_MAX_OFF; ++k) S.st_off[k] = R.st.off[k];
    S.inv = inv;
    S.b = b;
    S.x = x;
    S.r = r;
    S.rr = rr;
    S.d = dv;
    S.p = pv;
    S.q = qv;
    S.z = zv;
    S.rid = R.id;
    S.n = R.n;
    S.nb = nb;
    S.rows_per_wg = B->rows_per_wg;
    S.active = 1;
    S.status = ST_RUNNING;
    S.phase = PH_OUTER;
    S.action3 = A3_NONE;
    S.max_outer = R.max_outer;
    S.max_inner = R.max_inner;
    S.rtol = R.rtol;
    S.atol = 1e-50;
    S.dtol = 1e5;
    S.inner_rtol = R.inner_rtol;
    B->h_new[B->stage][slot] = S;
    new_slots.push_back(slot);
    HostSlot &H = B->slots[slot];
    H.id = R.id;
    H.dev = base;
    H.bytes = bytes;
    H.chunk = chunk;
    H.n = R.n;
    H.nb = nb;
    H.x = x;
    used += nb;
    B->queue.pop_front();  // the host arrays are staged: R may go
    ++admitted;
  }
  // the uploads run on; the next admission waits for them before it reuses
  // the staging, and the main stream waits for them (publish_tables) before
  // the next step
  B->uploads_in_flight = in_flight;
  return admitted;
}

// Stages and enqueues (main stream, after everything already queued there)
// the new slot records and the rebuilt work table.
void publish_tables(mtjg_batch B, const std::vector<int> &new_slots, bool uploads) {
  hipStream_t st = B->ctx->stream;
  const int i = B->stage;
  if (uploads) {
    HIPCHK(hipEventRecord(B->ev_copy, B->copy));
    HIPCHK(hipStreamWaitEvent(st, B->ev_copy, 0));
  }
  for (int s : new_slots) {
    if (s < 0 || s >= B->max_slots) fail(MTJG_E_DEVICE, "batch: new slot outside the slot table");
    HIPCHK(hipMemcpyAsync(B->d_slots + s, &B->h_new[i][s], sizeof(BSlot), hipMemcpyHostToDevice, st));
  }
  BWork *w = B->h_work[i];
  int nw = 0;
  int64_t rows = 0;
  int off = 0;
  for (int s = 0; s < B->max_slots; ++s) {
    const HostSlot &H = B->slots[s];
    if (H.id < 0) continue;
    if (nw + H.nb > B->max_work) fail(MTJG_E_DEVICE, "batch: work table overflow");
    for (int t = 0; t < H.nb; ++t) w[nw++] = BWork{s, t, off, 0};
    off += H.nb;
    rows += H.n;
  }
  // the count rides behind the entries in the same pinned buffer
  int32_t *cnt = reinterpret_cast<int32_t *>(w + B->max_work);
  *cnt = nw;
  if (nw) HIPCHK(hipMemcpyAsync(B->d_work, w, sizeof(BWork) * nw, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(B->d_nwork, cnt, sizeof(int32_t), hipMemcpyHostToDevice, st));
  HIPCHK(hipEventRecord(B->ev_stage[i], st));
  B->table_work = nw;
  B->table_rows = rows;
  B->stage ^= 1;
  // the next stage's buffers must be free before the host writes them
  HIPCHK(hipEventSynchronize(B->ev_stage[B->stage]));
}

}  // namespace

// ===========================================================================
// C ABI

extern "C" int mtjg_batch_create(mtjg_ctx ctx, int32_t max_slots, int32_t max_workgroups,
                                 int64_t arena_bytes, int32_t rows_per_workgroup, int32_t iters_per_step,
                                 mtjg_batch *out) {
  MTJG_API_BEGIN
  if (!ctx || !out) fail(MTJG_E_NULL, "null argument");
  if (max_slots <= 0 || max_workgroups <= 0 || arena_bytes < 0 || iters_per_step <= 0)
    fail(MTJG_E_INVALID_ARGUMENT, "batch sizes must be positive");
  if (rows_per_workgroup < BS || rows_per_workgroup % BS)
    fail(MTJG_E_INVALID_ARGUMENT, "rows per workgroup must be a positive multiple of 256");
  auto *B = new mtjg_batch_s();
  B->ctx = ctx;
  B->max_slots = max_slots;
  B->max_work = max_workgroups;
  B->iters = iters_per_step;
  B->rows_per_wg = rows_per_workgroup;
  if (arena_bytes > 0) {
    B->arena.init_fixed((size_t)arena_bytes);
  } else {
    size_t fr = 0;
    HIPCHK(hipMemGetInfo(&fr, &B->arena.total));
  }
  B->workers.reset(new Workers(host_workers()));
  B->up.grow((size_t)64 << 20);
  B->down.grow((size_t)32 << 20);
  HIPCHK(hipStreamCreateWithFlags(&B->copy, hipStreamNonBlocking));
  HIPCHK(hipStreamCreateWithFlags(&B->dl, hipStreamNonBlocking));
  HIPCHK(hipMalloc(&B->d_slots, sizeof(BSlot) * max_slots));
  HIPCHK(hipMemsetAsync(B->d_slots, 0, sizeof(BSlot) * max_slots, ctx->stream));
  HIPCHK(hipMalloc(&B->d_ticket, sizeof(uint32_t) * TICKET_STRIDE * (size_t)max_slots));
  HIPCHK(hipMemsetAsync(B->d_ticket, 0, sizeof(uint32_t) * TICKET_STRIDE * (size_t)max_slots, ctx->stream));
  HIPCHK(hipMalloc(&B->d_work, sizeof(BWork) * max_workgroups));
  HIPCHK(hipMalloc(&B->d_nwork, sizeof(int32_t)));
  HIPCHK(hipMemsetAsync(B->d_nwork, 0, sizeof(int32_t), ctx->stream));
  HIPCHK(hipMalloc(&B->d_partials, sizeof(double) * 4 * (size_t)max_workgroups));
  if (const char *e = getenv("MTJG_BATCH_PERSISTENT")) B->persistent = atoi(e) != 0;
  if (B->persistent) {
    int dev = 0, cus = 0, per_cu = 0, coop = 0;
    HIPCHK(hipGetDevice(&dev));
    HIPCHK(hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, dev));
    HIPCHK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)k_batch_persist, BS, 0));
    if (!coop || cus <= 0 || per_cu <= 0) fail(MTJG_E_UNSUPPORTED, "batch: no cooperative launch for the persistent step");
    B->max_coresident = cus * per_cu;
    HIPCHK(hipMalloc(&B->d_bar, sizeof(GridBar)));
    HIPCHK(hipMemsetAsync(B->d_bar, 0, sizeof(GridBar), ctx->stream));
  }
  for (int i = 0; i < 2; ++i) {
    HIPCHK(hipHostMalloc(&B->h_work[i], sizeof(BWork) * (max_workgroups + 1), hipHostMallocDefault));
    HIPCHK(hipHostMalloc(&B->h_new[i], sizeof(BSlot) * max_slots, hipHostMallocDefault));
    HIPCHK(hipHostMalloc(&B->h_snap[i], sizeof(BSlot) * max_slots, hipHostMallocDefault));
    HIPCHK(hipHostMalloc(&B->h_bar_err[i], sizeof(int32_t), hipHostMallocDefault));
    *B->h_bar_err[i] = 0;
    HIPCHK(hipEventCreateWithFlags(&B->ev_stage[i], hipEventDisableTiming));
    HIPCHK(hipEventCreateWithFlags(&B->ev_snap[i], hipEventDisableTiming));
    HIPCHK(hipEventCreate(&B->ev_g0[i]));
    HIPCHK(hipEventCreate(&B->ev_g1[i]));
  }
  HIPCHK(hipEventCreateWithFlags(&B->ev_copy, hipEventDisableTiming));
  B->slots.assign(max_slots, HostSlot{});
  // capture, instantiate and launch once (with an empty work table, so every
  // workgroup returns at once) every grid size a step can use, so that no
  // step pays for a graph's first launch
  for (int grid = 64;; grid *= 2) {
    HIPCHK(hipGraphLaunch(graph_for(B, std::min(grid, max_workgroups)), ctx->stream));
    if (grid >= max_workgroups) break;
  }
  HIPCHK(hipStreamSynchronize(ctx->stream));
  *out = B;
  MTJG_API_END
}

extern "C" int mtjg_batch_destroy(mtjg_batch B) {
  MTJG_API_BEGIN
  if (!B) return MTJG_OK;
  HIPCHK(hipStreamSynchronize(B->ctx->stream));
  HIPCHK(hipStreamSynchronize(B->copy));
  HIPCHK(hipStreamSynchronize(B->dl));
  for (auto &kv : B->graphs) (void)hipGraphExecDestroy(kv.second);
  B->arena.free_all();
  B->up.release();
  B->down.release();
  (void)hipFree(B->d_slots);
  (void)hipFree(B->d_ticket);
  (void)hipFree(B->d_work);
  (void)hipFree(B->d_nwork);
  (void)hipFree(B->d_partials);
  (void)hipFree(B->d_bar);
  for (int i = 0; i < 2; ++i) {
    (void)hipHostFree(B->h_work[i]);
    (void)hipHostFree(B->h_new[i]);
    (void)hipHostFree(B->h_snap[i]);
    (void)hipHostFree(B->h_bar_err[i]);
    (void)hipEventDestroy(B->ev_stage[i]);
    (void)hipEventDestroy(B->ev_snap[i]);
    (void)hipEventDestroy(B->ev_g0[i]);
    (void)hipEventDestroy(B->ev_g1[i]);
  }
  (void)hipEventDestroy(B->ev_copy);
  (void)hipStreamDestroy(B->copy);
  (void)hipStreamDestroy(B->dl);
  delete B;
  MTJG_API_END
}

extern "C" int mtjg_batch_submit(mtjg_batch B, int32_t n, int64_t nnz, const int32_t *row_pointer,
                                 const int32_t *column_index, const double *data, const double *b,
                                 double rtol, int32_t max_outer, double inner_rtol, int32_t max_inner,
                                 int64_t *id) {
  MTJG_API_BEGIN
  if (!B || !row_pointer || !b || !id || (nnz > 0 && (!column_index || !data))) fail(MTJG_E_NULL, "null argument");
  const double t0 = now_ms();
  if (n <= 0 || nnz < n) fail(MTJG_E_INVALID_ARGUMENT, "system must be square with a stored diagonal");
  if (row_pointer[0] != 0 || row_pointer[n] != nnz) fail(MTJG_E_INVALID_ARGUMENT, "row pointers must run from 0 to nnz");
  Request R;
  R.n = n;
  R.nnz = nnz;
  R.mem = B->pool.get(req_host_bytes((size_t)n, (size_t)nnz));
  place_arrays(R);
  check_and_copy(*B->workers, R, row_pointer, column_index, data, b);
  build_sell(*B->workers, R, B->rows_per_wg);
  if (R.nb > B->max_work || req_bytes(R) > B->arena.capacity())
    fail(MTJG_E_OUT_OF_MEMORY, "system larger than the batch can ever hold");
  R.id = B->next_id++;
  R.rtol = rtol;
  R.max_outer = max_outer;
  R.inner_rtol = inner_rtol;
  R.max_inner = max_inner;
  *id = R.id;
  B->queue.push_back(std::move(R));
  B->host_ms[0] += now_ms() - t0;
  MTJG_API_END
}

extern "C" int mtjg_batch_step(mtjg_batch B, int32_t *n_active, int32_t *n_completed) {
  MTJG_API_BEGIN
  if (!B) fail(MTJG_E_NULL, "null batch");
  hipStream_t st = B->ctx->stream;
  // 1. run the current table (published by the previous call) and snapshot
  //    the slot table behind it
  if (B->table_work > 0) {
    int grid = 64;
    while (grid < B->table_work) grid *= 2;
    grid = std::min(grid, B->max_work);
    const int i = B->steps & 1;
    const double tl = now_ms();
    if (B->persistent) {
      BDev dv = bdev(B);
      int iters = B->iters;
      GridBar *bar = B->d_bar;
      void *args[] = {&dv, &iters, &bar};
      const int pg = std::min(B->table_work, B->max_coresident);
      HIPCHK(hipEventRecord(B->ev_g0[i], st));
      HIPCHK(hipLaunchCooperativeKernel((const void *)k_batch_persist, dim3(pg), dim3(BS), args, 0, st));
      HIPCHK(hipEventRecord(B->ev_g1[i], st));
      HIPCHK(hipMemcpyAsync(B->h_bar_err[i], &B->d_bar->err, sizeof(int32_t), hipMemcpyDeviceToHost, st));
    } else {
      hipGraphExec_t g = graph_for(B, grid);
      HIPCHK(hipEventRecord(B->ev_g0[i], st));
      HIPCHK(hipGraphLaunch(g, st));
      HIPCHK(hipEventRecord(B->ev_g1[i], st));
    }
    int hi = 0;
    for (int s = 0; s < B->max_slots; ++s)
      if (B->slots[s].id >= 0) hi = s + 1;
    if (hi > B->max_slots) fail(MTJG_E_DEVICE, "batch: snapshot larger than the slot table");
    HIPCHK(hipMemcpyAsync(B->h_snap[i], B->d_slots, sizeof(BSlot) * hi, hipMemcpyDeviceToHost, st));
    HIPCHK(hipEventRecord(B->ev_snap[i], st));
    B->host_ms[7] += now_ms() - tl;
    // 2. while it runs: retire what the previous snapshot shows finished
    const int completed = retire(B);
    B->snap_pending = i;
    B->snap_hi = hi;
    if (n_completed) *n_completed = completed;
    ++B->steps;
  } else {
    // nothing queued on the device: consume a leftover snapshot
    const int completed = retire(B);
    if (n_completed) *n_completed = completed;
  }
  // 3. admit, then publish the next table behind the running graph
  std::vector<int> new_slots;
  const double ta = now_ms();
  const int admitted = admit(B, new_slots);
  B->host_ms[1] += now_ms() - ta;
  bool changed = admitted > 0;
  if (!changed) {
    // a retirement also changes the table
    changed = used_work(B) != B->table_work;
  }
  if (changed) {
    const double tp = now_ms();
    publish_tables(B, new_slots, admitted > 0);
    B->host_ms[6] += now_ms() - tp;
  }
  if (n_active) *n_active = active_slots(B);
  MTJG_API_END
}

extern "C" int mtjg_batch_result(mtjg_batch B, int64_t id, double *x, mtjg_result *result, int32_t *inner_iterations) {
  MTJG_API_BEGIN
  if (!B) fail(MTJG_E_NULL, "null batch");
  auto it = B->done.find(id);
  if (it == B->done.end() || !it->second.ready) fail(MTJG_E_INVALID_ARGUMENT, "request not finished (or unknown)");
  const Done &D = it->second;
  if (x) memcpy(x, D.x.p, sizeof(double) * (size_t)D.n);
  int code = D.status == ST_CONVERGED ? MTJG_OK
             : D.status == ST_ITERATIONS ? MTJG_NOT_CONVERGED_ITERATIONS
                                         : MTJG_NOT_CONVERGED_DIVERGENCE;
  if (result) {
    result->iterations = D.outer;
    result->residual = D.residual;
    result->reason = code;
  }
  if (inner_iterations) *inner_iterations = D.inner;
  B->done.erase(it);
  return code;
  MTJG_API_END
}

extern "C" int mtjg_batch_completed(mtjg_batch B, int64_t *ids, int32_t max_ids, int32_t *count) {
  MTJG_API_BEGIN
  if (!B || !count || (max_ids > 0 && !ids)) fail(MTJG_E_NULL, "null argument");
  const int32_t c = (int32_t)std::min<size_t>(B->finished.size(), (size_t)std::max(max_ids, 0));
  for (int32_t k = 0; k < c; ++k) ids[k] = B->finished[(size_t)k];
  B->finished.erase(B->finished.begin(), B->finished.begin() + c);
  *count = c;
  MTJG_API_END
}

extern "C" int mtjg_batch_info(mtjg_batch B, int32_t *queued, int32_t *active, int64_t *arena_used,
                               int64_t *arena_size) {
  MTJG_API_BEGIN
  if (!B) fail(MTJG_E_NULL, "null batch");
  if (queued) *queued = (int32_t)B->queue.size();
  if (active) *active = active_slots(B);
  if (arena_used) *arena_used = (int64_t)B->arena.used;
  if (arena_size) *arena_size = (int64_t)B->arena.size;
  MTJG_API_END
}

extern "C" int mtjg_batch_stats(mtjg_batch B, double *busy_ms, int64_t *steps, double *host_ms) {
  MTJG_API_BEGIN
  if (!B) fail(MTJG_E_NULL, "null batch");
  if (busy_ms) *busy_ms = B->busy_ms;
  if (steps) *steps = B->steps;
  if (host_ms)
    for (int k = 0; k < 8; ++k) host_ms[k] = B->host_ms[k];
  MTJG_API_END
}
