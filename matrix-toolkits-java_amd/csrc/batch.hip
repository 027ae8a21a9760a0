// Batched, continuously admitted mixed-precision CG for many independent SPD
// systems (the serving path).  No MTJ counterpart: each system is solved the
// way solve_mixed solves one (FP64 residual r = b - A x with FP64 values, FP32
// inner Jacobi-CG on r / ||r|| to a relative tolerance, FP64 update
// x += ||r|| d), but every active system advances in the same launches.
//
// Work layout: system s is cut into nb_s row tiles of rows_per_wg rows; a
// device work table maps each workgroup of a launch to (slot, tile).  One
// batch iteration is three launches over the whole table:
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
// batch.  Between steps the host retires finished slots and admits queued
// requests while slots, work-table entries and arena memory are free; each
// step replays a hipGraph of `iters` batch iterations captured once per grid
// size.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <deque>
#include <map>
#include <vector>

#include "mtjg_internal.hpp"
#include "reduce.hpp"

namespace mtjg {
namespace {

enum : int32_t { PH_OUTER = 0, PH_START = 1, PH_INNER = 2 };
enum : int32_t { A3_NONE = 0, A3_PINIT = 1, A3_PUPDATE = 2, A3_PROMOTE = 3 };

struct BSlot {
  // system and work vectors (device pointers into the arena)
  const int32_t *rp;
  const int32_t *ci;
  const double *v64;
  const float *v32;
  const float *inv;
  const double *b;
  double *x, *r;
  float *rr, *d, *p, *q, *z;
  int32_t n, nb, rows_per_wg, part_off;
  // control
  int32_t active, status, phase, action3;
  int32_t outer, inner, inner_total, max_outer, max_inner;
  double rtol, atol, dtol, init_r, rnorm;
  double rho, alpha, beta, inner_init, inner_rtol;
  uint32_t ticket;
  int32_t pad;
};

struct BWork {
  int32_t slot, tile;
};

struct BDev {
  BSlot *slots;
  const BWork *work;
  const int32_t *nwork;
  double *partials;  // 4 arrays of `cap`: hi/lo of two sums
  int cap;
};

__device__ __forceinline__ bool slot_live(const BSlot &S) { return S.active && S.status == ST_RUNNING; }

// Per-slot two-sum reduction; returns true in the slot's last workgroup with
// the totals in out[].
template <int K>
__device__ bool slot_sum(dd (&v)[K], BSlot &S, int tile, const BDev &d, double (&out)[K]) {
  __shared__ int last;
  block_sum_dd<K>(v);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < K; ++k) {
      store_agent(&d.partials[(2 * k) * d.cap + S.part_off + tile], v[k].hi);
      store_agent(&d.partials[(2 * k + 1) * d.cap + S.part_off + tile], v[k].lo);
    }
    stores_complete();
    uint32_t t = take_ticket(&S.ticket);
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
      acc[k] = dd_add(acc[k], dd{load_agent(&d.partials[(2 * k) * d.cap + S.part_off + j]),
                                 load_agent(&d.partials[(2 * k + 1) * d.cap + S.part_off + j])});
  }
  block_sum_dd<K>(acc);
#pragma unroll
  for (int k = 0; k < K; ++k) out[k] = acc[k].hi + acc[k].lo;
  if (threadIdx.x == 0) __hip_atomic_store(&S.ticket, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  return true;
}

// Rows of tile `tile`: thread t owns rows r0 + t + k*BS.
#define TILE_ROWS(i, S, tile)                                                      \
  for (int32_t i = (tile) * (S).rows_per_wg + threadIdx.x,                         \
               i##_end = min((S).n, ((tile) + 1) * (S).rows_per_wg);               \
       i < i##_end; i += BS)

__global__ __launch_bounds__(BS) void k_batch_b1(BDev d) {
  if ((int)blockIdx.x >= *d.nwork) return;
  const BWork w = d.work[blockIdx.x];
  BSlot &S = d.slots[w.slot];
  if (!slot_live(S)) return;
  const int phase = S.phase;
  if (phase == PH_START) return;
  dd v[1] = {{0.0, 0.0}};
  if (phase == PH_INNER) {
    TILE_ROWS(i, S, w.tile) {
      float s = 0.0f;
      for (int32_t k = S.rp[i]; k < S.rp[i + 1]; ++k) s += S.v32[k] * S.p[S.ci[k]];
      S.q[i] = s;
      dd_fma(v[0], (double)S.p[i], (double)s);
    }
  } else {
    // r = b - A x in MTJ's multAdd order: dot sequential, then b + (-1) dot
    TILE_ROWS(i, S, w.tile) {
      double s = 0.0;
      for (int32_t k = S.rp[i]; k < S.rp[i + 1]; ++k) s += S.v64[k] * S.x[S.ci[k]];
      const double ri = S.b[i] + (-1.0) * s;
      S.r[i] = ri;
      dd_fma(v[0], ri, ri);
    }
  }
  double tot[1];
  if (slot_sum<1>(v, S, w.tile, d, tot) && threadIdx.x == 0) {
    if (phase == PH_INNER) {
      S.alpha = S.rho / tot[0];
    } else {
      const double rn = sqrt(tot[0]);
      S.rnorm = rn;
      if (S.outer == 0) S.init_r = rn;
      if (rn < java_max(S.rtol * S.init_r, S.atol)) S.status = ST_CONVERGED;
      else if (rn > S.dtol * S.init_r) S.status = ST_DIVERGENCE;
      else if (S.outer >= S.max_outer) S.status = ST_ITERATIONS;
      else if (rn != rn) S.status = ST_DIVERGENCE;
      else S.phase = PH_START;
    }
  }
}

__global__ __launch_bounds__(BS) void k_batch_b2(BDev d) {
  if ((int)blockIdx.x >= *d.nwork) return;
  const BWork w = d.work[blockIdx.x];
  BSlot &S = d.slots[w.slot];
  if (!slot_live(S)) return;
  const int phase = S.phase;
  if (phase == PH_OUTER) return;
  dd v[2] = {{0.0, 0.0}, {0.0, 0.0}};
  if (phase == PH_START) {
    const double sc = 1.0 / S.rnorm;
    TILE_ROWS(i, S, w.tile) {
      const float ri = (float)(S.r[i] * sc);
      S.rr[i] = ri;
      S.d[i] = 0.0f;
      const float zi = ri * S.inv[i];
      S.z[i] = zi;
      dd_fma(v[0], (double)ri, (double)zi);
      dd_fma(v[1], (double)ri, (double)ri);
    }
  } else {
    const float a = (float)S.alpha;
    TILE_ROWS(i, S, w.tile) {
      S.d[i] = S.d[i] + a * S.p[i];
      const float ri = S.rr[i] + (-a) * S.q[i];
      S.rr[i] = ri;
      const float zi = ri * S.inv[i];
      S.z[i] = zi;
      dd_fma(v[0], (double)ri, (double)zi);
      dd_fma(v[1], (double)ri, (double)ri);
    }
  }
  double tot[2];
  if (slot_sum<2>(v, S, w.tile, d, tot) && threadIdx.x == 0) {
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
      if (rn < S.inner_rtol * S.inner_init || S.inner >= S.max_inner || rn != rn) {
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

__global__ __launch_bounds__(BS) void k_batch_b3(BDev d) {
  if ((int)blockIdx.x >= *d.nwork) return;
  const BWork w = d.work[blockIdx.x];
  BSlot &S = d.slots[w.slot];
  if (!slot_live(S)) return;
  const int act = S.action3;
  if (act == A3_PINIT) {
    TILE_ROWS(i, S, w.tile) S.p[i] = S.z[i];
  } else if (act == A3_PUPDATE) {
    const float b = (float)S.beta;
    TILE_ROWS(i, S, w.tile) {
      const float pb = S.p[i] * b;
      S.p[i] = pb + S.z[i];
    }
  } else if (act == A3_PROMOTE) {
    const double sc = S.rnorm;
    TILE_ROWS(i, S, w.tile) S.x[i] = S.x[i] + sc * (double)S.d[i];
  }
  // the outer counter advances once per refinement step, in the slot's
  // first tile only (single writer; read next by B1's epilogue)
  if (act == A3_PROMOTE && w.tile == 0 && threadIdx.x == 0) S.outer++;
}

// Clears action3 after B3 so a slot that starts its outer check does not
// repeat the last action (single thread per slot).
__global__ void k_batch_clear(BDev d, int nslots) {
  int s = blockIdx.x * BS + threadIdx.x;
  if (s < nslots) d.slots[s].action3 = A3_NONE;
}

__global__ void k_invdiag_f32(int32_t n, const int32_t *rp, const int32_t *ci, const double *v, float *inv,
                              float *v32, int64_t nnz) {
  int64_t t = blockIdx.x * (int64_t)BS + threadIdx.x;
  for (int64_t k = t; k < nnz; k += (int64_t)gridDim.x * BS) v32[k] = (float)v[k];
  for (int64_t i = t; i < n; i += (int64_t)gridDim.x * BS) {
    double dg = 0;
    for (int32_t k = rp[i]; k < rp[i + 1]; ++k)
      if (ci[k] == (int32_t)i) dg = v[k];
    inv[i] = (float)(1.0 / dg);
  }
}

// ---- host side ---------------------------------------------------------

// Simple first-fit allocator over one device arena.
struct Arena {
  char *base = nullptr;
  size_t size = 0, used = 0;
  std::map<size_t, size_t> free_;  // offset -> length
  void init(size_t bytes) {
    HIPCHK(hipMalloc(&base, bytes));
    size = bytes;
    free_.clear();
    free_[0] = bytes;
  }
  // returns offset or SIZE_MAX
  size_t alloc(size_t bytes) {
    bytes = (bytes + 255) & ~(size_t)255;
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second >= bytes) {
        size_t off = it->first, len = it->second;
        free_.erase(it);
        if (len > bytes) free_[off + bytes] = len - bytes;
        used += bytes;
        return off;
      }
    }
    return SIZE_MAX;
  }
  void release(size_t off, size_t bytes) {
    bytes = (bytes + 255) & ~(size_t)255;
    used -= bytes;
    auto it = free_.emplace(off, bytes).first;
    auto nx = std::next(it);
    if (nx != free_.end() && it->first + it->second == nx->first) {
      it->second += nx->second;
      free_.erase(nx);
    }
    if (it != free_.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        free_.erase(it);
      }
    }
  }
};

struct Request {
  int64_t id;
  int32_t n;
  std::vector<int32_t> rp, ci;
  std::vector<double> val, b;
  double rtol;
  int32_t max_outer, max_inner;
  double inner_rtol;
};

struct Done {
  int32_t status, outer, inner;
  double residual;
  std::vector<double> x;
};

size_t req_bytes(int32_t n, int64_t nnz) {
  auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
  return al(4 * (size_t)(n + 1)) + al(4 * (size_t)nnz) + al(8 * (size_t)nnz) + al(4 * (size_t)nnz) +
         al(4 * (size_t)n) + 3 * al(8 * (size_t)n) + 5 * al(4 * (size_t)n);
}

}  // namespace
}  // namespace mtjg

using namespace mtjg;

struct mtjg_batch_s {
  mtjg_ctx ctx = nullptr;
  int max_slots = 0, max_work = 0, iters = 32;
  int rows_per_wg = 1024;
  Arena arena;
  BSlot *d_slots = nullptr;     // device slot table
  BSlot *h_slots = nullptr;     // pinned mirror
  BWork *d_work = nullptr;
  int32_t *d_nwork = nullptr;
  double *d_partials = nullptr;
  std::vector<BWork> work;
  std::vector<int64_t> slot_req;   // request id per slot (-1 free)
  std::vector<size_t> slot_off;    // arena offset per slot
  std::vector<size_t> slot_bytes;
  std::deque<Request> queue;
  std::map<int64_t, Done> done;
  std::vector<int64_t> finished;  // ids retired since the last mtjg_batch_completed
  int64_t next_id = 1;
  std::map<int, hipGraphExec_t> graphs;  // per grid size
  int64_t steps = 0;
};

namespace {

BDev bdev(mtjg_batch B) {
  return BDev{B->d_slots, B->d_work, B->d_nwork, B->d_partials, B->max_work};
}

// Rebuilds the work table from the active slots (tile-major per slot).
void rebuild_work(mtjg_batch B) {
  B->work.clear();
  int off = 0;
  for (int s = 0; s < B->max_slots; ++s) {
    if (B->slot_req[s] < 0) continue;
    BSlot &S = B->h_slots[s];
    S.part_off = off;
    for (int t = 0; t < S.nb; ++t) B->work.push_back(BWork{s, t});
    off += S.nb;
  }
}

int used_work(mtjg_batch B) {
  int u = 0;
  for (int s = 0; s < B->max_slots; ++s)
    if (B->slot_req[s] >= 0) u += B->h_slots[s].nb;
  return u;
}

void upload_tables(mtjg_batch B) {
  hipStream_t st = B->ctx->stream;
  rebuild_work(B);
  int32_t nw = (int32_t)B->work.size();
  if (nw) HIPCHK(hipMemcpyAsync(B->d_work, B->work.data(), sizeof(BWork) * nw, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(B->d_nwork, &nw, sizeof(int32_t), hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(B->d_slots, B->h_slots, sizeof(BSlot) * B->max_slots, hipMemcpyHostToDevice, st));
  HIPCHK(hipStreamSynchronize(st));
}

// Admits queued requests while a slot, work-table room and arena memory
// are free (first come, first served; a request that does not fit blocks
// the queue, so large requests are not starved).
int admit(mtjg_batch B) {
  hipStream_t st = B->ctx->stream;
  int admitted = 0;
  while (!B->queue.empty()) {
    Request &R = B->queue.front();
    int slot = -1;
    for (int s = 0; s < B->max_slots; ++s)
      if (B->slot_req[s] < 0) {
        slot = s;
        break;
      }
    if (slot < 0) break;
    const int nb = (R.n + B->rows_per_wg - 1) / B->rows_per_wg;
    if (used_work(B) + nb > B->max_work) break;
    const int64_t nnz = (int64_t)R.ci.size();
    const size_t bytes = req_bytes(R.n, nnz);
    const size_t off = B->arena.alloc(bytes);
    if (off == SIZE_MAX) break;
    char *p = B->arena.base + off;
    auto take = [&](size_t b) {
      char *q = p;
      p += (b + 255) & ~(size_t)255;
      return q;
    };
    int32_t *rp = (int32_t *)take(4 * (size_t)(R.n + 1));
    int32_t *ci = (int32_t *)take(4 * (size_t)nnz);
    double *v64 = (double *)take(8 * (size_t)nnz);
    float *v32 = (float *)take(4 * (size_t)nnz);
    float *inv = (float *)take(4 * (size_t)R.n);
    double *b = (double *)take(8 * (size_t)R.n);
    double *x = (double *)take(8 * (size_t)R.n);
    double *r = (double *)take(8 * (size_t)R.n);
    float *rr = (float *)take(4 * (size_t)R.n);
    float *dv = (float *)take(4 * (size_t)R.n);
    float *pv = (float *)take(4 * (size_t)R.n);
    float *qv = (float *)take(4 * (size_t)R.n);
    float *zv = (float *)take(4 * (size_t)R.n);
    HIPCHK(hipMemcpyAsync(rp, R.rp.data(), 4 * (size_t)(R.n + 1), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(ci, R.ci.data(), 4 * (size_t)nnz, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(v64, R.val.data(), 8 * (size_t)nnz, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(b, R.b.data(), 8 * (size_t)R.n, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemsetAsync(x, 0, 8 * (size_t)R.n, st));
    k_invdiag_f32<<<grid_for(std::max<int64_t>(nnz, R.n)), BS, 0, st>>>(R.n, rp, ci, v64, inv, v32, nnz);
    HIPCHK(hipGetLastError());
    BSlot S{};
    S.rp = rp;
    S.ci = ci;
    S.v64 = v64;
    S.v32 = v32;
    S.inv = inv;
    S.b = b;
    S.x = x;
    S.r = r;
    S.rr = rr;
    S.d = dv;
    S.p = pv;
    S.q = qv;
    S.z = zv;
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
    B->h_slots[slot] = S;
    B->slot_req[slot] = R.id;
    B->slot_off[slot] = off;
    B->slot_bytes[slot] = bytes;
    B->queue.pop_front();
    ++admitted;
  }
  return admitted;
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
    k_batch_clear<<<(B->max_slots + BS - 1) / BS, BS, 0, st>>>(d, B->max_slots);
  }
  HIPCHK(hipStreamEndCapture(st, &g));
  hipGraphExec_t ex;
  HIPCHK(hipGraphInstantiate(&ex, g, nullptr, nullptr, 0));
  HIPCHK(hipGraphDestroy(g));
  B->graphs[grid] = ex;
  return ex;
}

// Copies finished slots' results out and frees their slots.
int retire(mtjg_batch B) {
  hipStream_t st = B->ctx->stream;
  HIPCHK(hipMemcpyAsync(B->h_slots, B->d_slots, sizeof(BSlot) * B->max_slots, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  int n = 0;
  for (int s = 0; s < B->max_slots; ++s) {
    if (B->slot_req[s] < 0) continue;
    BSlot &S = B->h_slots[s];
    if (S.status == ST_RUNNING) continue;
    Done D;
    D.status = S.status;
    D.outer = S.outer;
    D.inner = S.inner_total;
    D.residual = S.rnorm;
    D.x.resize((size_t)S.n);
    HIPCHK(hipMemcpyAsync(D.x.data(), S.x, 8 * (size_t)S.n, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    B->done[B->slot_req[s]] = std::move(D);
    B->finished.push_back(B->slot_req[s]);
    B->arena.release(B->slot_off[s], B->slot_bytes[s]);
    B->slot_req[s] = -1;
    S = BSlot{};
    ++n;
  }
  return n;
}

}  // namespace

// ===========================================================================
// C ABI

extern "C" int mtjg_batch_create(mtjg_ctx ctx, int32_t max_slots, int32_t max_workgroups,
                                 int64_t arena_bytes, int32_t rows_per_workgroup, int32_t iters_per_step,
                                 mtjg_batch *out) {
  MTJG_API_BEGIN
  if (!ctx || !out) fail(MTJG_E_NULL, "null argument");
  if (max_slots <= 0 || max_workgroups <= 0 || arena_bytes <= 0 || iters_per_step <= 0)
    fail(MTJG_E_INVALID_ARGUMENT, "batch sizes must be positive");
  if (rows_per_workgroup < BS || rows_per_workgroup % BS)
    fail(MTJG_E_INVALID_ARGUMENT, "rows per workgroup must be a positive multiple of 256");
  auto *B = new mtjg_batch_s();
  B->ctx = ctx;
  B->max_slots = max_slots;
  B->max_work = max_workgroups;
  B->iters = iters_per_step;
  B->rows_per_wg = rows_per_workgroup;
  B->arena.init((size_t)arena_bytes);
  HIPCHK(hipMalloc(&B->d_slots, sizeof(BSlot) * max_slots));
  HIPCHK(hipHostMalloc(&B->h_slots, sizeof(BSlot) * max_slots, hipHostMallocDefault));
  memset(B->h_slots, 0, sizeof(BSlot) * max_slots);
  HIPCHK(hipMalloc(&B->d_work, sizeof(BWork) * max_workgroups));
  HIPCHK(hipMalloc(&B->d_nwork, sizeof(int32_t)));
  HIPCHK(hipMalloc(&B->d_partials, sizeof(double) * 4 * (size_t)max_workgroups));
  B->slot_req.assign(max_slots, -1);
  B->slot_off.assign(max_slots, 0);
  B->slot_bytes.assign(max_slots, 0);
  upload_tables(B);
  *out = B;
  MTJG_API_END
}

extern "C" int mtjg_batch_destroy(mtjg_batch B) {
  MTJG_API_BEGIN
  if (!B) return MTJG_OK;
  HIPCHK(hipStreamSynchronize(B->ctx->stream));
  for (auto &kv : B->graphs) (void)hipGraphExecDestroy(kv.second);
  (void)hipFree(B->arena.base);
  (void)hipFree(B->d_slots);
  (void)hipHostFree(B->h_slots);
  (void)hipFree(B->d_work);
  (void)hipFree(B->d_nwork);
  (void)hipFree(B->d_partials);
  delete B;
  MTJG_API_END
}

extern "C" int mtjg_batch_submit(mtjg_batch B, int32_t n, int64_t nnz, const int32_t *row_pointer,
                                 const int32_t *column_index, const double *data, const double *b,
                                 double rtol, int32_t max_outer, double inner_rtol, int32_t max_inner,
                                 int64_t *id) {
  MTJG_API_BEGIN
  if (!B || !row_pointer || !b || !id || (nnz > 0 && (!column_index || !data))) fail(MTJG_E_NULL, "null argument");
  if (n <= 0 || nnz < n) fail(MTJG_E_INVALID_ARGUMENT, "system must be square with a stored diagonal");
  if (row_pointer[0] != 0 || row_pointer[n] != nnz) fail(MTJG_E_INVALID_ARGUMENT, "row pointers must run from 0 to nnz");
  for (int32_t i = 0; i < n; ++i) {
    if (row_pointer[i + 1] < row_pointer[i]) fail(MTJG_E_INVALID_ARGUMENT, "row pointers must be non-decreasing");
    bool diag = false;
    for (int32_t k = row_pointer[i]; k < row_pointer[i + 1]; ++k) {
      if (column_index[k] < 0 || column_index[k] >= n)
        fail(MTJG_E_INVALID_ARGUMENT, "column index out of range in row " + std::to_string(i));
      if (column_index[k] == i) {
        if (data[k] == 0.0) fail(MTJG_E_ZERO_DIAGONAL, "zero diagonal in row " + std::to_string(i));
        diag = true;
      }
    }
    if (!diag) fail(MTJG_E_MISSING_DIAGONAL, "missing diagonal in row " + std::to_string(i));
  }
  const int nb = (n + B->rows_per_wg - 1) / B->rows_per_wg;
  if (nb > B->max_work || req_bytes(n, nnz) > B->arena.size)
    fail(MTJG_E_OUT_OF_MEMORY, "system larger than the batch can ever hold");
  Request R;
  R.id = B->next_id++;
  R.n = n;
  R.rp.assign(row_pointer, row_pointer + n + 1);
  R.ci.assign(column_index, column_index + nnz);
  R.val.assign(data, data + nnz);
  R.b.assign(b, b + n);
  R.rtol = rtol;
  R.max_outer = max_outer;
  R.inner_rtol = inner_rtol;
  R.max_inner = max_inner;
  *id = R.id;
  B->queue.push_back(std::move(R));
  MTJG_API_END
}

extern "C" int mtjg_batch_step(mtjg_batch B, int32_t *n_active, int32_t *n_completed) {
  MTJG_API_BEGIN
  if (!B) fail(MTJG_E_NULL, "null batch");
  int admitted = admit(B);
  int active = 0;
  for (int s = 0; s < B->max_slots; ++s) active += B->slot_req[s] >= 0;
  if (admitted) upload_tables(B);
  int completed = 0;
  if (active) {
    const int need = (int)B->work.size();
    int grid = 64;
    while (grid < need) grid *= 2;
    grid = std::min(grid, B->max_work);
    HIPCHK(hipGraphLaunch(graph_for(B, grid), B->ctx->stream));
    completed = retire(B);
    if (completed) upload_tables(B);
    ++B->steps;
  }
  if (n_active) {
    int a = 0;
    for (int s = 0; s < B->max_slots; ++s) a += B->slot_req[s] >= 0;
    *n_active = a;
  }
  if (n_completed) *n_completed = completed;
  MTJG_API_END
}

extern "C" int mtjg_batch_result(mtjg_batch B, int64_t id, double *x, mtjg_result *result, int32_t *inner_iterations) {
  MTJG_API_BEGIN
  if (!B) fail(MTJG_E_NULL, "null batch");
  auto it = B->done.find(id);
  if (it == B->done.end()) fail(MTJG_E_INVALID_ARGUMENT, "request not finished (or unknown)");
  const Done &D = it->second;
  if (x) memcpy(x, D.x.data(), sizeof(double) * D.x.size());
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
  int a = 0;
  for (int s = 0; s < B->max_slots; ++s) a += B->slot_req[s] >= 0;
  if (queued) *queued = (int32_t)B->queue.size();
  if (active) *active = a;
  if (arena_used) *arena_used = (int64_t)B->arena.used;
  if (arena_size) *arena_size = (int64_t)B->arena.size;
  MTJG_API_END
}
