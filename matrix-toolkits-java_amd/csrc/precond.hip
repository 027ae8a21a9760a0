// Preconditioners on the device: DiagonalPreconditioner (Jacobi), ILU(0),
// SSOR and ICC (S/DiagonalPreconditioner.java, S/ILU.java,
// S/UnitLowerCompRowMatrix.java, S/UpperCompRowMatrix.java, S/SSOR.java,
// S/ICC.java), and the entry points of the algebraic multigrid preconditioner
// (S/AMG.java), whose cycle lives in amg.hip.  SSOR's two sweeps and ICC's two
// triangular solves use the same level sets as ILU's triangular solves.
//
// ILU(0) is level scheduled.  Rows are grouped on the host into dependency
// levels (a row's level is one more than the highest level among the rows its
// strictly-lower entries reference), sorted by level, and each level is one
// launch with one thread per row.  A thread does exactly MTJ's sequential
// per-row work (the IKJ elimination of S/ILU.java:98-121, or the per-row sum
// of the triangular solves), and every row it reads was finished by an
// earlier launch, so factors and solves are bitwise identical to MTJ's.
// In the solves and SSOR sweeps, stretches of narrow levels (at most 512 rows
// each) run in a single 512-thread workgroup that walks them with a barrier
// between levels: a level launch costs 2.8-4.5 us, a barrier-separated level
// of one pass about 1.6 us.  Wider levels stay separate launches: one compute
// unit gathering a thousand rows is slower than the whole device.
//
// One deliberate difference: MTJ's merge walk (S/ILU.java:115-119) tests
// colind[l] even after l has run past the end of row k, and can then update
// the first entry of row k+1; this implementation stops at the end of row k.
// The two agree on every pattern where row k+1 has a strictly-lower entry,
// which includes all structurally symmetric patterns.
//
// ICC's factorization is not level scheduled: MTJ eliminates each row k
// against every earlier row through a dense work row that keeps the fill
// produced inside row k and drops it only when the row is stored
// (S/ICC.java:97-133), so row k depends on every row its fill reaches, which
// on a mesh is the whole preceding band, and the rows form one chain.  The
// factor is computed on the host during setMatrix, visiting only the rows
// that can contribute (icc_factor_host), and both triangular solves of every
// apply run on the device.
#include <algorithm>
#include <cstring>

#include "amg_setup.hpp"
#include "icc_factor.hpp"
#include "mtjg_internal.hpp"
#include "reduce.hpp"
#include "rowdot.hpp"

// A run of levels [l0, l1) launched together: `fused` runs go to one
// workgroup, either as an LDS chunk (`lds`, slab levels, see k_sweep_lds) or
// as a walk over narrow row-major levels with a barrier between them; other
// runs are a single level launched over the whole device.
struct LevelRun {
  int32_t l0, l1;
  bool fused;
  bool slab = false;  // levels stored as slabs
  // LDS chunks: positions [t0, t1), imports imp[imp0, imp0 + nimp), widest
  // row wmax, steps[s0, s0 + nsteps)
  bool lds = false;
  int32_t t0 = 0, t1 = 0, nimp = 0, wmax = 0, nsteps = 0;
  bool grouped = false;  // some level's rows spread over several lanes
  int64_t imp0 = 0, s0 = 0;
};

// One tile of a tiled sweep: its LDS chunk (steps [s0, s0 + nsteps),
// positions [t0, t1), imports [imp0, imp0 + nimp)) and the tiles it imports
// from, deps[dep0, dep0 + ndep).
struct TileMeta {
  int32_t nsteps, t0, t1, nimp;
  int64_t s0, imp0;
  int32_t dep0, ndep;
};

// The entries one sweep over a level set reads, copied in level order.  Level
// l owns the entry region [eoff[l], eoff[l+1]).  A level whose rows hold at
// most one load batch (ROWDOT_K entries) is stored as a column-major slab
// (ELLPACK): width[l] entry slots per row, slot k of position t at
// eoff[l] + k * cnt_l + (t - ptr[l]), a row's entries in its first slots in
// the order it consumes them and column -1 in the rest; the thread of a row
// then loads its row index, diagonal and entries in one coalesced stage.
// Levels of longer rows inside LDS chunks are grouped slabs: a row spreads
// over G lanes, lane (i, g) of V = cnt_l * G holding slot kk at
// eoff[l] + kk * V + i * G + g for the row's entry g + kk * G (width[l] slots
// per lane).  Other levels (width[l] = -1) keep each row's entries
// contiguous (row-major) and find them through a 16-byte descriptor, which
// measured faster for long rows.  Per position: the row index, diagonal
// value and (SSOR) number of strictly-lower entries.  Values are gathered
// from the factor through emap / pdmap and gathered again whenever it
// changes.
struct Packed {
  int form = -1;             // TRSV_* form or PACK_SSOR; -1: not built
  uint64_t version = ~0ull;  // factor values gathered at this version
  int64_t slots = 0;
  std::vector<int64_t> eoff;   // entry region of each level (host)
  std::vector<int32_t> width;  // entry slots per row (per lane when grouped), -1 row-major (host)
  int32_t *prow = nullptr, *pnlo = nullptr, *pdmap = nullptr;
  double *pdv = nullptr;
  int4 *desc = nullptr;  // row-major levels: {row, first entry, count, lower count}
  std::vector<LevelRun> runs;  // launches of one sweep, in order
  std::vector<uint8_t> slab;   // per level: stored as a slab
  int32_t *eci = nullptr, *emap = nullptr;
  double *eval = nullptr;
  // LDS chunks: per slot of their levels the LDS index of the value the slot
  // multiplies (-1 none, -2 a fixed SSOR operand whose product the sweep's
  // prep stores in fval), the imported global indices, the fixed slots, and
  // per-apply level-order copies of b and of SSOR's own operand
  bool any_lds = false;
  int32_t *lci = nullptr, *imp = nullptr;
  // per step {first position - t0, lanes | width << 16 | barrier << 24 |
  // padding << 25 | log2(lanes per row) << 26, slot base, stride}
  int4 *steps = nullptr;
  int64_t *fixslot = nullptr;
  int64_t nfix = 0;
  double *fval = nullptr, *bp = nullptr, *own = nullptr;
  // tiled sweeps (every run an LDS chunk, run k = tile k): tile table, the
  // tiles each imports from, completion flags stamped with the sweep's epoch,
  // ticket and exit counters, the largest LDS need without grouped-row
  // scratch, whether any tile has grouped rows, widest lane
  int32_t ntiles = 0;
  TileMeta *tiles = nullptr;
  int32_t *tdeps = nullptr;
  uint32_t *tflags = nullptr, *tctr = nullptr;
  uint32_t epoch = 0;
  size_t tile_lds = 0;
  bool tile_grouped = false;
  int32_t tile_wmax = 0;
};

struct Levels {
  bool forward = true;  // dependencies have smaller row indices
  int tiling = 0;       // tiled sweeps: 0 not yet decided, 1 in use, -1 not used
  Levels *tiled = nullptr;  // the tiled order of this level set (tile_levels)
  std::vector<int32_t> tile_lev;  // tiled order: first level of each tile [ntiles + 1]
  int32_t nlev = 0;
  std::vector<int32_t> ptr;     // host copy of level offsets [nlev + 1]
  int32_t *dptr = nullptr;      // device copy of the level offsets
  int32_t *rows = nullptr;      // device: rows sorted by level
  std::vector<int32_t> h_rows;  // host copy of rows
  Packed pk;
};

// A factor whose pattern is a natural-order 5- or 7-point grid stencil (see
// k_grid_sweep): row (x, y, z) is x + nx (y + ny z) and its off-diagonal
// entries are among the six grid neighbours.  Slots are skewed: the rows
// y = 64 yb + l (lane l) of plane z at skew column t (x = t - l) fill slot
// ((z YB + yb) TW + t) 64 + l, so the rows a wavefront of 64 lanes solves at
// one step (x + y + z = T, so t = T - z - 64 yb for all of them) are 64
// consecutive slots.
struct GridPlan {
  int32_t nx = 0, ny = 0, nz = 0;
  int32_t YB = 0;    // 64-row blocks of a plane (one wavefront each)
  int32_t P = 0;     // planes of a sweep workgroup
  int32_t G = 0;     // workgroups of a sweep: plane groups [P g, P g + P)
  int32_t TW = 0;    // skew columns: nx + 64 (the last holds no grid row in any lane)
  int32_t lead = 0;  // steps a workgroup starts behind the group it imports from
  int64_t slab = 0, nslot = 0;  // slots per plane, slots in all
  int64_t poll_budget = 0;      // polls of one pending value before the error flag (<= 0: none)
};
// One allocation holds a factor's grid copy and the sweeps' vectors, at these
// byte offsets (nslot slots each; all below 2^32, so one buffer descriptor
// with 32-bit offsets reaches everything): the seven factor values
// (L z-1, y-1, x-1, U x+1, y+1, z+1, diagonal), b, y and x, then the
// neighbour masks (one byte per slot).
enum { GRID_B = 7, GRID_Y = 8, GRID_X = 9, GRID_MASK = 10 };
__host__ __device__ inline int64_t grid_off(const GridPlan &g, int k) { return (int64_t)k * 8 * g.nslot; }
__host__ __device__ inline int64_t grid_bytes(const GridPlan &g) { return 81 * g.nslot; }
// SSOR's grid copy appends, per slot, a sweep's three fixed-side products and
// the value it relaxes from (doubles, nslot each, from byte grid_bytes on).
__host__ __device__ inline int64_t grid_ssor_f(const GridPlan &g) { return 81 * g.nslot; }
__host__ __device__ inline int64_t grid_ssor_own(const GridPlan &g) { return 105 * g.nslot; }
__host__ __device__ inline int64_t grid_ssor_bytes(const GridPlan &g) { return 113 * g.nslot; }
// The transposed copy (U' and L' solves: ILU's transApply, ICC's R' solve),
// a second allocation: U(k-nx ny, k), U(k-nx, k), U(k-1, k), the diagonal,
// L(k+nx ny, k), L(k+nx, k), L(k+1, k) (doubles, nslot each), then the
// GM_T* mask bytes.
enum { GRIDT_U = 0, GRIDT_D = 3, GRIDT_L = 4, GRIDT_MASK = 7 };
__host__ __device__ inline int64_t gridt_bytes(const GridPlan &g) { return 57 * g.nslot; }

struct mtjg_prec_s {
  mtjg_ctx ctx = nullptr;
  int kind = mtjg::PREC_IDENTITY;
  int32_t n = 0;
  // Jacobi
  double *invdiag = nullptr;
  int32_t *bad = nullptr;  // device scratch: first offending row
  // ILU(0): factor storage in A's pattern
  bool factored = false;
  int64_t nnz = 0;
  int32_t *rp = nullptr, *ci = nullptr, *diag = nullptr;
  double *lu = nullptr;
  double *y = nullptr;  // temporary of apply
  std::vector<int32_t> h_rp, h_ci, h_diag;
  Levels lev_l, lev_u;
  uint64_t vals_version = 0;  // bumped whenever lu or tval change
  // transposed factor (transApply), built on first use
  bool trans_ready = false;
  int32_t *trp = nullptr, *tci = nullptr, *tdiag = nullptr, *tperm = nullptr;
  double *tval = nullptr;
  std::vector<int32_t> h_trp, h_tci, h_tdiag;
  Levels lev_ut, lev_lt;
  // single-launch solves: ticket counters [4] and the poll-budget error flag
  uint32_t *sf_next = nullptr;
  int32_t *sf_err = nullptr;
  bool level_launch = true;   // false: single-launch solves (MTJG_ILU_SYNCFREE=1)
  // ILU(0) grid sweeps (stencil patterns in natural order): plan, the grid
  // copy of the factor with the skewed b, y and x (laid out as grid_off
  // gives), ticket / exit counters; grid_vals_version: values copied at this
  // version
  bool grid_allowed = true;  // false: MTJG_ILU_GRID=0
  bool grid = false;
  GridPlan gp;
  char *g_buf = nullptr;
  uint32_t *g_ctr = nullptr;
  uint64_t grid_vals_version = ~0ull;
  char *g_tbuf = nullptr;  // transposed copy (gridt_bytes), built on first use
  uint64_t grid_t_version = ~0ull;
  // SSOR: relaxations, reverse sweep flag, half iterate (values in lu)
  double omega_f = 1, omega_r = 1;
  bool reverse = true;
  double *xx = nullptr;
  // AMG: parameters and the device hierarchy built by setMatrix
  mtjg::amg::Params amg_params;
  mtjg::AmgDevice *amg = nullptr;
  // FP32 inner CG: its workspace and the matrix set by setMatrix (referenced,
  // not copied, as SSOR references F)
  mtjg_solver inner = nullptr;
  mtjg_csr A = nullptr;
};

namespace mtjg {

// --- Jacobi -----------------------------------------------------------------

// invdiag[i] = 1 / A.get(i, i); the first zero or structurally missing
// diagonal is recorded (S/DiagonalPreconditioner.java:69-75).
__global__ __launch_bounds__(BS) void k_jacobi_setup(int32_t n, const int32_t *rp, const int32_t *ci,
                                                     const double *val, double *invdiag,
                                                     int32_t *bad) {
  int32_t i = blockIdx.x * BS + threadIdx.x;
  if (i >= n) return;
  int32_t lo = rp[i], hi = rp[i + 1] - 1;
  double d = 0.0;
  while (lo <= hi) {
    int32_t mid = (lo + hi) >> 1;
    int32_t c = ci[mid];
    if (c < i) lo = mid + 1;
    else if (c > i) hi = mid - 1;
    else {
      d = val[mid];
      break;
    }
  }
  if (d == 0) atomicMin(bad, i);
  invdiag[i] = 1 / d;
}

// x_i = b_i * invdiag_i (S/DiagonalPreconditioner.java:54-55).
__global__ __launch_bounds__(BS) void k_jacobi_apply(int64_t n, const double *b, const double *invdiag,
                                                     double *x, const KState *st) {
  if (st && st->status != ST_RUNNING) return;
  for (int64_t i = blockIdx.x * (int64_t)BS + threadIdx.x; i < n; i += (int64_t)gridDim.x * BS)
    x[i] = b[i] * invdiag[i];
}

__global__ __launch_bounds__(BS) void k_copy_if_running(int64_t n, const double *b, double *x,
                                                        const KState *st) {
  if (st && st->status != ST_RUNNING) return;
  for (int64_t i = blockIdx.x * (int64_t)BS + threadIdx.x; i < n; i += (int64_t)gridDim.x * BS)
    x[i] = b[i];
}

// --- ILU(0) -----------------------------------------------------------------

// Diagonal positions (S/ILU.java:127-140); missing ones are recorded.
__global__ __launch_bounds__(BS) void k_find_diag(int32_t n, const int32_t *rp, const int32_t *ci,
                                                  int32_t *diag, int32_t *bad) {
  int32_t i = blockIdx.x * BS + threadIdx.x;
  if (i >= n) return;
  int32_t lo = rp[i], hi = rp[i + 1] - 1, pos = -1;
  while (lo <= hi) {
    int32_t mid = (lo + hi) >> 1;
    int32_t c = ci[mid];
    if (c < i) lo = mid + 1;
    else if (c > i) hi = mid - 1;
    else {
      pos = mid;
      break;
    }
  }
  diag[i] = pos;
  if (pos < 0) atomicMin(bad, i);
}

// IKJ elimination of the rows of one level (S/ILU.java:98-121).
__global__ __launch_bounds__(BS) void k_ilu_factor_level(int32_t cnt, const int32_t *rows,
                                                         const int32_t *rp, const int32_t *ci,
                                                         const int32_t *diag, double *lu,
                                                         int32_t *bad) {
  int32_t t = blockIdx.x * BS + threadIdx.x;
  if (t >= cnt) return;
  const int32_t k = rows[t];
  const int32_t kend = rp[k + 1];
  for (int32_t i = rp[k]; i < diag[k]; ++i) {
    const int32_t index = ci[i];
    const double LUii = lu[diag[index]];
    if (LUii == 0) {
      atomicMin(bad, k);
      return;
    }
    const double LUki = (lu[i] /= LUii);
    int32_t l = rp[k] + 1;
    for (int32_t j = diag[index] + 1; j < rp[index + 1]; ++j) {
      const int32_t cj = ci[j];
      while (l < kend && ci[l] < cj) l++;
      if (l < kend && ci[l] == cj) lu[l] -= LUki * lu[j];
    }
  }
}

enum TrsvForm : int {
  TRSV_ROW_UNIT = 0,     // x_i = b_i - sum_{beg..end} a x        (L.solve)
  TRSV_ROW_DIV = 1,      // x_i = (b_i - sum) / d_i                (U.solve)
  TRSV_COL_DIV = 2,      // s = b_i; s -= a x ascending; x_i = s/d  (U'.solve)
  TRSV_COL_UNIT_REV = 3  // s = b_i; s -= a x descending; x_i = s   (L'.solve)
};

constexpr int PACK_SSOR = 4;      // forward sweeps (lower level set)
constexpr int PACK_SSOR_BWD = 5;  // reverse sweeps (upper level set)

// Sums over the entry slots of one packed row (slot k at base + k * stride,
// w slots, column -1 after the row's last entry) in slot order, loads issued
// ROWDOT_K slots at a time: s (+ or -)= v * x[c].
template <bool SUB>
__device__ __forceinline__ double ell_row(double s, int64_t base, int32_t stride, int32_t w,
                                          const int32_t *__restrict__ eci, const double *__restrict__ eval,
                                          const double *x) {
  for (int32_t k0 = 0; k0 < w; k0 += ROWDOT_K) {
    int32_t c[ROWDOT_K];
    double v[ROWDOT_K], xv[ROWDOT_K];
#pragma unroll
    for (int u = 0; u < ROWDOT_K; ++u) {
      c[u] = -1;
      if (k0 + u < w) {
        const int64_t e = base + (int64_t)(k0 + u) * stride;
        c[u] = eci[e];
        v[u] = eval[e];
      }
    }
#pragma unroll
    for (int u = 0; u < ROWDOT_K; ++u)
      if (c[u] >= 0) xv[u] = x[c[u]];
#pragma unroll
    for (int u = 0; u < ROWDOT_K; ++u)
      if (c[u] >= 0) {
        if (SUB) s -= v[u] * xv[u];
        else s += v[u] * xv[u];
      }
  }
  return s;
}

// Position t0 + tl of a triangular solve over a level stored as a slab (cnt
// rows, slab at base, w slots), in MTJ's per-row order (the transposed
// unit-lower form's entries are packed in its descending order).
template <int FORM>
__device__ __forceinline__ void trsv_row_slab(int32_t t0, int32_t tl, int32_t cnt, int64_t base, int32_t w,
                                              const int32_t *__restrict__ prow, const double *__restrict__ pdv,
                                              const int32_t *__restrict__ eci, const double *__restrict__ eval,
                                              const double *b, double *x) {
  const int32_t t = t0 + tl;
  const int32_t i = prow[t];
  const int64_t e0 = base + tl;
  if (FORM == TRSV_ROW_UNIT) {
    const double sum = ell_row<false>(0.0, e0, cnt, w, eci, eval, x);
    x[i] = b[i] - sum;
  } else if (FORM == TRSV_ROW_DIV) {
    const double d = pdv[t];
    const double sum = ell_row<false>(0.0, e0, cnt, w, eci, eval, x);
    x[i] = (b[i] - sum) / d;
  } else if (FORM == TRSV_COL_DIV) {
    const double d = pdv[t];
    x[i] = ell_row<true>(b[i], e0, cnt, w, eci, eval, x) / d;
  } else {
    x[i] = ell_row<true>(b[i], e0, cnt, w, eci, eval, x);
  }
}

// Position t of a triangular solve over a row-major level (the row's entries
// contiguous, found through its descriptor).
template <int FORM>
__device__ __forceinline__ void trsv_row_desc(int32_t t, const int4 *__restrict__ desc,
                                              const double *__restrict__ pdv, const int32_t *__restrict__ eci,
                                              const double *__restrict__ eval, const double *b, double *x) {
  const int4 d = desc[t];
  const int32_t i = d.x;
  if (FORM == TRSV_ROW_UNIT) {
    const double sum = rowdot<false>(0.0, d.y, d.z, eci, eval, x);
    x[i] = b[i] - sum;
  } else if (FORM == TRSV_ROW_DIV) {
    const double sum = rowdot<false>(0.0, d.y, d.z, eci, eval, x);
    x[i] = (b[i] - sum) / pdv[t];
  } else if (FORM == TRSV_COL_DIV) {
    x[i] = rowdot<true>(b[i], d.y, d.z, eci, eval, x) / pdv[t];
  } else {
    x[i] = rowdot<true>(b[i], d.y, d.z, eci, eval, x);
  }
}

// One launched level of a triangular solve; a thread per row.  w >= 0: the
// level is a slab of w slots per row; w < 0: row-major.
template <int FORM, bool SLAB>
__global__ __launch_bounds__(BS) void k_trsv_level(int32_t cnt, int32_t t0, int64_t base, int32_t w,
                                                   const int4 *desc, const int32_t *prow, const double *pdv,
                                                   const int32_t *eci, const double *eval, const double *b,
                                                   double *x) {
  const int32_t tl = blockIdx.x * BS + threadIdx.x;
  if (tl >= cnt) return;
  if (SLAB) trsv_row_slab<FORM>(t0, tl, cnt, base, w, prow, pdv, eci, eval, b, x);
  else trsv_row_desc<FORM>(t0 + tl, desc, pdv, eci, eval, b, x);
}

constexpr int FUSE_BS = 512;
constexpr int FUSE_MAX = FUSE_BS;  // widest level a fused run takes: one pass
// A run of narrow row-major levels on one workgroup, a barrier between
// levels (slab levels run in LDS chunks instead, see k_sweep_lds).
template <int FORM>
__global__ __launch_bounds__(FUSE_BS) void k_trsv_fused(int32_t l0, int32_t l1, const int32_t *lptr,
                                                        const int4 *desc, const double *pdv, const int32_t *eci,
                                                        const double *eval, const double *b, double *x) {
  for (int32_t l = l0; l < l1; ++l) {
    const int32_t t0 = lptr[l], cnt = lptr[l + 1] - t0;
    if ((int32_t)threadIdx.x < cnt) trsv_row_desc<FORM>(t0 + (int32_t)threadIdx.x, desc, pdv, eci, eval, b, x);
    __syncthreads();
  }
}

// Position t0 + tl of an SSOR sweep (S/SSOR.java:143-194) over a packed
// level, in MTJ's order: sigma adds the strictly-lower entries against xx and
// the strictly-upper ones against xd, then sigma = (b_i - sigma) / a_ii.  The
// forward sweep writes xx_i = xd_i + omega (sigma - xd_i) (levels of the lower
// triangle; xd still holds x's incoming values), the backward sweep
// xd_i = xx_i + omega (sigma - xx_i) (levels of the upper triangle, from the
// last row up).
template <bool FORWARD>
__device__ __forceinline__ void ssor_row_slab(int32_t t0, int32_t tl, int32_t cnt, int64_t base, int32_t w,
                                         const int32_t *__restrict__ prow, const int32_t *__restrict__ pnlo,
                                         const double *__restrict__ pdv, const int32_t *__restrict__ eci,
                                         const double *__restrict__ eval, const double *b, double *xd, double *xx,
                                         double omega) {
  const int32_t t = t0 + tl;
  const int32_t i = prow[t], nlo = pnlo[t];
  const double d = pdv[t];
  const int64_t e0 = base + tl;
  double sigma = 0;
  for (int32_t k0 = 0; k0 < w; k0 += ROWDOT_K) {
    int32_t c[ROWDOT_K];
    double v[ROWDOT_K], xv[ROWDOT_K];
#pragma unroll
    for (int u = 0; u < ROWDOT_K; ++u) {
      c[u] = -1;
      if (k0 + u < w) {
        const int64_t e = e0 + (int64_t)(k0 + u) * cnt;
        c[u] = eci[e];
        v[u] = eval[e];
      }
    }
#pragma unroll
    for (int u = 0; u < ROWDOT_K; ++u)
      if (c[u] >= 0) xv[u] = (k0 + u < nlo) ? xx[c[u]] : xd[c[u]];
#pragma unroll
    for (int u = 0; u < ROWDOT_K; ++u)
      if (c[u] >= 0) sigma += v[u] * xv[u];
  }
  sigma = (b[i] - sigma) / d;
  if (FORWARD) xx[i] = xd[i] + omega * (sigma - xd[i]);
  else xd[i] = xx[i] + omega * (sigma - xx[i]);
}

template <bool FORWARD>
__device__ __forceinline__ void ssor_row_desc(int32_t t, const int4 *__restrict__ desc, const double *__restrict__ pdv,
                                              const int32_t *__restrict__ eci, const double *__restrict__ eval,
                                              const double *b, double *xd, double *xx, double omega) {
  const int4 d = desc[t];
  const int32_t i = d.x;
  double sigma = rowdot<false>(0.0, d.y, d.w, eci, eval, xx);
  sigma = rowdot<false>(sigma, d.y + d.w, d.z - d.w, eci, eval, xd);
  sigma = (b[i] - sigma) / pdv[t];
  if (FORWARD) xx[i] = xd[i] + omega * (sigma - xd[i]);
  else xd[i] = xx[i] + omega * (sigma - xx[i]);
}

template <bool FORWARD, bool SLAB>
__global__ __launch_bounds__(BS) void k_ssor_level(int32_t cnt, int32_t t0, int64_t base, int32_t w,
                                                   const int4 *desc, const int32_t *prow, const int32_t *pnlo,
                                                   const double *pdv, const int32_t *eci, const double *eval,
                                                   const double *b, double *xd, double *xx, double omega) {
  const int32_t tl = blockIdx.x * BS + threadIdx.x;
  if (tl >= cnt) return;
  if (SLAB) ssor_row_slab<FORWARD>(t0, tl, cnt, base, w, prow, pnlo, pdv, eci, eval, b, xd, xx, omega);
  else ssor_row_desc<FORWARD>(t0 + tl, desc, pdv, eci, eval, b, xd, xx, omega);
}

template <bool FORWARD>
__global__ __launch_bounds__(FUSE_BS) void k_ssor_fused(int32_t l0, int32_t l1, const int32_t *lptr, const int4 *desc,
                                                        const double *pdv, const int32_t *eci, const double *eval,
                                                        const double *b, double *xd, double *xx, double omega) {
  for (int32_t l = l0; l < l1; ++l) {
    const int32_t t0 = lptr[l], cnt = lptr[l + 1] - t0;
    if ((int32_t)threadIdx.x < cnt)
      ssor_row_desc<FORWARD>(t0 + (int32_t)threadIdx.x, desc, pdv, eci, eval, b, xd, xx, omega);
    __syncthreads();
  }
}

// ---- fused sweeps through LDS -----------------------------------------------
//
// A chunk of consecutive levels runs on one workgroup whose LDS holds the
// chunk's results (LDS index t - t0 for position t) followed by the values
// it reads from outside the chunk (its imports, loaded once at the start).
// Every operand a row waits for is then an LDS read, and each step's own
// loads (the row's slots, its diagonal, b and, for SSOR, its own incoming
// value and the precomputed products of the fixed operands) are issued
// LDS_DEPTH steps ahead, so a level costs an LDS round trip and a barrier
// rather than the L2 store-to-load round trip.  The results leave LDS at the
// end of the chunk.  Every row still sums its slots in MTJ's order with the
// same operations, so the output is bitwise the launched kernels'.
//
// Widest level of at most four slots an LDS chunk takes, in steps of FUSE_BS
// rows (wider levels of up to eight slots: FUSE_MAX).  One compute unit walks
// a chunk, so a wide level is better launched over the device: measured on
// 3-D Poisson 64^3, ILU's apply took 1.27 / 1.02 / 1.39 ms with 512 / 2048 /
// 8192-row levels in chunks, SSOR's (six slots) 1.33 / 1.40 / 2.12 ms.
constexpr int LDS_LEVEL_MAX = 2048;
constexpr int LDS_PAD = LDS_LEVEL_MAX;  // slots a stage may read past the last level
// Steps whose loads are in flight (step lists are padded to a multiple of
// it).  Eight measured no faster than four on 2-D ILU (0.573 ms either way):
// one compute unit's L1 misses, not the latency of a step's loads, set the
// pace.
constexpr int LDS_DEPTH = 4;
constexpr int SWEEP_GROUP_MAX = 64;    // lanes per row at most (one wavefront)
constexpr int LDS_GROUP_LANES = 1024;  // lanes of a level of grouped rows an LDS chunk takes

// One step's operands for the calling lane, W slot registers.  A row of a
// grouped level (more than ROWDOT_K entries) spreads over G = 2^lg lanes:
// lane g of the row holds its slots g, g + G, g + 2G, ...
template <int W>
struct SweepStage {
  int32_t tb, cnt, act, w, bar, lg;  // first position (relative to the chunk), lanes, lanes computed,
                                     // slots per lane, barrier after, log2 lanes per row
  double bp, d, own;
  int32_t c[W];
  double v[W];
  int32_t row;  // the row of the lane's position (tiled sweeps only)
};

// Step li of the chunk: its meta from LDS, then the lane's operands.  Lanes
// past the step's lanes load the last lane's, and slots past the level's
// width load the last slot (or, for a level without slots, a slot no further
// than LDS_PAD past the end), so every load is in bounds and the loads of a
// stage are the same on every path.
template <int KIND, int W, bool ROW = false>
__device__ __forceinline__ void sweep_stage_load(SweepStage<W> &S, const int4 *lm, int32_t li, int32_t t0,
                                                 const int32_t *lci, const double *vals, const double *pdv,
                                                 const double *bp, const double *own,
                                                 const int32_t *prow = nullptr) {
  const int4 m = lm[li];
  S.tb = __builtin_amdgcn_readfirstlane(m.x);
  const int32_t y = __builtin_amdgcn_readfirstlane(m.y);
  S.cnt = y & 0xFFFF;
  S.w = (y >> 16) & 0xFF;
  S.bar = (y >> 24) & 1;
  S.act = ((y >> 25) & 1) ? 0 : S.cnt;  // padding steps repeat the last step's loads and compute nothing
  S.lg = (y >> 26) & 7;
  const int32_t base = __builtin_amdgcn_readfirstlane(m.z), stride = __builtin_amdgcn_readfirstlane(m.w);
  const int32_t tl = min((int32_t)threadIdx.x, S.cnt - 1);
  const int32_t t = t0 + S.tb + (tl >> S.lg);
  // with ROW the lane's row is loaded and bp is the right-hand side in row
  // order (no level-order copy)
  if (ROW) {
    S.row = prow[t];
    S.bp = bp[S.row];
  } else {
    S.bp = bp[t];
  }
  S.d = (KIND == TRSV_ROW_UNIT || KIND == TRSV_COL_UNIT_REV) ? 1.0 : pdv[t];
  S.own = (KIND >= PACK_SSOR) ? own[t] : 0.0;
  const int32_t klast = max(S.w - 1, 0);
#pragma unroll
  for (int k = 0; k < W; ++k) {
    const int64_t e = (int64_t)base + (int64_t)min(k, klast) * stride + tl;
    S.c[k] = lci[e];
    S.v[k] = vals[e];
  }
}

// The row's result from its sum s, in the form's operation order.
template <int KIND>
__device__ __forceinline__ double sweep_finish(double s, double bp, double d, double own, double omega) {
  if (KIND >= PACK_SSOR) {
    const double sigma = (bp - s) / d;
    return own + omega * (sigma - own);
  }
  if (KIND == TRSV_ROW_UNIT) return bp - s;
  if (KIND == TRSV_ROW_DIV) return (bp - s) / d;
  if (KIND == TRSV_COL_DIV) return s / d;
  return s;
}

// With GSTORE the result also goes straight to x (agent scope) as soon as it
// is computed (tiled sweeps).
template <int KIND, int W, bool GSTORE = false>
__device__ __forceinline__ void sweep_stage_row(const SweepStage<W> &S, double *lx, double *scr, double omega,
                                                double *xdep = nullptr) {
  constexpr bool SUB = (KIND == TRSV_COL_DIV || KIND == TRSV_COL_UNIT_REV);  // s = b, s -= a x
  if ((int32_t)threadIdx.x >= S.act) return;
  if (S.lg == 0) {
    double xv[W];
#pragma unroll
    for (int k = 0; k < W; ++k)
      if (k < S.w && S.c[k] >= 0) xv[k] = lx[S.c[k]];
    double s = SUB ? S.bp : 0.0;
#pragma unroll
    for (int k = 0; k < W; ++k) {
      if (k >= S.w) break;
      if (S.c[k] >= 0) {
        if (SUB) s -= S.v[k] * xv[k];
        else s += S.v[k] * xv[k];
      } else if (KIND >= PACK_SSOR && S.c[k] == -2) {
        s += S.v[k];
      }
    }
    const double r = sweep_finish<KIND>(s, S.bp, S.d, S.own, omega);
    lx[S.tb + (int32_t)threadIdx.x] = r;
    if (GSTORE) store_agent(&xdep[S.row], r);
    return;
  }
  // grouped row: every lane stores its products (slot g + kk * G at index
  // g + kk * G of the row's scratch), then the row's first lane adds them in
  // slot order.  The row's lanes share a wavefront, whose LDS accesses are
  // performed in order, so no barrier is needed between the two.  An empty
  // slot stores the zero that leaves every sum bitwise unchanged: s + (-0.0)
  // and s - (+0.0) are s for every s, signed zeros included.
  const int32_t G = 1 << S.lg, g = (int32_t)threadIdx.x & (G - 1);
  double *pr = scr + ((int32_t)threadIdx.x - g) * W;
#pragma unroll
  for (int kk = 0; kk < W; ++kk) {
    if (kk >= S.w) break;
    const int32_t c = S.c[kk];
    double p = SUB ? 0.0 : -0.0;
    if (c >= 0) p = S.v[kk] * lx[c];
    else if (c == -2) p = S.v[kk];
    pr[g + kk * G] = p;
  }
  // the row's first lane reads the other lanes' products: a wavefront-scope
  // release/acquire around a wave barrier orders those LDS stores before its
  // reads (it generates no instruction for LDS within one wavefront, but
  // keeps the compiler from sinking the stores into the lanes' exit path)
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  if (g != 0) return;
  // products read two at a time (16-byte LDS reads: the row's scratch and
  // every even slot are 16-byte aligned, and S.w * G is even since G >= 2):
  // AMG V-cycle 4.64 ms against 4.73 ms read one at a time on 2-D.
  double s = SUB ? S.bp : 0.0;
  const int32_t np = S.w * G;
#pragma unroll 4
  for (int32_t k = 0; k < np; k += 2) {
    const double2 p2 = *reinterpret_cast<const double2 *>(pr + k);
    if (SUB) {
      s -= p2.x;
      s -= p2.y;
    } else {
      s += p2.x;
      s += p2.y;
    }
  }
  const double r = sweep_finish<KIND>(s, S.bp, S.d, S.own, omega);
  lx[S.tb + ((int32_t)threadIdx.x >> S.lg)] = r;
  if (GSTORE) store_agent(&xdep[S.row], r);
}

// A chunk is a list of steps of at most FUSE_BS lanes: a level is cut into
// consecutive steps of whole rows, the last of which ends with the barrier
// (rows of one level never read each other).  LDS: the chunk's results, its
// imports, the steps' meta (at an even double offset), then FUSE_BS * W doubles of
// product scratch for grouped rows.  D = LDS_DEPTH stage buffers are used in
// rotation: stage j always holds a step congruent to j, so no register
// copies, which would wait for the loads in flight; a step's buffer is
// refilled with the step D ahead right after the step is computed.
// The step list is padded to a multiple of D, so the unrolled body
// has no exit: an exit inside it would give the loop head a path on which
// some stages were not refilled, and the compiler would then wait for every
// load at the head.
// The body of one chunk.  With AGENT the imports are read and the results
// written with agent-scope atomics (sc1: served from and written through to
// the device-wide coherence point), for chunks whose imports other
// workgroups of the same launch produced (k_sweep_tiles).
template <int KIND, int W, bool AGENT>
__device__ __forceinline__ void sweep_chunk(double *lx, int32_t nsteps, int32_t t0, int32_t t1, const int4 *steps,
                                            const int32_t *lci, const double *vals, const double *pdv,
                                            const double *bp, const double *own, const int32_t *imp, int32_t nimp,
                                            const int32_t *prow, double *xdep, double omega) {
  const int32_t nrows = t1 - t0;
  const int32_t moff = (nrows + nimp + 1) & ~1;
  int4 *lm = (int4 *)(lx + moff);
  double *scr = lx + moff + 2 * nsteps;
  for (int32_t k = threadIdx.x; k < nimp; k += FUSE_BS) lx[nrows + k] = AGENT ? load_agent(&xdep[imp[k]]) : xdep[imp[k]];
  for (int32_t k = threadIdx.x; k < nsteps; k += FUSE_BS) lm[k] = steps[k];
  __syncthreads();
  constexpr int D = LDS_DEPTH;
  SweepStage<W> st[D];
#pragma unroll
  for (int j = 0; j < D; ++j)
    sweep_stage_load<KIND, W>(st[j], lm, min(j, nsteps - 1), t0, lci, vals, pdv, bp, own);
  // The prologue's loads are issued in whatever order the scheduler picks;
  // waiting for them once here keeps that order from leaking into the waits
  // the compiler places inside the loop.
  __builtin_amdgcn_s_waitcnt(0);
  for (int32_t li = 0; li < nsteps; li += D) {
#pragma unroll
    for (int j = 0; j < D; ++j) {
      const int32_t bar = st[j].bar;
      sweep_stage_row<KIND, W>(st[j], lx, scr, omega);
      sweep_stage_load<KIND, W>(st[j], lm, min(li + j + D, nsteps - 1), t0, lci, vals, pdv, bp, own);
      if (bar) __syncthreads();
    }
  }
  __syncthreads();
  for (int32_t t = t0 + threadIdx.x; t < t1; t += FUSE_BS) {
    if (AGENT) store_agent(&xdep[prow[t]], lx[t - t0]);
    else xdep[prow[t]] = lx[t - t0];
  }
}

template <int KIND, int W>
__global__ __launch_bounds__(FUSE_BS) void k_sweep_lds(int32_t nsteps, int32_t t0, int32_t t1, const int4 *steps,
                                                   const int32_t *lci, const double *vals, const double *pdv,
                                                   const double *bp, const double *own, const int32_t *imp,
                                                   int32_t nimp, const int32_t *prow, double *xdep, double omega) {
  extern __shared__ double lx[];
  sweep_chunk<KIND, W, false>(lx, nsteps, t0, t1, steps, lci, vals, pdv, bp, own, imp, nimp, prow, xdep, omega);
}

// ---- sentinel of the single-launch and tiled solves -----------------------------

constexpr uint64_t TRSV_PENDING = 0x7FF4DEAD5EA1ED00ULL;
constexpr long long TRSV_POLL_BUDGET = 1LL << 20;

__device__ __forceinline__ uint64_t load_bits(const double *p) {
  return __hip_atomic_load((const uint64_t *)p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// A finished value never equals the sentinel: arithmetic produces only quiet
// NaNs, and a value copied unchanged from the input (b in the transposed
// unit-lower solve) that carries the sentinel's payload is stored as the
// default quiet NaN instead, which every later operation treats the same way.
__device__ __forceinline__ uint64_t finished_bits(double v) {
  const uint64_t b = (uint64_t)__double_as_longlong(v);
  return b == TRSV_PENDING ? 0x7FF8000000000000ULL : b;
}

__global__ __launch_bounds__(BS) void k_fill_pending(int64_t n, double *x) {
  for (int64_t i = blockIdx.x * (int64_t)BS + threadIdx.x; i < n; i += (int64_t)gridDim.x * BS)
    *(uint64_t *)&x[i] = TRSV_PENDING;
}

// ---- tiled sweeps ------------------------------------------------------------
//
// A whole sweep in one launch over tiles (see tile_levels): each tile is an
// LDS chunk, and tiles are numbered so that every tile a tile imports from
// has a smaller number.  The launch has one workgroup per tile; each takes
// its tile number from a ticket counter when it starts, stages everything
// that does not depend on other tiles (the step meta in LDS, the first
// stages' operands in flight), waits until the tiles it imports from are
// complete (one lane of wavefront 0 per source tile), runs the chunk, and
// stamps the tile's flag with the sweep's epoch (so flags never need
// clearing).  A waiting workgroup only ever waits on smaller numbers, taken
// by workgroups that started earlier and are running or done, so the lowest
// unfinished tile can always run and the launch cannot deadlock, whatever
// the hardware's dispatch order.  Results and imports go through the
// device-wide coherence point (agent-scope stores and loads), and a tile's
// flag is written only after its result stores have completed (s_waitcnt
// vmcnt(0) in every thread, then the barrier): the publication protocol of
// the grid reductions in reduce.hpp.  (Polling the imported values instead,
// with the output prefilled with the sentinel, measured slower: 1.6 against
// 1.1 ms per 128^3 sweep, every lane of a waiting tile polling.)  A bounded
// poll budget turns a stall into an error flag: the remaining tiles are
// skipped (still stamped) and the host reports MTJG_E_DEVICE.  Every lane
// loads its row index with its operands, reads b (passed as bp, in row
// order) through it and writes its result to x as soon as it is computed,
// so only the last step's stores are waited for before the flag.  On the
// 128^3 convection-diffusion sweeps (per-tile timestamps): hand-off from a
// source tile's flag to the waiting tile ~0.6 us, imports ~1.8 us, eight
// steps ~3.5 us, final store acknowledgement ~0.5 us; a sweep 0.68 ms.
constexpr long long TILE_POLL_BUDGET = 1LL << 22;  // polls of one dependency (about 0.5 s)
constexpr size_t TILE_STATIC_LDS = 64;             // bytes of LDS the tile kernel declares itself
constexpr int TILE_DEPTH = 8;                      // steps in flight in a tile (slots of at most four)
constexpr int TILE_IMP_PER = 4;                    // import indices a lane loads before the wait
constexpr int TILE_FLAG_STRIDE = 32;               // uint32 words between two tiles' flags (128 B)

template <int KIND, int W>
__global__ __launch_bounds__(FUSE_BS) void k_sweep_tiles(int32_t ntiles, const TileMeta *tiles, const int32_t *deps,
                                                     uint32_t *flags, uint32_t epoch, uint32_t *ctr, int32_t *err,
                                                     const int4 *steps, const int32_t *lci, const double *vals,
                                                     const double *pdv, const double *bp, const double *own,
                                                     const int32_t *imp, const int32_t *prow, double *xdep,
                                                     double omega) {
  extern __shared__ double lx[];
  __shared__ int32_t s_tile, s_skip;
  // One tile per workgroup, no loop: a persistent loop whose body starts and
  // ends with lane-0 work (ticket, flag) was restructured by the compiler
  // into a loop that lane 0 leaves after every tile while the other lanes
  // keep circling the barriers, which never ends.  Values read from LDS go
  // through readfirstlane so every branch on them is uniform.
  if (threadIdx.x == 0) s_tile = (int32_t)take_ticket(&ctr[0]);
  __syncthreads();
  const int32_t tile = __builtin_amdgcn_readfirstlane(s_tile);
  if (tile < ntiles) {
    const TileMeta tm = tiles[tile];
    const int32_t nrows = tm.t1 - tm.t0, nsteps = tm.nsteps;
    const int32_t moff = (nrows + tm.nimp + 1) & ~1;
    int4 *lm = (int4 *)(lx + moff);
    double *scr = lx + moff + 2 * nsteps;
    for (int32_t k = threadIdx.x; k < nsteps; k += FUSE_BS) lm[k] = steps[tm.s0 + k];
    __syncthreads();
    // deeper than the chunks' pipeline: a tile of up to D steps has all its
    // operands in flight before it waits (step lists are padded to a
    // multiple of TILE_DEPTH)
    constexpr int D = W <= 4 ? TILE_DEPTH : LDS_DEPTH;
    SweepStage<W> st[D];
#pragma unroll
    for (int j = 0; j < D; ++j)
      sweep_stage_load<KIND, W, true>(st[j], lm, min(j, nsteps - 1), tm.t0, lci, vals, pdv, bp, own, prow);
    // the import indices do not depend on other tiles either: load them now,
    // so that after the wait the imported values take one round trip
    // (128^3 convection-diffusion: imports 1.44 -> 1.04 us per tile)
    int32_t ii[TILE_IMP_PER];
#pragma unroll
    for (int u = 0; u < TILE_IMP_PER; ++u) {
      const int32_t k = (int32_t)threadIdx.x + u * FUSE_BS;
      ii[u] = k < tm.nimp ? imp[tm.imp0 + k] : 0;
    }
    // wait for the source tiles: wavefront 0, one lane per source tile
    if (threadIdx.x < 64) {
      const int32_t lane = (int32_t)threadIdx.x;
      int32_t skip = __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      for (int32_t base = 0; base < tm.ndep && !skip; base += 64) {
        const bool mine = base + lane < tm.ndep;
        const int32_t dt = mine ? deps[tm.dep0 + base + lane] : 0;
        long long budget = TILE_POLL_BUDGET;
        for (;;) {
          const bool wait = mine && __hip_atomic_load(&flags[(size_t)dt * TILE_FLAG_STRIDE], __ATOMIC_RELAXED,
                                                      __HIP_MEMORY_SCOPE_AGENT) != epoch;
          if (__builtin_amdgcn_ballot_w64(wait) == 0ULL) break;
          if (--budget < 0) {
            if (lane == 0) __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            skip = 1;
            break;
          }
          if (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
            skip = 1;
            break;
          }
          __builtin_amdgcn_s_sleep(1);
        }
      }
      if (lane == 0) s_skip = skip;
    }
    __syncthreads();
    if (!__builtin_amdgcn_readfirstlane(s_skip)) {
      double iv[TILE_IMP_PER];
#pragma unroll
      for (int u = 0; u < TILE_IMP_PER; ++u)
        if ((int32_t)threadIdx.x + u * FUSE_BS < tm.nimp) iv[u] = load_agent(&xdep[ii[u]]);
#pragma unroll
      for (int u = 0; u < TILE_IMP_PER; ++u) {
        const int32_t k = (int32_t)threadIdx.x + u * FUSE_BS;
        if (k < tm.nimp) lx[nrows + k] = iv[u];
      }
      for (int32_t k = threadIdx.x + TILE_IMP_PER * FUSE_BS; k < tm.nimp; k += FUSE_BS)
        lx[nrows + k] = load_agent(&xdep[imp[tm.imp0 + k]]);
      __syncthreads();
      // every result is written out as it is computed, so that only the
      // last step's stores are in flight at the end; no refill is issued
      // past the last step (its loads would be waited for at the end)
      for (int32_t li = 0; li < nsteps; li += D) {
#pragma unroll
        for (int j = 0; j < D; ++j) {
          const int32_t bar = st[j].bar;
          sweep_stage_row<KIND, W, true>(st[j], lx, scr, omega, xdep);
          if (li + j + D < nsteps)
            sweep_stage_load<KIND, W, true>(st[j], lm, li + j + D, tm.t0, lci, vals, pdv, bp, own, prow);
          if (bar) __syncthreads();
        }
      }
    }
    stores_complete();
    __syncthreads();
    if (threadIdx.x == 0)
      __hip_atomic_store(&flags[(size_t)tile * TILE_FLAG_STRIDE], epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  // the last workgroup out leaves both counters at zero for the next launch
  if (threadIdx.x == 0 && take_ticket(&ctr[TICKET_STRIDE]) == gridDim.x - 1) {
    __hip_atomic_store(&ctr[TICKET_STRIDE], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(&ctr[0], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

// Per-sweep preparation of the LDS chunks: b and (SSOR) the row's own fixed
// operand in level order, and the products of the fixed SSOR slots.
__global__ void k_sweep_prep(int32_t n, const int32_t *prow, const double *b, double *bp, const double *fixed,
                             double *own, int64_t nfix, const int64_t *fixslot, const int32_t *eci,
                             const double *eval, double *fval) {
  const int64_t stride = (int64_t)gridDim.x * BS;
  for (int64_t t = blockIdx.x * (int64_t)BS + threadIdx.x; t < n; t += stride) {
    const int32_t i = prow[t];
    bp[t] = b[i];
    if (own) own[t] = fixed[i];
  }
  for (int64_t j = blockIdx.x * (int64_t)BS + threadIdx.x; j < nfix; j += stride) {
    const int64_t e = fixslot[j];
    fval[e] = eval[e] * fixed[eci[e]];
  }
}

// ---- single-launch triangular solve ----------------------------------------
//
// Rows are taken in level order in chunks of 64 (one per wave) from an atomic
// ticket counter, so every row a chunk depends on sits in a chunk with a
// smaller ticket, which a running (or finished) wave has already claimed:
// some unfinished row always has all of its inputs, and the solve cannot
// deadlock.  The output vector starts filled with TRSV_PENDING, a signalling
// NaN that IEEE arithmetic never produces; a finished row stores its value
// with a 64-bit agent-scope atomic store, and readers load dependencies with
// agent-scope atomic loads until the value is no longer pending, so the value
// itself is the completion flag.  Each lane sums its row's entries strictly
// in storage order (pausing while an input is pending), which keeps MTJ's
// operation order.  A bounded poll budget turns any violation of these
// assumptions into an error flag instead of a hang.
template <int FORM>
__global__ __launch_bounds__(BS) void k_trsv_syncfree(int32_t n, const int32_t *__restrict__ rows,
                                                      const int32_t *__restrict__ rp,
                                                      const int32_t *__restrict__ ci,
                                                      const int32_t *__restrict__ diag,
                                                      const double *__restrict__ a,
                                                      const double *__restrict__ b, double *x,
                                                      uint32_t *next, int32_t *err) {
  const int lane = threadIdx.x & 63;
  for (;;) {
    int base = 0;
    if (lane == 0) base = (int)atomicAdd(next, 64u);
    base = __shfl(base, 0, 64);
    if (base >= n) return;
    const int t = base + lane;
    bool done = t >= n;
    int i = 0, j = 0, end = 0;
    double s = 0.0;
    if (!done) {
      i = rows[t];
      if (FORM == TRSV_ROW_UNIT) { j = rp[i]; end = diag[i]; s = 0.0; }
      else if (FORM == TRSV_ROW_DIV) { j = diag[i] + 1; end = rp[i + 1]; s = 0.0; }
      else if (FORM == TRSV_COL_DIV) { j = rp[i]; end = diag[i]; s = b[i]; }
      else { j = rp[i + 1] - 1; end = diag[i]; s = b[i]; }
    }
    long long budget = TRSV_POLL_BUDGET;
    int backoff = 1;
    while (__builtin_amdgcn_ballot_w64(!done) != 0ULL) {
      if (!done) {
        // consume entries in order while their inputs are available
        for (;;) {
          const bool more = (FORM == TRSV_COL_UNIT_REV) ? (j > end) : (j < end);
          if (!more) break;
          const uint64_t bits = load_bits(&x[ci[j]]);
          if (bits == TRSV_PENDING) break;
          const double xv = __longlong_as_double((long long)bits);
          if (FORM == TRSV_ROW_UNIT || FORM == TRSV_ROW_DIV) s += a[j] * xv;
          else s -= a[j] * xv;
          if (FORM == TRSV_COL_UNIT_REV) --j; else ++j;
        }
        const bool more = (FORM == TRSV_COL_UNIT_REV) ? (j > end) : (j < end);
        if (!more) {
          double v;
          if (FORM == TRSV_ROW_UNIT) v = b[i] - s;
          else if (FORM == TRSV_ROW_DIV) v = (b[i] - s) / a[diag[i]];
          else if (FORM == TRSV_COL_DIV) v = s / a[diag[i]];
          else v = s;
          __hip_atomic_store((uint64_t *)&x[i], finished_bits(v), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          done = true;
        } else if (--budget < 0) {
          atomicExch(err, 1);
          done = true;
        }
      }
      // wave-uniform back-off while any lane still waits
      if (__builtin_amdgcn_ballot_w64(!done) != 0ULL) {
        for (int k = 0; k < backoff; ++k) __builtin_amdgcn_s_sleep(2);
        backoff = backoff < 32 ? 2 * backoff : 32;
      }
    }
  }
}

__global__ void k_gather(int64_t n, const int32_t *perm, const double *src, double *dst) {
  for (int64_t i = blockIdx.x * (int64_t)BS + threadIdx.x; i < n; i += (int64_t)gridDim.x * BS)
    dst[i] = src[perm[i]];
}

// ---- grid sweeps --------------------------------------------------------------
//
// ILU(0) of a 5- or 7-point stencil in natural order has the wavefront
// structure of the grid itself: row (x, y, z) of L depends on (x-1, y, z),
// (x, y-1, z) and (x, y, z-1) only, all on the anti-diagonal x + y + z one
// lower (U: the three on the one higher).  A sweep walks the anti-diagonals
// T in order.  A workgroup owns P consecutive planes, one wavefront per
// plane and 64-row block, lane l of block yb holding row y = 64 yb + l of its
// plane; at step T the lane solves x = T - y - z.  Its x-neighbour is the
// value the lane itself produced one step earlier (a register), its
// y-neighbour the neighbouring lane's and its z-neighbour the plane below's,
// both read from a two-step LDS ring after one barrier per step.  Operands
// (mask, factor values, right-hand side) do not depend on the solve and are
// loaded GRID_DEPTH steps ahead; in the skewed layout a wavefront's operands
// of one step are 64 consecutive slots.
//
// Plane groups run concurrently: the bottom plane of group g imports the top
// plane of group g - 1 (U: the reverse) from global memory.  Workgroups take
// their group from a ticket counter, so a group only waits on groups taken
// earlier by running workgroups, and the launch cannot deadlock.  The
// imported plane is filled with TRSV_PENDING before the sweep; the producing
// wavefronts store it with agent-scope 64-bit stores and the importing lanes
// load it with agent-scope loads, so the value itself says whether it is
// ready, and a load may be issued early with the other operands (a pending
// one is polled again, within GRID_POLL_BUDGET polls, before an error flag is
// raised).  A workgroup starts once the group it imports from is `lead`
// steps ahead, which keeps those early loads from finding pending values.
// Each lane adds its row's terms in storage order (ascending column, as
// MTJ's row loops), so the result is bitwise that of the level schedule.
constexpr int GRID_WAVES = 16;                 // wavefronts of a sweep workgroup at most
constexpr int GRID_DEPTH = 4;                  // steps whose operands are in flight
constexpr long long GRID_POLL_BUDGET = 1LL << 22;
// mask bits: present L terms (z-1, y-1, x-1), present U terms (x+1, y+1, z+1)
constexpr uint32_t GM_ZL = 1, GM_YL = 2, GM_XL = 4, GM_XU = 8, GM_YU = 16, GM_ZU = 32;
// mask bits of the transposed copy: present U(k-nx ny, k), U(k-nx, k),
// U(k-1, k) (U' solve) and L(k+nx ny, k), L(k+nx, k), L(k+1, k) (L' solve)
constexpr uint32_t GM_TUZ = 1, GM_TUY = 2, GM_TUX = 4, GM_TLZ = 8, GM_TLY = 16, GM_TLX = 32;

// Decodes the slot of lane l in plane-row q = (z YB + yb) TW + t.
struct GridSlot {
  int32_t x, y, z;
  bool valid;
};
__device__ __forceinline__ GridSlot grid_slot(const GridPlan &g, int64_t q, int32_t l) {
  const int64_t zy = q / g.TW;
  const int32_t t = (int32_t)(q - zy * g.TW);
  const int32_t z = (int32_t)(zy / g.YB), yb = (int32_t)(zy - (int64_t)z * g.YB);
  GridSlot r;
  r.y = 64 * yb + l;
  r.x = t - l;
  r.z = z;
  r.valid = r.y < g.ny && r.x >= 0 && r.x < g.nx;
  return r;
}

// Copies the factor (CSR, MTJ's pattern) into the skewed slots: per slot the
// mask of present neighbour terms and the values
// val[k * nslot + slot], k = 0..6: L z-1, y-1, x-1, U x+1, y+1, z+1, diagonal
// (zeros in slots without a grid row and for absent terms).  One wavefront
// per plane-row.
__global__ __launch_bounds__(BS) void k_grid_fill(const GridPlan g, const int32_t *__restrict__ rp,
                                                  const int32_t *__restrict__ ci, const double *__restrict__ lu,
                                                  char *buf) {
  uint8_t *mask = (uint8_t *)(buf + grid_off(g, GRID_MASK));
  double *val = (double *)buf;
  const int64_t q = (int64_t)blockIdx.x * (BS / 64) + (threadIdx.x >> 6);
  const int32_t l = threadIdx.x & 63;
  if (q >= (int64_t)g.nz * g.YB * g.TW) return;
  const GridSlot c = grid_slot(g, q, l);
  const int64_t s = q * 64 + l;
  uint32_t m = 0;
  double v[7] = {0, 0, 0, 0, 0, 0, 0};
  if (c.valid) {
    const int64_t nxny = (int64_t)g.nx * g.ny;
    const int64_t i = c.x + (int64_t)g.nx * (c.y + (int64_t)g.ny * c.z);
    for (int32_t j = rp[i]; j < rp[i + 1]; ++j) {
      const int64_t d = (int64_t)ci[j] - i;
      const double a = lu[j];
      if (d == 0) v[6] = a;
      else if (d == -nxny) { v[0] = a; m |= GM_ZL; }
      else if (d == -g.nx) { v[1] = a; m |= GM_YL; }
      else if (d == -1) { v[2] = a; m |= GM_XL; }
      else if (d == 1) { v[3] = a; m |= GM_XU; }
      else if (d == g.nx) { v[4] = a; m |= GM_YU; }
      else if (d == nxny) { v[5] = a; m |= GM_ZU; }
    }
  }
  mask[s] = (uint8_t)m;
#pragma unroll
  for (int k = 0; k < 7; ++k) val[k * g.nslot + s] = v[k];
}

// Value of entry (row, col) of the factor (CSR, at most seven entries a
// row), and whether the pattern holds it.
__device__ __forceinline__ bool grid_entry(const int32_t *__restrict__ rp, const int32_t *__restrict__ ci,
                                           const double *__restrict__ lu, int64_t row, int64_t col, double &v) {
  for (int32_t j = rp[row]; j < rp[row + 1]; ++j)
    if (ci[j] == col) {
      v = lu[j];
      return true;
    }
  return false;
}

// The transposed copy: per slot of row k the U entries of column k above the
// diagonal (rows k - nx ny, k - nx, k - 1), the diagonal and the L entries of
// column k below it (rows k + nx ny, k + nx, k + 1), with their mask bits.
__global__ __launch_bounds__(BS) void k_grid_fill_t(const GridPlan g, const int32_t *__restrict__ rp,
                                                    const int32_t *__restrict__ ci, const double *__restrict__ lu,
                                                    char *tbuf) {
  const int64_t q = (int64_t)blockIdx.x * (BS / 64) + (threadIdx.x >> 6);
  const int32_t l = threadIdx.x & 63;
  if (q >= (int64_t)g.nz * g.YB * g.TW) return;
  const GridSlot c = grid_slot(g, q, l);
  const int64_t s = q * 64 + l;
  uint32_t m = 0;
  double v[7] = {0, 0, 0, 0, 0, 0, 0};
  if (c.valid) {
    const int64_t nxny = (int64_t)g.nx * g.ny;
    const int64_t k = c.x + (int64_t)g.nx * (c.y + (int64_t)g.ny * c.z);
    if (c.z > 0 && grid_entry(rp, ci, lu, k - nxny, k, v[0])) m |= GM_TUZ;
    if (c.y > 0 && grid_entry(rp, ci, lu, k - g.nx, k, v[1])) m |= GM_TUY;
    if (c.x > 0 && grid_entry(rp, ci, lu, k - 1, k, v[2])) m |= GM_TUX;
    grid_entry(rp, ci, lu, k, k, v[3]);
    if (c.z < g.nz - 1 && grid_entry(rp, ci, lu, k + nxny, k, v[4])) m |= GM_TLZ;
    if (c.y < g.ny - 1 && grid_entry(rp, ci, lu, k + g.nx, k, v[5])) m |= GM_TLY;
    if (c.x < g.nx - 1 && grid_entry(rp, ci, lu, k + 1, k, v[6])) m |= GM_TLX;
  }
  ((uint8_t *)(tbuf + 8 * GRIDT_MASK * g.nslot))[s] = (uint8_t)m;
#pragma unroll
  for (int j = 0; j < 7; ++j) ((double *)tbuf)[j * g.nslot + s] = v[j];
}

// b into the skewed slots, and TRSV_PENDING into the imported planes of the
// two sweeps: the top plane of every group but the last in y (lower sweep),
// the bottom plane of every group but the first in x (upper sweep).
__global__ __launch_bounds__(BS) void k_grid_gather(const GridPlan g, const double *__restrict__ b, char *buf) {
  double *bs = (double *)(buf + grid_off(g, GRID_B)), *ys = (double *)(buf + grid_off(g, GRID_Y)),
         *xs = (double *)(buf + grid_off(g, GRID_X));
  const int64_t q = (int64_t)blockIdx.x * (BS / 64) + (threadIdx.x >> 6);
  const int32_t l = threadIdx.x & 63;
  if (q >= (int64_t)g.nz * g.YB * g.TW) return;
  const GridSlot c = grid_slot(g, q, l);
  const int64_t s = q * 64 + l;
  double v = 0.0;
  if (c.valid) v = b[c.x + (int64_t)g.nx * (c.y + (int64_t)g.ny * c.z)];
  bs[s] = v;
  if (c.z % g.P == g.P - 1 && c.z + 1 < g.nz) *(uint64_t *)&ys[s] = TRSV_PENDING;
  if (c.z % g.P == 0 && c.z > 0) *(uint64_t *)&xs[s] = TRSV_PENDING;
}

// x from the skewed slots (of array src).
__global__ __launch_bounds__(BS) void k_grid_scatter(const GridPlan g, const char *__restrict__ buf, double *x,
                                                     int src = GRID_X) {
  const double *xs = (const double *)(buf + grid_off(g, src));
  const int64_t q = (int64_t)blockIdx.x * (BS / 64) + (threadIdx.x >> 6);
  const int32_t l = threadIdx.x & 63;
  if (q >= (int64_t)g.nz * g.YB * g.TW) return;
  const GridSlot c = grid_slot(g, q, l);
  if (c.valid) x[c.x + (int64_t)g.nx * (c.y + (int64_t)g.ny * c.z)] = xs[q * 64 + l];
}

// Slot of grid row (x, y, z).
__device__ __forceinline__ int64_t grid_slot_of(const GridPlan &g, int32_t x, int32_t y, int32_t z) {
  const int32_t yb = y >> 6, l = y & 63;
  return (((int64_t)z * g.YB + yb) * g.TW + x + l) * 64 + l;
}

// SSOR's forward sweep: b and x (the value relaxed from) into the slots, the
// fixed-side products of the upper entries (x+1, y+1, z+1 in storage order)
// with x, and TRSV_PENDING into the sweeps' imported planes (as
// k_grid_gather).  An absent entry's product is +0.0, which leaves the
// row's sum unchanged.
__global__ __launch_bounds__(BS) void k_grid_ssor_prep_fwd(const GridPlan g, const double *__restrict__ b,
                                                           const double *__restrict__ x, char *buf) {
  const int64_t q = (int64_t)blockIdx.x * (BS / 64) + (threadIdx.x >> 6);
  const int32_t l = threadIdx.x & 63;
  if (q >= (int64_t)g.nz * g.YB * g.TW) return;
  const GridSlot c = grid_slot(g, q, l);
  const int64_t s = q * 64 + l;
  const double *val = (const double *)buf;
  double *f = (double *)(buf + grid_ssor_f(g)), *own = (double *)(buf + grid_ssor_own(g));
  double bv = 0.0, xv = 0.0, f0 = 0.0, f1 = 0.0, f2 = 0.0;
  if (c.valid) {
    const uint32_t m = ((const uint8_t *)(buf + grid_off(g, GRID_MASK)))[s];
    const int64_t nxny = (int64_t)g.nx * g.ny;
    const int64_t i = c.x + (int64_t)g.nx * (c.y + (int64_t)g.ny * c.z);
    bv = b[i];
    xv = x[i];
    if (m & GM_XU) f0 = val[3 * g.nslot + s] * x[i + 1];
    if (m & GM_YU) f1 = val[4 * g.nslot + s] * x[i + g.nx];
    if (m & GM_ZU) f2 = val[5 * g.nslot + s] * x[i + nxny];
  }
  ((double *)(buf + grid_off(g, GRID_B)))[s] = bv;
  own[s] = xv;
  f[s] = f0;
  f[g.nslot + s] = f1;
  f[2 * g.nslot + s] = f2;
  if (c.z % g.P == g.P - 1 && c.z + 1 < g.nz) ((uint64_t *)(buf + grid_off(g, GRID_Y)))[s] = TRSV_PENDING;
  if (c.z % g.P == 0 && c.z > 0) ((uint64_t *)(buf + grid_off(g, GRID_X)))[s] = TRSV_PENDING;
}

// SSOR's backward sweep: the fixed-side products of the lower entries (z-1,
// y-1, x-1 in storage order) with the forward sweep's result (y slots).
__global__ __launch_bounds__(BS) void k_grid_ssor_prep_bwd(const GridPlan g, char *buf) {
  const int64_t q = (int64_t)blockIdx.x * (BS / 64) + (threadIdx.x >> 6);
  const int32_t l = threadIdx.x & 63;
  if (q >= (int64_t)g.nz * g.YB * g.TW) return;
  const GridSlot c = grid_slot(g, q, l);
  const int64_t s = q * 64 + l;
  const double *val = (const double *)buf, *xx = (const double *)(buf + grid_off(g, GRID_Y));
  double *f = (double *)(buf + grid_ssor_f(g));
  double f0 = 0.0, f1 = 0.0, f2 = 0.0;
  if (c.valid) {
    const uint32_t m = ((const uint8_t *)(buf + grid_off(g, GRID_MASK)))[s];
    if (m & GM_ZL) f0 = val[s] * xx[grid_slot_of(g, c.x, c.y, c.z - 1)];
    if (m & GM_YL) f1 = val[g.nslot + s] * xx[grid_slot_of(g, c.x, c.y - 1, c.z)];
    if (m & GM_XL) f2 = val[2 * g.nslot + s] * xx[grid_slot_of(g, c.x - 1, c.y, c.z)];
  }
  f[s] = f0;
  f[g.nslot + s] = f1;
  f[2 * g.nslot + s] = f2;
}

// A barrier that orders LDS only: global loads stay in flight across it.
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Operands of one step of one lane.
struct GridOp {
  double v0, v1, v2, v3, b;
  double f0, f1, f2, own;  // SSOR: the fixed-side products and the value relaxed from
  uint64_t imp;  // imported z-neighbour (bits), importing wavefronts only
  uint32_t m;
};

typedef uint32_t grid_u32x2 __attribute__((ext_vector_type(2)));  // the 64-bit buffer access type
constexpr int GRID_SC1 = 16;  // cache policy of agent-scope (device-coherent) accesses

__device__ __forceinline__ double grid_ld(__amdgpu_buffer_rsrc_t r, uint32_t vo, uint32_t so) {
  return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, vo, so, 0));
}

// One triangular solve of the grid copy, in one of the four forms of the
// level-scheduled solves (TRSV_*): ROW_UNIT L.solve and ROW_DIV U.solve
// (S/UnitLowerCompRowMatrix.java:56-64, S/UpperCompRowMatrix.java:56-64:
// s = sum of a_ij x_j in ascending column, then b - s, divided by the
// diagonal for U), COL_DIV U'.solve and COL_UNIT_REV L'.solve (their
// transSolve: s = b, minus a_ji x_j in the column algorithm's order, divided
// by the diagonal for U').  ROW_UNIT and COL_DIV walk the anti-diagonals up
// (dependencies at x-1, y-1, z-1), the other two down.  vbuf holds the three
// values of each row in summation order at v0, v0 + 8 nslot, v0 + 16 nslot,
// the diagonal at vd and the mask bytes at vm; xbuf holds the right-hand
// side at rhs and receives the result at out (byte offsets).
// blockDim = 64 P YB, grid = G.
// Wavefront w = zl YB + yb; its lanes' ring index is the thread index.  The
// skew column t of a step is the same for all lanes of a wavefront, so every
// load and store is a buffer access at a wavefront-uniform (scalar) offset
// plus a constant per lane; a step whose column lies outside the row (and
// every step of a wavefront past the last plane) uses the padding column
// TW - 1, so every access is in bounds.  Every lane computes at every step:
// a lane whose slot holds no grid row computes a value nobody reads into its
// own padding slot (a slot holds a grid row exactly when its lane is active
// at that step), and a present neighbour term always refers to a value its
// lane produced while active.  The step has no divergent branch, whose
// exec-mask bookkeeping costs scalar instructions that the sixteen
// wavefronts of a compute unit issue through one scalar unit.
template <int FORM>
__global__ __launch_bounds__(GRID_WAVES * 64) void k_grid_sweep(const GridPlan g, const char *vbuf, uint32_t vbytes,
                                                                uint32_t v0, uint32_t vd, uint32_t vm, char *xbuf,
                                                                uint32_t xbytes, uint32_t rhs, uint32_t out,
                                                                uint32_t fo, uint32_t owo, double omega,
                                                                uint32_t *ctr, int32_t *err) {
  constexpr bool SSOR = FORM == PACK_SSOR || FORM == PACK_SSOR_BWD;
  constexpr int DIR = (FORM == TRSV_ROW_UNIT || FORM == TRSV_COL_DIV || FORM == PACK_SSOR) ? 1 : -1;
  constexpr bool ROW = FORM == TRSV_ROW_UNIT || FORM == TRSV_ROW_DIV;
  constexpr bool DIV = FORM == TRSV_ROW_DIV || FORM == TRSV_COL_DIV || SSOR;
  constexpr bool XFIRST = FORM == TRSV_ROW_DIV || FORM == PACK_SSOR_BWD;  // terms x, y, z (else z, y, x)
  constexpr bool UPPER = FORM == TRSV_ROW_DIV || FORM == PACK_SSOR_BWD;   // row copy's U terms and bits
  // mask bits of the z-, y- and x-terms
  constexpr uint32_t mz = UPPER ? GM_ZU : FORM == TRSV_COL_UNIT_REV ? GM_TLZ : FORM == TRSV_COL_DIV ? GM_TUZ : GM_ZL;
  constexpr uint32_t my = UPPER ? GM_YU : FORM == TRSV_COL_UNIT_REV ? GM_TLY : FORM == TRSV_COL_DIV ? GM_TUY : GM_YL;
  constexpr uint32_t mx = UPPER ? GM_XU : FORM == TRSV_COL_UNIT_REV ? GM_TLX : FORM == TRSV_COL_DIV ? GM_TUX : GM_XL;
  __shared__ double ring[2][GRID_WAVES * 64];
  __shared__ int32_t s_ticket;
  const int32_t tid = threadIdx.x, l = tid & 63;
  const int32_t w = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int32_t zl = w / g.YB, yb = w - zl * g.YB;
  if (tid == 0) s_ticket = (int32_t)atomicAdd(ctr, 1u);
  __syncthreads();
  const int32_t tk = __builtin_amdgcn_readfirstlane(s_ticket);
  const int32_t gi = DIR > 0 ? tk : g.G - 1 - tk;
  const int32_t z0 = gi * g.P, z1 = min(z0 + g.P, g.nz) - 1;
  const int32_t z = z0 + zl;
  const bool wave_ok = z <= z1;  // planes past the grid's last (last group only) stay idle
  const int32_t Tbeg = z0, Tend = z1 + g.ny + g.nx - 2, nsteps = Tend - Tbeg + 1;
  // t = t0 + DIR p; an idle wavefront's columns all fall outside the row
  const int32_t t0 = wave_ok ? (DIR > 0 ? Tbeg : Tend) - z - 64 * yb : -(1 << 30);
  const uint32_t q0 = (uint32_t)(((wave_ok ? z : z1) * g.YB + yb) * g.TW);  // plane-row of column 0
  const bool importer = DIR > 0 ? zl == 0 : zl == g.P - 1;
  const bool exporter = DIR > 0 ? zl == g.P - 1 : zl == 0;
  const bool imports = importer && (DIR > 0 ? gi > 0 : gi < g.G - 1);  // the imported plane exists
  const uint32_t ioff = (uint32_t)(DIR > 0 ? -g.slab : g.slab) * 8u;  // the imported plane (bytes, mod 2^32)
  const __amdgpu_buffer_rsrc_t rv = __builtin_amdgcn_make_buffer_rsrc(const_cast<char *>(vbuf), (short)0,
                                                                      (int)vbytes, 0x00020000);
  const __amdgpu_buffer_rsrc_t rx = __builtin_amdgcn_make_buffer_rsrc(xbuf, (short)0, (int)xbytes, 0x00020000);
  // byte offsets of the operand arrays (scalar) and of the lane
  const uint32_t ns8 = (uint32_t)g.nslot * 8u, l8 = (uint32_t)l * 8u;
  const uint32_t a0 = v0, a1 = a0 + ns8, a2 = a1 + ns8, a3 = vd, am = vm;
  double *const outp = (double *)(xbuf + out);
  // ring indices of the y- and z-neighbours (clamped: used only when present)
  const int32_t iy = DIR > 0 ? max(tid - 1, 0) : min(tid + 1, GRID_WAVES * 64 - 1);
  const int32_t iz = DIR > 0 ? max(tid - 64 * g.YB, 0) : min(tid + 64 * g.YB, GRID_WAVES * 64 - 1);
  // steps at which this lane holds a grid row: x = t - l in [0, nx)
  int32_t pa = INT32_MAX, pb = -1;
  if (wave_ok && 64 * yb + l < g.ny) {
    if (DIR > 0) {
      pa = l - t0;
      pb = pa + g.nx - 1;
    } else {
      pb = t0 - l;
      pa = pb - g.nx + 1;
    }
  }

  // head start: wait until the imported group has solved the row of its
  // plane that this group's step `lead` imports
  if (tid == 0 && (DIR > 0 ? gi > 0 : gi < g.G - 1)) {
    int64_t ws;
    if (DIR > 0) {
      const int32_t xw = min(g.lead, g.nx - 1);
      ws = (int64_t)((z0 - 1) * g.YB) * g.TW * 64 + (int64_t)xw * 64;
    } else {
      const int32_t xw = max(g.nx - 1 - g.lead, 0), yw = g.ny - 1, ybw = yw >> 6, lw = yw & 63;
      ws = (int64_t)((z0 + g.P) * g.YB + ybw) * g.TW * 64 + (int64_t)(xw + lw) * 64 + lw;
    }
    // a budget of 0 or less admits no wait at all (the host forces the
    // error path that way)
    long long budget = g.poll_budget;
    while (budget <= 0 || load_bits(&outp[ws]) == TRSV_PENDING) {
      if (--budget < 0) {
        atomicExch(err, 1);
        break;
      }
      __builtin_amdgcn_s_sleep(2);
    }
  }
  __syncthreads();

  // wavefront-uniform byte offset of the slots of step p's column (lane 0)
  auto step_off = [&](int32_t p) -> uint32_t {
    const int32_t t = t0 + DIR * p;
    const uint32_t tc = (uint32_t)t < (uint32_t)(g.TW - 1) ? (uint32_t)t : (uint32_t)(g.TW - 1);
    return (q0 + tc) * 512u;
  };
  // A lane reads and writes only while it holds a grid row: otherwise its
  // buffer offset lies past the buffer's end, and the access returns 0 (a
  // store is dropped) without touching memory.  No branch: every lane issues
  // the same loads at every step, since with loads under divergent branches
  // the compiler cannot count the loads in flight and waits for all of them
  // before each use.
  auto lane_off = [&](int32_t p) -> uint32_t { return (p >= pa && p <= pb) ? l8 : 0x80000000u; };
  auto load_op = [&](int32_t p, GridOp &o) {
    const uint32_t so = step_off(p), vo = lane_off(p);
    o.m = __builtin_amdgcn_raw_buffer_load_b8(rv, vo >> 3, am + (so >> 3), 0);
    o.v0 = grid_ld(rv, vo, a0 + so);
    o.v1 = grid_ld(rv, vo, a1 + so);
    o.v2 = grid_ld(rv, vo, a2 + so);
    if (DIV) o.v3 = grid_ld(rv, vo, a3 + so);
    o.b = grid_ld(rx, vo, rhs + so);
    if (SSOR) {
      o.f0 = grid_ld(rx, vo, fo + so);
      o.f1 = grid_ld(rx, vo, fo + ns8 + so);
      o.f2 = grid_ld(rx, vo, fo + 2 * ns8 + so);
      o.own = grid_ld(rx, vo, owo + so);
    }
    if (imports)
      o.imp = __builtin_bit_cast(uint64_t, __builtin_amdgcn_raw_buffer_load_b64(rx, vo, out + so + ioff, GRID_SC1));
  };

  GridOp op[GRID_DEPTH];
#pragma unroll
  for (int u = 0; u < GRID_DEPTH; ++u) load_op(u, op[u]);
  double prev = 0.0;  // this lane's value of the previous step
  for (int32_t p0 = 0; p0 < nsteps; p0 += GRID_DEPTH) {
#pragma unroll
    for (int u = 0; u < GRID_DEPTH; ++u) {
      // steps past the last are computed like the others (into padding): an
      // early exit here would give the loop paths that issue fewer loads,
      // and the compiler would then wait for nearly all loads in flight
      const int32_t p = p0 + u;
      GridOp &o = op[u];
      const uint32_t so = step_off(p);
      const double *rg = ring[(p + 1) & 1];  // the previous step's values
      double zn = rg[iz];
      const double yn = rg[iy];
      if (importer) {
        uint64_t bits = o.imp;
        if (p >= pa && p <= pb && (o.m & mz) && bits == TRSV_PENDING) {
          const double *src = outp + ((so + ioff) >> 3) + l;  // the plane below (above)
          long long budget = g.poll_budget;
          do {
            if (--budget < 0) {
              atomicExch(err, 1);
              break;
            }
            __builtin_amdgcn_s_sleep(1);
            bits = load_bits(src);
          } while (bits == TRSV_PENDING);
        }
        zn = __longlong_as_double((long long)bits);
      }
      // absent terms change nothing: a present term adds (subtracts)
      // v * neighbour; an absent one (v = 0) adds 0 * 0 = +0, which leaves
      // every value unchanged (row sums start at +0.0 and never become
      // -0.0; -0.0 - +0.0 is -0.0)
      const double zx = (o.m & mz) ? zn : 0.0, yx = (o.m & my) ? yn : 0.0, xx = (o.m & mx) ? prev : 0.0;
      const double n0 = XFIRST ? xx : zx, n2 = XFIRST ? zx : xx;
      double r;
      if (SSOR) {
        // the row's entries in storage order (lower, then upper): the
        // forward sweep's lower terms and the backward sweep's upper terms
        // are this sweep's values, the others the fixed-side products; then
        // MTJ's relaxation of the value the sweep starts from
        double sum = 0.0;
        if (FORM == PACK_SSOR_BWD) {
          sum += o.f0;
          sum += o.f1;
          sum += o.f2;
        }
        sum += o.v0 * n0;
        sum += o.v1 * yx;
        sum += o.v2 * n2;
        if (FORM == PACK_SSOR) {
          sum += o.f0;
          sum += o.f1;
          sum += o.f2;
        }
        const double sigma = (o.b - sum) / o.v3;
        r = o.own + omega * (sigma - o.own);
      } else if (ROW) {
        double sum = 0.0;
        sum += o.v0 * n0;
        sum += o.v1 * yx;
        sum += o.v2 * n2;
        r = DIV ? (o.b - sum) / o.v3 : o.b - sum;
      } else {
        double sum = o.b;
        sum -= o.v0 * n0;
        sum -= o.v1 * yx;
        sum -= o.v2 * n2;
        r = DIV ? sum / o.v3 : sum;
      }
      if (exporter)
        __hip_atomic_store((uint64_t *)outp + (so >> 3) + l, finished_bits(r), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
      else
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(grid_u32x2, r), rx, lane_off(p), out + so, 0);
      ring[p & 1][tid] = r;
      prev = r;
      load_op(p + GRID_DEPTH, o);
      lds_barrier();
    }
  }
  if (tid == 0 && atomicAdd(&ctr[1], 1u) == (uint32_t)g.G - 1) {
    __hip_atomic_store(&ctr[0], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(&ctr[1], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

// Level sets for a triangular dependency structure given, per row, the
// index range of its dependency entries.  `forward` levels rows from the
// first row down (dependencies have smaller indices); otherwise from the last
// row up.
static void build_levels(mtjg_ctx ctx, int32_t n, const std::vector<int32_t> &rp,
                         const std::vector<int32_t> &ci, const std::vector<int32_t> &dbeg,
                         const std::vector<int32_t> &dend, bool forward, Levels &L) {
  std::vector<int32_t> lev((size_t)n, 0);
  int32_t maxlev = 0;
  for (int32_t t = 0; t < n; ++t) {
    int32_t i = forward ? t : n - 1 - t;
    int32_t l = 0;
    for (int32_t j = dbeg[i]; j < dend[i]; ++j) l = std::max(l, lev[(size_t)ci[j]] + 1);
    lev[(size_t)i] = l;
    maxlev = std::max(maxlev, l);
  }
  (void)rp;
  L.forward = forward;
  L.nlev = n > 0 ? maxlev + 1 : 0;
  L.ptr.assign((size_t)L.nlev + 1, 0);
  for (int32_t i = 0; i < n; ++i) L.ptr[(size_t)lev[(size_t)i] + 1]++;
  for (int32_t l = 0; l < L.nlev; ++l) L.ptr[l + 1] += L.ptr[l];
  std::vector<int32_t> pos(L.ptr.begin(), L.ptr.end() - (L.nlev > 0 ? 1 : 0));
  std::vector<int32_t> rows((size_t)std::max(n, 1));
  for (int32_t i = 0; i < n; ++i) rows[(size_t)pos[(size_t)lev[(size_t)i]]++] = i;
  L.h_rows.assign(rows.begin(), rows.begin() + n);
  if (L.rows) (void)hipFree(L.rows);
  if (L.dptr) (void)hipFree(L.dptr);
  HIPCHK(hipMalloc(&L.rows, sizeof(int32_t) * rows.size()));
  HIPCHK(hipMemcpyAsync(L.rows, rows.data(), sizeof(int32_t) * rows.size(), hipMemcpyHostToDevice,
                        ctx->stream));
  HIPCHK(hipMalloc(&L.dptr, sizeof(int32_t) * L.ptr.size()));
  HIPCHK(hipMemcpyAsync(L.dptr, L.ptr.data(), sizeof(int32_t) * L.ptr.size(), hipMemcpyHostToDevice,
                        ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
}

__global__ void k_gather_pad(int64_t n, const int32_t *map, const double *src, double *dst) {
  for (int64_t i = blockIdx.x * (int64_t)BS + threadIdx.x; i < n; i += (int64_t)gridDim.x * BS) {
    const int32_t m = map[i];
    dst[i] = m >= 0 ? src[m] : 0.0;
  }
}

static void free_pack(Packed &P) {
  for (void *q : {(void *)P.prow, (void *)P.pnlo, (void *)P.pdmap, (void *)P.pdv,
                  (void *)P.desc, (void *)P.eci, (void *)P.emap, (void *)P.eval, (void *)P.lci, (void *)P.imp,
                  (void *)P.fixslot, (void *)P.fval, (void *)P.bp, (void *)P.own, (void *)P.steps,
                  (void *)P.tiles, (void *)P.tdeps, (void *)P.tflags, (void *)P.tctr})
    (void)hipFree(q);
  P = Packed();
}

template <typename T>
static T *to_device(const std::vector<T> &v, hipStream_t s) {
  T *d = nullptr;
  HIPCHK(hipMalloc(&d, sizeof(T) * std::max<size_t>(v.size(), 1)));
  if (!v.empty()) HIPCHK(hipMemcpyAsync(d, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice, s));
  return d;
}

// Doubles of LDS one fused chunk may use: the device's per-workgroup limit,
// at most 160 KiB.
static int32_t lds_cap_doubles() {
  static int32_t cap = 0;
  if (!cap) {
    int dev = 0, v = 0;
    HIPCHK(hipGetDevice(&dev));
    HIPCHK(hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    cap = std::max(ROWDOT_K * FUSE_BS + FUSE_BS, std::min(v, 160 * 1024) / (int)sizeof(double));
  }
  return cap;
}

// LDS bytes of a chunk: results and imports (rounded to an even count of
// doubles), 16 bytes of meta per step, then the grouped rows' product scratch.
static size_t chunk_lds_bytes(int64_t rows, int64_t nimp, int64_t nsteps, int64_t scratch) {
  return sizeof(double) * (size_t)(((rows + nimp + 1) & ~1) + 2 * nsteps + scratch);
}

// Builds (once per pattern and form) and refreshes (whenever the values
// changed) the packed copy of level set L over the structure (rp, ci, diag),
// whose values are `a`.  With tile_lev (a tiled level set, see tile_levels)
// every level goes to an LDS chunk and tile k's levels [tile_lev[k],
// tile_lev[k + 1]) form chunk k, with the tiles it imports from; returns
// false, building nothing, when a tile does not fit one chunk.
static bool ensure_pack(mtjg_prec M, Levels &L, int form, const std::vector<int32_t> &rp,
                        const std::vector<int32_t> &ci, const std::vector<int32_t> &diag, const double *a,
                        hipStream_t s, const std::vector<int32_t> *tile_lev = nullptr) {
  Packed &P = L.pk;
  const int32_t n = (int32_t)L.h_rows.size();
  if (P.form != form) {
    free_pack(P);
    // the row's entries in consumption order, as indices into the value array
    auto entries = [&](int32_t i, std::vector<int32_t> &out, int32_t &nlo) {
      const int32_t r0 = rp[(size_t)i], r1 = rp[(size_t)i + 1], d = diag[(size_t)i];
      out.clear();
      nlo = 0;
      if (form == TRSV_ROW_UNIT || form == TRSV_COL_DIV) {
        for (int32_t j = r0; j < d; ++j) out.push_back(j);
      } else if (form == TRSV_ROW_DIV) {
        for (int32_t j = d + 1; j < r1; ++j) out.push_back(j);
      } else if (form == TRSV_COL_UNIT_REV) {
        for (int32_t j = r1 - 1; j > d; --j) out.push_back(j);
      } else {  // PACK_SSOR: the strictly-lower, then the strictly-upper entries
        for (int32_t j = r0; j < d; ++j) out.push_back(j);
        nlo = d - r0;
        for (int32_t j = d + 1; j < r1; ++j) out.push_back(j);
      }
    };
    std::vector<int32_t> prow((size_t)n), pnlo((size_t)n), pdmap((size_t)n), tmp, pos((size_t)n);
    std::vector<int4> desc((size_t)n);
    for (int32_t t = 0; t < n; ++t) pos[(size_t)L.h_rows[(size_t)t]] = t;
    auto is_dep = [&](int32_t k, int32_t nlo) {
      return form == PACK_SSOR ? k < nlo : form == PACK_SSOR_BWD ? k >= nlo : true;
    };
    const size_t nl = (size_t)std::max(L.nlev, 1);
    auto cnt_of = [&](int32_t l) { return L.ptr[l + 1] - L.ptr[l]; };
    // per level: widest row, entries, and lanes per row in an LDS chunk (the
    // fewest, a power of two, that leave each lane at most ROWDOT_K slots)
    std::vector<int32_t> wrow(nl, 0), grp(nl, 1);
    std::vector<int64_t> tot(nl, 0);
    for (int32_t l = 0; l < L.nlev; ++l) {
      int32_t nlo = 0;
      for (int32_t t = L.ptr[l]; t < L.ptr[l + 1]; ++t) {
        entries(L.h_rows[(size_t)t], tmp, nlo);
        wrow[(size_t)l] = std::max(wrow[(size_t)l], (int32_t)tmp.size());
        tot[(size_t)l] += (int64_t)tmp.size();
      }
      while (wrow[(size_t)l] > ROWDOT_K * grp[(size_t)l] && grp[(size_t)l] < SWEEP_GROUP_MAX) grp[(size_t)l] *= 2;
    }
    auto lanes_of = [&](int32_t l) { return (int64_t)cnt_of(l) * grp[(size_t)l]; };
    auto lane_width = [&](int32_t l) { return (wrow[(size_t)l] + grp[(size_t)l] - 1) / grp[(size_t)l]; };
    auto nsteps_of = [](int64_t lanes) { return (int32_t)((lanes + FUSE_BS - 1) / FUSE_BS); };
    const int32_t cap = lds_cap_doubles();
    const size_t cap_bytes = (size_t)cap * sizeof(double);
    constexpr int64_t SCRATCH_MAX = (int64_t)FUSE_BS * ROWDOT_K;  // product scratch of grouped rows, doubles
    std::vector<int32_t> mark((size_t)n, -1), impidx((size_t)n, -1);
    int32_t stamp = 0;
    // distinct operands of level l at positions before t0 (marked with stamp)
    auto new_operands = [&](int32_t l, int32_t t0, std::vector<int32_t> &out) {
      int32_t nlo = 0;
      for (int32_t t = L.ptr[l]; t < L.ptr[l + 1]; ++t) {
        entries(L.h_rows[(size_t)t], tmp, nlo);
        for (int32_t k = 0; k < (int32_t)tmp.size(); ++k) {
          if (!is_dep(k, nlo)) continue;
          const int32_t col = ci[(size_t)tmp[(size_t)k]];
          if (pos[(size_t)col] < t0 && mark[(size_t)col] != stamp) {
            mark[(size_t)col] = stamp;
            out.push_back(col);
          }
        }
      }
    };
    // Each level is launched on its own, joins a fused row-major run (rows
    // of more than ROWDOT_K entries, at most FUSE_MAX rows) or an LDS chunk:
    // rows of at most ROWDOT_K entries (a slab) on up to LDS_LEVEL_MAX lanes,
    // or FUSE_MAX lanes when its rows hold more than four entries; longer
    // rows (grouped) on up to LDS_GROUP_LANES lanes; in both cases only if
    // its rows, operands and steps fit one chunk by themselves.  Stretches of
    // at least two fused or LDS levels run together.
    enum { CAT_LAUNCH, CAT_FUSED, CAT_LDS };
    std::vector<uint8_t> cat(nl, CAT_LAUNCH);
    std::vector<int32_t> tmpimp;
    for (int32_t l = 0; l < L.nlev; ++l) {
      const int32_t w = wrow[(size_t)l], g = grp[(size_t)l];
      const int64_t lanes = lanes_of(l);
      const bool fits_lanes = w <= ROWDOT_K * g &&
                              lanes <= (g > 1 ? LDS_GROUP_LANES : w <= 4 ? LDS_LEVEL_MAX : FUSE_MAX);
      if (tile_lev) {
        if (!fits_lanes) return false;
        cat[(size_t)l] = CAT_LDS;
        continue;
      }
      if (fits_lanes) {
        tmpimp.clear();
        ++stamp;
        new_operands(l, L.ptr[l], tmpimp);
        if (chunk_lds_bytes(cnt_of(l), (int64_t)tmpimp.size(), nsteps_of(lanes) + LDS_DEPTH - 1,
                            g > 1 ? SCRATCH_MAX : 0) <= cap_bytes)
          cat[(size_t)l] = CAT_LDS;
      }
      if (cat[(size_t)l] != CAT_LDS && w > ROWDOT_K && cnt_of(l) <= FUSE_MAX) cat[(size_t)l] = CAT_FUSED;
    }
    // final layout: a slab (rows of at most ROWDOT_K entries), grouped slab
    // (longer rows in an LDS chunk: lane (row i, g) of V = cnt * G lanes holds
    // slot kk * V + i * G + g for the row's entries g + kk * G), or row-major
    std::vector<uint8_t> in_lds(nl, 0);
    for (int32_t l = 0; l < L.nlev;) {
      int32_t e = l + 1;
      while (e < L.nlev && cat[(size_t)e] == cat[(size_t)l]) ++e;
      if (cat[(size_t)l] == CAT_LDS && (e - l >= 2 || tile_lev))
        for (int32_t lv = l; lv < e; ++lv) in_lds[(size_t)lv] = 1;
      l = e;
    }
    enum { LAY_SLAB, LAY_GROUP, LAY_ROWS };
    std::vector<uint8_t> lay(nl, LAY_ROWS), slab(nl, 0);
    P.eoff.assign((size_t)L.nlev + 1, 0);
    P.width.assign(nl, 0);
    for (int32_t l = 0; l < L.nlev; ++l) {
      const int32_t w = wrow[(size_t)l], cnt = cnt_of(l);
      int64_t region;
      if (w <= ROWDOT_K) {
        lay[(size_t)l] = LAY_SLAB;
        slab[(size_t)l] = 1;
        P.width[(size_t)l] = w;
        region = (int64_t)w * cnt;
      } else if (in_lds[(size_t)l]) {
        lay[(size_t)l] = LAY_GROUP;
        P.width[(size_t)l] = lane_width(l);
        region = (int64_t)lane_width(l) * lanes_of(l);
      } else {
        P.width[(size_t)l] = -1;
        region = tot[(size_t)l];
      }
      P.eoff[(size_t)l + 1] = P.eoff[(size_t)l] + region;
    }
    P.slots = P.eoff[(size_t)L.nlev];
    if (P.slots > INT32_MAX - LDS_PAD) fail(MTJG_E_UNSUPPORTED, "triangular sweep with more than 2^31 packed entries");
    // slot of entry k of the row at position t (row-major levels: `next` is
    // the row's first slot)
    auto slot = [&](int32_t l, int32_t k, int32_t t, int64_t next) -> int64_t {
      const int32_t i = t - L.ptr[l];
      if (lay[(size_t)l] == LAY_SLAB) return P.eoff[(size_t)l] + (int64_t)k * cnt_of(l) + i;
      if (lay[(size_t)l] == LAY_GROUP) {
        const int32_t g = grp[(size_t)l];
        return P.eoff[(size_t)l] + (int64_t)(k / g) * lanes_of(l) + (int64_t)i * g + (k % g);
      }
      return next + k;
    };
    std::vector<int32_t> eci((size_t)P.slots, -1), emap((size_t)P.slots, -1);
    for (int32_t l = 0; l < L.nlev; ++l) {
      int64_t next = P.eoff[(size_t)l];
      for (int32_t t = L.ptr[l]; t < L.ptr[l + 1]; ++t) {
        const int32_t i = L.h_rows[(size_t)t];
        int32_t nlo = 0;
        entries(i, tmp, nlo);
        prow[(size_t)t] = i;
        pnlo[(size_t)t] = nlo;
        pdmap[(size_t)t] = diag[(size_t)i];
        desc[(size_t)t] = make_int4(i, (int32_t)next, (int32_t)tmp.size(), nlo);
        for (int32_t k = 0; k < (int32_t)tmp.size(); ++k) {
          const int64_t e = slot(l, k, t, next);
          eci[(size_t)e] = ci[(size_t)tmp[(size_t)k]];
          emap[(size_t)e] = tmp[(size_t)k];
        }
        if (lay[(size_t)l] == LAY_ROWS) next += (int64_t)tmp.size();
      }
    }
    P.slab.assign(slab.begin(), slab.end());
    std::vector<int32_t> lci((size_t)P.slots + LDS_PAD, -1), imps;
    std::vector<int64_t> fixslot;
    std::vector<int4> steps;
    P.runs.clear();
    // tiled: the tile of every position, and per tile its import sources
    size_t tk = 0;
    std::vector<int32_t> pos_tile, tmark, tdeps;
    std::vector<TileMeta> tiles;
    size_t tile_lds = 0;
    if (tile_lev) {
      pos_tile.assign((size_t)n, -1);
      const int32_t nt = (int32_t)tile_lev->size() - 1;
      for (int32_t k = 0; k < nt; ++k)
        for (int32_t t = L.ptr[(size_t)(*tile_lev)[k]]; t < L.ptr[(size_t)(*tile_lev)[k + 1]]; ++t)
          pos_tile[(size_t)t] = k;
      tmark.assign((size_t)std::max(nt, 1), -1);
    }
    for (int32_t l = 0; l < L.nlev;) {
      int32_t e = l + 1;
      while (e < L.nlev && cat[(size_t)e] == cat[(size_t)l]) ++e;
      if (cat[(size_t)l] == CAT_FUSED && e - l >= 2) {
        P.runs.push_back({l, e, true, false});
        l = e;
        continue;
      }
      if (!in_lds[(size_t)l]) {
        for (int32_t lv = l; lv < e; ++lv) P.runs.push_back({lv, lv + 1, false, slab[(size_t)lv] != 0});
        l = e;
        continue;
      }
      // LDS stretch [l, e): chunks of whole levels within the LDS budget
      while (l < e) {
        const int32_t c0 = l, t0 = L.ptr[l];
        const size_t i0 = imps.size();
        int32_t rows = 0, nst = 0;
        bool grouped = false;
        ++stamp;
        const int32_t lend = tile_lev ? (*tile_lev)[tk + 1] : e;
        while (l < lend) {
          const int32_t cnt = cnt_of(l);
          const size_t before = imps.size();
          new_operands(l, t0, imps);
          const bool gr = grouped || grp[(size_t)l] > 1;
          if (!tile_lev && l > c0 &&
              chunk_lds_bytes((int64_t)rows + cnt, (int64_t)(imps.size() - i0),
                              nst + nsteps_of(lanes_of(l)) + LDS_DEPTH - 1, gr ? SCRATCH_MAX : 0) > cap_bytes) {
            imps.resize(before);  // the level starts the next chunk
            break;
          }
          rows += cnt;
          nst += nsteps_of(lanes_of(l));
          grouped = gr;
          ++l;
        }
        if (tile_lev) {
          const size_t need = chunk_lds_bytes(rows, (int64_t)(imps.size() - i0), nst + TILE_DEPTH - 1,
                                              grouped ? SCRATCH_MAX : 0);
          if (need + TILE_STATIC_LDS > cap_bytes) return false;
          tile_lds = std::max(tile_lds, chunk_lds_bytes(rows, (int64_t)(imps.size() - i0), nst + TILE_DEPTH - 1, 0));
          P.tile_grouped |= grouped;
          // the tiles this one imports from, each once; every one must be an
          // earlier tile (what keeps the launch free of deadlock)
          TileMeta tm{};
          tm.dep0 = (int32_t)tdeps.size();
          for (size_t q = i0; q < imps.size(); ++q) {
            const int32_t src = pos_tile[(size_t)pos[(size_t)imps[q]]];
            if (src < 0 || src >= (int32_t)tk) fail(MTJG_E_DEVICE, "tiled sweep: an import outside the earlier tiles");
            if (tmark[(size_t)src] != (int32_t)tk) {
              tmark[(size_t)src] = (int32_t)tk;
              tdeps.push_back(src);
            }
          }
          tm.ndep = (int32_t)tdeps.size() - tm.dep0;
          tiles.push_back(tm);
          ++tk;
        }
        LevelRun c{c0, l, true, true};
        c.lds = true;
        c.t0 = t0;
        c.t1 = L.ptr[l];
        c.imp0 = (int64_t)i0;
        c.nimp = (int32_t)(imps.size() - i0);
        c.s0 = (int64_t)steps.size();
        c.grouped = grouped;
        // steps of FUSE_BS lanes also for narrow levels: 256-lane blocks made
        // the 2-D ILU apply 0.61 ms against 0.57 ms
        for (int32_t lv = c0; lv < l; ++lv) c.wmax = std::max(c.wmax, P.width[(size_t)lv]);
        for (int32_t lv = c0; lv < l; ++lv) {
          const int32_t g = grp[(size_t)lv], lg = __builtin_ctz((unsigned)g);
          const int64_t lanes = lanes_of(lv);
          const int32_t nq = nsteps_of(lanes);
          for (int32_t q = 0; q < nq; ++q) {
            const int32_t vq = (int32_t)std::min<int64_t>(FUSE_BS, lanes - (int64_t)q * FUSE_BS);
            const int32_t bar = (q == nq - 1) ? 1 : 0;
            steps.push_back(make_int4(L.ptr[lv] - t0 + q * (FUSE_BS / g),
                                      vq | (P.width[(size_t)lv] << 16) | (bar << 24) | (lg << 26),
                                      (int32_t)(P.eoff[(size_t)lv] + (int64_t)q * FUSE_BS), (int32_t)lanes));
          }
        }
        // padding steps: the last step's loads again, flagged to compute nothing
        while (((int64_t)steps.size() - c.s0) % (tile_lev ? TILE_DEPTH : LDS_DEPTH)) {
          int4 pad = steps.back();
          pad.y = (pad.y & ~(1 << 24)) | (1 << 25);
          steps.push_back(pad);
        }
        c.nsteps = (int32_t)((int64_t)steps.size() - c.s0);
        for (size_t q = i0; q < imps.size(); ++q) impidx[(size_t)imps[q]] = rows + (int32_t)(q - i0);
        for (int32_t lv = c0; lv < l; ++lv)
          for (int32_t t = L.ptr[lv]; t < L.ptr[lv + 1]; ++t) {
            int32_t nlo = 0;
            entries(L.h_rows[(size_t)t], tmp, nlo);
            for (int32_t k = 0; k < (int32_t)tmp.size(); ++k) {
              const int64_t se = slot(lv, k, t, 0);
              if (!is_dep(k, nlo)) {
                lci[(size_t)se] = -2;
                fixslot.push_back(se);
                continue;
              }
              const int32_t col = ci[(size_t)tmp[(size_t)k]], pc = pos[(size_t)col];
              lci[(size_t)se] = pc >= t0 ? pc - t0 : impidx[(size_t)col];
            }
          }
        for (size_t q = i0; q < imps.size(); ++q) impidx[(size_t)imps[q]] = -1;
        if (tile_lev) {
          TileMeta &tm = tiles.back();
          tm.nsteps = c.nsteps;
          tm.t0 = c.t0;
          tm.t1 = c.t1;
          tm.nimp = c.nimp;
          tm.s0 = c.s0;
          tm.imp0 = c.imp0;
          P.tile_wmax = std::max(P.tile_wmax, c.wmax);
        }
        P.runs.push_back(c);
      }
    }
    if (tile_lev) {
      P.ntiles = (int32_t)tiles.size();
      P.tile_lds = tile_lds;
      P.tiles = to_device(tiles, s);
      P.tdeps = to_device(tdeps, s);
      HIPCHK(hipMalloc(&P.tflags, sizeof(uint32_t) * TILE_FLAG_STRIDE * (size_t)std::max(P.ntiles, 1)));
      HIPCHK(hipMemsetAsync(P.tflags, 0, sizeof(uint32_t) * TILE_FLAG_STRIDE * (size_t)std::max(P.ntiles, 1), s));
      // the dispatch and exit counters, each on a memory line of its own
      HIPCHK(hipMalloc(&P.tctr, sizeof(uint32_t) * 2 * TICKET_STRIDE));
      HIPCHK(hipMemsetAsync(P.tctr, 0, sizeof(uint32_t) * 2 * TICKET_STRIDE, s));
      P.epoch = 0;
    }
    P.any_lds = false;
    for (const LevelRun &r : P.runs) P.any_lds |= r.lds;
    P.nfix = (int64_t)fixslot.size();
    P.lci = to_device(lci, s);
    P.imp = to_device(imps, s);
    P.fixslot = to_device(fixslot, s);
    P.steps = to_device(steps, s);
    if (P.any_lds) {
      HIPCHK(hipMalloc(&P.bp, sizeof(double) * (size_t)std::max(n, 1)));
      if (form >= PACK_SSOR) {
        HIPCHK(hipMalloc(&P.own, sizeof(double) * (size_t)std::max(n, 1)));
        HIPCHK(hipMalloc(&P.fval, sizeof(double) * (size_t)(P.slots + LDS_PAD)));
      }
    }
    P.desc = to_device(desc, s);
    P.prow = to_device(prow, s);
    P.pnlo = to_device(pnlo, s);
    P.pdmap = to_device(pdmap, s);
    P.eci = to_device(eci, s);
    P.emap = to_device(emap, s);
    HIPCHK(hipMalloc(&P.pdv, sizeof(double) * (size_t)std::max(n, 1)));
    HIPCHK(hipMalloc(&P.eval, sizeof(double) * (size_t)(P.slots + LDS_PAD)));
    HIPCHK(hipStreamSynchronize(s));  // host staging vectors go out of scope
    P.form = form;
    P.version = ~0ull;
  }
  if (P.version != M->vals_version) {
    if (P.slots) k_gather_pad<<<grid_for(P.slots), BS, 0, s>>>(P.slots, P.emap, a, P.eval);
    if (P.slots && P.fval) k_gather_pad<<<grid_for(P.slots), BS, 0, s>>>(P.slots, P.emap, a, P.fval);
    if (n) k_gather<<<grid_for(n), BS, 0, s>>>(n, P.pdmap, a, P.pdv);
    HIPCHK(hipGetLastError());
    P.version = M->vals_version;
  }
  return true;
}

template <int KIND, int W>
static void launch_sweep_lds_w(const LevelRun &r, const Packed &P, const double *vals, double *xdep, double omega,
                               hipStream_t s) {
  static bool attr_set = false;  // per instantiation: allow LDS beyond the 64 KiB default
  if (!attr_set) {
    HIPCHK(hipFuncSetAttribute((const void *)k_sweep_lds<KIND, W>, hipFuncAttributeMaxDynamicSharedMemorySize,
                               lds_cap_doubles() * (int)sizeof(double)));
    attr_set = true;
  }
  const size_t bytes = chunk_lds_bytes(r.t1 - r.t0, r.nimp, r.nsteps, r.grouped ? (int64_t)FUSE_BS * W : 0);
  k_sweep_lds<KIND, W><<<1, FUSE_BS, bytes, s>>>(r.nsteps, r.t0, r.t1, P.steps + r.s0, P.lci, vals, P.pdv, P.bp,
                                                 P.own, P.imp + r.imp0, r.nimp, P.prow, xdep, omega);
}

// Launches one LDS chunk with W slot registers for its widest lane.
template <int KIND>
static void launch_sweep_lds(const LevelRun &r, const Packed &P, const double *vals, double *xdep, double omega,
                             hipStream_t s) {
  if (r.wmax <= 2) launch_sweep_lds_w<KIND, 2>(r, P, vals, xdep, omega, s);
  else if (r.wmax <= 4) launch_sweep_lds_w<KIND, 4>(r, P, vals, xdep, omega, s);
  else launch_sweep_lds_w<KIND, ROWDOT_K>(r, P, vals, xdep, omega, s);
}

// Cuts level set L into tiles: rows by their sweep index (r in a forward
// set, n - 1 - r in a backward one) into chunks of R rows, and by level into
// windows of W levels; tile (w, c) holds window w's rows of chunk c.  Every
// row a row reads has a smaller level and a smaller sweep index, so it lies
// in the same tile at an earlier level or in a tile (w', c') with w' <= w and
// c' <= c.  Tiles are ordered by w + c, then c, which puts every tile after
// the tiles it reads.  T receives the rows in that order as levels of its own
// (one tile's rows of one level each, rows ascending, as in L), and
// T.tile_lev the first of each tile's levels.
static void tile_levels(const Levels &L, int32_t R, int32_t W, Levels &T) {
  const int32_t n = (int32_t)L.h_rows.size();
  const int32_t nc = (n + R - 1) / R, nw = (L.nlev + W - 1) / W;
  T.forward = L.forward;
  T.nlev = 0;
  T.ptr.assign(1, 0);
  T.h_rows.clear();
  T.h_rows.reserve((size_t)n);
  T.tile_lev.clear();
  for (int32_t d = 0; d <= nw + nc - 2; ++d) {
    for (int32_t c = std::max(0, d - (nw - 1)); c <= std::min(d, nc - 1); ++c) {
      const int32_t w = d - c;
      const int32_t first = T.nlev;
      // rows of chunk c: [lo, hi) in row index
      const int32_t lo = L.forward ? c * R : std::max(0, n - (c + 1) * R);
      const int32_t hi = L.forward ? std::min(n, (c + 1) * R) : n - c * R;
      for (int32_t l = w * W; l < std::min((w + 1) * W, L.nlev); ++l) {
        const int32_t *b = L.h_rows.data() + L.ptr[(size_t)l], *e = L.h_rows.data() + L.ptr[(size_t)l + 1];
        const int32_t *p0 = std::lower_bound(b, e, lo), *p1 = std::lower_bound(p0, e, hi);
        if (p0 == p1) continue;
        T.h_rows.insert(T.h_rows.end(), p0, p1);
        T.ptr.push_back((int32_t)T.h_rows.size());
        ++T.nlev;
      }
      if (T.nlev > first) T.tile_lev.push_back(first);
    }
  }
  T.tile_lev.push_back(T.nlev);
  if ((int32_t)T.h_rows.size() != n) fail(MTJG_E_DEVICE, "tiled sweep: tiles do not cover the level set");
}

static int32_t env_int(const char *name, int32_t dflt) {
  const char *e = getenv(name);
  return e && *e ? (int32_t)atoi(e) : dflt;
}

static void free_levels(Levels &L);

// Whether sweeps over L run tiled (k_sweep_tiles).  By default they do when
// some level is wider than an LDS chunk takes, so that the untiled plan would
// launch levels one by one; MTJG_SWEEP_TILES=0 or 1 forces the choice.
// Tiles are R rows by W levels: R = n / ceil(widest level / 384), so that a
// tile's share of the widest level is about 384 rows (under one 512-lane
// step), and W = 8 (MTJG_TILE_ROWS and MTJG_TILE_LEVELS override them).  On
// the 128^3 convection-diffusion matrix this gives 4 z-planes by 8 levels.
// Falls back to the untiled plan when a tile does not fit one LDS chunk.
static void decide_tiling(mtjg_prec M, Levels &L, int form, const std::vector<int32_t> &rp,
                          const std::vector<int32_t> &ci, const std::vector<int32_t> &diag, const double *a,
                          hipStream_t s) {
  L.tiling = -1;
  const int32_t n = (int32_t)L.h_rows.size();
  int32_t maxw = 0;
  for (int32_t l = 0; l < L.nlev; ++l) maxw = std::max(maxw, L.ptr[(size_t)l + 1] - L.ptr[(size_t)l]);
  const int32_t mode = env_int("MTJG_SWEEP_TILES", -1);
  if (n == 0 || mode == 0 || (mode < 0 && maxw <= LDS_LEVEL_MAX)) return;
  const int32_t nc = std::max(1, (maxw + 383) / 384);
  const int32_t R = std::max(1, env_int("MTJG_TILE_ROWS", (int32_t)(((int64_t)n + nc - 1) / nc)));
  const int32_t W = std::max(1, env_int("MTJG_TILE_LEVELS", 8));
  auto *T = new Levels();
  tile_levels(L, R, W, *T);
  const size_t cap = (size_t)lds_cap_doubles() * sizeof(double) - TILE_STATIC_LDS;
  if (!ensure_pack(M, *T, form, rp, ci, diag, a, s, &T->tile_lev) ||
      T->pk.tile_lds + (T->pk.tile_grouped ? sizeof(double) * FUSE_BS * ROWDOT_K : 0) > cap) {
    free_levels(*T);
    delete T;
    return;
  }
  L.tiled = T;
  L.tiling = 1;
}

template <int KIND, int W>
static void launch_sweep_tiles_w(Packed &P, const double *vals, const double *b, double *xdep, double omega,
                                 int32_t *err, hipStream_t s) {
  static bool attr_set = false;  // per instantiation: allow LDS beyond the 64 KiB default
  if (!attr_set) {
    HIPCHK(hipFuncSetAttribute((const void *)k_sweep_tiles<KIND, W>, hipFuncAttributeMaxDynamicSharedMemorySize,
                               lds_cap_doubles() * (int)sizeof(double) - (int)TILE_STATIC_LDS));
    attr_set = true;
  }
  const size_t bytes = P.tile_lds + (P.tile_grouped ? sizeof(double) * FUSE_BS * W : 0);
  const int grid = P.ntiles;
  if (++P.epoch == 0) {  // flags hold epochs; after 2^32 sweeps start again from clear flags
    HIPCHK(hipMemsetAsync(P.tflags, 0, sizeof(uint32_t) * TILE_FLAG_STRIDE * (size_t)std::max(P.ntiles, 1), s));
    P.epoch = 1;
  }
  k_sweep_tiles<KIND, W><<<grid, FUSE_BS, bytes, s>>>(P.ntiles, P.tiles, P.tdeps, P.tflags, P.epoch, P.tctr, err,
                                                      P.steps, P.lci, vals, P.pdv, b, P.own, P.imp, P.prow, xdep,
                                                      omega);
}

// One tiled sweep, W slot registers for the widest lane of any tile.
// b: the right-hand side in row order.
template <int KIND>
static void launch_sweep_tiles(Packed &P, const double *vals, const double *b, double *xdep, double omega,
                               int32_t *err, hipStream_t s) {
  if (P.ntiles == 0) return;
  if (P.tile_wmax <= 2) launch_sweep_tiles_w<KIND, 2>(P, vals, b, xdep, omega, err, s);
  else if (P.tile_wmax <= 4) launch_sweep_tiles_w<KIND, 4>(P, vals, b, xdep, omega, err, s);
  else launch_sweep_tiles_w<KIND, ROWDOT_K>(P, vals, b, xdep, omega, err, s);
}

template <int FORM>
static void run_levels(mtjg_prec M, Levels &L, hipStream_t s, const std::vector<int32_t> &rp,
                       const std::vector<int32_t> &ci, const std::vector<int32_t> &diag, const double *a,
                       const double *b, double *x) {
  if (L.tiling == 0) decide_tiling(M, L, FORM, rp, ci, diag, a, s);
  if (L.tiling > 0) {
    Levels &T = *L.tiled;
    ensure_pack(M, T, FORM, rp, ci, diag, a, s, &T.tile_lev);
    Packed &P = T.pk;
    launch_sweep_tiles<FORM>(P, P.eval, b, x, 1.0, M->sf_err, s);
    HIPCHK(hipGetLastError());
    return;
  }
  ensure_pack(M, L, FORM, rp, ci, diag, a, s);
  const Packed &P = L.pk;
  if (P.any_lds)
    k_sweep_prep<<<grid_for(M->n), BS, 0, s>>>(M->n, P.prow, b, P.bp, nullptr, nullptr, 0, nullptr, nullptr, nullptr,
                                               nullptr);
  for (const LevelRun &r : P.runs) {
    if (r.lds) {
      launch_sweep_lds<FORM>(r, P, P.eval, x, 1.0, s);
    } else if (r.fused) {
      k_trsv_fused<FORM><<<1, FUSE_BS, 0, s>>>(r.l0, r.l1, L.dptr, P.desc, P.pdv, P.eci, P.eval, b, x);
    } else {
      const int32_t l = r.l0, cnt = L.ptr[l + 1] - L.ptr[l];
      if (r.slab)
        k_trsv_level<FORM, true><<<(cnt + BS - 1) / BS, BS, 0, s>>>(cnt, L.ptr[l], P.eoff[(size_t)l],
                                                                    P.width[(size_t)l], P.desc, P.prow, P.pdv, P.eci,
                                                                    P.eval, b, x);
      else
        k_trsv_level<FORM, false><<<(cnt + BS - 1) / BS, BS, 0, s>>>(cnt, L.ptr[l], P.eoff[(size_t)l],
                                                                     P.width[(size_t)l], P.desc, P.prow, P.pdv,
                                                                     P.eci, P.eval, b, x);
    }
  }
  HIPCHK(hipGetLastError());
}

// One triangular solve as a single launch (see k_trsv_syncfree); `slot`
// selects the ticket counter so consecutive solves never share one.
template <int FORM>
static void run_syncfree(mtjg_prec M, const Levels &L, int slot, hipStream_t s, const int32_t *rp,
                         const int32_t *ci, const int32_t *diag, const double *a, const double *b,
                         double *x) {
  const int32_t n = M->n;
  if (n == 0) return;
  k_fill_pending<<<grid_for(n), BS, 0, s>>>(n, x);
  HIPCHK(hipMemsetAsync(M->sf_next + slot, 0, sizeof(uint32_t), s));
  // one workgroup per CU bounds the number of polling lanes
  const int grid = std::min((n + BS - 1) / BS, 256);
  k_trsv_syncfree<FORM><<<grid, BS, 0, s>>>(n, L.rows, rp, ci, diag, a, b, x, M->sf_next + slot, M->sf_err);
  HIPCHK(hipGetLastError());
}

// Recognises a natural-order 5- or 7-point grid stencil: every row's columns
// strictly ascending, the off-diagonal distances |col - row| drawn from
// {1, nx} (2-D, n = nx ny) or {1, nx, nx ny} (3-D, n = nx ny nz), and every
// entry a neighbour inside the grid (no wrap-around at a line or plane end).
// Rows may miss neighbours (the mask records which are present).
static bool detect_grid(int32_t n, const std::vector<int32_t> &rp, const std::vector<int32_t> &ci, GridPlan &g) {
  int64_t offs[3] = {0, 0, 0};
  int no = 0;
  for (int32_t i = 0; i < n; ++i)
    for (int32_t j = rp[(size_t)i]; j < rp[(size_t)i + 1]; ++j) {
      if (j > rp[(size_t)i] && ci[(size_t)j] <= ci[(size_t)j - 1]) return false;
      const int64_t d = std::llabs((int64_t)ci[(size_t)j] - i);
      if (d == 0 || d == offs[0] || d == offs[1] || d == offs[2]) continue;
      if (no == 3) return false;
      offs[no++] = d;
    }
  std::sort(offs, offs + no);
  if (no < 2 || offs[0] != 1) return false;
  const int64_t nx = offs[1];
  int64_t ny, nz;
  if (no == 2) {
    if (n % nx) return false;
    ny = n / nx;
    nz = 1;
  } else {
    if (offs[2] % nx || n % offs[2]) return false;
    ny = offs[2] / nx;
    nz = n / offs[2];
  }
  if (nx < 2 || ny < 2 || ny > 64 * GRID_WAVES) return false;
  for (int32_t i = 0; i < n; ++i) {
    const int64_t x = i % nx, y = (i / nx) % ny;
    for (int32_t j = rp[(size_t)i]; j < rp[(size_t)i + 1]; ++j) {
      const int64_t d = (int64_t)ci[(size_t)j] - i;
      if ((d == -1 && x == 0) || (d == 1 && x == nx - 1) || (d == -nx && y == 0) || (d == nx && y == ny - 1))
        return false;
    }
  }
  g.nx = (int32_t)nx;
  g.ny = (int32_t)ny;
  g.nz = (int32_t)nz;
  g.YB = (int32_t)((ny + 63) / 64);
  // about four wavefronts per workgroup: a step's cost grows with the
  // wavefronts that share a compute unit (0.32, 0.44 and 0.72 us at 4, 8 and
  // 16 on a 128^3 sweep), while each further plane group adds one hand-off
  // (about 3.4 us at two planes per group)
  g.P = std::max(1, std::min<int32_t>(4 / g.YB, g.nz));
  g.P = std::max(1, std::min({env_int("MTJG_ILU_GRID_PLANES", g.P), GRID_WAVES / g.YB, g.nz}));
  g.G = (g.nz + g.P - 1) / g.P;
  g.TW = g.nx + 64;
  // a head start of 3 steps: a larger one idles every workgroup longer than
  // the polls it saves (128^3 sweeps 242 / 261 us at 3, 277 / 306 at 5)
  g.lead = std::max(0, env_int("MTJG_ILU_GRID_LEAD", 3));
  g.poll_budget = GRID_POLL_BUDGET;
  if (const char *e = getenv("MTJG_ILU_GRID_POLL_BUDGET")) g.poll_budget = atoll(e);
  g.slab = (int64_t)g.YB * g.TW * 64;
  g.nslot = g.slab * g.nz;
  // one buffer descriptor with 32-bit offsets; below 2^31 bytes, an offset of
  // 2^31 or more is out of range whatever the scalar offset added to it
  return grid_bytes(g) < (int64_t)1 << 31;
}

static void grid_free(mtjg_prec M) {
  (void)hipFree(M->g_buf);
  (void)hipFree(M->g_tbuf);
  (void)hipFree(M->g_ctr);
  M->g_buf = M->g_tbuf = nullptr;
  M->g_ctr = nullptr;
  M->grid_t_version = ~0ull;
  M->grid = false;
  M->grid_vals_version = ~0ull;
}

static void grid_setup(mtjg_prec M) {
  GridPlan g;
  if (!M->grid_allowed || !detect_grid(M->n, M->h_rp, M->h_ci, g)) return;
  hipStream_t s = M->ctx->stream;
  const int64_t bytes = M->kind == PREC_SSOR ? grid_ssor_bytes(g) : grid_bytes(g);
  if (bytes >= (int64_t)1 << 31) return;  // one buffer descriptor with 32-bit offsets
  HIPCHK(hipMalloc(&M->g_buf, (size_t)bytes));
  HIPCHK(hipMalloc(&M->g_ctr, sizeof(uint32_t) * 2));
  HIPCHK(hipMemsetAsync(M->g_ctr, 0, sizeof(uint32_t) * 2, s));
  M->gp = g;
  M->grid = true;
  M->grid_vals_version = ~0ull;
}

// Two triangular solves as grid sweeps (see k_grid_sweep): a forward solve
// into y, then a backward solve into x.  FWD / BWD: TRSV_ROW_UNIT /
// TRSV_ROW_DIV for L.solve(b, y); U.solve(y, x) (ILU's apply), TRSV_COL_DIV /
// TRSV_ROW_DIV for R'.solve, R.solve (ICC's apply), TRSV_COL_DIV /
// TRSV_COL_UNIT_REV for U'.solve, L'.solve (ILU's transApply).
template <int FWD, int BWD>
static void grid_solves(mtjg_prec M, hipStream_t s, const double *b, double *x) {
  const GridPlan &g = M->gp;
  const int64_t nq = (int64_t)g.nz * g.YB * g.TW;
  const int blocks = (int)((nq + BS / 64 - 1) / (BS / 64));
  constexpr bool TRANS = FWD == TRSV_COL_DIV || BWD == TRSV_COL_UNIT_REV;
  if (M->grid_vals_version != M->vals_version) {
    k_grid_fill<<<blocks, BS, 0, s>>>(g, M->rp, M->ci, M->lu, M->g_buf);
    M->grid_vals_version = M->vals_version;
  }
  if (TRANS) {
    if (!M->g_tbuf) HIPCHK(hipMalloc(&M->g_tbuf, (size_t)gridt_bytes(g)));
    if (M->grid_t_version != M->vals_version) {
      k_grid_fill_t<<<blocks, BS, 0, s>>>(g, M->rp, M->ci, M->lu, M->g_tbuf);
      M->grid_t_version = M->vals_version;
    }
  }
  k_grid_gather<<<blocks, BS, 0, s>>>(g, b, M->g_buf);
  const int threads = 64 * g.P * g.YB;
  const uint32_t ns8 = (uint32_t)g.nslot * 8u;
  const uint32_t rb = (uint32_t)grid_bytes(g), tb = (uint32_t)gridt_bytes(g);
  const uint32_t ob = (uint32_t)grid_off(g, GRID_B), oy = (uint32_t)grid_off(g, GRID_Y),
                 ox = (uint32_t)grid_off(g, GRID_X), om = (uint32_t)grid_off(g, GRID_MASK);
  const uint32_t tm = GRIDT_MASK * ns8;
  if (FWD == TRSV_ROW_UNIT)
    k_grid_sweep<TRSV_ROW_UNIT><<<g.G, threads, 0, s>>>(g, M->g_buf, rb, 0, 0, om, M->g_buf, rb, ob, oy, 0, 0, 0.0,
                                                         M->g_ctr, M->sf_err);
  else
    k_grid_sweep<TRSV_COL_DIV><<<g.G, threads, 0, s>>>(g, M->g_tbuf, tb, GRIDT_U * ns8, GRIDT_D * ns8, tm, M->g_buf,
                                                        rb, ob, oy, 0, 0, 0.0, M->g_ctr, M->sf_err);
  if (BWD == TRSV_ROW_DIV)
    k_grid_sweep<TRSV_ROW_DIV><<<g.G, threads, 0, s>>>(g, M->g_buf, rb, 3 * ns8, 6 * ns8, om, M->g_buf, rb, oy, ox,
                                                        0, 0, 0.0, M->g_ctr, M->sf_err);
  else
    k_grid_sweep<TRSV_COL_UNIT_REV><<<g.G, threads, 0, s>>>(g, M->g_tbuf, tb, GRIDT_L * ns8, 0, tm, M->g_buf, rb,
                                                             oy, ox, 0, 0, 0.0, M->g_ctr, M->sf_err);
  k_grid_scatter<<<blocks, BS, 0, s>>>(g, M->g_buf, x);
  HIPCHK(hipGetLastError());
}

// SSOR.apply as grid sweeps (see k_grid_sweep): the forward sweep from x
// into the y slots, then (reverse) the backward sweep from them into the x
// slots; each sweep's fixed-side products are formed first.
static void grid_ssor(mtjg_prec M, hipStream_t s, const double *b, double *x) {
  const GridPlan &g = M->gp;
  const int64_t nq = (int64_t)g.nz * g.YB * g.TW;
  const int blocks = (int)((nq + BS / 64 - 1) / (BS / 64));
  if (M->grid_vals_version != M->vals_version) {
    k_grid_fill<<<blocks, BS, 0, s>>>(g, M->rp, M->ci, M->lu, M->g_buf);
    M->grid_vals_version = M->vals_version;
  }
  k_grid_ssor_prep_fwd<<<blocks, BS, 0, s>>>(g, b, x, M->g_buf);
  const int threads = 64 * g.P * g.YB;
  const uint32_t ns8 = (uint32_t)g.nslot * 8u, rb = (uint32_t)grid_ssor_bytes(g);
  const uint32_t ob = (uint32_t)grid_off(g, GRID_B), oy = (uint32_t)grid_off(g, GRID_Y),
                 ox = (uint32_t)grid_off(g, GRID_X), om = (uint32_t)grid_off(g, GRID_MASK);
  const uint32_t of = (uint32_t)grid_ssor_f(g), oo = (uint32_t)grid_ssor_own(g);
  k_grid_sweep<PACK_SSOR><<<g.G, threads, 0, s>>>(g, M->g_buf, rb, 0, 6 * ns8, om, M->g_buf, rb, ob, oy, of, oo,
                                                   M->omega_f, M->g_ctr, M->sf_err);
  if (M->reverse) {
    k_grid_ssor_prep_bwd<<<blocks, BS, 0, s>>>(g, M->g_buf);
    k_grid_sweep<PACK_SSOR_BWD><<<g.G, threads, 0, s>>>(g, M->g_buf, rb, 3 * ns8, 6 * ns8, om, M->g_buf, rb, ob, ox,
                                                         of, oy, M->omega_r, M->g_ctr, M->sf_err);
  }
  k_grid_scatter<<<blocks, BS, 0, s>>>(g, M->g_buf, x, M->reverse ? GRID_X : GRID_Y);
  HIPCHK(hipGetLastError());
}

// SSOR.apply (S/SSOR.java:143-194): forward sweep into xx, then the reverse
// sweep into x, or x = xx when the reverse sweep is off.  MTJ's initial copy
// xx = x is not needed: the forward sweep writes every xx_i before any row
// reads it.
template <bool FORWARD>
static void ssor_sweep(mtjg_prec M, Levels &L, hipStream_t s, const double *b, double *x, double omega) {
  constexpr int KIND = FORWARD ? PACK_SSOR : PACK_SSOR_BWD;
  ensure_pack(M, L, KIND, M->h_rp, M->h_ci, M->h_diag, M->lu, s);
  const Packed &P = L.pk;
  // the forward sweep reads x (as xd) and writes xx; the reverse one reads xx
  double *fixed = FORWARD ? x : M->xx, *xdep = FORWARD ? M->xx : x;
  if (P.any_lds)
    k_sweep_prep<<<grid_for(std::max<int64_t>(M->n, P.nfix)), BS, 0, s>>>(M->n, P.prow, b, P.bp, fixed, P.own,
                                                                         P.nfix, P.fixslot, P.eci, P.eval, P.fval);
  for (const LevelRun &r : P.runs) {
    if (r.lds) {
      launch_sweep_lds<KIND>(r, P, P.fval, xdep, omega, s);
    } else if (r.fused) {
      k_ssor_fused<FORWARD><<<1, FUSE_BS, 0, s>>>(r.l0, r.l1, L.dptr, P.desc, P.pdv, P.eci, P.eval, b, x, M->xx,
                                                  omega);
    } else {
      const int32_t l = r.l0, cnt = L.ptr[l + 1] - L.ptr[l];
      if (r.slab)
        k_ssor_level<FORWARD, true><<<(cnt + BS - 1) / BS, BS, 0, s>>>(
            cnt, L.ptr[l], P.eoff[(size_t)l], P.width[(size_t)l], P.desc, P.prow, P.pnlo, P.pdv, P.eci, P.eval, b, x,
            M->xx, omega);
      else
        k_ssor_level<FORWARD, false><<<(cnt + BS - 1) / BS, BS, 0, s>>>(
            cnt, L.ptr[l], P.eoff[(size_t)l], P.width[(size_t)l], P.desc, P.prow, P.pnlo, P.pdv, P.eci, P.eval, b, x,
            M->xx, omega);
    }
  }
}

static void run_ssor(mtjg_prec M, hipStream_t s, const double *b, double *x) {
  Levels &L = M->lev_l, &U = M->lev_u;
  ssor_sweep<true>(M, L, s, b, x, M->omega_f);
  if (!M->reverse) {
    if (M->n) HIPCHK(hipMemcpyAsync(x, M->xx, sizeof(double) * (size_t)M->n, hipMemcpyDeviceToDevice, s));
    HIPCHK(hipGetLastError());
    return;
  }
  ssor_sweep<false>(M, U, s, b, x, M->omega_r);
  HIPCHK(hipGetLastError());
}

static void free_levels(Levels &L) {
  if (L.tiled) {
    free_levels(*L.tiled);
    delete L.tiled;
    L.tiled = nullptr;
  }
  L.tiling = 0;
  if (L.rows) (void)hipFree(L.rows);
  if (L.dptr) (void)hipFree(L.dptr);
  L.rows = nullptr;
  L.dptr = nullptr;
  L.h_rows.clear();
  free_pack(L.pk);
  L.nlev = 0;
  L.ptr.clear();
}

static void ilu_free_storage(mtjg_prec M) {
  (void)hipFree(M->rp);
  (void)hipFree(M->ci);
  (void)hipFree(M->diag);
  (void)hipFree(M->lu);
  (void)hipFree(M->trp);
  (void)hipFree(M->tci);
  (void)hipFree(M->tdiag);
  (void)hipFree(M->tperm);
  (void)hipFree(M->tval);
  M->rp = M->ci = M->diag = M->trp = M->tci = M->tdiag = M->tperm = nullptr;
  M->lu = M->tval = nullptr;
  M->h_trp.clear();
  M->h_tci.clear();
  M->h_tdiag.clear();
  free_levels(M->lev_l);
  free_levels(M->lev_u);
  free_levels(M->lev_ut);
  free_levels(M->lev_lt);
  grid_free(M);
  M->trans_ready = false;
  M->factored = false;
}

static int32_t read_bad(mtjg_prec M) {
  int32_t h = INT32_MAX;
  HIPCHK(hipMemcpyAsync(&h, M->bad, sizeof(int32_t), hipMemcpyDeviceToHost, M->ctx->stream));
  HIPCHK(hipStreamSynchronize(M->ctx->stream));
  return h;
}

static void reset_bad(mtjg_prec M) {
  int32_t big = INT32_MAX;
  HIPCHK(hipMemcpyAsync(M->bad, &big, sizeof(int32_t), hipMemcpyHostToDevice, M->ctx->stream));
}

// LU.set(A) / F.set(A) (S/CompRowMatrix.java:449-471): takes A's pattern
// (rebuilding the diagonal positions and the lower / upper level sets when it
// differs from the stored one) and values.  A missing diagonal fails with
// `missing` + the 1-based row, MTJ's message.
static void take_pattern(mtjg_prec M, mtjg_csr A, int32_t *bad_row, const char *missing) {
  mtjg_ctx ctx = M->ctx;
  hipStream_t s = ctx->stream;
  const int32_t n = A->nr;
  bool same_pattern = M->rp && M->nnz == A->nnz && M->h_rp == A->h_rp;
  if (same_pattern) {
    std::vector<int32_t> ci((size_t)A->nnz);
    if (A->nnz)
      HIPCHK(hipMemcpyAsync(ci.data(), A->ci, sizeof(int32_t) * (size_t)A->nnz, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    same_pattern = (ci == M->h_ci);
  }
  if (!same_pattern) {
    ilu_free_storage(M);
    M->nnz = A->nnz;
    M->h_rp = A->h_rp;
    M->h_ci.resize((size_t)A->nnz);
    size_t nnz1 = (size_t)std::max<int64_t>(A->nnz, 1);
    HIPCHK(hipMalloc(&M->rp, sizeof(int32_t) * (size_t)(n + 1)));
    HIPCHK(hipMalloc(&M->ci, sizeof(int32_t) * nnz1));
    HIPCHK(hipMalloc(&M->lu, sizeof(double) * nnz1));
    HIPCHK(hipMalloc(&M->diag, sizeof(int32_t) * (size_t)std::max(n, 1)));
    HIPCHK(hipMemcpyAsync(M->rp, A->rp, sizeof(int32_t) * (size_t)(n + 1), hipMemcpyDeviceToDevice, s));
    if (A->nnz) {
      HIPCHK(hipMemcpyAsync(M->ci, A->ci, sizeof(int32_t) * (size_t)A->nnz, hipMemcpyDeviceToDevice, s));
      HIPCHK(hipMemcpyAsync(M->h_ci.data(), A->ci, sizeof(int32_t) * (size_t)A->nnz,
                            hipMemcpyDeviceToHost, s));
    }
    reset_bad(M);
    if (n > 0) k_find_diag<<<(n + BS - 1) / BS, BS, 0, s>>>(n, M->rp, M->ci, M->diag, M->bad);
    HIPCHK(hipGetLastError());
    int32_t b = read_bad(M);
    if (b != INT32_MAX) {
      if (bad_row) *bad_row = b;
      ilu_free_storage(M);
      M->h_rp.clear();
      fail(MTJG_E_MISSING_DIAGONAL, missing + std::to_string(b + 1));
    }
    std::vector<int32_t> &hd = M->h_diag;
    hd.assign((size_t)n, 0);
    if (n) HIPCHK(hipMemcpyAsync(hd.data(), M->diag, sizeof(int32_t) * (size_t)n, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    // L levels: dependencies are the strictly-lower entries [rp[i], diag[i])
    std::vector<int32_t> beg((size_t)n), end((size_t)n);
    for (int32_t i = 0; i < n; ++i) {
      beg[(size_t)i] = M->h_rp[i];
      end[(size_t)i] = hd[(size_t)i];
    }
    build_levels(ctx, n, M->h_rp, M->h_ci, beg, end, true, M->lev_l);
    // U levels: strictly-upper entries (diag[i], rp[i+1])
    for (int32_t i = 0; i < n; ++i) {
      beg[(size_t)i] = hd[(size_t)i] + 1;
      end[(size_t)i] = M->h_rp[i + 1];
    }
    build_levels(ctx, n, M->h_rp, M->h_ci, beg, end, false, M->lev_u);
    if (M->kind == PREC_ILU || M->kind == PREC_ICC || M->kind == PREC_SSOR) grid_setup(M);
  }
  if (A->nnz)
    HIPCHK(hipMemcpyAsync(M->lu, A->val, sizeof(double) * (size_t)A->nnz, hipMemcpyDeviceToDevice, s));
  M->vals_version++;
}

// ILU.setMatrix: LU.set(A), then factor() (S/ILU.java:86-125).
static void ilu_set_matrix(mtjg_prec M, mtjg_csr A, int32_t *bad_row) {
  hipStream_t s = M->ctx->stream;
  take_pattern(M, A, bad_row, "Missing diagonal entry on row ");
  reset_bad(M);
  // rows of the first L level have no strictly-lower entries: nothing to do
  for (int32_t l = 1; l < M->lev_l.nlev; ++l) {
    int32_t cnt = M->lev_l.ptr[l + 1] - M->lev_l.ptr[l];
    k_ilu_factor_level<<<(cnt + BS - 1) / BS, BS, 0, s>>>(cnt, M->lev_l.rows + M->lev_l.ptr[l], M->rp,
                                                          M->ci, M->diag, M->lu, M->bad);
  }
  HIPCHK(hipGetLastError());
  int32_t b = read_bad(M);
  M->factored = false;
  if (b != INT32_MAX) {
    if (bad_row) *bad_row = b;
    fail(MTJG_E_ZERO_PIVOT, "Zero pivot encountered on row " + std::to_string(b + 1) +
                                " during ILU process");
  }
  M->factored = true;
  if (M->trans_ready && M->tval)
    k_gather<<<grid_for(M->nnz), BS, 0, s>>>(M->nnz, M->tperm, M->lu, M->tval);
  M->vals_version++;
}

static void build_transpose(mtjg_prec M, bool lower_levels);

// ICC.setMatrix: R.set(A), then factor() (S/ICC.java:80-84).
static void icc_set_matrix(mtjg_prec M, mtjg_csr A, int32_t *bad_row) {
  hipStream_t s = M->ctx->stream;
  M->factored = false;
  take_pattern(M, A, bad_row, "Missing diagonal entry on row ");
  std::vector<double> data((size_t)M->nnz);
  if (M->nnz)
    HIPCHK(hipMemcpyAsync(data.data(), M->lu, sizeof(double) * (size_t)M->nnz, hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  int32_t bad = -1;
  int rc = icc_factor_host(M->n, M->h_rp, M->h_ci, M->h_diag, data, &bad);
  if (rc) {
    if (bad_row) *bad_row = bad;
    if (rc == 1)
      fail(MTJG_E_ZERO_PIVOT, "Zero pivot encountered on row " + std::to_string(bad + 1) + " during ICC process");
    fail(MTJG_E_ZERO_DIAGONAL,
         "Zero diagonal entry encountered on row " + std::to_string(bad + 1) + " during ICC process");
  }
  if (M->nnz) {
    HIPCHK(hipMemcpyAsync(M->lu, data.data(), sizeof(double) * (size_t)M->nnz, hipMemcpyHostToDevice, s));
    HIPCHK(hipStreamSynchronize(s));
  }
  if (!M->trans_ready) build_transpose(M, false);
  else if (M->nnz) k_gather<<<grid_for(M->nnz), BS, 0, s>>>(M->nnz, M->tperm, M->lu, M->tval);
  HIPCHK(hipGetLastError());
  M->vals_version++;
  M->factored = true;
}

// Transposed factor for ILU's transApply (S/ILU.java:72-78) and ICC's apply
// (S/ICC.java:65-67): CSR of the factor's transpose with each transposed row
// in ascending original-row order, a gather map from the factor storage, and
// the level sets of U' (forward) and, for ILU, L' (backward).
static void build_transpose(mtjg_prec M, bool lower_levels) {
  mtjg_ctx ctx = M->ctx;
  hipStream_t s = ctx->stream;
  const int32_t n = M->n;
  const int64_t nnz = M->nnz;
  std::vector<int32_t> trp((size_t)n + 1, 0), tci((size_t)nnz), tperm((size_t)nnz), tdiag((size_t)n, -1);
  for (int64_t j = 0; j < nnz; ++j) trp[(size_t)M->h_ci[(size_t)j] + 1]++;
  for (int32_t c = 0; c < n; ++c) trp[c + 1] += trp[c];
  std::vector<int32_t> pos(trp.begin(), trp.end() - 1);
  for (int32_t i = 0; i < n; ++i)
    for (int32_t j = M->h_rp[i]; j < M->h_rp[i + 1]; ++j) {
      int32_t c = M->h_ci[(size_t)j];
      int32_t p = pos[(size_t)c]++;
      tci[(size_t)p] = i;
      tperm[(size_t)p] = j;
      if (i == c) tdiag[(size_t)c] = p;
    }
  size_t nnz1 = (size_t)std::max<int64_t>(nnz, 1);
  HIPCHK(hipMalloc(&M->trp, sizeof(int32_t) * (size_t)(n + 1)));
  HIPCHK(hipMalloc(&M->tci, sizeof(int32_t) * nnz1));
  HIPCHK(hipMalloc(&M->tperm, sizeof(int32_t) * nnz1));
  HIPCHK(hipMalloc(&M->tval, sizeof(double) * nnz1));
  HIPCHK(hipMalloc(&M->tdiag, sizeof(int32_t) * (size_t)std::max(n, 1)));
  HIPCHK(hipMemcpyAsync(M->trp, trp.data(), sizeof(int32_t) * trp.size(), hipMemcpyHostToDevice, s));
  if (nnz) {
    HIPCHK(hipMemcpyAsync(M->tci, tci.data(), sizeof(int32_t) * (size_t)nnz, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(M->tperm, tperm.data(), sizeof(int32_t) * (size_t)nnz, hipMemcpyHostToDevice, s));
  }
  if (n) HIPCHK(hipMemcpyAsync(M->tdiag, tdiag.data(), sizeof(int32_t) * (size_t)n, hipMemcpyHostToDevice, s));
  std::vector<int32_t> beg((size_t)n), end((size_t)n);
  for (int32_t c = 0; c < n; ++c) {
    beg[(size_t)c] = trp[c];
    end[(size_t)c] = tdiag[(size_t)c];
  }
  build_levels(ctx, n, trp, tci, beg, end, true, M->lev_ut);
  if (lower_levels) {
    for (int32_t c = 0; c < n; ++c) {
      beg[(size_t)c] = tdiag[(size_t)c] + 1;
      end[(size_t)c] = trp[c + 1];
    }
    build_levels(ctx, n, trp, tci, beg, end, false, M->lev_lt);
  }
  if (nnz) k_gather<<<grid_for(nnz), BS, 0, s>>>(nnz, M->tperm, M->lu, M->tval);
  HIPCHK(hipGetLastError());
  M->h_trp = std::move(trp);
  M->h_tci = std::move(tci);
  M->h_tdiag = std::move(tdiag);
  M->trans_ready = true;
  M->vals_version++;
}

void prec_apply_dev(mtjg_prec M, int64_t n, const double *b, double *x, hipStream_t s, KState *st) {
  if (!M || M->kind == PREC_IDENTITY) {
    // IdentityPreconditioner: x.set(b) (S/AbstractIterativeSolver.java:84-86)
    if (b != x && n > 0) k_copy_if_running<<<grid_for(n), BS, 0, s>>>(n, b, x, st);
  } else if (M->kind == PREC_DIAGONAL) {
    if (n > 0) k_jacobi_apply<<<grid_for(n), BS, 0, s>>>(n, b, M->invdiag, x, st);
  } else if (M->kind == PREC_SSOR) {
    if (!M->factored) fail(MTJG_E_INVALID_ARGUMENT, "SSOR.setMatrix must precede apply");
    if (M->grid) grid_ssor(M, s, b, x);
    else run_ssor(M, s, b, x);
  } else if (M->kind == PREC_AMG) {
    if (!M->amg) fail(MTJG_E_INVALID_ARGUMENT, "AMG.setMatrix must precede apply");
    amg_apply_device(M->amg, b, x, false, s);
  } else if (M->kind == PREC_CG32) {
    if (!M->A) fail(MTJG_E_INVALID_ARGUMENT, "setMatrix must precede apply");
    cg32_apply(M->inner, M->A, b, x, st);
  } else if (M->kind == PREC_ICC) {
    if (!M->factored) fail(MTJG_E_INVALID_ARGUMENT, "ICC.setMatrix must precede apply");
    // Rt.transSolve(b, y); Rt.solve(y, x) (S/ICC.java:65-71)
    if (M->level_launch && M->grid) {
      grid_solves<TRSV_COL_DIV, TRSV_ROW_DIV>(M, s, b, x);
    } else if (M->level_launch) {
      run_levels<TRSV_COL_DIV>(M, M->lev_ut, s, M->h_trp, M->h_tci, M->h_tdiag, M->tval, b, M->y);
      run_levels<TRSV_ROW_DIV>(M, M->lev_u, s, M->h_rp, M->h_ci, M->h_diag, M->lu, M->y, x);
    } else {
      run_syncfree<TRSV_COL_DIV>(M, M->lev_ut, 2, s, M->trp, M->tci, M->tdiag, M->tval, b, M->y);
      run_syncfree<TRSV_ROW_DIV>(M, M->lev_u, 1, s, M->rp, M->ci, M->diag, M->lu, M->y, x);
    }
  } else {
    if (!M->factored) fail(MTJG_E_INVALID_ARGUMENT, "ILU.setMatrix must precede apply");
    // L.solve(b, y); U.solve(y, x) (S/ILU.java:64-70)
    if (M->level_launch && M->grid) {
      grid_solves<TRSV_ROW_UNIT, TRSV_ROW_DIV>(M, s, b, x);
    } else if (M->level_launch) {
      run_levels<TRSV_ROW_UNIT>(M, M->lev_l, s, M->h_rp, M->h_ci, M->h_diag, M->lu, b, M->y);
      run_levels<TRSV_ROW_DIV>(M, M->lev_u, s, M->h_rp, M->h_ci, M->h_diag, M->lu, M->y, x);
    } else {
      run_syncfree<TRSV_ROW_UNIT>(M, M->lev_l, 0, s, M->rp, M->ci, M->diag, M->lu, b, M->y);
      run_syncfree<TRSV_ROW_DIV>(M, M->lev_u, 1, s, M->rp, M->ci, M->diag, M->lu, M->y, x);
    }
  }
  HIPCHK(hipGetLastError());
}

void prec_trans_apply_dev(mtjg_prec M, const double *b, double *x, hipStream_t s) {
  if (M->kind == PREC_AMG) {
    if (!M->amg) fail(MTJG_E_INVALID_ARGUMENT, "AMG.setMatrix must precede transApply");
    amg_apply_device(M->amg, b, x, true, s);
    return;
  }
  if (M->kind != PREC_ILU) {
    prec_apply_dev(M, M->n, b, x, s, nullptr);
    return;
  }
  if (!M->factored) fail(MTJG_E_INVALID_ARGUMENT, "ILU.setMatrix must precede transApply");
  if (!M->trans_ready) build_transpose(M, true);
  // U.transSolve(b, y); L.transSolve(y, x) (S/ILU.java:72-78)
  if (M->level_launch && M->grid) {
    grid_solves<TRSV_COL_DIV, TRSV_COL_UNIT_REV>(M, s, b, x);
  } else if (M->level_launch) {
    run_levels<TRSV_COL_DIV>(M, M->lev_ut, s, M->h_trp, M->h_tci, M->h_tdiag, M->tval, b, M->y);
    run_levels<TRSV_COL_UNIT_REV>(M, M->lev_lt, s, M->h_trp, M->h_tci, M->h_tdiag, M->tval, M->y, x);
  } else {
    run_syncfree<TRSV_COL_DIV>(M, M->lev_ut, 2, s, M->trp, M->tci, M->tdiag, M->tval, b, M->y);
    run_syncfree<TRSV_COL_UNIT_REV>(M, M->lev_lt, 3, s, M->trp, M->tci, M->tdiag, M->tval, M->y, x);
  }
}

// Raises if a single-launch triangular solve ran out of its poll budget
// since the last check (synchronises the stream).
void prec_check(mtjg_prec M) {
  if (M && M->kind == PREC_AMG && M->amg) {
    amg_check_device(M->amg);  // its smoothers' grid sweeps
    return;
  }
  if (!M || !M->sf_err) return;
  if (M->kind != PREC_ILU && M->kind != PREC_ICC && !(M->kind == PREC_SSOR && M->grid)) return;
  int32_t e = 0;
  HIPCHK(hipMemcpyAsync(&e, M->sf_err, sizeof(int32_t), hipMemcpyDeviceToHost, M->ctx->stream));
  HIPCHK(hipStreamSynchronize(M->ctx->stream));
  if (e) {
    HIPCHK(hipMemsetAsync(M->sf_err, 0, sizeof(int32_t), M->ctx->stream));
    fail(MTJG_E_DEVICE, "triangular solve exceeded its dependency poll budget");
  }
}

int prec_kind(mtjg_prec M) { return M ? M->kind : PREC_IDENTITY; }
const double *prec_invdiag(mtjg_prec M) { return M ? M->invdiag : nullptr; }
int32_t prec_n(mtjg_prec M) { return M ? M->n : 0; }

}  // namespace mtjg

using namespace mtjg;

// ===========================================================================
// C ABI

extern "C" int mtjg_prec_diagonal_create(mtjg_ctx ctx, int32_t n, mtjg_prec *out) {
  MTJG_API_BEGIN
  if (!ctx || !out) fail(MTJG_E_NULL, "null argument");
  if (n < 0) fail(MTJG_E_INVALID_ARGUMENT, "negative size");
  auto *M = new mtjg_prec_s();
  M->ctx = ctx;
  M->kind = PREC_DIAGONAL;
  M->n = n;
  HIPCHK(hipMalloc(&M->invdiag, sizeof(double) * (size_t)std::max(n, 1)));
  HIPCHK(hipMemsetAsync(M->invdiag, 0, sizeof(double) * (size_t)std::max(n, 1), ctx->stream));
  HIPCHK(hipMalloc(&M->bad, sizeof(int32_t)));
  *out = M;
  MTJG_API_END
}

extern "C" int mtjg_prec_cg32_create(mtjg_ctx ctx, int32_t n, double inner_rtol, int32_t inner_max_iter,
                                     mtjg_prec *out) {
  MTJG_API_BEGIN
  if (!ctx || !out) fail(MTJG_E_NULL, "null argument");
  if (n < 1) fail(MTJG_E_INVALID_ARGUMENT, "size must be positive");
  if (!(inner_rtol > 0 && inner_rtol < 1) || inner_max_iter < 1)
    fail(MTJG_E_INVALID_ARGUMENT, "inner tolerance must be in (0, 1) and the iteration limit positive");
  auto *M = new mtjg_prec_s();
  M->ctx = ctx;
  M->kind = PREC_CG32;
  M->n = n;
  try {
    M->inner = cg32_create(ctx, n, inner_rtol, inner_max_iter);
  } catch (...) {
    delete M;
    throw;
  }
  *out = M;
  MTJG_API_END
}

extern "C" int mtjg_prec_inner_iterations(mtjg_prec M, int64_t *total) {
  MTJG_API_BEGIN
  if (!M || !total) fail(MTJG_E_NULL, "null argument");
  if (M->kind != PREC_CG32) fail(MTJG_E_INVALID_ARGUMENT, "not an FP32 inner-CG preconditioner");
  *total = cg32_iterations(M->inner);
  MTJG_API_END
}

extern "C" int mtjg_prec_ilu_create(mtjg_ctx ctx, mtjg_csr LU, mtjg_prec *out) {
  MTJG_API_BEGIN
  if (!ctx || !LU || !out) fail(MTJG_E_NULL, "null argument");
  if (LU->nr != LU->nc) fail(MTJG_E_INVALID_ARGUMENT, "ILU only applies to square matrices");
  auto *M = new mtjg_prec_s();
  M->ctx = ctx;
  M->kind = PREC_ILU;
  M->n = LU->nr;
  HIPCHK(hipMalloc(&M->y, sizeof(double) * (size_t)std::max(M->n, 1)));
  HIPCHK(hipMemsetAsync(M->y, 0, sizeof(double) * (size_t)std::max(M->n, 1), ctx->stream));
  HIPCHK(hipMalloc(&M->bad, sizeof(int32_t)));
  HIPCHK(hipMalloc(&M->sf_next, sizeof(uint32_t) * 4));
  HIPCHK(hipMalloc(&M->sf_err, sizeof(int32_t)));
  HIPCHK(hipMemsetAsync(M->sf_next, 0, sizeof(uint32_t) * 4, ctx->stream));
  HIPCHK(hipMemsetAsync(M->sf_err, 0, sizeof(int32_t), ctx->stream));
  M->level_launch = true;
  if (const char *e = getenv("MTJG_ILU_SYNCFREE")) M->level_launch = atoi(e) == 0;
  if (const char *e = getenv("MTJG_ILU_GRID")) M->grid_allowed = atoi(e) != 0;
  *out = M;
  MTJG_API_END
}

extern "C" int mtjg_prec_ilu_grid(mtjg_prec M, int32_t dims[3]) {
  MTJG_API_BEGIN
  if (!M || !dims) fail(MTJG_E_NULL, "null argument");
  if (M->kind != PREC_ILU && M->kind != PREC_ICC && M->kind != PREC_SSOR)
    fail(MTJG_E_INVALID_ARGUMENT, "not an ILU, ICC or SSOR preconditioner");
  dims[0] = M->grid ? M->gp.nx : 0;
  dims[1] = M->grid ? M->gp.ny : 0;
  dims[2] = M->grid ? M->gp.nz : 0;
  MTJG_API_END
}

extern "C" int mtjg_prec_icc_create(mtjg_ctx ctx, mtjg_csr R, mtjg_prec *out) {
  MTJG_API_BEGIN
  if (!ctx || !R || !out) fail(MTJG_E_NULL, "null argument");
  if (R->nr != R->nc) fail(MTJG_E_INVALID_ARGUMENT, "ICC only applies to square matrices");
  auto *M = new mtjg_prec_s();
  M->ctx = ctx;
  M->kind = PREC_ICC;
  M->n = R->nr;
  HIPCHK(hipMalloc(&M->y, sizeof(double) * (size_t)std::max(M->n, 1)));
  HIPCHK(hipMemsetAsync(M->y, 0, sizeof(double) * (size_t)std::max(M->n, 1), ctx->stream));
  HIPCHK(hipMalloc(&M->bad, sizeof(int32_t)));
  HIPCHK(hipMalloc(&M->sf_next, sizeof(uint32_t) * 4));
  HIPCHK(hipMalloc(&M->sf_err, sizeof(int32_t)));
  HIPCHK(hipMemsetAsync(M->sf_next, 0, sizeof(uint32_t) * 4, ctx->stream));
  HIPCHK(hipMemsetAsync(M->sf_err, 0, sizeof(int32_t), ctx->stream));
  M->level_launch = true;
  if (const char *e = getenv("MTJG_ILU_SYNCFREE")) M->level_launch = atoi(e) == 0;
  if (const char *e = getenv("MTJG_ILU_GRID")) M->grid_allowed = atoi(e) != 0;
  *out = M;
  MTJG_API_END
}

extern "C" int mtjg_prec_amg_create(mtjg_ctx ctx, double omega_pre_f, double omega_pre_r, double omega_post_f,
                                    double omega_post_r, int32_t nu1, int32_t nu2, int32_t gamma, int32_t min,
                                    double omega, int32_t reverse, mtjg_prec *out) {
  MTJG_API_BEGIN
  if (!ctx || !out) fail(MTJG_E_NULL, "null argument");
  auto *M = new mtjg_prec_s();
  M->ctx = ctx;
  M->kind = PREC_AMG;
  M->n = 0;
  amg::Params &p = M->amg_params;
  p.omega_pre_f = omega_pre_f;
  p.omega_pre_r = omega_pre_r;
  p.omega_post_f = omega_post_f;
  p.omega_post_r = omega_post_r;
  p.nu1 = nu1;
  p.nu2 = nu2;
  p.gamma = gamma;
  p.min = min;
  p.omega = omega;
  p.reverse = reverse != 0;
  *out = M;
  MTJG_API_END
}

extern "C" int mtjg_prec_amg_levels(mtjg_prec M, int32_t *levels, int32_t *sizes, int32_t max_sizes) {
  MTJG_API_BEGIN
  if (!M || !levels) fail(MTJG_E_NULL, "null argument");
  if (M->kind != PREC_AMG) fail(MTJG_E_INVALID_ARGUMENT, "not an AMG preconditioner");
  int32_t tmp = 0;
  *levels = M->amg ? amg_level_sizes(M->amg, sizes ? sizes : &tmp, sizes ? max_sizes : 0) : 0;
  MTJG_API_END
}

extern "C" int mtjg_prec_ssor_create(mtjg_ctx ctx, mtjg_csr F, int32_t reverse, double omega_f, double omega_r,
                                     mtjg_prec *out) {
  MTJG_API_BEGIN
  if (!ctx || !F || !out) fail(MTJG_E_NULL, "null argument");
  if (F->nr != F->nc) fail(MTJG_E_INVALID_ARGUMENT, "SSOR only applies to square matrices");
  if (omega_f < 0 || omega_f > 2) fail(MTJG_E_INVALID_ARGUMENT, "omegaF must be between 0 and 2");
  if (omega_r < 0 || omega_r > 2) fail(MTJG_E_INVALID_ARGUMENT, "omegaR must be between 0 and 2");
  auto *M = new mtjg_prec_s();
  M->ctx = ctx;
  M->kind = PREC_SSOR;
  M->n = F->nr;
  M->reverse = reverse != 0;
  M->omega_f = omega_f;
  M->omega_r = omega_r;
  HIPCHK(hipMalloc(&M->xx, sizeof(double) * (size_t)std::max(M->n, 1)));
  HIPCHK(hipMemsetAsync(M->xx, 0, sizeof(double) * (size_t)std::max(M->n, 1), ctx->stream));
  HIPCHK(hipMalloc(&M->bad, sizeof(int32_t)));
  HIPCHK(hipMalloc(&M->sf_err, sizeof(int32_t)));  // grid sweeps' poll-budget flag
  HIPCHK(hipMemsetAsync(M->sf_err, 0, sizeof(int32_t), ctx->stream));
  if (const char *e = getenv("MTJG_ILU_GRID")) M->grid_allowed = atoi(e) != 0;
  *out = M;
  MTJG_API_END
}

extern "C" int mtjg_prec_ssor_set_omega(mtjg_prec M, double omega_f, double omega_r) {
  MTJG_API_BEGIN
  if (!M) fail(MTJG_E_NULL, "null argument");
  if (M->kind != PREC_SSOR) fail(MTJG_E_INVALID_ARGUMENT, "not an SSOR preconditioner");
  if (omega_f < 0 || omega_f > 2) fail(MTJG_E_INVALID_ARGUMENT, "omegaF must be between 0 and 2");
  if (omega_r < 0 || omega_r > 2) fail(MTJG_E_INVALID_ARGUMENT, "omegaR must be between 0 and 2");
  M->omega_f = omega_f;
  M->omega_r = omega_r;
  MTJG_API_END
}

extern "C" int mtjg_prec_set_matrix(mtjg_prec M, mtjg_csr A, int32_t *bad_row) {
  MTJG_API_BEGIN
  if (!M || !A) fail(MTJG_E_NULL, "null argument");
  hipStream_t s = M->ctx->stream;
  if (M->kind == PREC_DIAGONAL) {
    if (A->nr != M->n) fail(MTJG_E_INVALID_ARGUMENT, "Matrix size differs from preconditioner size");
    if (A->nc < A->nr) fail(MTJG_E_INDEX_OUT_OF_BOUNDS, "matrix has fewer columns than rows");
    reset_bad(M);
    if (M->n > 0)
      k_jacobi_setup<<<(M->n + BS - 1) / BS, BS, 0, s>>>(M->n, A->rp, A->ci, A->val, M->invdiag, M->bad);
    HIPCHK(hipGetLastError());
    int32_t b = read_bad(M);
    if (b != INT32_MAX) {
      if (bad_row) *bad_row = b;
      fail(MTJG_E_ZERO_DIAGONAL, "Zero diagonal on row " + std::to_string(b + 1));
    }
  } else if (M->kind == PREC_ILU) {
    if (A->nr != M->n || A->nc != M->n) fail(MTJG_E_INDEX_OUT_OF_BOUNDS, "matrix size differs from ILU size");
    ilu_set_matrix(M, A, bad_row);
  } else if (M->kind == PREC_AMG) {
    if (A->nr != A->nc) fail(MTJG_E_INVALID_ARGUMENT, "AMG only applies to square matrices");
    amg_destroy_device(M->amg);
    M->amg = nullptr;
    M->n = 0;
    M->amg = amg_create_device(M->ctx, A, M->amg_params, bad_row);
    M->n = A->nr;
  } else if (M->kind == PREC_ICC) {
    if (A->nr != M->n || A->nc != M->n) fail(MTJG_E_INDEX_OUT_OF_BOUNDS, "matrix size differs from ICC size");
    icc_set_matrix(M, A, bad_row);
  } else if (M->kind == PREC_CG32) {
    if (A->nr != M->n || A->nc != M->n) fail(MTJG_E_INDEX_OUT_OF_BOUNDS, "matrix size differs from preconditioner size");
    M->A = nullptr;
    cg32_set_matrix(M->inner, A, bad_row);
    M->A = A;
  } else if (M->kind == PREC_SSOR) {
    if (A->nr != M->n || A->nc != M->n) fail(MTJG_E_INDEX_OUT_OF_BOUNDS, "matrix size differs from SSOR size");
    M->factored = false;
    take_pattern(M, A, bad_row, "Missing diagonal on row ");
    M->factored = true;
  }
  MTJG_API_END
}

static void check_apply(mtjg_prec M, mtjg_vec b, mtjg_vec x) {
  if (!M || !b || !x) fail(MTJG_E_NULL, "null argument");
  if (M->kind == PREC_AMG && !M->amg) fail(MTJG_E_INVALID_ARGUMENT, "AMG.setMatrix must precede apply");
  if (M->kind == PREC_CG32 && !M->A) fail(MTJG_E_INVALID_ARGUMENT, "setMatrix must precede apply");
  if (b->dtype != MTJG_F64 || x->dtype != MTJG_F64)
    fail(MTJG_E_INVALID_ARGUMENT, "Vector must be DenseVectors");
  if (b->n != M->n || x->n != M->n) fail(MTJG_E_INDEX_OUT_OF_BOUNDS, "vector size differs from preconditioner size");
  if ((M->kind == PREC_ILU || M->kind == PREC_SSOR) && b->ptr == x->ptr && M->n > 0)
    fail(MTJG_E_INVALID_ARGUMENT, "b and x must not alias");
}

extern "C" int mtjg_prec_apply(mtjg_prec M, mtjg_vec b, mtjg_vec x) {
  MTJG_API_BEGIN
  check_apply(M, b, x);
  prec_apply_dev(M, M->n, (const double *)b->ptr, (double *)x->ptr, M->ctx->stream, nullptr);
  prec_check(M);
  MTJG_API_END
}

extern "C" int mtjg_prec_trans_apply(mtjg_prec M, mtjg_vec b, mtjg_vec x) {
  MTJG_API_BEGIN
  check_apply(M, b, x);
  prec_trans_apply_dev(M, (const double *)b->ptr, (double *)x->ptr, M->ctx->stream);
  prec_check(M);
  MTJG_API_END
}

extern "C" int mtjg_prec_destroy(mtjg_prec M) {
  MTJG_API_BEGIN
  if (!M) return MTJG_OK;
  HIPCHK(hipStreamSynchronize(M->ctx->stream));
  (void)hipFree(M->invdiag);
  (void)hipFree(M->bad);
  (void)hipFree(M->y);
  (void)hipFree(M->sf_next);
  (void)hipFree(M->sf_err);
  (void)hipFree(M->xx);
  amg_destroy_device(M->amg);
  if (M->inner) cg32_destroy(M->inner);
  ilu_free_storage(M);
  delete M;
  MTJG_API_END
}
