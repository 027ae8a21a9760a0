// CompRowMatrix on the device and its CSR products (S/CompRowMatrix.java).
//
// Two device formats, both summing rows in MTJ's storage order:
//  * SELL-64 (chosen when rows in each group of 64 have similar lengths, the
//    padding being at most a quarter of the entries): a wavefront per slice
//    of 64 rows, a lane per row, entries column-major within the slice so
//    every load is coalesced and no LDS or barrier is needed; every row is
//    summed sequentially, so every product is bitwise MTJ's;
//  * row-block CSR (any pattern): one 256-thread workgroup per host-cut
//    row-block (spmv_rowblock.hpp), bitwise MTJ's for blocks of >= 128 rows;
//  * for irregular patterns whose columns scatter over an x larger than one
//    XCD's L2, XCD column blocks in front of the row-block kernel: a first
//    pass forms every product with its x gather confined to one XCD's slice
//    of x, and the row-block kernel then sums those products (cb_build).
// The kernels are HBM-bound (12 bytes per stored entry, 20 per row), so MFMA
// is not used.
#include <algorithm>
#include <cstring>
#include <numeric>
#include <type_traits>

#include "mtjg_internal.hpp"
#include "reduce.hpp"
#include "spmv_rowblock.hpp"

namespace mtjg {

constexpr int SPMV_DOT_GRID = 2048;  // workgroups of the fused product

// Order of the row-blocks over the XCDs.  The plain products take runs of 16
// consecutive row-blocks per XCD (xcd_run_block; -16 here): on the 256^3
// Poisson product 5.52 TB/s against 5.21 in plain row order, 5.35 / 5.42 /
// 5.44 with runs of 8 / 32 / 64, and 4.99 with 16 contiguous slabs.  The
// fused product walks its tickets grid-stride in plain row order.
// MTJG_SPMV_SLABS overrides the choice for experiments (1-32: slabs, -R:
// runs of R).
static int choose_slabs(int32_t nrb) {
  (void)nrb;
  if (const char *e = getenv("MTJG_SPMV_SLABS")) {
    int v = atoi(e);
    if (v == 1 || v == 2 || v == 4 || v == 8 || v == 16 || v == 32) return v;
    if (v == -4 || v == -8 || v == -16 || v == -32 || v == -64 || v == -256) return v;
  }
  return -16;
}

// Whether the plain products store y with the non-temporal hint: when y is
// at least NT_VECTOR_BYTES (MTJG_SPMV_NT=0|1 forces it for experiments).
static bool spmv_nt_store(int64_t rows, size_t elem) {
  if (const char *e = getenv("MTJG_SPMV_NT")) return atoi(e) != 0;
  return rows * (int64_t)elem >= NT_VECTOR_BYTES;
}

// Scalar epilogues of fused solver kernels: run by the last workgroup.
__device__ __forceinline__ void spmv_finish(KState *st, int fin, double s0, double s1) {
  if (fin == FIN_CG_ALPHA) {
    st->alpha = st->rho / s0;  // S/CG.java:78
    st->rho_1 = st->rho;       // S/CG.java:83
  } else if (fin == FIN_BICG_ALPHA) {
    st->alpha = st->rho_1 / s0;  // S/BiCGstab.java:99
  } else if (fin == FIN_BICG_OMEGA) {
    st->s[0] = s0;
    st->s[1] = s1;
  }
}

// Maps a launch-order ticket to a row-block.  Workgroups are dispatched
// round-robin over the XCDs, so with `slabs` = 8 the tickets of XCD k
// (t % 8 == k) sweep the k-th contiguous eighth of the row-blocks in order:
// the x entries a stencil row gathers from neighbouring planes were fetched
// into that XCD's own L2 by the rows just before it.  slabs = 1 keeps plain
// row order.  Returns -1 for tickets past the end of a slab.
__device__ __forceinline__ int ticket_to_block(int t, int nrb, int slabs, int per) {
  const int k = t % slabs, l = t / slabs;
  const int b = k * per + l;
  return (l < per && b < nrb) ? b : -1;
}

// Maps workgroup `blk` of a grid dealt round-robin over the eight XCDs so
// that each XCD takes runs of `run` consecutive row-blocks, the runs rotating
// over the XCDs (the tail past the last whole round keeps plain order): the
// x entries a row gathers from neighbouring rows were mostly fetched into the
// same XCD's L2 by the rows just before it.
__device__ __forceinline__ int xcd_run_block(int blk, int grid, int run) {
  const int full = (grid / (8 * run)) * (8 * run);
  if (blk < full) {
    const int xcd = blk % 8, j = blk / 8;
    blk = ((j / run) * 8 + xcd) * run + j % run;
  }
  return blk;
}

// Plain products: one workgroup per row-block, under the XCD run mapping
// (slabs = -run) or per slab-ordered ticket (slabs > 0).  NTY: y is stored
// with the non-temporal hint.
template <typename VT, typename XT, int MODE, bool NTY, bool PRE>
__global__ __launch_bounds__(BS) void k_spmv(int nrb, int slabs, const int2 *__restrict__ info,
                                             const int32_t *__restrict__ rp,
                                             const int32_t *__restrict__ ci,
                                             const VT *__restrict__ val, const XT *__restrict__ x,
                                             XT *y, const XT *y0, double alpha, double inv_alpha, CbView cb) {
  __shared__ XT prod[SPMV_CAP];
  int b;
  if (slabs < 0) {
    b = xcd_run_block(blockIdx.x, gridDim.x, -slabs);
    if (b >= nrb) return;
  } else {
    const int per = (nrb + slabs - 1) / slabs;
    b = ticket_to_block(blockIdx.x, nrb, slabs, per);
    if (b < 0) return;
  }
  dd acc0{0.0, 0.0}, acc1{0.0, 0.0};
  spmv_rowblock<VT, XT, MODE, NTY, PRE>(b, prod, info, rp, ci, val, x, y, y0, alpha, inv_alpha, nullptr, acc0,
                                        acc1, cb);
}

// Fused solver product q = A p with the grid sums w.q and q.q: at most
// SPMV_DOT_GRID co-resident workgroups walk the row-block tickets, so only
// that many take a reduction ticket; a two-level deterministic reduction
// ends in the solver epilogue in the last workgroup.  Every workgroup returns
// at once when the solve has left the running state.
template <typename VT, typename XT, bool PRE>
__global__ __launch_bounds__(BS) void k_spmv_dot(int nrb, int slabs, const int2 *__restrict__ info,
                                                 const int32_t *__restrict__ rp,
                                                 const int32_t *__restrict__ ci,
                                                 const VT *__restrict__ val, const XT *__restrict__ x,
                                                 XT *y, const XT *w, KState *st, int fin, GridRed red, CbView cb) {
  if (st->status != ST_RUNNING) return;
  __shared__ XT prod[SPMV_CAP];
  dd acc0{0.0, 0.0}, acc1{0.0, 0.0};
  const int per = (nrb + slabs - 1) / slabs;
  const int ntickets = per * slabs;
  for (int t = blockIdx.x; t < ntickets; t += gridDim.x) {
    const int b = ticket_to_block(t, nrb, slabs, per);
    if (b >= 0)
      spmv_rowblock<VT, XT, SPMV_DOT, false, PRE>(b, prod, info, rp, ci, val, x, y, nullptr, 1.0, 1.0, w, acc0,
                                                  acc1, cb);
    __syncthreads();  // prod is reused by the next row-block
  }
  dd v[2] = {acc0, acc1};
  double tot[2];
  if (grid_sum_dd<2>(v, red, tot) && threadIdx.x == 0) spmv_finish(st, fin, tot[0], tot[1]);
}

// ---- XCD column blocks ----------------------------------------------------
//
// On a matrix whose columns scatter at random over a vector of 32 MB (the
// 4 M-row power-law graph), nearly every x gather of the row-block kernel
// misses L2 and fills a 128-byte line for 8 useful bytes (L2 hit rate 13%,
// 2.0 GB of fills per product against 0.31 GB algorithmic).  Random 8-byte
// gathers from 32 MB run at 63 G/s, against 182 G/s when the gathered slice
// is 4 MB, the size of one XCD's L2.  So the products are formed in a pass of
// their own: the entries are kept a second time ordered by column block
// (CB_BLOCK_BYTES of x per block), the blocks dealt to the XCDs (block b to
// XCD b % 8, which runs workgroups w with w % 8 == b % 8), so each XCD
// gathers only from its own slices of x, which stay in its L2.  On the
// power-law graph (products pass + sums pass): 4 MB blocks and 8 entries per
// thread 163 + 95 us, 16 entries 147 + 94 us (32: 152 us), 2 MB blocks with
// 16 entries 133 + 98 us (1 MB: 130 + 110 us), against 298 us for the
// row-block kernel alone.  Within a
// block the entries keep CSR order, and the products are stored with the
// non-temporal hint so that they do not evict the slice.  The row-block
// kernel then stages each row-block's products segment by segment (PRE,
// spmv_rowblock.hpp) and sums the rows exactly as it sums products it forms
// itself: the result is bitwise the row-block product's.
constexpr int64_t CB_BLOCK_BYTES = 2ll << 20;  // x per column block: half of one XCD's L2
constexpr int CB_PER = 16;                      // entries per thread of the products pass
constexpr int CB_CHUNK = BS * CB_PER;           // entries per workgroup
constexpr int CB_NXCD = 8;

struct CbRanges {
  int32_t off[CB_NXCD + 1];  // XCD k's entries: [off[k], off[k+1])
};

// prod[k] = val[k] * x[col[k]] over the reordered entries: workgroup w takes
// the (w / 8)-th chunk of XCD (w % 8)'s range.  The entries are read once
// (non-temporal loads, so they do not evict the XCD's slice of x).
__global__ __launch_bounds__(BS) void k_cb_prod(CbRanges r, const int32_t *__restrict__ col,
                                                const double *__restrict__ val, const double *__restrict__ x,
                                                double *__restrict__ prod) {
  const int xcd = blockIdx.x % CB_NXCD;
  const int64_t k0 = r.off[xcd] + (int64_t)(blockIdx.x / CB_NXCD) * CB_CHUNK;
  const int64_t end = r.off[xcd + 1];
  if (k0 >= end) return;
  int c[CB_PER];
  double v[CB_PER], xv[CB_PER];
#pragma unroll
  for (int u = 0; u < CB_PER; ++u) {
    const int64_t k = k0 + threadIdx.x + u * BS;
    if (k < end) {
      c[u] = __builtin_nontemporal_load(&col[k]);
      v[u] = __builtin_nontemporal_load(&val[k]);
    }
  }
#pragma unroll
  for (int u = 0; u < CB_PER; ++u) {
    const int64_t k = k0 + threadIdx.x + u * BS;
    if (k < end) xv[u] = x[c[u]];
  }
#pragma unroll
  for (int u = 0; u < CB_PER; ++u) {
    const int64_t k = k0 + threadIdx.x + u * BS;
    if (k < end) __builtin_nontemporal_store(v[u] * xv[u], &prod[k]);
  }
}

// Copies CSR entry j's value (and, with col, its column) to its position
// perm[j] in the column-block order.
__global__ void k_cb_fill(int64_t nnz, const int32_t *__restrict__ perm, const int32_t *__restrict__ ci,
                          const double *__restrict__ val, int32_t *col, double *cval) {
  for (int64_t j = blockIdx.x * (int64_t)BS + threadIdx.x; j < nnz; j += (int64_t)gridDim.x * BS) {
    const int32_t p = perm[j];
    cval[p] = val[j];
    if (col) col[p] = ci[j];
  }
}

// Products from the SELL-64 copy: four slices (wavefronts) per workgroup.
constexpr int SELL_XCD_RUN = 64;

template <typename VT, typename XT, int MODE, bool NTY>
__global__ __launch_bounds__(BS) void k_sell(int nslice, int n, const int32_t *__restrict__ sptr,
                                             const int32_t *__restrict__ scol, const VT *__restrict__ sval,
                                             const XT *__restrict__ x, XT *y, const XT *y0, double alpha,
                                             double inv_alpha) {
  // Workgroups are dealt to the eight XCDs round-robin (workgroup w runs on
  // XCD w % 8).  Remap so that each XCD takes runs of SELL_XCD_RUN
  // consecutive row blocks, the runs rotating over the XCDs; the tail that
  // does not fill a whole rotation keeps its place.  Measured on the 256^3
  // Poisson product: 5.93 TB/s against 5.82 TB/s round-robin (runs of 2:
  // 5.91, of 4: 5.84, of 16-128: 5.87-5.88; one contiguous range per XCD:
  // 5.29).
  const int blk = xcd_run_block(blockIdx.x, gridDim.x, SELL_XCD_RUN);
  const int slice = blk * (BS / SELL_C) + (int)threadIdx.x / SELL_C;
  if (slice >= nslice) return;
  dd acc0{0.0, 0.0}, acc1{0.0, 0.0};
  sell_row<VT, XT, MODE, NTY>(slice, threadIdx.x % SELL_C, n, sptr, scol, sval, x, y, y0, alpha, inv_alpha,
                              nullptr, acc0, acc1);
}

// Fused solver product from the SELL-64 copy (see k_spmv_dot) for grids of
// at most SPMV_DOT_GRID workgroups (larger ones use k_sell_part + k_sell_fin):
// the workgroups walk the slices in grid-stride order.
template <typename VT, typename XT>
__global__ __launch_bounds__(BS) void k_sell_dot(int nslice, int n, const int32_t *__restrict__ sptr,
                                                 const int32_t *__restrict__ scol, const VT *__restrict__ sval,
                                                 const XT *__restrict__ x, XT *y, const XT *w, KState *st,
                                                 int fin, GridRed red) {
  if (st->status != ST_RUNNING) return;
  dd acc0{0.0, 0.0}, acc1{0.0, 0.0};
  constexpr int WPB = BS / SELL_C;
  // (the XCD run mapping of k_sell measured 1.4% slower on this grid-stride walk)
  for (int slice = blockIdx.x * WPB + (int)threadIdx.x / SELL_C; slice < nslice; slice += gridDim.x * WPB)
    sell_row<VT, XT, SPMV_DOT>(slice, threadIdx.x % SELL_C, n, sptr, scol, sval, x, y, nullptr, 1.0, 1.0, w,
                               acc0, acc1);
  dd v[2] = {acc0, acc1};
  double tot[2];
  if (grid_sum_dd<2>(v, red, tot) && threadIdx.x == 0) spmv_finish(st, fin, tot[0], tot[1]);
}

// Split form of k_sell_dot for grids above SPMV_DOT_GRID: the product runs
// k_sell's grid (one workgroup per four slices, XCD run mapping) and each
// workgroup stores its two compensated partial sums, without a ticket, to
// red.partials[j * red.cap + off + blockIdx.x]; k_sell_fin then adds them.
// A per-workgroup ticket waits on a device-scope atomic round trip while the
// workgroup holds its CU slot, which is what made the ticketed full grid slow.
template <typename VT, typename XT>
__global__ __launch_bounds__(BS) void k_sell_part(int nslice, int n, const int32_t *__restrict__ sptr,
                                                  const int32_t *__restrict__ scol, const VT *__restrict__ sval,
                                                  const XT *__restrict__ x, XT *y, const XT *w, const KState *st,
                                                  double *part, int cap) {
  if (st->status != ST_RUNNING) return;
  const int blk = xcd_run_block(blockIdx.x, gridDim.x, SELL_XCD_RUN);
  const int slice = blk * (BS / SELL_C) + (int)threadIdx.x / SELL_C;
  dd v[2] = {{0.0, 0.0}, {0.0, 0.0}};
  if (slice < nslice)
    sell_row<VT, XT, SPMV_DOT>(slice, threadIdx.x % SELL_C, n, sptr, scol, sval, x, y, nullptr, 1.0, 1.0, w, v[0],
                               v[1]);
  block_sum_dd<2>(v);
  if (threadIdx.x == 0) {
    part[0 * (size_t)cap + blockIdx.x] = v[0].hi;
    part[1 * (size_t)cap + blockIdx.x] = v[0].lo;
    part[2 * (size_t)cap + blockIdx.x] = v[1].hi;
    part[3 * (size_t)cap + blockIdx.x] = v[1].lo;
  }
}

// Adds k_sell_part's np partials (each thread those at its index and every
// gridDim.x * BS after it, in order; then the grid reduction, whose block
// partials sit below `part`) and runs the epilogue.  The grid is at most one
// reduction group (sell_fin_grid), so the sum takes one ticket round trip.
__global__ __launch_bounds__(BS) void k_sell_fin(int np, const double *part, KState *st, int fin, GridRed red) {
  if (st->status != ST_RUNNING) return;
  dd v[2] = {{0.0, 0.0}, {0.0, 0.0}};
  for (int i = blockIdx.x * BS + (int)threadIdx.x; i < np; i += gridDim.x * BS) {
    v[0] = dd_add(v[0], dd{part[0 * (size_t)red.cap + i], part[1 * (size_t)red.cap + i]});
    v[1] = dd_add(v[1], dd{part[2 * (size_t)red.cap + i], part[3 * (size_t)red.cap + i]});
  }
  double tot[2];
  if (grid_sum_dd<2>(v, red, tot) && threadIdx.x == 0) spmv_finish(st, fin, tot[0], tot[1]);
}

// ---- FP32 stencil copy --------------------------------------------------------
//
// The mixed-precision inner solves stream the FP32 matrix once per iteration;
// in SELL-64 an entry costs a 4-byte value and a 4-byte column.  When every
// entry lies on one of at most ST_MAX_OFF diagonals (stencil matrices:
// column - row in {0, +-1, +-nx, +-nx*ny}), row i's columns are i + off[k]
// for the diagonals its mask byte names, so the FP32 copy keeps only the
// values, diagonal-major (a lane per row, each diagonal's values of 64
// consecutive rows one contiguous 256-byte line, as in SELL-64), and the
// gathers x[i + off[k]] of neighbouring lanes are contiguous.  A row's
// products are added in diagonal order, which is ascending column order,
// the order of the SELL-64 row and of MTJ's loop: every sum is bitwise the
// SELL-64 FP32 product's.
struct StOff {
  int32_t off[ST_MAX_OFF];
};

// Split fused product (k_sell_part's protocol): a workgroup per 256 rows
// under the XCD run mapping, compensated partials stored without a ticket.
__global__ __launch_bounds__(BS) void k_st_part(int n, int m, StOff o, const float *__restrict__ val,
                                                const uint8_t *__restrict__ mask, const float *__restrict__ x,
                                                float *y, const float *w, const KState *st, double *part, int cap) {
  if (st->status != ST_RUNNING) return;
  const int blk = xcd_run_block(blockIdx.x, gridDim.x, SELL_XCD_RUN);
  dd v[2] = {{0.0, 0.0}, {0.0, 0.0}};
  st_row<SPMV_DOT>(blk * BS + (int)threadIdx.x, n, m, o.off, val, mask, x, y, w, v[0], v[1]);
  block_sum_dd<2>(v);
  if (threadIdx.x == 0) {
    part[0 * (size_t)cap + blockIdx.x] = v[0].hi;
    part[1 * (size_t)cap + blockIdx.x] = v[0].lo;
    part[2 * (size_t)cap + blockIdx.x] = v[1].hi;
    part[3 * (size_t)cap + blockIdx.x] = v[1].lo;
  }
}

// One-kernel fused product for grids of at most SPMV_DOT_GRID workgroups.
__global__ __launch_bounds__(BS) void k_st_dot(int n, int m, StOff o, const float *__restrict__ val,
                                               const uint8_t *__restrict__ mask, const float *__restrict__ x,
                                               float *y, const float *w, KState *st, int fin, GridRed red) {
  if (st->status != ST_RUNNING) return;
  dd acc0{0.0, 0.0}, acc1{0.0, 0.0};
  for (int row = blockIdx.x * BS + (int)threadIdx.x; row < n; row += gridDim.x * BS)
    st_row<SPMV_DOT>(row, n, m, o.off, val, mask, x, y, w, acc0, acc1);
  dd v[2] = {acc0, acc1};
  double tot[2];
  if (grid_sum_dd<2>(v, red, tot) && threadIdx.x == 0) spmv_finish(st, fin, tot[0], tot[1]);
}

// Fills the stencil copy from the CSR values: a thread per row, each entry
// on the diagonal of its offset (as (float)value, the SELL-64 FP32 copy's
// conversion), the other diagonals of the row 0 with their mask bit clear.
__global__ void k_st_fill32(int32_t n, int m, StOff o, const int32_t *__restrict__ rp,
                            const int32_t *__restrict__ ci, const double *__restrict__ val, float *sv,
                            uint8_t *mask) {
  const int32_t row = blockIdx.x * BS + threadIdx.x;
  if (row >= n) return;
  float out[ST_MAX_OFF];
  uint32_t mk = 0;
#pragma unroll
  for (int k = 0; k < ST_MAX_OFF; ++k) out[k] = 0.0f;
  for (int32_t j = rp[row]; j < rp[row + 1]; ++j) {
    const int32_t d = ci[j] - row;
#pragma unroll
    for (int k = 0; k < ST_MAX_OFF; ++k)
      if (k < m && o.off[k] == d) {
        out[k] = (float)val[j];
        mk |= 1u << k;
      }
  }
#pragma unroll
  for (int k = 0; k < ST_MAX_OFF; ++k)
    if (k < m) sv[(size_t)k * n + row] = out[k];
  mask[row] = (uint8_t)mk;
}

// CG's direction update folded into its fused product (S/CG.java:71-79):
// p' = z on the first iteration and p' = (p * beta) + z after it, formed at
// every gather from p and z (the operations of the separate update, so p' is
// bitwise the same), q = A p' summed as sell_row sums it, p' of the row
// stored to pn (a buffer other than p, which other rows still gather), and
// the sums p'.q and q.q accumulated as the fused product's epilogue does.
template <typename VT>
__device__ __forceinline__ void sell_row_cgp(int slice, int lane, int n, const int32_t *__restrict__ sptr,
                                             const int32_t *__restrict__ scol, const VT *__restrict__ sval,
                                             const double *__restrict__ p, const double *__restrict__ z,
                                             double *pn, double *q, bool first, double beta, dd &acc0,
                                             dd &acc1) {
  const int row = slice * SELL_C + lane;
  const int off = sptr[slice];
  const int width = (sptr[slice + 1] - off) / SELL_C;
  double s = 0;
  for (int k0 = 0; k0 < width; k0 += SELL_UNROLL) {
    int c[SELL_UNROLL];
    VT v[SELL_UNROLL];
    double xv[SELL_UNROLL];
#pragma unroll
    for (int u = 0; u < SELL_UNROLL; ++u) {
      c[u] = -1;
      if (k0 + u < width) {
        const int e = off + (k0 + u) * SELL_C + lane;
        c[u] = __builtin_nontemporal_load(&scol[e]);
        v[u] = __builtin_nontemporal_load(&sval[e]);
      }
    }
#pragma unroll
    for (int u = 0; u < SELL_UNROLL; ++u)
      if (c[u] >= 0) {
        const double zc = z[c[u]];
        xv[u] = first ? zc : (p[c[u]] * beta) + zc;
      }
#pragma unroll
    for (int u = 0; u < SELL_UNROLL; ++u)
      if (c[u] >= 0) s += (double)v[u] * xv[u];
  }
  if (row < n) {
    const double pr = first ? z[row] : (p[row] * beta) + z[row];
    pn[row] = pr;
    const double vq = (s != 0.0) ? s : 0.0;
    q[row] = vq;
    dd_fma(acc0, pr, vq);
    dd_fma(acc1, vq, vq);
  }
}

// k_sell_dot with the folded update.
template <typename VT>
__global__ __launch_bounds__(BS) void k_sell_dot_cgp(int nslice, int n, const int32_t *__restrict__ sptr,
                                                     const int32_t *__restrict__ scol, const VT *__restrict__ sval,
                                                     const double *p, const double *z, double *pn, double *q,
                                                     KState *st, GridRed red) {
  if (st->status != ST_RUNNING) return;
  const bool first = st->iter == 0;
  const double beta = st->beta;
  dd acc0{0.0, 0.0}, acc1{0.0, 0.0};
  constexpr int WPB = BS / SELL_C;
  for (int slice = blockIdx.x * WPB + (int)threadIdx.x / SELL_C; slice < nslice; slice += gridDim.x * WPB)
    sell_row_cgp<VT>(slice, threadIdx.x % SELL_C, n, sptr, scol, sval, p, z, pn, q, first, beta, acc0, acc1);
  dd v[2] = {acc0, acc1};
  double tot[2];
  if (grid_sum_dd<2>(v, red, tot) && threadIdx.x == 0) spmv_finish(st, FIN_CG_ALPHA, tot[0], tot[1]);
}

// Copies each row's values (and, with scol, its columns) from the CSR
// arrays into its lane of the SELL-64 copy (one thread per row).
__global__ void k_sell_fill(int32_t nr, const int32_t *rp, const int32_t *ci, const int32_t *sptr,
                            const double *val, int32_t *scol, double *sval) {
  const int32_t row = blockIdx.x * BS + threadIdx.x;
  if (row >= nr) return;
  const int32_t off = sptr[row / SELL_C] + row % SELL_C;
  for (int32_t k = 0, j = rp[row]; j < rp[row + 1]; ++k, ++j) {
    sval[off + k * SELL_C] = val[j];
    if (scol) scol[off + k * SELL_C] = ci[j];
  }
}

__global__ void k_to_f32(int64_t n, const double *a, float *b) {
  for (int64_t i = blockIdx.x * (int64_t)BS + threadIdx.x; i < n; i += (int64_t)gridDim.x * BS)
    b[i] = (float)a[i];
}

// Brings the SELL-64 values (and their FP32 copy when `fp32`) up to date.
static void sell_ready(mtjg_csr A, bool fp32, hipStream_t s) {
  if (A->sell_stale) {
    k_sell_fill<<<(A->nr + BS - 1) / BS, BS, 0, s>>>(A->nr, A->rp, A->ci, A->sptr, A->val, nullptr, A->sval);
    HIPCHK(hipGetLastError());
    A->sell_stale = false;
  }
  if (fp32 && !A->sval32) {
    const int64_t total = A->h_sptr_total;
    HIPCHK(hipMalloc(&A->sval32, sizeof(float) * (size_t)std::max<int64_t>(total, 1)));
    if (total) k_to_f32<<<grid_for(total), BS, 0, s>>>(total, A->sval, A->sval32);
    HIPCHK(hipGetLastError());
  }
}

// Workgroups of k_sell_fin for a k_sell_part grid.
static int sell_fin_grid(int grid) { return std::min(RED_GROUP, (grid + BS - 1) / BS); }

int64_t spmv_dot_red_cap(mtjg_csr A) {
  int64_t need = SPMV_DOT_GRID;
  if (A->sell) {
    const int grid = (A->nslice + BS / SELL_C - 1) / (BS / SELL_C);
    if (grid > SPMV_DOT_GRID) need = std::max<int64_t>(need, (int64_t)sell_fin_grid(grid) + grid);
  }
  return need;
}

template <typename VT, typename XT>
static void sell_launch(mtjg_csr A, const VT *sval, const XT *x, XT *y, const SpmvArgs &a, hipStream_t s) {
  const int grid = (A->nslice + BS / SELL_C - 1) / (BS / SELL_C);
  const XT *y0 = (const XT *)a.y0;
  const bool nt = spmv_nt_store(A->nr, sizeof(XT));
  auto plain = [&](auto mode, auto nty) {
    k_sell<VT, XT, decltype(mode)::value, decltype(nty)::value><<<grid, BS, 0, s>>>(
        A->nslice, A->nr, A->sptr, A->scol, sval, x, y, y0, a.alpha, a.inv_alpha);
  };
  switch (a.mode) {
    case SPMV_MULT:
      if (nt) plain(std::integral_constant<int, SPMV_MULT>{}, std::true_type{});
      else plain(std::integral_constant<int, SPMV_MULT>{}, std::false_type{});
      break;
    case SPMV_MULTADD:
      if (nt) plain(std::integral_constant<int, SPMV_MULTADD>{}, std::true_type{});
      else plain(std::integral_constant<int, SPMV_MULTADD>{}, std::false_type{});
      break;
    case SPMV_TMULTADD:
      if (nt) plain(std::integral_constant<int, SPMV_TMULTADD>{}, std::true_type{});
      else plain(std::integral_constant<int, SPMV_TMULTADD>{}, std::false_type{});
      break;
    case SPMV_DOT: {
      if (grid > SPMV_DOT_GRID) {  // 256^3 Poisson CG: 0.310 -> 0.288 s
        const int f = sell_fin_grid(grid);
        if ((int64_t)f + grid > a.red.cap) fail(MTJG_E_INVALID_ARGUMENT, "reduction workspace smaller than the grid");
        double *part = a.red.partials + f;
        k_sell_part<VT, XT><<<grid, BS, 0, s>>>(A->nslice, A->nr, A->sptr, A->scol, sval, x, y, (const XT *)a.w,
                                                a.st, part, a.red.cap);
        HIPCHK(hipGetLastError());
        k_sell_fin<<<f, BS, 0, s>>>(grid, part, a.st, a.fin, a.red);
        break;
      }
      const int g = std::min(grid, SPMV_DOT_GRID);
      if (g > a.red.cap) fail(MTJG_E_INVALID_ARGUMENT, "reduction workspace smaller than the grid");
      k_sell_dot<VT, XT><<<g, BS, 0, s>>>(A->nslice, A->nr, A->sptr, A->scol, sval, x, y, (const XT *)a.w, a.st,
                                          a.fin, a.red);
      break;
    }
    default:
      fail(MTJG_E_INVALID_ARGUMENT, "bad spmv mode");
  }
  HIPCHK(hipGetLastError());
}

// With PRE the products come from the column-block products pass (cb); val
// and x are not read.
template <typename VT, typename XT, bool PRE = false>
static void spmv_launch(mtjg_csr A, const VT *val, const XT *x, XT *y, const SpmvArgs &a,
                        hipStream_t s, const CbView cb = CbView{}) {
  const int32_t *ci = A->ci;
  if (A->nr == 0) {
    if (a.mode == SPMV_DOT) fail(MTJG_E_INVALID_ARGUMENT, "empty matrix in fused product");
    return;
  }
  // plain products: one workgroup per row-block (XCD run mapping) or per
  // slab-ordered ticket; the fused product walks plain-order tickets
  // grid-stride on at most SPMV_DOT_GRID workgroups
  const int sl = a.mode == SPMV_DOT ? std::max(A->slabs, 1) : A->slabs;
  const int per = sl > 0 ? (A->nrb + sl - 1) / sl : A->nrb;
  const int grid = (a.mode == SPMV_DOT) ? std::min(per * sl, SPMV_DOT_GRID) : sl > 0 ? per * sl : A->nrb;
  const XT *y0 = (const XT *)a.y0;
  const XT *w = (const XT *)a.w;
  const bool nt = spmv_nt_store(A->nr, sizeof(XT));
  auto plain = [&](auto mode, auto nty) {
    k_spmv<VT, XT, decltype(mode)::value, decltype(nty)::value, PRE><<<grid, BS, 0, s>>>(
        A->nrb, sl, A->rb, A->rp, ci, val, x, y, y0, a.alpha, a.inv_alpha, cb);
  };
  switch (a.mode) {
    case SPMV_MULT:
      if (nt) plain(std::integral_constant<int, SPMV_MULT>{}, std::true_type{});
      else plain(std::integral_constant<int, SPMV_MULT>{}, std::false_type{});
      break;
    case SPMV_MULTADD:
      if (nt) plain(std::integral_constant<int, SPMV_MULTADD>{}, std::true_type{});
      else plain(std::integral_constant<int, SPMV_MULTADD>{}, std::false_type{});
      break;
    case SPMV_TMULTADD:
      if (nt) plain(std::integral_constant<int, SPMV_TMULTADD>{}, std::true_type{});
      else plain(std::integral_constant<int, SPMV_TMULTADD>{}, std::false_type{});
      break;
    case SPMV_DOT:
      if (grid > a.red.cap) fail(MTJG_E_INVALID_ARGUMENT, "reduction workspace smaller than the grid");
      k_spmv_dot<VT, XT, PRE><<<grid, BS, 0, s>>>(A->nrb, sl, A->rb, A->rp, ci, val, x, y, w, a.st,
                                                  a.fin, a.red, cb);
      break;
    default:
      fail(MTJG_E_INVALID_ARGUMENT, "bad spmv mode");
  }
  HIPCHK(hipGetLastError());
}

// The column-block products pass (k_cb_prod) into A->cb_prod, the values
// refreshed first if they changed.  Products of one matrix share cb_prod, so
// they are ordered on one stream.
static void cb_products(mtjg_csr A, const double *x, hipStream_t s) {
  if (A->cb_stale) {
    k_cb_fill<<<grid_for(A->nnz), BS, 0, s>>>(A->nnz, A->cb_perm, A->ci, A->val, nullptr, A->cb_val);
    HIPCHK(hipGetLastError());
    A->cb_stale = false;
  }
  CbRanges r;
  int64_t chunks = 0;
  for (int k = 0; k <= CB_NXCD; ++k) r.off[k] = A->cb_off[k];
  for (int k = 0; k < CB_NXCD; ++k)
    chunks = std::max<int64_t>(chunks, (r.off[k + 1] - r.off[k] + CB_CHUNK - 1) / CB_CHUNK);
  if (chunks == 0) return;
  k_cb_prod<<<(int)(chunks * CB_NXCD), BS, 0, s>>>(r, A->cb_col, A->cb_val, x, A->cb_prod);
  HIPCHK(hipGetLastError());
}

void spmv_f64(mtjg_csr A, const double *x, double *y, const SpmvArgs &a, hipStream_t s) {
  if (A->sell && A->nr > 0) {
    sell_ready(A, false, s);
    sell_launch<double, double>(A, A->sval, x, y, a, s);
    return;
  }
  if (A->cb && A->nr > 0) {
    cb_products(A, x, s);
    CbView cb;
    cb.prod = A->cb_prod;
    cb.seg = A->cb_seg;
    cb.q = A->cb_q;
    cb.nblk = A->cb_nblk;
    spmv_launch<double, double, true>(A, nullptr, x, y, a, s, cb);
    return;
  }
  spmv_launch<double, double>(A, A->val, x, y, a, s);
}

// Larger grids keep the separate update: gathering p and z at every entry
// made the 256^3 Poisson CG 0.313 s against 0.289 s, while the one-kernel
// form saves a launch per iteration on latency-bound systems (2-D Poisson
// 256^2 CG 0.0112 -> 0.0104 s).
bool spmv_cg_fusable(mtjg_csr A) {
  return A->sell && A->nr > 0 && (A->nslice + BS / SELL_C - 1) / (BS / SELL_C) <= SPMV_DOT_GRID;
}

bool spmv_cg_fused(mtjg_csr A, const double *p, const double *z, double *pn, double *q, const SpmvArgs &a,
                   hipStream_t s) {
  if (!spmv_cg_fusable(A)) return false;
  const int grid = (A->nslice + BS / SELL_C - 1) / (BS / SELL_C);
  sell_ready(A, false, s);
  if (grid > a.red.cap) fail(MTJG_E_INVALID_ARGUMENT, "reduction workspace smaller than the grid");
  k_sell_dot_cgp<double><<<grid, BS, 0, s>>>(A->nslice, A->nr, A->sptr, A->scol, A->sval, p, z, pn, q, a.st,
                                             a.red);
  HIPCHK(hipGetLastError());
  return true;
}

// The FP32 stencil copy's fused product (SPMV_DOT), filled first if the
// values changed.
static bool st_use(mtjg_csr A, const SpmvArgs &a) {
  return A->st_m > 0 && A->sell && A->nr > 0 && a.mode == SPMV_DOT;
}

static void st_launch(mtjg_csr A, const float *x, float *y, const SpmvArgs &a, hipStream_t s) {
  StOff o;
  for (int k = 0; k < ST_MAX_OFF; ++k) o.off[k] = k < A->st_m ? A->st_off[k] : 0;
  if (!A->st_val32) {
    HIPCHK(hipMalloc(&A->st_val32, sizeof(float) * (size_t)A->st_m * (size_t)A->nr));
    HIPCHK(hipMalloc(&A->st_mask, (size_t)A->nr));
    A->st_stale = true;
  }
  if (A->st_stale) {
    k_st_fill32<<<(A->nr + BS - 1) / BS, BS, 0, s>>>(A->nr, A->st_m, o, A->rp, A->ci, A->val, A->st_val32,
                                                    A->st_mask);
    HIPCHK(hipGetLastError());
    A->st_stale = false;
  }
  const int grid = (A->nr + BS - 1) / BS;  // equals the SELL-64 product's grid (four slices per workgroup)
  if (grid > SPMV_DOT_GRID) {
    const int f = sell_fin_grid(grid);
    if ((int64_t)f + grid > a.red.cap) fail(MTJG_E_INVALID_ARGUMENT, "reduction workspace smaller than the grid");
    double *part = a.red.partials + f;
    k_st_part<<<grid, BS, 0, s>>>(A->nr, A->st_m, o, A->st_val32, A->st_mask, x, y, (const float *)a.w, a.st,
                                  part, a.red.cap);
    HIPCHK(hipGetLastError());
    k_sell_fin<<<f, BS, 0, s>>>(grid, part, a.st, a.fin, a.red);
  } else {
    if (grid > a.red.cap) fail(MTJG_E_INVALID_ARGUMENT, "reduction workspace smaller than the grid");
    k_st_dot<<<grid, BS, 0, s>>>(A->nr, A->st_m, o, A->st_val32, A->st_mask, x, y, (const float *)a.w, a.st,
                                 a.fin, a.red);
  }
  HIPCHK(hipGetLastError());
}

void spmv_f32(mtjg_csr A, const float *x, float *y, const SpmvArgs &a, hipStream_t s) {
  if (st_use(A, a)) {
    st_launch(A, x, y, a, s);
    return;
  }
  if (A->sell && A->nr > 0) {
    sell_ready(A, true, s);
    sell_launch<float, float>(A, A->sval32, x, y, a, s);
    return;
  }
  csr_ensure_fp32(A);
  spmv_launch<float, float>(A, A->val32, x, y, a, s);
}

void csr_ensure_fp32(mtjg_csr A) {
  if (A->val32 || A->nnz == 0) return;
  HIPCHK(hipMalloc(&A->val32, sizeof(float) * (size_t)A->nnz));
  k_to_f32<<<grid_for(A->nnz), BS, 0, A->ctx->stream>>>(A->nnz, A->val, A->val32);
  HIPCHK(hipGetLastError());
}

// ---------------------------------------------------------------------------
// construction

static void csr_free(mtjg_csr_s *A) {
  if (!A) return;
  if (A->trans) csr_free(A->trans);
  (void)hipFree(A->rp);
  (void)hipFree(A->ci);
  (void)hipFree(A->val);
  (void)hipFree(A->val32);
  (void)hipFree(A->rb);
  (void)hipFree(A->sptr);
  (void)hipFree(A->scol);
  (void)hipFree(A->sval);
  (void)hipFree(A->sval32);
  (void)hipFree(A->cb_perm);
  (void)hipFree(A->cb_col);
  (void)hipFree(A->cb_val);
  (void)hipFree(A->cb_prod);
  (void)hipFree(A->cb_seg);
  (void)hipFree(A->cb_q);
  (void)hipFree(A->st_val32);
  (void)hipFree(A->st_mask);
  delete A;
}

// Builds the SELL-64 copy when its padding is at most a quarter of the
// stored entries (stencils and other regular patterns; MTJG_SPMV_FORMAT=csr
// or =sell forces the choice).  Irregular row lengths keep the row-block
// products only.  The slice widths come from the host row pointers; the
// device fills the copy from its own CSR arrays.
static void sell_build(mtjg_csr_s *A) {
  const int32_t nr = A->nr;
  if (nr == 0) return;
  int mode = 0;  // 0 auto, 1 never, 2 always
  if (const char *e = getenv("MTJG_SPMV_FORMAT")) {
    if (!strcmp(e, "csr") || !strcmp(e, "xcd")) mode = 1;
    else if (!strcmp(e, "sell")) mode = 2;
  }
  if (mode == 1) return;
  const int32_t nslice = (nr + SELL_C - 1) / SELL_C;
  std::vector<int32_t> sptr((size_t)nslice + 1, 0);
  int64_t total = 0;
  for (int32_t sl = 0; sl < nslice; ++sl) {
    int32_t w = 0;
    const int32_t r1 = std::min(nr, (sl + 1) * SELL_C);
    for (int32_t r = sl * SELL_C; r < r1; ++r) w = std::max(w, A->h_rp[r + 1] - A->h_rp[r]);
    total += (int64_t)w * SELL_C;
    if (total > INT32_MAX) return;
    sptr[(size_t)sl + 1] = (int32_t)total;
  }
  if (mode == 0 && (double)total > 1.25 * (double)A->nnz + SELL_C) return;
  hipStream_t s = A->ctx->stream;
  const size_t t1 = (size_t)std::max<int64_t>(total, 1);
  HIPCHK(hipMalloc(&A->sptr, sizeof(int32_t) * sptr.size()));
  HIPCHK(hipMalloc(&A->scol, sizeof(int32_t) * t1));
  HIPCHK(hipMalloc(&A->sval, sizeof(double) * t1));
  HIPCHK(hipMemcpyAsync(A->sptr, sptr.data(), sizeof(int32_t) * sptr.size(), hipMemcpyHostToDevice, s));
  HIPCHK(hipMemsetAsync(A->scol, 0xFF, sizeof(int32_t) * t1, s));  // padding column -1
  HIPCHK(hipMemsetAsync(A->sval, 0, sizeof(double) * t1, s));
  k_sell_fill<<<(nr + BS - 1) / BS, BS, 0, s>>>(nr, A->rp, A->ci, A->sptr, A->val, A->scol, A->sval);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(s));
  A->nslice = nslice;
  A->h_sptr_total = total;
  A->sell = true;
}

// Fraction of a row-block's entries that gather a 128-byte line of x that no
// other entry of the block gathers, over up to 512 row-blocks spread over the
// matrix (the first 4096 entries of each): about 1 when the columns scatter
// at random, a few hundredths for stencils and other local patterns.
static double gather_scatter(const int32_t *ci, const std::vector<int2> &rb) {
  const size_t nrb = rb.size() - 1;
  const size_t step = std::max<size_t>(1, nrb / 512);
  std::vector<int32_t> lines;
  int64_t uniq = 0, tot = 0;
  for (size_t b = 0; b < nrb; b += step) {
    const int32_t j0 = rb[b].y, j1 = std::min(rb[b + 1].y, j0 + 4096);
    lines.assign(ci + j0, ci + j1);
    for (auto &l : lines) l >>= 4;
    std::sort(lines.begin(), lines.end());
    uniq += std::unique(lines.begin(), lines.end()) - lines.begin();
    tot += j1 - j0;
  }
  return tot ? (double)uniq / (double)tot : 0.0;
}

// Adds the XCD column blocks (see k_cb_prod) to a matrix without a SELL-64
// copy when x is larger than CB_MIN_X_BYTES and at least CB_SCATTER_MIN of
// the gathers fetch a line no neighbouring entry fetches (auto), or always
// with MTJG_SPMV_FORMAT=xcd.  The blocks are the fewest multiple of eight
// whose x slices are at most CB_BLOCK_BYTES, or CB_MAX_BLK blocks; the
// automatic choice leaves x of more than CB_MAX_BLK slices of twice that
// (128 MB) alone, since larger slices no longer stay in an XCD's L2.  XCD
// k's blocks k, k + 8, ... are laid out one after the other, each block's
// entries in CSR order.  Alongside: each row-block's
// segment starts per block and, in the products' order, each product's
// entry index within its row-block.  ci is the host copy of the (sorted)
// columns and rb the row-block schedule.
constexpr int64_t CB_MIN_X_BYTES = 16ll << 20;
constexpr double CB_SCATTER_MIN = 0.5;

static void cb_build(mtjg_csr_s *A, const int32_t *ci, const std::vector<int2> &rb) {
  if (A->sell || A->nnz == 0 || A->nc == 0) return;
  int mode = 0;  // 0 auto, 1 never, 2 always
  if (const char *e = getenv("MTJG_SPMV_FORMAT")) {
    if (!strcmp(e, "csr") || !strcmp(e, "sell")) mode = 1;
    else if (!strcmp(e, "xcd")) mode = 2;
  }
  if (mode == 1) return;
  const int64_t xbytes = (int64_t)A->nc * (int64_t)sizeof(double);
  if (mode == 0 && (xbytes <= CB_MIN_X_BYTES || gather_scatter(ci, rb) < CB_SCATTER_MIN)) return;
  int64_t blk_bytes = CB_BLOCK_BYTES;  // MTJG_CB_BLOCK_KB overrides it for experiments
  if (const char *e = getenv("MTJG_CB_BLOCK_KB")) blk_bytes = std::max<int64_t>(8, atoll(e)) << 10;
  int64_t nblk = (xbytes + blk_bytes - 1) / blk_bytes;
  nblk = std::max<int64_t>(CB_NXCD, (nblk + CB_NXCD - 1) / CB_NXCD * CB_NXCD);
  if (nblk > CB_MAX_BLK) {
    if (mode == 0 && xbytes > CB_MAX_BLK * 2 * CB_BLOCK_BYTES) return;
    nblk = CB_MAX_BLK;
  }
  nblk = std::min<int64_t>(nblk, std::max<int64_t>(CB_NXCD, (int64_t)A->nc / CB_NXCD * CB_NXCD));
  const int32_t width = (int32_t)(((int64_t)A->nc + nblk - 1) / nblk);
  nblk = ((int64_t)A->nc + width - 1) / width;  // blocks past the last column stay empty
  const size_t nnz = (size_t)A->nnz;
  std::vector<int64_t> start((size_t)nblk, 0);
  for (size_t j = 0; j < nnz; ++j) start[(size_t)(ci[j] / width)]++;
  int64_t pos = 0;
  for (int k = 0; k < CB_NXCD; ++k) {
    A->cb_off[k] = (int32_t)pos;
    for (int64_t b = k; b < nblk; b += CB_NXCD) {
      const int64_t c = start[(size_t)b];
      start[(size_t)b] = pos;
      pos += c;
    }
  }
  A->cb_off[CB_NXCD] = (int32_t)pos;
  const size_t nrb = rb.size() - 1, nb = (size_t)nblk;
  std::vector<int32_t> perm(nnz), seg((nrb + 1) * nb);
  std::vector<uint16_t> q(nnz, 0);
  for (size_t r = 0; r < nrb; ++r) {
    const int32_t j0 = rb[r].y, j1 = rb[r + 1].y;
    for (size_t k = 0; k < nb; ++k) seg[r * nb + k] = (int32_t)start[k];
    const bool staged = j1 - j0 <= SPMV_CAP;  // a longer row reads its segments in place
    for (int32_t j = j0; j < j1; ++j) {
      const int64_t p = start[(size_t)(ci[j] / width)]++;
      perm[(size_t)j] = (int32_t)p;
      if (staged) q[(size_t)p] = (uint16_t)(j - j0);
    }
  }
  for (size_t k = 0; k < nb; ++k) seg[nrb * nb + k] = (int32_t)start[k];
  hipStream_t s = A->ctx->stream;
  HIPCHK(hipMalloc(&A->cb_perm, sizeof(int32_t) * nnz));
  HIPCHK(hipMalloc(&A->cb_col, sizeof(int32_t) * nnz));
  HIPCHK(hipMalloc(&A->cb_val, sizeof(double) * nnz));
  HIPCHK(hipMalloc(&A->cb_prod, sizeof(double) * nnz));
  HIPCHK(hipMalloc(&A->cb_seg, sizeof(int32_t) * seg.size()));
  HIPCHK(hipMalloc(&A->cb_q, sizeof(uint16_t) * nnz));
  HIPCHK(hipMemcpyAsync(A->cb_perm, perm.data(), sizeof(int32_t) * nnz, hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(A->cb_seg, seg.data(), sizeof(int32_t) * seg.size(), hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(A->cb_q, q.data(), sizeof(uint16_t) * nnz, hipMemcpyHostToDevice, s));
  k_cb_fill<<<grid_for(A->nnz), BS, 0, s>>>(A->nnz, A->cb_perm, A->ci, A->val, A->cb_col, A->cb_val);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(s));
  A->cb_width = width;
  A->cb_nblk = (int32_t)nblk;
  A->cb = true;
}

// Records the diagonals of a pattern whose entries all lie on at most
// ST_MAX_OFF of them (column - row offsets, ascending), for the FP32 stencil
// copy; only patterns that got a SELL-64 copy (rows of similar length) are
// examined.  MTJG_SPMV_STENCIL=0 at creation keeps the SELL-64 FP32 product.
// ci is the host copy of the (sorted) columns.
static void stencil_detect(mtjg_csr_s *A, const int32_t *ci) {
  if (!A->sell || A->nnz == 0) return;
  if (const char *e = getenv("MTJG_SPMV_STENCIL"))
    if (atoi(e) == 0) return;
  int32_t off[ST_MAX_OFF];
  int m = 0;
  for (int32_t i = 0; i < A->nr; ++i) {
    for (int32_t j = A->h_rp[i]; j < A->h_rp[i + 1]; ++j) {
      const int32_t d = ci[j] - i;
      int k = 0;
      while (k < m && off[k] < d) ++k;
      if (k < m && off[k] == d) continue;
      if (m == ST_MAX_OFF) return;
      for (int t = m; t > k; --t) off[t] = off[t - 1];
      off[k] = d;
      ++m;
    }
  }
  A->st_m = m;
  for (int k = 0; k < m; ++k) A->st_off[k] = off[k];
}

// Uploads a validated host CSR (sorted unique columns) and builds its
// schedule.
static mtjg_csr_s *csr_upload(mtjg_ctx ctx, int32_t nr, int32_t nc, std::vector<int32_t> &&rp,
                              const int32_t *ci, const double *val) {
  auto *A = new mtjg_csr_s();
  A->ctx = ctx;
  A->nr = nr;
  A->nc = nc;
  A->nnz = rp[nr];
  A->h_rp = std::move(rp);
  try {
    size_t nnz1 = (size_t)std::max<int64_t>(A->nnz, 1);
    HIPCHK(hipMalloc(&A->rp, sizeof(int32_t) * (size_t)(nr + 1)));
    HIPCHK(hipMalloc(&A->ci, sizeof(int32_t) * nnz1));
    HIPCHK(hipMalloc(&A->val, sizeof(double) * nnz1));
    HIPCHK(hipMemcpyAsync(A->rp, A->h_rp.data(), sizeof(int32_t) * (size_t)(nr + 1),
                          hipMemcpyHostToDevice, ctx->stream));
    if (A->nnz > 0) {
      HIPCHK(hipMemcpyAsync(A->ci, ci, sizeof(int32_t) * (size_t)A->nnz, hipMemcpyHostToDevice,
                            ctx->stream));
      HIPCHK(hipMemcpyAsync(A->val, val, sizeof(double) * (size_t)A->nnz, hipMemcpyHostToDevice,
                            ctx->stream));
    }
    int cut = SPMV_CUT_ROWS;  // MTJG_SPMV_MAX_ROWS overrides it for experiments
    if (const char *e = getenv("MTJG_SPMV_MAX_ROWS")) cut = std::min(SPMV_MAX_ROWS, std::max(129, atoi(e)));
    std::vector<int2> rb = build_row_blocks(A->h_rp, nr, cut);
    A->nrb = (int32_t)rb.size() - 1;
    A->slabs = choose_slabs(A->nrb);
    HIPCHK(hipMalloc(&A->rb, sizeof(int2) * rb.size()));
    HIPCHK(hipMemcpyAsync(A->rb, rb.data(), sizeof(int2) * rb.size(), hipMemcpyHostToDevice,
                          ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    sell_build(A);
    stencil_detect(A, ci);
    cb_build(A, ci, rb);
  } catch (...) {
    csr_free(A);
    throw;
  }
  return A;
}

// CSR of A' with each column's entries in ascending row order, so that the
// sequential sum of a transposed row adds the scatter contributions of
// S/CompRowMatrix.java:361-363 in MTJ's order.
mtjg_csr_s *csr_transpose(mtjg_csr A) {
  if (A->trans) return A->trans;
  mtjg_ctx ctx = A->ctx;
  std::vector<int32_t> ci((size_t)A->nnz);
  std::vector<double> val((size_t)A->nnz);
  if (A->nnz > 0) {
    HIPCHK(hipMemcpyAsync(ci.data(), A->ci, sizeof(int32_t) * (size_t)A->nnz, hipMemcpyDeviceToHost,
                          ctx->stream));
    HIPCHK(hipMemcpyAsync(val.data(), A->val, sizeof(double) * (size_t)A->nnz,
                          hipMemcpyDeviceToHost, ctx->stream));
  }
  HIPCHK(hipStreamSynchronize(ctx->stream));
  std::vector<int32_t> trp((size_t)A->nc + 1, 0);
  for (int64_t j = 0; j < A->nnz; ++j) trp[(size_t)ci[j] + 1]++;
  for (int32_t c = 0; c < A->nc; ++c) trp[c + 1] += trp[c];
  std::vector<int32_t> pos(trp.begin(), trp.end() - 1);
  std::vector<int32_t> tci((size_t)A->nnz);
  std::vector<double> tval((size_t)A->nnz);
  for (int32_t i = 0; i < A->nr; ++i)
    for (int32_t j = A->h_rp[i]; j < A->h_rp[i + 1]; ++j) {
      int32_t p = pos[ci[j]]++;
      tci[p] = i;
      tval[p] = val[j];
    }
  A->trans = csr_upload(ctx, A->nc, A->nr, std::move(trp), tci.data(), tval.data());
  return A->trans;
}

}  // namespace mtjg

using namespace mtjg;

// ===========================================================================
// C ABI

extern "C" int mtjg_csr_create(mtjg_ctx ctx, int32_t n_rows, int32_t n_cols, int64_t nnz,
                               const int32_t *row_pointer, const int32_t *column_index,
                               const double *data, mtjg_csr *out) {
  MTJG_API_BEGIN
  if (!ctx || !out || !row_pointer) fail(MTJG_E_NULL, "null argument");
  if (n_rows < 0 || n_cols < 0 || nnz < 0) fail(MTJG_E_INVALID_ARGUMENT, "negative size");
  if (nnz > INT32_MAX) fail(MTJG_E_INVALID_ARGUMENT, "nnz exceeds int32 indexing");
  if (nnz > 0 && (!column_index || !data)) fail(MTJG_E_NULL, "null argument");
  std::vector<int32_t> rp(row_pointer, row_pointer + (size_t)n_rows + 1);
  if (rp[0] != 0 || rp[n_rows] != nnz)
    fail(MTJG_E_INVALID_ARGUMENT, "row pointers must run from 0 to nnz");
  bool sorted = true;
  for (int32_t i = 0; i < n_rows; ++i) {
    if (rp[i + 1] < rp[i]) fail(MTJG_E_INVALID_ARGUMENT, "row pointers must be non-decreasing");
    for (int32_t j = rp[i]; j < rp[i + 1]; ++j) {
      int32_t c = column_index[j];
      if (c < 0 || c >= n_cols)
        fail(MTJG_E_INVALID_ARGUMENT, "nz[" + std::to_string(i) + "][" + std::to_string(j - rp[i]) +
                                          "]=" + std::to_string(c) +
                                          ", which is not a valid column index");
      if (j > rp[i] && column_index[j - 1] >= c) sorted = false;
    }
  }
  const int32_t *ci = column_index;
  const double *val = data;
  std::vector<int32_t> sci;
  std::vector<double> sval;
  if (!sorted) {
    // CompRowMatrix keeps columns sorted within each row
    // (S/CompRowMatrix.java:206); the values follow their columns.
    sci.resize((size_t)nnz);
    sval.resize((size_t)nnz);
    std::vector<int32_t> perm;
    for (int32_t i = 0; i < n_rows; ++i) {
      int32_t b = rp[i], e = rp[i + 1];
      perm.resize((size_t)(e - b));
      std::iota(perm.begin(), perm.end(), b);
      std::stable_sort(perm.begin(), perm.end(),
                       [&](int32_t p, int32_t q) { return column_index[p] < column_index[q]; });
      for (int32_t k = 0; k < e - b; ++k) {
        sci[(size_t)(b + k)] = column_index[perm[k]];
        sval[(size_t)(b + k)] = data[perm[k]];
        if (k > 0 && sci[(size_t)(b + k)] == sci[(size_t)(b + k - 1)])
          fail(MTJG_E_INVALID_ARGUMENT,
               "duplicate column " + std::to_string(sci[(size_t)(b + k)]) + " in row " + std::to_string(i));
      }
    }
    ci = sci.data();
    val = sval.data();
  }
  HIPCHK(hipSetDevice(ctx->device));
  *out = csr_upload(ctx, n_rows, n_cols, std::move(rp), ci, val);
  MTJG_API_END
}

extern "C" int mtjg_csr_destroy(mtjg_csr A) {
  MTJG_API_BEGIN
  if (!A) return MTJG_OK;
  HIPCHK(hipStreamSynchronize(A->ctx->stream));
  csr_free(A);
  MTJG_API_END
}

extern "C" int mtjg_csr_format(mtjg_csr A, int32_t *format, int64_t *stored) {
  MTJG_API_BEGIN
  if (!A) fail(MTJG_E_NULL, "null matrix");
  if (format) *format = A->sell ? MTJG_FORMAT_SELL : A->cb ? MTJG_FORMAT_CSR_XCD : MTJG_FORMAT_CSR;
  if (stored) *stored = A->sell ? A->h_sptr_total : A->nnz;
  MTJG_API_END
}

extern "C" int mtjg_csr_fp32_diagonals(mtjg_csr A, int32_t *count) {
  MTJG_API_BEGIN
  if (!A || !count) fail(MTJG_E_NULL, "null argument");
  *count = A->sell ? A->st_m : 0;
  MTJG_API_END
}

extern "C" int mtjg_csr_shape(mtjg_csr A, int32_t *n_rows, int32_t *n_cols, int64_t *nnz) {
  MTJG_API_BEGIN
  if (!A) fail(MTJG_E_NULL, "null matrix");
  if (n_rows) *n_rows = A->nr;
  if (n_cols) *n_cols = A->nc;
  if (nnz) *nnz = A->nnz;
  MTJG_API_END
}

extern "C" int mtjg_csr_download(mtjg_csr A, int32_t *row_pointer, int32_t *column_index,
                                 double *data) {
  MTJG_API_BEGIN
  if (!A) fail(MTJG_E_NULL, "null matrix");
  hipStream_t s = A->ctx->stream;
  if (row_pointer) memcpy(row_pointer, A->h_rp.data(), sizeof(int32_t) * (size_t)(A->nr + 1));
  if (column_index && A->nnz)
    HIPCHK(hipMemcpyAsync(column_index, A->ci, sizeof(int32_t) * (size_t)A->nnz, hipMemcpyDeviceToHost, s));
  if (data && A->nnz)
    HIPCHK(hipMemcpyAsync(data, A->val, sizeof(double) * (size_t)A->nnz, hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  MTJG_API_END
}

static void csr_values_changed(mtjg_csr A) {
  // derived copies of the values are rebuilt on next use
  if (A->sell) A->sell_stale = true;
  if (A->cb) A->cb_stale = true;
  A->st_stale = true;
  if (A->val32 || A->sval32) {
    HIPCHK(hipStreamSynchronize(A->ctx->stream));
    (void)hipFree(A->val32);
    (void)hipFree(A->sval32);
    A->val32 = nullptr;
    A->sval32 = nullptr;
  }
  if (A->trans) {
    HIPCHK(hipStreamSynchronize(A->ctx->stream));
    csr_free(A->trans);
    A->trans = nullptr;
  }
}

extern "C" int mtjg_csr_set_values(mtjg_csr A, const double *data) {
  MTJG_API_BEGIN
  if (!A || (!data && A->nnz)) fail(MTJG_E_NULL, "null argument");
  csr_values_changed(A);
  if (A->nnz)
    HIPCHK(hipMemcpyAsync(A->val, data, sizeof(double) * (size_t)A->nnz, hipMemcpyHostToDevice,
                          A->ctx->stream));
  HIPCHK(hipStreamSynchronize(A->ctx->stream));
  MTJG_API_END
}

// Position of (row, column) in the pattern or -1 (S/CompRowMatrix.java:
// 406-431 through S/Arrays.java:119-121); host binary search over a
// downloaded row.
static int64_t csr_find(mtjg_csr A, int32_t row, int32_t column) {
  if (row < 0 || row >= A->nr || column < 0 || column >= A->nc)
    fail(MTJG_E_INDEX_OUT_OF_BOUNDS, "index out of range (" + std::to_string(row) + ", " +
                                         std::to_string(column) + ")");
  int32_t b = A->h_rp[row], e = A->h_rp[row + 1];
  if (e == b) return -1;
  std::vector<int32_t> cols((size_t)(e - b));
  HIPCHK(hipMemcpyAsync(cols.data(), A->ci + b, sizeof(int32_t) * cols.size(), hipMemcpyDeviceToHost,
                        A->ctx->stream));
  HIPCHK(hipStreamSynchronize(A->ctx->stream));
  auto it = std::lower_bound(cols.begin(), cols.end(), column);
  if (it == cols.end() || *it != column) return -1;
  return b + (it - cols.begin());
}

extern "C" int mtjg_csr_get(mtjg_csr A, int32_t row, int32_t column, double *out) {
  MTJG_API_BEGIN
  if (!A || !out) fail(MTJG_E_NULL, "null argument");
  int64_t p = csr_find(A, row, column);
  if (p < 0) {
    *out = 0;
  } else {
    HIPCHK(hipMemcpyAsync(out, A->val + p, sizeof(double), hipMemcpyDeviceToHost, A->ctx->stream));
    HIPCHK(hipStreamSynchronize(A->ctx->stream));
  }
  MTJG_API_END
}

extern "C" int mtjg_csr_set(mtjg_csr A, int32_t row, int32_t column, double value) {
  MTJG_API_BEGIN
  if (!A) fail(MTJG_E_NULL, "null matrix");
  int64_t p = csr_find(A, row, column);
  if (p < 0)
    fail(MTJG_E_INDEX_OUT_OF_BOUNDS, "Entry (" + std::to_string(row + 1) + ", " +
                                         std::to_string(column + 1) +
                                         ") is not in the matrix structure");
  csr_values_changed(A);
  HIPCHK(hipMemcpyAsync(A->val + p, &value, sizeof(double), hipMemcpyHostToDevice, A->ctx->stream));
  HIPCHK(hipStreamSynchronize(A->ctx->stream));
  MTJG_API_END
}

// checkMultAdd / checkTransMultAdd (M/AbstractMatrix.java:169-176, 204-211).
static void check_mult(mtjg_csr A, mtjg_vec x, mtjg_vec y, bool trans) {
  if (!A || !x || !y) fail(MTJG_E_NULL, "null argument");
  int64_t xin = trans ? A->nr : A->nc, yout = trans ? A->nc : A->nr;
  if (x->n != xin)
    fail(MTJG_E_INDEX_OUT_OF_BOUNDS, std::string(trans ? "A.numRows" : "A.numColumns") +
                                         " != x.size (" + std::to_string(xin) + " != " +
                                         std::to_string(x->n) + ")");
  if (y->n != yout)
    fail(MTJG_E_INDEX_OUT_OF_BOUNDS, std::string(trans ? "A.numColumns" : "A.numRows") +
                                         " != y.size (" + std::to_string(yout) + " != " +
                                         std::to_string(y->n) + ")");
  if (x->dtype != MTJG_F64 || y->dtype != MTJG_F64)
    fail(MTJG_E_UNSUPPORTED, "CompRowMatrix products take FP64 vectors");
  if (x->ptr == y->ptr && (x->n > 0))
    fail(MTJG_E_INVALID_ARGUMENT, "x and y must not alias");
}

extern "C" int mtjg_csr_mult(mtjg_csr A, mtjg_vec x, mtjg_vec y) {
  MTJG_API_BEGIN
  check_mult(A, x, y, false);
  SpmvArgs a;
  a.mode = SPMV_MULT;
  spmv_f64(A, (const double *)x->ptr, (double *)y->ptr, a, A->ctx->stream);
  MTJG_API_END
}

extern "C" int mtjg_csr_mult_add(mtjg_csr A, double alpha, mtjg_vec x, mtjg_vec y) {
  MTJG_API_BEGIN
  check_mult(A, x, y, false);
  SpmvArgs a;
  a.mode = SPMV_MULTADD;
  a.alpha = alpha;
  a.y0 = y->ptr;
  spmv_f64(A, (const double *)x->ptr, (double *)y->ptr, a, A->ctx->stream);
  MTJG_API_END
}

extern "C" int mtjg_csr_mult_scaled(mtjg_csr A, double alpha, mtjg_vec x, mtjg_vec y) {
  MTJG_API_BEGIN
  check_mult(A, x, y, false);
  if (y->n) HIPCHK(hipMemsetAsync(y->ptr, 0, sizeof(double) * (size_t)y->n, A->ctx->stream));
  SpmvArgs a;
  a.mode = SPMV_MULTADD;
  a.alpha = alpha;
  a.y0 = y->ptr;
  spmv_f64(A, (const double *)x->ptr, (double *)y->ptr, a, A->ctx->stream);
  MTJG_API_END
}

extern "C" int mtjg_csr_trans_mult(mtjg_csr A, mtjg_vec x, mtjg_vec y) {
  MTJG_API_BEGIN
  check_mult(A, x, y, true);
  mtjg_csr_s *T = csr_transpose(A);
  SpmvArgs a;
  a.mode = SPMV_MULT;
  spmv_f64(T, (const double *)x->ptr, (double *)y->ptr, a, A->ctx->stream);
  MTJG_API_END
}

extern "C" int mtjg_csr_trans_mult_add(mtjg_csr A, double alpha, mtjg_vec x, mtjg_vec y) {
  MTJG_API_BEGIN
  check_mult(A, x, y, true);
  mtjg_csr_s *T = csr_transpose(A);
  SpmvArgs a;
  a.mode = SPMV_TMULTADD;
  a.alpha = alpha;
  a.inv_alpha = 1. / alpha;
  a.y0 = y->ptr;
  spmv_f64(T, (const double *)x->ptr, (double *)y->ptr, a, A->ctx->stream);
  MTJG_API_END
}

extern "C" int mtjg_csr_trans_mult_scaled(mtjg_csr A, double alpha, mtjg_vec x, mtjg_vec y) {
  MTJG_API_BEGIN
  check_mult(A, x, y, true);
  if (y->n) HIPCHK(hipMemsetAsync(y->ptr, 0, sizeof(double) * (size_t)y->n, A->ctx->stream));
  return mtjg_csr_trans_mult_add(A, alpha, x, y);
  MTJG_API_END
}

extern "C" int mtjg_csr_row_histogram(mtjg_csr A, int64_t *bins, int nbins) {
  MTJG_API_BEGIN
  if (!A || !bins) fail(MTJG_E_NULL, "null argument");
  for (int b = 0; b < nbins; ++b) bins[b] = 0;
  for (int32_t i = 0; i < A->nr; ++i) {
    int32_t len = A->h_rp[i + 1] - A->h_rp[i];
    int b = 0;
    while (len > 0) {
      ++b;
      len >>= 1;
    }
    if (b >= nbins) b = nbins - 1;
    if (nbins > 0) bins[b]++;
  }
  MTJG_API_END
}
