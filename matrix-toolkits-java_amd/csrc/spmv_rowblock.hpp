// The row-block SpMV building block shared by the CSR products (spmv.hip)
// and the batched serving engine (batch.hip): the host-side row-block cutter
// and the device routine one workgroup runs per row-block.  The rows are cut
// into row-blocks holding at most SPMV_CAP stored entries (and at most
// SPMV_MAX_ROWS rows); a workgroup
//   1. streams the block's column indices and values with unit-stride,
//      fully coalesced loads (SPMV_PER entries per thread, all issued before
//      any use), gathers x at those columns, and stages the products
//      val[j] * x[col[j]] in LDS;
//   2. reduces each row from LDS.  When the block holds at least 128 rows a
//      thread owns a row of at most SPMV_LONG_ROW entries and adds its
//      products sequentially in storage order, exactly as MTJ's loop
//      (S/CompRowMatrix.java:304-312) does, and a longer row is then summed
//      by a whole wavefront; blocks with fewer, longer rows give each row a
//      power-of-two group of lanes.  Lane groups and wavefronts add strided
//      partials and combine them with an xor butterfly.
// A row longer than SPMV_CAP gets a row-block of its own, reduced by the
// whole workgroup straight from global memory.  Blocks are cut at 256 rows
// (SPMV_CUT_ROWS) so that on short-row matrices every thread sums one row in
// a single pass: on the 3D Poisson 256^3 product 335.2 -> 315.6 us against
// blocks of up to 512 rows, both with the XCD run mapping of spmv.hip (either
// change alone: 333-335 us, plain order and 512 rows: 331-334 us).
//
// With PRE the products were formed by an earlier pass (the XCD column-block
// products of spmv.hip), which writes them grouped by column block: a
// row-block's products of block k are one contiguous segment of its output.
// Step 1 then reads the block's segments one after the other with coalesced
// loads (thread slot s of the segments' order reads the segment that holds
// s), together with each product's entry index within the row-block (q,
// stored alongside the products), and stores each product at its entry's
// place in LDS: step 2 then runs unchanged.  Each product is the value
// val[j] * x[col[j]] the multiplication here would give, so the result is
// bitwise the plain form's.  A row longer than SPMV_CAP is one row, whose
// entries' column order is the segments' order.
#pragma once

#include <vector>

#include "mtjg_internal.hpp"
#include "reduce.hpp"

namespace mtjg {

constexpr int SPMV_CAP = 2048;           // products staged per row-block
constexpr int SPMV_PER = SPMV_CAP / BS;  // entries per thread per row-block
constexpr int SPMV_MAX_ROWS = 512;       // rows a row-block may hold (2 per thread)
constexpr int SPMV_CUT_ROWS = 256;       // rows the cutter puts in a row-block
constexpr int SPMV_LONG_ROW = 64;        // longer rows of one-lane blocks are summed by a wavefront

// Row-block schedule: greedy cut so every block holds <= SPMV_CAP entries
// and <= max_rows (<= SPMV_MAX_ROWS) rows, except single rows longer than
// SPMV_CAP.  Entry b is (first row, first stored entry) of row-block b; entry
// nrb closes the last block.
inline std::vector<int2> build_row_blocks(const std::vector<int32_t> &rp, int32_t nr,
                                          int max_rows = SPMV_CUT_ROWS) {
  std::vector<int2> rb;
  rb.reserve((size_t)nr / 64 + 2);
  rb.push_back(make_int2(0, 0));
  int32_t r = 0;
  while (r < nr) {
    int32_t len = rp[r + 1] - rp[r];
    if (len > SPMV_CAP) {
      ++r;
      rb.push_back(make_int2(r, rp[r]));
      continue;
    }
    int32_t acc = 0, rows = 0;
    while (r < nr && rows < max_rows) {
      int32_t l = rp[r + 1] - rp[r];
      if (acc + l > SPMV_CAP) break;
      acc += l;
      ++r;
      ++rows;
    }
    rb.push_back(make_int2(r, rp[r]));
  }
  return rb;
}

__device__ __forceinline__ int pow2_floor(int v) { return 1 << (31 - __clz(v)); }

// The products pass's output as the PRE form reads it: row-block r's
// products of column block k at [seg[r * nblk + k], seg[(r + 1) * nblk + k])
// of prod, and q[p] the index within its row-block of the entry whose
// product is prod[p] (row-blocks of at most SPMV_CAP entries).
constexpr int CB_MAX_BLK = 32;
struct CbView {
  const double *prod = nullptr;
  const int32_t *seg = nullptr;
  const uint16_t *q = nullptr;
  int nblk = 0;
};

// Segment starts and staging offsets of row-block b (lo[k] = entries of the
// segments before k, lo[nblk] = all of them) into LDS, by wavefront 0.
__device__ __forceinline__ void cb_segments(int b, const CbView &cb, int32_t *st, int32_t *lo) {
  if (threadIdx.x < 64) {
    const int k = threadIdx.x;
    int32_t s0 = 0, len = 0;
    if (k < cb.nblk) {
      s0 = cb.seg[(size_t)b * cb.nblk + k];
      len = cb.seg[(size_t)(b + 1) * cb.nblk + k] - s0;
    }
    int32_t incl = len;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const int32_t t = __shfl_up(incl, off, 64);
      if (k >= off) incl += t;
    }
    if (k < cb.nblk) {
      st[k] = s0;
      lo[k] = incl - len;
    }
    if (k == cb.nblk - 1) lo[cb.nblk] = incl;
  }
}

// Position in prod of staged slot i: the segment k holding it found by a
// forward walk from the caller's k (slots visited in increasing order).
__device__ __forceinline__ int32_t cb_slot_pos(int i, int &k, const int32_t *st, const int32_t *lo) {
  while (lo[k + 1] <= i) ++k;
  return st[k] + (i - lo[k]);
}

// Store of a result entry; NTY: with the non-temporal hint (y is written
// once and is too large to stay cached for the next kernel).
template <bool NTY, typename XT>
__device__ __forceinline__ void y_store(XT *p, XT v) {
  if (NTY) __builtin_nontemporal_store(v, p);
  else *p = v;
}

// Writes row `row`'s result and accumulates the epilogue dots.
template <typename XT, int MODE, bool NTY = false>
__device__ __forceinline__ void spmv_store(int row, XT s, XT *__restrict__ y, const XT *y0,
                                           double alpha, double inv_alpha, const XT *w,
                                           dd &acc0, dd &acc1) {
  if (MODE == SPMV_MULT || MODE == SPMV_DOT) {
    XT v = (s != XT(0)) ? s : XT(0);
    y_store<NTY>(&y[row], v);
    if (MODE == SPMV_DOT) {
      dd_fma(acc0, (double)w[row], (double)v);
      dd_fma(acc1, (double)v, (double)v);
    }
  } else if (MODE == SPMV_MULTADD) {
    y_store<NTY>(&y[row], y0[row] + (XT)alpha * s);
  } else if (MODE == SPMV_RESID) {
    const XT r = y0[row] + (XT)alpha * s;
    y[row] = r;
    dd_fma(acc1, (double)r, (double)r);
  } else {  // SPMV_TMULTADD: the initial value was folded into s
    y_store<NTY>(&y[row], s * (XT)alpha);
  }
}

template <typename XT, int MODE>
__device__ __forceinline__ XT spmv_init(int row, const XT *y0, double inv_alpha) {
  if (MODE == SPMV_TMULTADD) return y0[row] * (XT)inv_alpha;
  return XT(0);
}

// One row-block of y = A x (and the mode's epilogue) by one workgroup.
template <typename VT, typename XT, int MODE, bool NTY = false, bool PRE = false>
__device__ __forceinline__ void spmv_rowblock(int b, XT *prod, const int2 *__restrict__ info,
                                              const int32_t *__restrict__ rp,
                                              const int32_t *__restrict__ ci,
                                              const VT *__restrict__ val, const XT *__restrict__ x,
                                              XT *y, const XT *y0, double alpha, double inv_alpha,
                                              const XT *w, dd &acc0, dd &acc1, const CbView cb = CbView{}) {
  constexpr int RPER = SPMV_MAX_ROWS / BS;
  __shared__ int32_t seg_st[PRE ? CB_MAX_BLK : 1], seg_lo[PRE ? CB_MAX_BLK + 1 : 1];
  {
    const int2 s0 = info[b], s1 = info[b + 1];
    const int r0 = s0.x, r1 = s1.x;
    const int j0 = s0.y;
    const int nnzb = s1.y - j0;
    const int nrows = r1 - r0;
    if (PRE) {
      cb_segments(b, cb, seg_st, seg_lo);
      __syncthreads();
    }
    if (nnzb > SPMV_CAP) {
      // one long row: strided partials from global memory, fixed tree
      double v[1] = {0.0};
      XT s = 0;
      if (PRE) {
        int kseg = 0;
        for (int k = threadIdx.x; k < nnzb; k += BS) s += (XT)cb.prod[cb_slot_pos(k, kseg, seg_st, seg_lo)];
      } else {
        for (int k = threadIdx.x; k < nnzb; k += BS) s += (XT)val[j0 + k] * x[ci[j0 + k]];
      }
      v[0] = (double)s;
      block_sum<1>(v);
      if (threadIdx.x == 0) {
        XT tot = (XT)v[0];
        if (MODE == SPMV_TMULTADD) tot = spmv_init<XT, MODE>(r0, y0, inv_alpha) + tot;
        spmv_store<XT, MODE, NTY>(r0, tot, y, y0, alpha, inv_alpha, w, acc0, acc1);
      }
      return;
    }
    // 1. stage products in LDS: coalesced col/val streaming, the block's row
    //    pointers alongside, then the x gathers
    int c[SPMV_PER];
    VT a[SPMV_PER];
    XT xv[SPMV_PER];
    int qv[SPMV_PER];
    if (PRE) {
      // the position in prod of slot k of the segments' order
      int kseg = 0;
#pragma unroll
      for (int u = 0; u < SPMV_PER; ++u) {
        const int k = threadIdx.x + u * BS;
        if (k < nnzb) c[u] = cb_slot_pos(k, kseg, seg_st, seg_lo);
      }
    } else {
#pragma unroll
      for (int u = 0; u < SPMV_PER; ++u) {
        const int k = threadIdx.x + u * BS;
        if (k < nnzb) {  // (non-temporal loads here measured 5.22 -> 4.88 TB/s on 256^3)
          c[u] = ci[j0 + k];
          a[u] = val[j0 + k];
        }
      }
    }
    const bool one_lane = nrows * 2 > BS;  // uniform over the workgroup
    int rs[RPER], re[RPER];
    if (one_lane) {
#pragma unroll
      for (int u = 0; u < RPER; ++u) {
        const int g = threadIdx.x + u * BS;
        if (g < nrows) {
          rs[u] = rp[r0 + g] - j0;
          re[u] = rp[r0 + g + 1] - j0;
        }
      }
    }
#pragma unroll
    for (int u = 0; u < SPMV_PER; ++u) {
      const int k = threadIdx.x + u * BS;
      if (k < nnzb) {
        if (PRE) {
          xv[u] = (XT)cb.prod[c[u]];
          qv[u] = cb.q[c[u]];
        } else {
          xv[u] = x[c[u]];
        }
      }
    }
#pragma unroll
    for (int u = 0; u < SPMV_PER; ++u) {
      const int k = threadIdx.x + u * BS;
      if (k < nnzb) prod[PRE ? qv[u] : k] = PRE ? xv[u] : (XT)a[u] * xv[u];
    }
    __syncthreads();
    // 2. per-row reduction from LDS
    if (one_lane) {
      // a thread owns a row: sequential sum in storage order (MTJ's order).
      // A row of more than SPMV_LONG_ROW entries would hold the whole
      // workgroup while one lane walks it, so those rows are listed (in
      // (slot, wavefront, lane) order, from ballots) and summed afterwards
      // by a wavefront each: strided partials, then the xor butterfly.
      __shared__ int32_t lcnt[RPER * (BS / 64)];
      __shared__ int16_t lbeg[SPMV_MAX_ROWS], lend[SPMV_MAX_ROWS], lrow[SPMV_MAX_ROWS];
      const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
      uint64_t lmask[RPER];
#pragma unroll
      for (int u = 0; u < RPER; ++u) {
        const int g = threadIdx.x + u * BS;
        const bool lng = g < nrows && re[u] - rs[u] > SPMV_LONG_ROW;
        lmask[u] = __ballot(lng);
        if (lane == 0) lcnt[u * (BS / 64) + wv] = __popcll(lmask[u]);
        if (g < nrows && !lng) {
          const int row = r0 + g;
          XT s = spmv_init<XT, MODE>(row, y0, inv_alpha);
          for (int k = rs[u]; k < re[u]; ++k) s += prod[k];
          spmv_store<XT, MODE, NTY>(row, s, y, y0, alpha, inv_alpha, w, acc0, acc1);
        }
      }
      __syncthreads();
      int nlong = 0;
#pragma unroll
      for (int u = 0; u < RPER; ++u) {
        int base = nlong;
        for (int v = 0; v < BS / 64; ++v) {
          const int c = lcnt[u * (BS / 64) + v];
          if (v < wv) base += c;
          nlong += c;
        }
        if ((lmask[u] >> lane) & 1ull) {
          const int pos = base + __popcll(lmask[u] & ((1ull << lane) - 1ull));
          lbeg[pos] = (int16_t)rs[u];  // < SPMV_CAP
          lend[pos] = (int16_t)re[u];
          lrow[pos] = (int16_t)(threadIdx.x + u * BS);
        }
      }
      if (nlong > 0) {
        __syncthreads();
        for (int i = wv; i < nlong; i += BS / 64) {
          const int row = r0 + lrow[i];
          XT s = (lane == 0) ? spmv_init<XT, MODE>(row, y0, inv_alpha) : XT(0);
          for (int k = lbeg[i] + lane; k < lend[i]; k += 64) s += prod[k];
#pragma unroll
          for (int off = 32; off > 0; off >>= 1) s += lane_xor_rt(s, off);
          if (lane == 0) spmv_store<XT, MODE, NTY>(row, s, y, y0, alpha, inv_alpha, w, acc0, acc1);
        }
      }
    } else {
      const int tpr = min(64, pow2_floor(BS / max(nrows, 1)));
      const int groups = BS / tpr;
      const int lane = threadIdx.x & (tpr - 1);
      for (int g = threadIdx.x / tpr; g < nrows; g += groups) {
        const int row = r0 + g;
        const int beg = rp[row] - j0, end = rp[row + 1] - j0;
        XT s = (lane == 0) ? spmv_init<XT, MODE>(row, y0, inv_alpha) : XT(0);
        for (int k = beg + lane; k < end; k += tpr) s += prod[k];
        for (int off = tpr >> 1; off > 0; off >>= 1) s += lane_xor_rt(s, off);
        if (lane == 0) spmv_store<XT, MODE, NTY>(row, s, y, y0, alpha, inv_alpha, w, acc0, acc1);
      }
    }
  }
}

// ---- FP32 stencil copy --------------------------------------------------------

// Row `row` of y = A x (and the mode's epilogue) from an FP32 stencil copy:
// the values of an n-row matrix whose entries all lie on the m <= ST_MAX_OFF
// diagonals off[0..m) (ascending), diagonal-major (row i's value on
// diagonal k at val[k * n + i]), and one mask byte per row whose bit k says
// that row i has an entry on diagonal k (column i + off[k]).  The products
// are added in diagonal order, which is ascending column order: the order
// sell_row adds them in, so the sums are bitwise the SELL-64 FP32 ones.
constexpr int ST_MAX_OFF = 8;
template <int MODE>
__device__ __forceinline__ void st_row(int row, int n, int m, const int32_t *off, const float *__restrict__ val,
                                       const uint8_t *__restrict__ mask, const float *__restrict__ x, float *y,
                                       const float *w, dd &acc0, dd &acc1) {
  if (row >= n) return;
  const uint32_t mk = mask[row];
  float v[ST_MAX_OFF], xv[ST_MAX_OFF];
#pragma unroll
  for (int k = 0; k < ST_MAX_OFF; ++k) {
    if (k < m && ((mk >> k) & 1u)) {
      v[k] = __builtin_nontemporal_load(&val[(size_t)k * n + row]);
      xv[k] = x[row + off[k]];
    }
  }
  float s = 0.0f;
#pragma unroll
  for (int k = 0; k < ST_MAX_OFF; ++k)
    if (k < m && ((mk >> k) & 1u)) s += v[k] * xv[k];
  spmv_store<float, MODE>(row, s, y, nullptr, 1.0, 1.0, w, acc0, acc1);
}

// ---- SELL-64 -----------------------------------------------------------------

constexpr int SELL_C = 64;       // rows per slice: one wavefront, a lane per row
constexpr int SELL_UNROLL = 8;   // entries of a row in flight per lane

// Row slice*64 + lane of y = A x (and the mode's epilogue) from the SELL-64
// copy: the lane adds its row's products in storage order, exactly MTJ's
// sequential loop (S/CompRowMatrix.java:304-312), loading eight entries and
// their x values before using any; padding entries (column -1) are skipped.
// The matrix entries are read once per product, so they are loaded with the
// non-temporal hint: they stream past L2 instead of evicting the lines of x
// that neighbouring rows gather again (256^3 Poisson: 5.94 -> 6.43 TB/s).
template <typename VT, typename XT, int MODE, bool NTY = false>
__device__ __forceinline__ void sell_row(int slice, int lane, int n, const int32_t *__restrict__ sptr,
                                         const int32_t *__restrict__ scol, const VT *__restrict__ sval,
                                         const XT *__restrict__ x, XT *y, const XT *y0, double alpha,
                                         double inv_alpha, const XT *w, dd &acc0, dd &acc1) {
  const int row = slice * SELL_C + lane;
  const int off = sptr[slice];
  const int width = (sptr[slice + 1] - off) / SELL_C;
  XT s = (row < n) ? spmv_init<XT, MODE>(row, y0, inv_alpha) : XT(0);
  for (int k0 = 0; k0 < width; k0 += SELL_UNROLL) {
    int c[SELL_UNROLL];
    VT v[SELL_UNROLL];
    XT xv[SELL_UNROLL];
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
      if (c[u] >= 0) xv[u] = x[c[u]];
#pragma unroll
    for (int u = 0; u < SELL_UNROLL; ++u)
      if (c[u] >= 0) s += (XT)v[u] * xv[u];
  }
  if (row < n) spmv_store<XT, MODE, NTY>(row, s, y, y0, alpha, inv_alpha, w, acc0, acc1);
}

}  // namespace mtjg
