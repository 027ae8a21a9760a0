// Internal declarations shared by the HIP translation units of libmtjgpu.
//
// Device memory layout: every CompRowMatrix lives on the device as three
// arrays (int32 row pointers, int32 column indices sorted within each row,
// FP64 values; an FP32 copy of the values is added on demand for the mixed
// precision path) plus a row-block schedule for the SpMV kernel.  Vectors are
// plain contiguous device arrays.  Solver scalars (rho, alpha, omega, the
// iteration monitor state) live in a small device struct so that a whole
// Krylov iteration runs without returning scalars to the host.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "mtjgpu.h"

namespace mtjg {

struct Error {
  int code;
  std::string msg;
};

[[noreturn]] void fail(int code, const std::string &msg);
void hip_check(hipError_t e, const char *what, const char *file, int line);
#define HIPCHK(x) ::mtjg::hip_check((x), #x, __FILE__, __LINE__)

void set_last_error(const std::string &msg);

// Workgroup size of every kernel in this library: four 64-wide wavefronts.
constexpr int BS = 256;
// Vectors of at least this many bytes are streamed with the non-temporal
// hint (CG's vector kernels, the result vector of the plain products): too
// large for the 256 MB Infinity Cache to pass them from one kernel to the next.
constexpr int64_t NT_VECTOR_BYTES = 96ll << 20;
// Upper bound on the grid of any kernel that reduces across the grid; the
// per-block partial sums of such a kernel fit in RED_MAX_BLOCKS slots.
constexpr int RED_MAX_BLOCKS = 4096;
// Number of independent sums one reduction kernel may carry.
constexpr int RED_MAX_K = 4;

// Solver status values kept in KState::status.
enum : int {
  ST_RUNNING = 0,
  ST_CONVERGED = 1,
  ST_ITERATIONS = 2,   // == 1 + MTJG_NOT_CONVERGED_ITERATIONS
  ST_DIVERGENCE = 3,
  ST_BREAKDOWN = 4,
  ST_SINGULAR = 5,
  ST_DEVICE = 6,  // a device-side wait ran out of its poll budget
};

// Device-resident scalar state of one Krylov solve.  All fields are written
// only by the last workgroup of a reducing kernel (or by single-thread
// kernels) and read by later kernels on the same stream.
struct KState {
  // monitor (DefaultIterationMonitor)
  double rtol, atol, dtol;
  double init_r;
  double residual;
  int32_t max_iter;
  int32_t iter;
  int32_t status;
  int32_t flags;
  // CG / BiCGstab scalars
  double rho, rho_1, rho_2, alpha, beta, omega;
  double s[8];  // scratch scalars (dot results)
  // per-kernel grid tickets (self-resetting)
  uint32_t ticket[8];
  // GMRES small state
  int32_t inner_i;     // index of the current Arnoldi step
  int32_t restart;
  int32_t cycle_done;  // inner loop of this cycle has ended
  int32_t pad;
  double normr;
};

// Workspace of the two-level grid reduction (grid_sum_h in reduce.hpp) for
// grids of up to `cap` workgroups in groups of RED_GROUP.  Atomic updates of
// one memory line are serialised, so each ticket counter sits on a line of
// its own, TICKET_STRIDE counters apart.  A compensated sum over 2^21
// doubles on 2048 workgroups: 29.8 us with the counters of groups of 256
// packed in one line, 11.4 us with them 256 B apart, 8.6 us with groups of
// 64 (groups of 32: 8.9 us, of 128: 9.3 us); the loads alone take 2.8 us.
constexpr int RED_GROUP = 64;
constexpr int TICKET_STRIDE = 64;  // uint32 counters, 256 B
struct GridRed {
  double *partials = nullptr;   // [RED_MAX_K * cap]
  double *gpart = nullptr;      // [RED_MAX_K * gcap]
  uint32_t *gticket = nullptr;  // group g's counter at g * TICKET_STRIDE, zero between launches
  uint32_t *fticket = nullptr;  // [1] on a line of its own, zero between launches
  int cap = 0, gcap = 0;
};
// uint32 words of a GridRed's counters (group counters, then the final one).
inline size_t red_ticket_words(int gcap) { return (size_t)(gcap + 1) * TICKET_STRIDE; }

// Deterministic-reduction workspace of the context's BLAS-1 calls: a
// two-level reduction workspace for RED_MAX_BLOCKS workgroups and the
// device-side results.
struct RedWS {
  GridRed g;
  double *out;  // [8]
};

// Grid size for an elementwise or reducing kernel over n items.
inline int grid_for(int64_t n, int per_thread = 4, int cap = 2048) {
  int64_t g = (n + (int64_t)BS * per_thread - 1) / ((int64_t)BS * per_thread);
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return (int)g;
}

}  // namespace mtjg

struct mtjg_ctx_s {
  int device = 0;
  hipStream_t stream = nullptr;
  mtjg::RedWS red{};
  double *h_out = nullptr;  // pinned host scalars [8]
};

struct mtjg_vec_s {
  mtjg_ctx ctx = nullptr;
  int64_t n = 0;
  int dtype = MTJG_F64;
  void *ptr = nullptr;
  bool owned = true;
};

struct mtjg_csr_s {
  mtjg_ctx ctx = nullptr;
  int32_t nr = 0, nc = 0;
  int64_t nnz = 0;
  int32_t *rp = nullptr;
  int32_t *ci = nullptr;
  double *val = nullptr;
  float *val32 = nullptr;  // FP32 copy of val (mixed precision), lazily built
  // SpMV row-block schedule: row-block b covers rows [rb[b].x, rb[b+1].x)
  // and stored entries [rb[b].y, rb[b+1].y); `slabs` orders the blocks over
  // the XCDs (> 0: that many contiguous slabs; -R: runs of R row-blocks per
  // XCD, the plain products only).
  int32_t nrb = 0;
  int32_t slabs = 1;
  int2 *rb = nullptr;
  // SELL-64 copy used by the products when its padding is small: rows in
  // slices of 64 (one wavefront, a lane per row), each slice stored column-
  // major with the width of its longest row, padding columns -1 and values
  // 0.  Slice s holds entries [sptr[s], sptr[s+1]).  `sell_stale` marks
  // values changed since the copy was filled.
  bool sell = false;
  bool sell_stale = false;
  int32_t nslice = 0;
  int64_t h_sptr_total = 0;  // entries of the copy, padding included
  int32_t *sptr = nullptr;
  int32_t *scol = nullptr;
  double *sval = nullptr;
  float *sval32 = nullptr;
  // XCD column blocks, added for irregular patterns whose x gathers scatter
  // over more than one XCD's L2 (spmv.hip, cb_build): the entries reordered
  // by column block (block b, columns [b*cb_width, (b+1)*cb_width), is
  // gathered on XCD b % 8), XCD k's blocks filling [cb_off[k], cb_off[k+1]);
  // cb_perm[j] is CSR entry j's position in that order and cb_prod the
  // products pass's output.  Row-block r's entries of block k sit at
  // [cb_seg[r*nblk + k], cb_seg[(r+1)*nblk + k]) of that order, and cb_q[p]
  // is the index within its row-block of the entry at position p of that
  // order.  `cb_stale` marks values changed since cb_val was filled.
  bool cb = false;
  bool cb_stale = false;
  int32_t cb_width = 0, cb_nblk = 0;
  int32_t cb_off[9] = {};
  int32_t *cb_perm = nullptr;
  int32_t *cb_col = nullptr;
  double *cb_val = nullptr;
  double *cb_prod = nullptr;
  int32_t *cb_seg = nullptr;
  uint16_t *cb_q = nullptr;
  // FP32 stencil copy for the FP32 products (mixed precision), added when
  // every entry lies on one of at most 8 diagonals (column - row offsets
  // st_off[0..st_m), ascending; spmv.hip, stencil_detect): st_val32[k*nr + i]
  // is row i's value on diagonal k (0 where the row has none) and bit k of
  // st_mask[i] says whether it has one.  Columns are implicit, so an entry
  // costs 4 bytes instead of 8.  Filled on first use; `st_stale` marks values
  // changed since.
  int32_t st_m = 0;
  int32_t st_off[8] = {};
  bool st_stale = true;
  float *st_val32 = nullptr;
  uint8_t *st_mask = nullptr;
  // Host copy of the row pointers (schedules, level analysis, validation).
  std::vector<int32_t> h_rp;
  // Transpose (CSR of A'), built on first transposed product.
  mtjg_csr_s *trans = nullptr;
};

namespace mtjg {

// ---- vector helpers (vector.hip) ----
void vec_check_same(mtjg_vec x, mtjg_vec y);
void launch_fill(hipStream_t s, int64_t n, double *x, double v);
void launch_scale(hipStream_t s, int64_t n, double *x, double alpha);
void launch_set_scaled(hipStream_t s, int64_t n, double *x, double alpha, const double *y);
void launch_add_scaled(hipStream_t s, int64_t n, double *x, double alpha, const double *y);
void launch_add(hipStream_t s, int64_t n, double *x, const double *y);
double read_scalar(mtjg_ctx ctx, const double *dev);

// ---- SpMV (spmv.hip) ----
enum SpmvMode : int {
  SPMV_MULT = 0,      // y = A x (zero results stored as +0.0)
  SPMV_MULTADD = 1,   // y = y0 + alpha * (A x)
  SPMV_TMULTADD = 2,  // y = ((y0 * inv_alpha) + sum) * alpha (transposed storage)
  SPMV_DOT = 3,       // y = A x, plus grid sums w.y and y.y, then fin()
  SPMV_RESID = 4,     // y = y0 + alpha * (A x), plus y.y (batched residual)
};

// Solver epilogue selector for SPMV_DOT.
enum Fin : int {
  FIN_NONE = 0,
  FIN_CG_ALPHA = 1,      // alpha = rho / (p.q); rho_1 = rho
  FIN_BICG_ALPHA = 2,    // alpha = rho_1 / (rtilde.v)
  FIN_BICG_OMEGA = 3,    // omega = (t.s) / (t.t)
};

struct SpmvArgs {
  int mode = SPMV_MULT;
  double alpha = 1.0;
  double inv_alpha = 1.0;
  const void *y0 = nullptr;
  const void *w = nullptr;   // SPMV_DOT: first dot operand
  KState *st = nullptr;      // SPMV_DOT: solver state (status checked on entry)
  int fin = FIN_NONE;
  GridRed red{};             // SPMV_DOT: two-level reduction workspace
};

// Reduction workspace columns (GridRed::cap) an SPMV_DOT product of A needs.
int64_t spmv_dot_red_cap(mtjg_csr A);

void spmv_f64(mtjg_csr A, const double *x, double *y, const SpmvArgs &a, hipStream_t s);
// CG's p' = (first ? z : p * beta + z) into pn, q = A p', and the fused
// product's sums and FIN_CG_ALPHA epilogue, in one pass over the SELL-64
// copy; false (nothing launched) when A has none or its grid exceeds the
// persistent fused product's.
bool spmv_cg_fused(mtjg_csr A, const double *p, const double *z, double *pn, double *q, const SpmvArgs &a,
                   hipStream_t s);
bool spmv_cg_fusable(mtjg_csr A);  // whether spmv_cg_fused launches for A
void spmv_f32(mtjg_csr A, const float *x, float *y, const SpmvArgs &a, hipStream_t s);
mtjg_csr_s *csr_transpose(mtjg_csr A);
void csr_ensure_fp32(mtjg_csr A);

// ---- preconditioners (precond.hip) ----
enum PrecKind : int { PREC_IDENTITY = 0, PREC_DIAGONAL = 1, PREC_ILU = 2, PREC_SSOR = 3, PREC_ICC = 4,
                      PREC_AMG = 5, PREC_CG32 = 6 };
// Applies M (identity when M is null) to n entries; kernels skip work once st's
// status leaves running (st may be null).
void prec_apply_dev(mtjg_prec M, int64_t n, const double *b, double *x, hipStream_t s, KState *st);
void prec_trans_apply_dev(mtjg_prec M, const double *b, double *x, hipStream_t s);
// Raises MTJG_E_DEVICE if a single-launch triangular solve hit its poll budget
// (synchronises the stream).
void prec_check(mtjg_prec M);
int prec_kind(mtjg_prec M);

// ---- FP32 inner-CG preconditioner (solvers.hip) ----
// Workspace of MTJG_PREC_CG32: an internal solver object over n rows.
mtjg_solver cg32_create(mtjg_ctx ctx, int32_t n, double rtol, int32_t max_iter);
// FP32 values of A and the FP32 inverse diagonal (zero diagonal: fails with
// MTJG_E_ZERO_DIAGONAL and the row).
void cg32_set_matrix(mtjg_solver S, mtjg_csr A, int32_t *bad_row);
// x = ||b|| d, d from FP32 Jacobi-CG on A32 d = b / ||b|| (d0 = 0); does
// nothing once outer's status has left running (outer may be null).
// Synchronises the stream (the inner loop reads its status word).
void cg32_apply(mtjg_solver S, mtjg_csr A, const double *b, double *x, KState *outer);
int64_t cg32_iterations(mtjg_solver S);
void cg32_destroy(mtjg_solver S);

// ---- algebraic multigrid (amg.hip) ----
namespace amg {
struct Params;
}
struct AmgDevice;
// Builds the hierarchy of A on the host and uploads it (AMG.setMatrix).
AmgDevice *amg_create_device(mtjg_ctx ctx, mtjg_csr A, const amg::Params &p, int32_t *bad_row);
// One cycle from x's incoming values: x = cycle(b, x); transpose selects the
// transposed coarse solve (transApply).
void amg_apply_device(AmgDevice *H, const double *b, double *x, bool transpose, hipStream_t s);
void amg_destroy_device(AmgDevice *H);
// Raises MTJG_E_DEVICE if a smoother's grid sweep ran out of its poll budget
// since the last check (prec_check of every level's SSOR).
void amg_check_device(AmgDevice *H);
// Level sizes (finest first); returns the number of levels.
int32_t amg_level_sizes(const AmgDevice *H, int32_t *sizes, int32_t cap);
const double *prec_invdiag(mtjg_prec M);
int32_t prec_n(mtjg_prec M);

}  // namespace mtjg

// C-ABI guard: converts internal failures into return codes.
#define MTJG_API_BEGIN try {
#define MTJG_API_END                                         \
  }                                                          \
  catch (const ::mtjg::Error &e) {                           \
    ::mtjg::set_last_error(e.msg);                           \
    return e.code;                                           \
  }                                                          \
  catch (const std::bad_alloc &) {                           \
    ::mtjg::set_last_error("out of host memory");            \
    return MTJG_E_OUT_OF_MEMORY;                             \
  }                                                          \
  catch (...) {                                              \
    ::mtjg::set_last_error("unexpected internal error");     \
    return MTJG_E_DEVICE;                                    \
  }                                                          \
  return MTJG_OK;
