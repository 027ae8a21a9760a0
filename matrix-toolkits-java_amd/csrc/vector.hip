// Context, error plumbing and DenseVector operations (M/DenseVector.java) on
// device-resident vectors.
//
// Elementwise kernels use the same operation per element as MTJ's loops
// (x[i] += alpha * y[i] with no FMA contraction: the library is compiled with
// -ffp-contract=off), so they are bitwise identical to MTJ.  Dot products and
// norms use the fixed-order reductions of reduce.hpp; they differ from MTJ's
// sequential sums only in summation order.
#include <cmath>
#include <cstdio>
#include <cstring>

#include "mtjg_internal.hpp"
#include "reduce.hpp"

namespace mtjg {

static thread_local std::string g_last_error;

void set_last_error(const std::string &msg) { g_last_error = msg; }

void fail(int code, const std::string &msg) { throw Error{code, msg}; }

void hip_check(hipError_t e, const char *what, const char *file, int line) {
  if (e == hipSuccess) return;
  char buf[512];
  snprintf(buf, sizeof buf, "HIP error %s (%s) at %s:%d in %s", hipGetErrorName(e),
           hipGetErrorString(e), file, line, what);
  if (e == hipErrorOutOfMemory) fail(MTJG_E_OUT_OF_MEMORY, buf);
  fail(MTJG_E_DEVICE, buf);
}

// ---------------------------------------------------------------------------
// elementwise kernels (grid-stride, one element per step)

template <typename T>
__global__ __launch_bounds__(BS) void k_fill(int64_t n, T *x, T v) {
  for (int64_t i = blockIdx.x * (int64_t)BS + threadIdx.x; i < n; i += (int64_t)gridDim.x * BS) x[i] = v;
}

template <typename T>
__global__ __launch_bounds__(BS) void k_scale(int64_t n, T *x, T alpha) {
  for (int64_t i = blockIdx.x * (int64_t)BS + threadIdx.x; i < n; i += (int64_t)gridDim.x * BS)
    x[i] *= alpha;
}

template <typename T>
__global__ __launch_bounds__(BS) void k_set_scaled(int64_t n, T *x, T alpha, const T *y) {
  for (int64_t i = blockIdx.x * (int64_t)BS + threadIdx.x; i < n; i += (int64_t)gridDim.x * BS)
    x[i] = alpha * y[i];
}

template <typename T>
__global__ __launch_bounds__(BS) void k_add_scaled(int64_t n, T *x, T alpha, const T *y) {
  for (int64_t i = blockIdx.x * (int64_t)BS + threadIdx.x; i < n; i += (int64_t)gridDim.x * BS)
    x[i] += alpha * y[i];
}

template <typename T>
__global__ __launch_bounds__(BS) void k_add(int64_t n, T *x, const T *y) {
  for (int64_t i = blockIdx.x * (int64_t)BS + threadIdx.x; i < n; i += (int64_t)gridDim.x * BS)
    x[i] += y[i];
}

// ---------------------------------------------------------------------------
// reductions

enum RedOp : int { RED_DOT = 0, RED_SUMSQ = 1, RED_SUMABS = 2, RED_MAXABS = 3 };

template <typename T, int OP>
__global__ __launch_bounds__(BS) void k_reduce(int64_t n, const T *x, const T *y, GridRed red, double *out) {
  double v[1] = {0.0};
  for (int64_t i = blockIdx.x * (int64_t)BS + threadIdx.x; i < n; i += (int64_t)gridDim.x * BS) {
    double a = (double)x[i];
    if (OP == RED_DOT) v[0] += a * (double)y[i];
    else if (OP == RED_SUMSQ) v[0] += a * a;
    else v[0] += fabs(a);
  }
  if (grid_sum_h<1>(v, red) && threadIdx.x == 0) *out = v[0];
}

// Maximum of |x_i| with Math.max semantics (NaN propagates), fixed order.
__device__ __forceinline__ double nan_max(double a, double b) {
  if (a != a) return a;
  if (b != b) return b;
  return a >= b ? a : b;
}

template <typename T>
__global__ __launch_bounds__(BS) void k_maxabs(int64_t n, const T *x, GridRed red, double *out) {
  __shared__ double sh[BS];
  double m = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)BS + threadIdx.x; i < n; i += (int64_t)gridDim.x * BS)
    m = nan_max(fabs((double)x[i]), m);
  sh[threadIdx.x] = m;
  __syncthreads();
  for (int s = BS / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sh[threadIdx.x] = nan_max(sh[threadIdx.x], sh[threadIdx.x + s]);
    __syncthreads();
  }
  double v[1] = {sh[0]};
  auto comb = [](double (&a)[1], const double (&b)[1]) { a[0] = nan_max(b[0], a[0]); };
  if (grid_combine_h<1>(v, red, comb) && threadIdx.x == 0) *out = v[0];
}

// Scaled sum of squares (LAPACK dnrm2 style), as M/DenseVector.java:297-310
// computes for Norm.TwoRobust: per-thread (scale, ssq) pairs are combined in a
// fixed tree.
__device__ __forceinline__ void ssq_combine(double &sa, double &qa, double sb, double qb) {
  if (sb == 0) return;
  if (sa < sb) {
    qa = qb + qa * (sa / sb) * (sa / sb);
    sa = sb;
  } else {
    qa += qb * (sb / sa) * (sb / sa);
  }
}

template <typename T>
__global__ __launch_bounds__(BS) void k_norm2_robust(int64_t n, const T *x, GridRed red, double *out) {
  __shared__ double shs[BS], shq[BS];
  double scale = 0, ssq = 1;
  bool any = false;
  for (int64_t i = blockIdx.x * (int64_t)BS + threadIdx.x; i < n; i += (int64_t)gridDim.x * BS) {
    double xi = (double)x[i];
    if (xi != 0) {
      double a = fabs(xi);
      if (!any) {
        scale = a;
        ssq = 1;
        any = true;
      } else if (scale < a) {
        ssq = 1 + ssq * (scale / a) * (scale / a);
        scale = a;
      } else {
        ssq += (a / scale) * (a / scale);
      }
    }
  }
  if (!any) { scale = 0; ssq = 0; }
  shs[threadIdx.x] = scale;
  shq[threadIdx.x] = ssq;
  __syncthreads();
  for (int s = BS / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      double sa = shs[threadIdx.x], qa = shq[threadIdx.x];
      ssq_combine(sa, qa, shs[threadIdx.x + s], shq[threadIdx.x + s]);
      shs[threadIdx.x] = sa;
      shq[threadIdx.x] = qa;
    }
    __syncthreads();
  }
  double v[2] = {shs[0], shq[0]};
  auto comb = [](double (&a)[2], const double (&b)[2]) { ssq_combine(a[0], a[1], b[0], b[1]); };
  if (grid_combine_h<2>(v, red, comb) && threadIdx.x == 0) *out = v[0] == 0 ? 0.0 : v[0] * sqrt(v[1]);
}

// ---------------------------------------------------------------------------
// host launchers

void launch_fill(hipStream_t s, int64_t n, double *x, double v) {
  if (n <= 0) return;
  k_fill<double><<<grid_for(n), BS, 0, s>>>(n, x, v);
  HIPCHK(hipGetLastError());
}
void launch_scale(hipStream_t s, int64_t n, double *x, double alpha) {
  if (n <= 0) return;
  k_scale<double><<<grid_for(n), BS, 0, s>>>(n, x, alpha);
  HIPCHK(hipGetLastError());
}
void launch_set_scaled(hipStream_t s, int64_t n, double *x, double alpha, const double *y) {
  if (n <= 0) return;
  k_set_scaled<double><<<grid_for(n), BS, 0, s>>>(n, x, alpha, y);
  HIPCHK(hipGetLastError());
}
void launch_add_scaled(hipStream_t s, int64_t n, double *x, double alpha, const double *y) {
  if (n <= 0) return;
  k_add_scaled<double><<<grid_for(n), BS, 0, s>>>(n, x, alpha, y);
  HIPCHK(hipGetLastError());
}
void launch_add(hipStream_t s, int64_t n, double *x, const double *y) {
  if (n <= 0) return;
  k_add<double><<<grid_for(n), BS, 0, s>>>(n, x, y);
  HIPCHK(hipGetLastError());
}
double read_scalar(mtjg_ctx ctx, const double *dev) {
  HIPCHK(hipMemcpyAsync(ctx->h_out, dev, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  return ctx->h_out[0];
}

void vec_check_same(mtjg_vec x, mtjg_vec y) {
  if (!x || !y) fail(MTJG_E_NULL, "null vector");
  if (x->n != y->n) {
    fail(MTJG_E_INDEX_OUT_OF_BOUNDS, "x.size != y.size (" + std::to_string(x->n) +
                                         " != " + std::to_string(y->n) + ")");
  }
  if (x->dtype != y->dtype) fail(MTJG_E_UNSUPPORTED, "vectors of different element types");
}

template <typename F>
static void dispatch(mtjg_vec v, F &&f) {
  if (v->dtype == MTJG_F64) f((double *)nullptr);
  else f((float *)nullptr);
}

}  // namespace mtjg

using namespace mtjg;

// ===========================================================================
// C ABI

static const char *kStatus[] = {"ok", "not converged: iterations", "not converged: divergence",
                                "not converged: breakdown"};
static const char *kErr[] = {"",
                             "invalid argument",
                             "index out of bounds",
                             "zero diagonal",
                             "missing diagonal entry",
                             "zero pivot",
                             "matrix singular",
                             "unsupported operation",
                             "out of memory",
                             "device error",
                             "null handle",
                             "input/output or format error"};

extern "C" const char *mtjg_strerror(int code) {
  if (code >= 0 && code <= 3) return kStatus[code];
  if (code < 0 && code >= -11) return kErr[-code];
  return "unknown status";
}

extern "C" const char *mtjg_last_error(void) { return g_last_error.c_str(); }

extern "C" int mtjg_ctx_create(int device, mtjg_ctx *out) {
  MTJG_API_BEGIN
  if (!out) fail(MTJG_E_NULL, "null output");
  int ndev = 0;
  HIPCHK(hipGetDeviceCount(&ndev));
  if (device < 0 || device >= ndev) fail(MTJG_E_INVALID_ARGUMENT, "no such device");
  HIPCHK(hipSetDevice(device));
  auto *c = new mtjg_ctx_s();
  c->device = device;
  HIPCHK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  GridRed &g = c->red.g;
  g.cap = RED_MAX_BLOCKS;
  g.gcap = (RED_MAX_BLOCKS + RED_GROUP - 1) / RED_GROUP;
  static_assert((RED_MAX_BLOCKS + RED_GROUP - 1) / RED_GROUP <= RED_GROUP, "grid_combine_h's second level");
  HIPCHK(hipMalloc(&g.partials, sizeof(double) * RED_MAX_K * (size_t)g.cap));
  HIPCHK(hipMalloc(&g.gpart, sizeof(double) * RED_MAX_K * (size_t)g.gcap));
  HIPCHK(hipMalloc(&g.gticket, sizeof(uint32_t) * red_ticket_words(g.gcap)));
  HIPCHK(hipMemset(g.gticket, 0, sizeof(uint32_t) * red_ticket_words(g.gcap)));
  g.fticket = g.gticket + (size_t)g.gcap * TICKET_STRIDE;
  HIPCHK(hipMalloc(&c->red.out, sizeof(double) * 8));
  HIPCHK(hipHostMalloc(&c->h_out, sizeof(double) * 8, hipHostMallocDefault));
  *out = c;
  MTJG_API_END
}

extern "C" int mtjg_ctx_destroy(mtjg_ctx ctx) {
  MTJG_API_BEGIN
  if (!ctx) return MTJG_OK;
  HIPCHK(hipSetDevice(ctx->device));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  (void)hipFree(ctx->red.g.partials);
  (void)hipFree(ctx->red.g.gpart);
  (void)hipFree(ctx->red.g.gticket);
  (void)hipFree(ctx->red.out);
  (void)hipHostFree(ctx->h_out);
  (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  MTJG_API_END
}

extern "C" int mtjg_ctx_synchronize(mtjg_ctx ctx) {
  MTJG_API_BEGIN
  if (!ctx) fail(MTJG_E_NULL, "null context");
  HIPCHK(hipStreamSynchronize(ctx->stream));
  MTJG_API_END
}

extern "C" void *mtjg_ctx_stream(mtjg_ctx ctx) { return ctx ? (void *)ctx->stream : nullptr; }

static size_t esize(int dtype) { return dtype == MTJG_F64 ? 8 : 4; }

extern "C" int mtjg_vec_create(mtjg_ctx ctx, int64_t n, int dtype, mtjg_vec *out) {
  MTJG_API_BEGIN
  if (!ctx || !out) fail(MTJG_E_NULL, "null argument");
  if (n < 0) fail(MTJG_E_INVALID_ARGUMENT, "negative size");
  if (dtype != MTJG_F64 && dtype != MTJG_F32) fail(MTJG_E_INVALID_ARGUMENT, "bad dtype");
  auto *v = new mtjg_vec_s();
  v->ctx = ctx;
  v->n = n;
  v->dtype = dtype;
  size_t bytes = (size_t)(n > 0 ? n : 1) * esize(dtype);
  hipError_t e = hipMalloc(&v->ptr, bytes);
  if (e != hipSuccess) {
    delete v;
    HIPCHK(e);
  }
  HIPCHK(hipMemsetAsync(v->ptr, 0, bytes, ctx->stream));
  *out = v;
  MTJG_API_END
}

extern "C" int mtjg_vec_wrap(mtjg_ctx ctx, int64_t n, int dtype, void *device_ptr, mtjg_vec *out) {
  MTJG_API_BEGIN
  if (!ctx || !out || (!device_ptr && n > 0)) fail(MTJG_E_NULL, "null argument");
  if (n < 0) fail(MTJG_E_INVALID_ARGUMENT, "negative size");
  if (dtype != MTJG_F64 && dtype != MTJG_F32) fail(MTJG_E_INVALID_ARGUMENT, "bad dtype");
  auto *v = new mtjg_vec_s();
  v->ctx = ctx;
  v->n = n;
  v->dtype = dtype;
  v->ptr = device_ptr;
  v->owned = false;
  *out = v;
  MTJG_API_END
}

extern "C" int mtjg_vec_destroy(mtjg_vec v) {
  MTJG_API_BEGIN
  if (!v) return MTJG_OK;
  if (v->owned && v->ptr) {
    HIPCHK(hipStreamSynchronize(v->ctx->stream));
    (void)hipFree(v->ptr);
  }
  delete v;
  MTJG_API_END
}

extern "C" int64_t mtjg_vec_size(mtjg_vec v) { return v ? v->n : -1; }
extern "C" void *mtjg_vec_device_ptr(mtjg_vec v) { return v ? v->ptr : nullptr; }

extern "C" int mtjg_vec_upload(mtjg_vec v, const void *host) {
  MTJG_API_BEGIN
  if (!v || (!host && v->n > 0)) fail(MTJG_E_NULL, "null argument");
  if (v->n > 0)
    HIPCHK(hipMemcpyAsync(v->ptr, host, (size_t)v->n * esize(v->dtype), hipMemcpyHostToDevice,
                          v->ctx->stream));
  HIPCHK(hipStreamSynchronize(v->ctx->stream));
  MTJG_API_END
}

extern "C" int mtjg_vec_download(mtjg_vec v, void *host) {
  MTJG_API_BEGIN
  if (!v || (!host && v->n > 0)) fail(MTJG_E_NULL, "null argument");
  if (v->n > 0)
    HIPCHK(hipMemcpyAsync(host, v->ptr, (size_t)v->n * esize(v->dtype), hipMemcpyDeviceToHost,
                          v->ctx->stream));
  HIPCHK(hipStreamSynchronize(v->ctx->stream));
  MTJG_API_END
}

extern "C" int mtjg_vec_zero(mtjg_vec x) {
  MTJG_API_BEGIN
  if (!x) fail(MTJG_E_NULL, "null vector");
  if (x->n > 0) HIPCHK(hipMemsetAsync(x->ptr, 0, (size_t)x->n * esize(x->dtype), x->ctx->stream));
  MTJG_API_END
}

extern "C" int mtjg_vec_scale(mtjg_vec x, double alpha) {
  MTJG_API_BEGIN
  if (!x) fail(MTJG_E_NULL, "null vector");
  if (x->n > 0) {
    if (x->dtype == MTJG_F64)
      k_scale<double><<<grid_for(x->n), BS, 0, x->ctx->stream>>>(x->n, (double *)x->ptr, alpha);
    else
      k_scale<float><<<grid_for(x->n), BS, 0, x->ctx->stream>>>(x->n, (float *)x->ptr, (float)alpha);
    HIPCHK(hipGetLastError());
  }
  MTJG_API_END
}

extern "C" int mtjg_vec_set(mtjg_vec x, mtjg_vec y) {
  MTJG_API_BEGIN
  vec_check_same(x, y);
  if (x->n > 0 && x->ptr != y->ptr)
    HIPCHK(hipMemcpyAsync(x->ptr, y->ptr, (size_t)x->n * esize(x->dtype), hipMemcpyDeviceToDevice,
                          x->ctx->stream));
  MTJG_API_END
}

extern "C" int mtjg_vec_set_scaled(mtjg_vec x, double alpha, mtjg_vec y) {
  MTJG_API_BEGIN
  vec_check_same(x, y);
  if (alpha == 0) return mtjg_vec_zero(x);
  if (x->n > 0) {
    if (x->dtype == MTJG_F64)
      k_set_scaled<double><<<grid_for(x->n), BS, 0, x->ctx->stream>>>(x->n, (double *)x->ptr, alpha,
                                                                     (const double *)y->ptr);
    else
      k_set_scaled<float><<<grid_for(x->n), BS, 0, x->ctx->stream>>>(x->n, (float *)x->ptr, (float)alpha,
                                                                    (const float *)y->ptr);
    HIPCHK(hipGetLastError());
  }
  MTJG_API_END
}

extern "C" int mtjg_vec_add(mtjg_vec x, mtjg_vec y) {
  MTJG_API_BEGIN
  vec_check_same(x, y);
  if (x->n > 0) {
    if (x->dtype == MTJG_F64)
      k_add<double><<<grid_for(x->n), BS, 0, x->ctx->stream>>>(x->n, (double *)x->ptr,
                                                              (const double *)y->ptr);
    else
      k_add<float><<<grid_for(x->n), BS, 0, x->ctx->stream>>>(x->n, (float *)x->ptr,
                                                             (const float *)y->ptr);
    HIPCHK(hipGetLastError());
  }
  MTJG_API_END
}

extern "C" int mtjg_vec_add_scaled(mtjg_vec x, double alpha, mtjg_vec y) {
  MTJG_API_BEGIN
  vec_check_same(x, y);
  if (alpha == 0) return MTJG_OK;
  if (x->n > 0) {
    if (x->dtype == MTJG_F64)
      k_add_scaled<double><<<grid_for(x->n), BS, 0, x->ctx->stream>>>(x->n, (double *)x->ptr, alpha,
                                                                     (const double *)y->ptr);
    else
      k_add_scaled<float><<<grid_for(x->n), BS, 0, x->ctx->stream>>>(x->n, (float *)x->ptr, (float)alpha,
                                                                    (const float *)y->ptr);
    HIPCHK(hipGetLastError());
  }
  MTJG_API_END
}

extern "C" int mtjg_vec_dot(mtjg_vec x, mtjg_vec y, double *out) {
  MTJG_API_BEGIN
  vec_check_same(x, y);
  if (!out) fail(MTJG_E_NULL, "null output");
  mtjg_ctx c = x->ctx;
  if (x->dtype == MTJG_F64)
    k_reduce<double, RED_DOT><<<grid_for(x->n), BS, 0, c->stream>>>(
        x->n, (const double *)x->ptr, (const double *)y->ptr, c->red.g, c->red.out);
  else
    k_reduce<float, RED_DOT><<<grid_for(x->n), BS, 0, c->stream>>>(
        x->n, (const float *)x->ptr, (const float *)y->ptr, c->red.g, c->red.out);
  HIPCHK(hipGetLastError());
  *out = read_scalar(c, c->red.out);
  MTJG_API_END
}

extern "C" int mtjg_vec_norm(mtjg_vec x, int norm_type, double *out) {
  MTJG_API_BEGIN
  if (!x || !out) fail(MTJG_E_NULL, "null argument");
  mtjg_ctx c = x->ctx;
  int g = grid_for(x->n);
  hipStream_t s = c->stream;
  auto run = [&](auto *tag) {
    using T = std::remove_pointer_t<decltype(tag)>;
    const T *p = (const T *)x->ptr;
    switch (norm_type) {
      case MTJG_NORM_ONE:
        k_reduce<T, RED_SUMABS><<<g, BS, 0, s>>>(x->n, p, p, c->red.g, c->red.out);
        break;
      case MTJG_NORM_TWO:
        k_reduce<T, RED_SUMSQ><<<g, BS, 0, s>>>(x->n, p, p, c->red.g, c->red.out);
        break;
      case MTJG_NORM_TWO_ROBUST:
        k_norm2_robust<T><<<g, BS, 0, s>>>(x->n, p, c->red.g, c->red.out);
        break;
      case MTJG_NORM_INFINITY:
        k_maxabs<T><<<g, BS, 0, s>>>(x->n, p, c->red.g, c->red.out);
        break;
      default:
        fail(MTJG_E_INVALID_ARGUMENT, "bad norm type");
    }
  };
  dispatch(x, run);
  HIPCHK(hipGetLastError());
  double v = read_scalar(c, c->red.out);
  *out = (norm_type == MTJG_NORM_TWO) ? std::sqrt(v) : v;
  MTJG_API_END
}
