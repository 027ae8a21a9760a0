// Matrix Market input for the CSR path: CompRowMatrix(MatrixVectorReader)
// (S/CompRowMatrix.java:86-166) and DenseVector(MatrixVectorReader)
// (M/DenseVector.java:52-96) with MatrixVectorReader's header, tokenizer and
// number rules (M/io/MatrixVectorReader.java:68-87, :121-172, :174-211,
// :240-271, :492-503, :606-657).  Host code only: the text is tokenized and
// converted by up to 16 threads over newline-aligned chunks, then the CSR
// pattern and values are assembled exactly as MTJ's constructor leaves them.
//
// Rules kept from MTJ:
//  * a first line whose first space-separated word is "%%MatrixMarket"
//    (any case) is the header, and must have exactly five words (four for a
//    vector); without it the data is read as "coordinate real general";
//  * tokens are maximal runs of characters other than whitespace (space and
//    0x09-0x0E) and '%', which starts a comment running to the end of the
//    line; tokens are lower-cased, so "NaN" and "Infinity" are not numbers;
//  * an index is read as a Java double and truncated to int; a value must
//    satisfy Java's Double.parseDouble grammar (optional f/d suffix, hex
//    floats with a binary exponent);
//  * pattern, array (dense) and complex matrices are rejected; the matrix
//    pattern is the set of (row, column) pairs (plus mirrored pairs for
//    symmetric and skew-symmetric files) with columns sorted in each row;
//    values are assigned in file order, so the last duplicate wins, and the
//    mirrored values of a symmetric file are assigned after all direct ones.
#include <sys/mman.h>
#include <sys/stat.h>
#include <fcntl.h>
#include <unistd.h>

#include <algorithm>
#include <atomic>
#include <cerrno>
#include <climits>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <locale.h>
#include <string>
#include <thread>
#include <vector>

#include "mtjg_internal.hpp"

struct mtjg_mm_s {
  int kind = 0;  // 0 matrix, 1 vector
  int symmetry = MTJG_MM_GENERAL;
  int32_t rows = 0, cols = 0;
  int64_t nnz = 0;
  std::vector<int32_t> rp, ci;
  std::vector<double> val;
};

namespace mtjg {
namespace {

constexpr int64_t PARALLEL_MIN_BYTES = 1 << 20;

int worker_count(int64_t bytes) {
  if (bytes < PARALLEL_MIN_BYTES) return 1;
  unsigned hc = std::thread::hardware_concurrency();
  int t = hc ? (int)hc : 4;
  return std::max(1, std::min(t, 16));
}

template <typename F>
void parallel_for(int workers, F f) {
  if (workers <= 1) {
    f(0);
    return;
  }
  std::vector<std::thread> th;
  th.reserve(workers - 1);
  for (int w = 1; w < workers; ++w) th.emplace_back(f, w);
  f(0);
  for (auto &t : th) t.join();
}

inline bool is_space(unsigned char c) { return c == ' ' || (c >= 0x09 && c <= 0x0e); }

// StreamTokenizer with MTJ's syntax table over [p, end).
struct Tokenizer {
  const char *p, *end;
  // Returns false at end of input; otherwise [*b, *e) is the next token.
  bool next(const char **b, const char **e) {
    for (;;) {
      while (p < end && is_space((unsigned char)*p)) ++p;
      if (p >= end) return false;
      if (*p == '%') {
        while (p < end && *p != '\n' && *p != '\r') ++p;
        continue;
      }
      *b = p;
      while (p < end && !is_space((unsigned char)*p) && *p != '%') ++p;
      *e = p;
      return true;
    }
  }
};

locale_t c_locale() {
  static locale_t loc = newlocale(LC_ALL_MASK, "C", (locale_t)0);
  return loc;
}

inline char lower(char c) { return (c >= 'A' && c <= 'Z') ? (char)(c - 'A' + 'a') : c; }

// Exact path for decimal literals (Clinger, "How to read floating point
// numbers accurately", 1990): when the significant digits form an integer
// w <= 2^53 and the decimal exponent e has |e| <= 22, w and 10^|e| are exact
// doubles, so one IEEE multiply or divide is the correctly rounded value --
// the same double strtod and Java's parseDouble return.  Returns false
// (caller falls back to strtod) for anything else, including hex literals.
// [s, q) is a validated, lower-cased literal without its f/d suffix.
bool fast_decimal(const char *s, const char *q, double *out) {
  static const double p10[23] = {1e0,  1e1,  1e2,  1e3,  1e4,  1e5,  1e6,  1e7,  1e8,  1e9,  1e10, 1e11,
                                 1e12, 1e13, 1e14, 1e15, 1e16, 1e17, 1e18, 1e19, 1e20, 1e21, 1e22};
  bool neg = false;
  if (s < q && (*s == '+' || *s == '-')) {
    neg = *s == '-';
    ++s;
  }
  uint64_t w = 0;
  int nd = 0, e = 0;
  auto dig = [](char c) { return c >= '0' && c <= '9'; };
  for (; s < q && dig(*s); ++s) {
    if (w == 0 && *s == '0') continue;
    if (++nd > 19) return false;
    w = w * 10 + (uint64_t)(*s - '0');
  }
  if (s < q && *s == '.') {
    for (++s; s < q && dig(*s); ++s) {
      --e;
      if (w == 0 && *s == '0') continue;
      if (++nd > 19) return false;
      w = w * 10 + (uint64_t)(*s - '0');
    }
  }
  if (s < q && *s == 'e') {
    ++s;
    bool eneg = false;
    if (s < q && (*s == '+' || *s == '-')) eneg = *s++ == '-';
    int x = 0;
    for (; s < q && dig(*s); ++s)
      if (x < 100000) x = x * 10 + (*s - '0');
    e += eneg ? -x : x;
  }
  if (s != q) return false;
  if (w == 0) {
    *out = neg ? -0.0 : 0.0;
    return true;
  }
  if (w > (1ull << 53) || e < -22 || e > 22) return false;
  double v = (double)w;
  v = e < 0 ? v / p10[-e] : v * p10[e];
  *out = neg ? -v : v;
  return true;
}

// Double.parseDouble on a lower-cased token (java.lang.Double.valueOf
// grammar).  Returns false if the token is not a number.
bool java_double(const char *b, const char *e, double *out) {
  // String.trim(): drop characters <= ' ' at both ends
  while (b < e && (unsigned char)*b <= ' ') ++b;
  while (e > b && (unsigned char)e[-1] <= ' ') --e;
  const size_t len = (size_t)(e - b);
  if (len == 0 || len > 512) return false;
  char buf[520];
  for (size_t k = 0; k < len; ++k) buf[k] = lower(b[k]);
  buf[len] = 0;
  const char *s = buf, *q = buf + len;
  if (*s == '+' || *s == '-') ++s;
  if (q > s && (q[-1] == 'f' || q[-1] == 'd')) {
    // a suffix is only legal after a complete literal (not after hex digits
    // without the binary exponent, which the hex branch checks)
    --q;
  }
  auto digits = [](const char *&c, const char *lim, bool hex) {
    const char *c0 = c;
    while (c < lim && (std::isdigit((unsigned char)*c) || (hex && *c >= 'a' && *c <= 'f'))) ++c;
    return c > c0;
  };
  const char *c = s;
  if (q - c >= 2 && c[0] == '0' && c[1] == 'x') {
    c += 2;
    bool any = digits(c, q, true);
    if (c < q && *c == '.') {
      ++c;
      any = digits(c, q, true) || any;
    }
    if (!any || c >= q || *c != 'p') return false;
    ++c;
    if (c < q && (*c == '+' || *c == '-')) ++c;
    if (!digits(c, q, false) || c != q) return false;
  } else {
    bool any = digits(c, q, false);
    if (c < q && *c == '.') {
      ++c;
      any = digits(c, q, false) || any;
    }
    if (!any) return false;
    if (c < q && *c == 'e') {
      ++c;
      if (c < q && (*c == '+' || *c == '-')) ++c;
      if (!digits(c, q, false)) return false;
    }
    if (c != q) return false;
  }
  buf[q - buf] = 0;
  if (fast_decimal(buf, q, out)) return true;
  *out = strtod_l(buf, nullptr, c_locale());
  return true;
}

// Double.valueOf(token).intValue(): truncation toward zero, saturating.
bool java_int(const char *b, const char *e, int32_t *out) {
  // fast path: plain decimal digits
  if (e - b > 0 && e - b <= 9) {
    int32_t v = 0;
    const char *c = b;
    for (; c < e && *c >= '0' && *c <= '9'; ++c) v = v * 10 + (*c - '0');
    if (c == e) {
      *out = v;
      return true;
    }
  }
  double d;
  if (!java_double(b, e, &d)) return false;
  if (d != d) *out = 0;
  else if (d >= 2147483647.0) *out = INT32_MAX;
  else if (d <= -2147483648.0) *out = INT32_MIN;
  else *out = (int32_t)d;
  return true;
}

// Splits a header line into space-separated words (String.split(" +") on the
// trimmed line).
std::vector<std::string> header_words(const char *b, const char *e) {
  while (b < e && (unsigned char)*b <= ' ') ++b;
  while (e > b && (unsigned char)e[-1] <= ' ') --e;
  std::vector<std::string> w;
  const char *c = b;
  while (c < e) {
    const char *s = c;
    while (c < e && *c != ' ') ++c;
    w.emplace_back(s, c);
    while (c < e && *c == ' ') ++c;
  }
  if (w.empty()) w.emplace_back();
  return w;
}

bool ieq(const std::string &a, const char *b) {
  size_t n = strlen(b);
  if (a.size() != n) return false;
  for (size_t k = 0; k < n; ++k)
    if (lower(a[k]) != lower(b[k])) return false;
  return true;
}

// First line of the text (BufferedReader.readLine terminators) and where the
// rest begins.
void first_line(const char *t, const char *end, const char **le, const char **rest) {
  const char *c = t;
  while (c < end && *c != '\n' && *c != '\r') ++c;
  *le = c;
  if (c < end) {
    if (*c == '\r') {
      ++c;
      if (c < end && *c == '\n') ++c;
    } else {
      ++c;
    }
  }
  *rest = c;
}

struct Info {
  bool sparse = true;
  int field = 0;  // 0 real, 1 integer, 2 complex, 3 pattern
  int sym = MTJG_MM_GENERAL;  // + 3 = hermitian
};

// hasInfo() + readMatrixInfo() / readVectorInfo(); returns where the
// tokenizer starts.
const char *read_info(const char *t, const char *end, bool vector, Info *info) {
  if (t >= end) fail(MTJG_E_IO, "end of file");
  const char *le, *rest;
  first_line(t, end, &le, &rest);
  auto w = header_words(t, le);
  if (!ieq(w[0], "%%MatrixMarket")) return t;  // no header: defaults
  const size_t want = vector ? 4 : 5;
  if (w.size() != want)
    fail(MTJG_E_IO, std::string("Current line unparsable. It must consist of ") + (vector ? "4" : "5") + " tokens");
  if (!ieq(w[1], vector ? "vector" : "matrix"))
    fail(MTJG_E_IO, std::string("Expected \"") + (vector ? "vector" : "matrix") + "\", got " + w[1]);
  if (ieq(w[2], "coordinate")) info->sparse = true;
  else if (ieq(w[2], "array")) info->sparse = false;
  else fail(MTJG_E_IO, "Unknown layout " + w[2]);
  if (ieq(w[3], "real")) info->field = 0;
  else if (ieq(w[3], "integer")) info->field = 1;
  else if (ieq(w[3], "complex")) info->field = 2;
  else if (ieq(w[3], "pattern")) info->field = 3;
  else fail(MTJG_E_IO, "Unknown field specification " + w[3]);
  if (!vector) {
    if (ieq(w[4], "general")) info->sym = MTJG_MM_GENERAL;
    else if (ieq(w[4], "symmetric")) info->sym = MTJG_MM_SYMMETRIC;
    else if (ieq(w[4], "skew-symmetric")) info->sym = MTJG_MM_SKEW_SYMMETRIC;
    else if (ieq(w[4], "hermitian")) info->sym = 3;
    else fail(MTJG_E_IO, "Unknown symmetry specification " + w[4]);
    // MatrixInfo.validate (M/io/MatrixInfo.java:118-131)
    if (!info->sparse && info->field == 3)
      fail(MTJG_E_INVALID_ARGUMENT, "Matrix cannot be dense with pattern storage");
    if (info->sym == 3 && info->field != 2)
      fail(MTJG_E_INVALID_ARGUMENT, "Data must be complex with hermitian symmetry");
    if (info->field == 3 && info->sym == MTJG_MM_SKEW_SYMMETRIC)
      fail(MTJG_E_INVALID_ARGUMENT, "Storage cannot be pattern and skew symmetrical");
  } else if (!info->sparse && info->field == 3) {
    fail(MTJG_E_INVALID_ARGUMENT, "Vector cannot be dense with pattern storage");
  }
  return rest;
}

int32_t next_int(Tokenizer &tk) {
  const char *b, *e;
  if (!tk.next(&b, &e)) fail(MTJG_E_IO, "End-of-File encountered during parsing");
  int32_t v;
  if (!java_int(b, e, &v)) fail(MTJG_E_INVALID_ARGUMENT, "For input string: \"" + std::string(b, e) + "\"");
  return v;
}

// Parses `per` tokens per record (int, [int,] double) for `count` records
// from [p, end) into the output arrays, in parallel.
void read_records(const char *p, const char *end, int64_t count, int per, int32_t *i0, int32_t *i1,
                  double *v) {
  const int64_t need = count * per;
  if (need == 0) return;
  const int workers = worker_count(end - p);
  // newline-aligned chunk starts
  std::vector<const char *> cut(workers + 1);
  cut[0] = p;
  cut[workers] = end;
  for (int w = 1; w < workers; ++w) {
    const char *c = p + (end - p) * w / workers;
    if (c < cut[w - 1]) c = cut[w - 1];
    while (c < end && *c != '\n') ++c;
    cut[w] = (c < end) ? c + 1 : end;
  }
  std::vector<int64_t> ntok(workers + 1, 0);
  parallel_for(workers, [&](int w) {
    Tokenizer tk{cut[w], cut[w + 1]};
    const char *b, *e;
    int64_t n = 0;
    while (tk.next(&b, &e)) ++n;
    ntok[w + 1] = n;
  });
  for (int w = 0; w < workers; ++w) ntok[w + 1] += ntok[w];
  // first failing token (global index) and its text
  std::atomic<int64_t> bad{INT64_MAX};
  std::vector<std::string> bad_text(workers);
  std::vector<int64_t> bad_at(workers, INT64_MAX);
  parallel_for(workers, [&](int w) {
    int64_t t = ntok[w];
    if (t >= need) return;
    Tokenizer tk{cut[w], cut[w + 1]};
    const char *b, *e;
    while (t < need && tk.next(&b, &e)) {
      const int64_t r = t / per;
      const int k = (int)(t % per);
      bool ok;
      if (k == per - 1) {
        ok = java_double(b, e, &v[r]);
      } else {
        ok = java_int(b, e, k == 0 ? &i0[r] : &i1[r]);
      }
      if (!ok) {
        bad_at[w] = t;
        bad_text[w].assign(b, e);
        int64_t cur = bad.load();
        while (t < cur && !bad.compare_exchange_weak(cur, t)) {
        }
        return;
      }
      ++t;
    }
  });
  if (bad.load() != INT64_MAX) {
    for (int w = 0; w < workers; ++w)
      if (bad_at[w] == bad.load()) fail(MTJG_E_INVALID_ARGUMENT, "For input string: \"" + bad_text[w] + "\"");
  }
  if (ntok[workers] < need) fail(MTJG_E_IO, "End-of-File encountered during parsing");
}

struct Mapped {
  const char *data = nullptr;
  size_t size = 0;
  int fd = -1;
  void *map = nullptr;
  explicit Mapped(const char *path) {
    fd = open(path, O_RDONLY);
    if (fd < 0) fail(MTJG_E_IO, std::string("cannot open ") + path + ": " + strerror(errno));
    struct stat sb;
    if (fstat(fd, &sb) != 0) fail(MTJG_E_IO, std::string("cannot stat ") + path);
    size = (size_t)sb.st_size;
    if (size) {
      map = mmap(nullptr, size, PROT_READ, MAP_PRIVATE, fd, 0);
      if (map == MAP_FAILED) {
        map = nullptr;
        fail(MTJG_E_IO, std::string("cannot map ") + path);
      }
      data = (const char *)map;
    }
  }
  ~Mapped() {
    if (map) munmap(map, size);
    if (fd >= 0) close(fd);
  }
};

mtjg_mm parse_matrix(const char *t, int64_t len) {
  const char *end = t + len;
  Info info;
  const char *body = read_info(t, end, false, &info);
  if (info.field == 3) fail(MTJG_E_UNSUPPORTED, "Pattern matrices are not supported");
  if (!info.sparse) fail(MTJG_E_UNSUPPORTED, "Dense matrices are not supported");
  if (info.field == 2) fail(MTJG_E_UNSUPPORTED, "Complex matrices are not supported");
  Tokenizer tk{body, end};
  const int32_t rows = next_int(tk), cols = next_int(tk), ne = next_int(tk);
  if (rows < 0 || cols < 0 || ne < 0) fail(MTJG_E_INVALID_ARGUMENT, "negative matrix size");
  std::vector<int32_t> r(ne), c(ne);
  std::vector<double> v(ne);
  read_records(tk.p, end, ne, 3, r.data(), c.data(), v.data());
  const bool mirror = info.sym == MTJG_MM_SYMMETRIC || info.sym == MTJG_MM_SKEW_SYMMETRIC;
  // 1-based to 0-based, then the checks of the pattern build in MTJ's order:
  // rows of the direct entries, rows of the mirrored ones, then columns.
  for (int32_t i = 0; i < ne; ++i) {
    r[i] -= 1;
    c[i] -= 1;
  }
  for (int32_t i = 0; i < ne; ++i)
    if (r[i] < 0 || r[i] >= rows)
      fail(MTJG_E_INDEX_OUT_OF_BOUNDS, "Index: " + std::to_string(r[i]) + ", Size: " + std::to_string(rows));
  if (mirror)
    for (int32_t i = 0; i < ne; ++i)
      if (r[i] != c[i] && (c[i] < 0 || c[i] >= rows))
        fail(MTJG_E_INDEX_OUT_OF_BOUNDS, "Index: " + std::to_string(c[i]) + ", Size: " + std::to_string(rows));
  for (int32_t i = 0; i < ne; ++i)
    if (c[i] < 0 || c[i] >= cols || (mirror && r[i] != c[i] && r[i] >= cols))
      fail(MTJG_E_INVALID_ARGUMENT, "not a valid column index in row " + std::to_string(r[i]));

  // Writes in assignment order: direct entries, then mirrored ones.
  int64_t nw = ne;
  if (mirror)
    for (int32_t i = 0; i < ne; ++i) nw += (r[i] != c[i]);
  std::vector<int64_t> start((size_t)rows + 1, 0);
  for (int32_t i = 0; i < ne; ++i) start[(size_t)r[i] + 1]++;
  if (mirror)
    for (int32_t i = 0; i < ne; ++i)
      if (r[i] != c[i]) start[(size_t)c[i] + 1]++;
  for (int32_t i = 0; i < rows; ++i) start[i + 1] += start[i];
  std::vector<int32_t> wc((size_t)nw);
  std::vector<double> wv((size_t)nw);
  {
    std::vector<int64_t> fill(start.begin(), start.end() - 1);
    for (int32_t i = 0; i < ne; ++i) {
      const int64_t k = fill[r[i]]++;
      wc[k] = c[i];
      wv[k] = v[i];
    }
    if (mirror) {
      const double sgn = info.sym == MTJG_MM_SKEW_SYMMETRIC ? -1.0 : 1.0;
      for (int32_t i = 0; i < ne; ++i)
        if (r[i] != c[i]) {
          const int64_t k = fill[c[i]]++;
          wc[k] = r[i];
          wv[k] = sgn == 1.0 ? v[i] : -v[i];
        }
    }
  }
  r.clear();
  r.shrink_to_fit();
  c.clear();
  c.shrink_to_fit();
  v.clear();
  v.shrink_to_fit();

  // Per row: stable sort by column; the last write of each column wins.
  std::vector<int32_t> cnt((size_t)rows, 0);
  const int workers = worker_count(nw * 12);
  parallel_for(workers, [&](int w) {
    const int64_t r0 = (int64_t)rows * w / workers, r1 = (int64_t)rows * (w + 1) / workers;
    std::vector<std::pair<int32_t, double>> tmp;
    for (int64_t i = r0; i < r1; ++i) {
      const int64_t b = start[i], e = start[i + 1];
      if (e == b) continue;
      // the row's writes in assignment order; a stable sort keeps that
      // order among writes to one column
      tmp.clear();
      for (int64_t k = b; k < e; ++k) tmp.emplace_back(wc[k], wv[k]);
      std::stable_sort(tmp.begin(), tmp.end(),
                       [](const std::pair<int32_t, double> &a, const std::pair<int32_t, double> &z) {
                         return a.first < z.first;
                       });
      // unique columns, each with the value of its last write
      int64_t o = b;
      for (size_t k = 0; k < tmp.size(); ++k) {
        if (k + 1 < tmp.size() && tmp[k + 1].first == tmp[k].first) continue;
        wc[o] = tmp[k].first;
        wv[o] = tmp[k].second;
        ++o;
      }
      cnt[i] = (int32_t)(o - b);
    }
  });
  auto *M = new mtjg_mm_s();
  M->kind = 0;
  M->symmetry = info.sym;
  M->rows = rows;
  M->cols = cols;
  M->rp.assign((size_t)rows + 1, 0);
  for (int32_t i = 0; i < rows; ++i) M->rp[i + 1] = M->rp[i] + cnt[i];
  M->nnz = M->rp[rows];
  M->ci.resize((size_t)M->nnz);
  M->val.resize((size_t)M->nnz);
  parallel_for(workers, [&](int w) {
    const int64_t r0 = (int64_t)rows * w / workers, r1 = (int64_t)rows * (w + 1) / workers;
    for (int64_t i = r0; i < r1; ++i) {
      std::copy(wc.begin() + start[i], wc.begin() + start[i] + cnt[i], M->ci.begin() + M->rp[i]);
      std::copy(wv.begin() + start[i], wv.begin() + start[i] + cnt[i], M->val.begin() + M->rp[i]);
    }
  });
  return M;
}

mtjg_mm parse_vector(const char *t, int64_t len) {
  const char *end = t + len;
  Info info;
  const char *body = read_info(t, end, true, &info);
  Tokenizer tk{body, end};
  const int32_t size = next_int(tk);
  const int32_t ne = info.sparse ? next_int(tk) : size;
  if (size < 0 || ne < 0) fail(MTJG_E_INVALID_ARGUMENT, "negative vector size");
  if (info.field == 3) fail(MTJG_E_UNSUPPORTED, "Pattern vectors are not supported");
  if (info.field == 2) fail(MTJG_E_UNSUPPORTED, "Complex vectors are not supported");
  auto *V = new mtjg_mm_s();
  try {
    V->kind = 1;
    V->rows = size;
    V->cols = 1;
    V->val.assign((size_t)size, 0.0);
    if (info.sparse) {
      std::vector<int32_t> idx(ne);
      std::vector<double> v(ne);
      read_records(tk.p, end, ne, 2, idx.data(), nullptr, v.data());
      for (int32_t i = 0; i < ne; ++i) {
        const int32_t k = idx[i] - 1;
        if (k < 0 || k >= size)
          fail(MTJG_E_INDEX_OUT_OF_BOUNDS, "Index " + std::to_string(k) + " out of bounds for length " +
                                               std::to_string(size));
        V->val[k] = v[i];
      }
      V->nnz = ne;
    } else {
      read_records(tk.p, end, size, 1, nullptr, nullptr, V->val.data());
      V->nnz = size;
    }
  } catch (...) {
    delete V;
    throw;
  }
  return V;
}

}  // namespace
}  // namespace mtjg

using namespace mtjg;

extern "C" int mtjg_mm_parse_matrix(const char *text, int64_t length, mtjg_mm *out) {
  MTJG_API_BEGIN
  if (!out || (length > 0 && !text)) fail(MTJG_E_NULL, "null argument");
  if (length < 0) fail(MTJG_E_INVALID_ARGUMENT, "negative length");
  *out = parse_matrix(text, length);
  MTJG_API_END
}

extern "C" int mtjg_mm_read_matrix(const char *path, mtjg_mm *out) {
  MTJG_API_BEGIN
  if (!path || !out) fail(MTJG_E_NULL, "null argument");
  Mapped f(path);
  *out = parse_matrix(f.data, (int64_t)f.size);
  MTJG_API_END
}

extern "C" int mtjg_mm_parse_vector(const char *text, int64_t length, mtjg_mm *out) {
  MTJG_API_BEGIN
  if (!out || (length > 0 && !text)) fail(MTJG_E_NULL, "null argument");
  if (length < 0) fail(MTJG_E_INVALID_ARGUMENT, "negative length");
  *out = parse_vector(text, length);
  MTJG_API_END
}

extern "C" int mtjg_mm_read_vector(const char *path, mtjg_mm *out) {
  MTJG_API_BEGIN
  if (!path || !out) fail(MTJG_E_NULL, "null argument");
  Mapped f(path);
  *out = parse_vector(f.data, (int64_t)f.size);
  MTJG_API_END
}

extern "C" int mtjg_mm_shape(mtjg_mm m, int32_t *rows, int32_t *cols, int64_t *nnz, int32_t *symmetry) {
  MTJG_API_BEGIN
  if (!m) fail(MTJG_E_NULL, "null handle");
  if (rows) *rows = m->rows;
  if (cols) *cols = m->cols;
  if (nnz) *nnz = m->nnz;
  if (symmetry) *symmetry = m->symmetry;
  MTJG_API_END
}

extern "C" int mtjg_mm_copy_csr(mtjg_mm m, int32_t *row_pointer, int32_t *column_index, double *data) {
  MTJG_API_BEGIN
  if (!m) fail(MTJG_E_NULL, "null handle");
  if (m->kind != 0) fail(MTJG_E_INVALID_ARGUMENT, "not a matrix");
  if (row_pointer) memcpy(row_pointer, m->rp.data(), sizeof(int32_t) * m->rp.size());
  if (column_index) memcpy(column_index, m->ci.data(), sizeof(int32_t) * m->ci.size());
  if (data) memcpy(data, m->val.data(), sizeof(double) * m->val.size());
  MTJG_API_END
}

extern "C" int mtjg_mm_copy_vector(mtjg_mm m, double *x) {
  MTJG_API_BEGIN
  if (!m || !x) fail(MTJG_E_NULL, "null argument");
  if (m->kind != 1) fail(MTJG_E_INVALID_ARGUMENT, "not a vector");
  memcpy(x, m->val.data(), sizeof(double) * m->val.size());
  MTJG_API_END
}

extern "C" int mtjg_mm_to_csr(mtjg_ctx ctx, mtjg_mm m, mtjg_csr *out) {
  MTJG_API_BEGIN
  if (!ctx || !m || !out) fail(MTJG_E_NULL, "null argument");
  if (m->kind != 0) fail(MTJG_E_INVALID_ARGUMENT, "not a matrix");
  return mtjg_csr_create(ctx, m->rows, m->cols, m->nnz, m->rp.data(), m->ci.data(), m->val.data(), out);
  MTJG_API_END
}

extern "C" int mtjg_mm_to_vec(mtjg_ctx ctx, mtjg_mm m, mtjg_vec *out) {
  MTJG_API_BEGIN
  if (!ctx || !m || !out) fail(MTJG_E_NULL, "null argument");
  if (m->kind != 1) fail(MTJG_E_INVALID_ARGUMENT, "not a vector");
  mtjg_vec v = nullptr;
  int rc = mtjg_vec_create(ctx, m->rows, MTJG_F64, &v);
  if (rc != MTJG_OK) return rc;
  rc = mtjg_vec_upload(v, m->val.data());
  if (rc != MTJG_OK) {
    mtjg_vec_destroy(v);
    return rc;
  }
  *out = v;
  MTJG_API_END
}

extern "C" int mtjg_mm_destroy(mtjg_mm m) {
  delete m;
  return MTJG_OK;
}
