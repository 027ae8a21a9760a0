/*
 * mtj_oracle_mm.c -- CPU restatement of MTJ's Matrix Market input path.
 *
 * TEST INFRASTRUCTURE ONLY (see mtj_oracle.h): the checker for the library's
 * mtjg_mm_* readers.  One thread, literal order of MTJ's constructors:
 *
 *   CompRowMatrix(MatrixVectorReader)   S/CompRowMatrix.java:86-166
 *   DenseVector(MatrixVectorReader)     M/DenseVector.java:52-96
 *   MatrixVectorReader                  M/io/MatrixVectorReader.java
 *       tokenizer syntax :68-87, hasInfo :229-237, readMatrixInfo :121-172,
 *       readVectorInfo :174-211, sizes :258-303, readCoordinate :393-402,
 *       :492-503, getInt/getDouble :606-645
 *   MatrixInfo.validate                 M/io/MatrixInfo.java:118-131
 *   VectorInfo.validate                 M/io/VectorInfo.java:82-86
 *
 * Numbers follow java.lang.Double.valueOf's grammar, matched here with POSIX
 * extended regular expressions on the lower-cased token.
 */
#include <ctype.h>
#include <regex.h>
#include <stdlib.h>
#include <string.h>

#include "mtj_oracle.h"

typedef struct {
    const char *s;
    long n, pos;
} reader;

/* BufferedReader.readLine: returns 0 at end of input. */
static int read_line(reader *r, long *b, long *e) {
    if (r->pos >= r->n) return 0;
    *b = r->pos;
    while (r->pos < r->n && r->s[r->pos] != '\n' && r->s[r->pos] != '\r') r->pos++;
    *e = r->pos;
    if (r->pos < r->n) {
        if (r->s[r->pos] == '\r' && r->pos + 1 < r->n && r->s[r->pos + 1] == '\n') r->pos += 2;
        else r->pos += 1;
    }
    return 1;
}

static int ws(int c) { return c == ' ' || (c >= 9 && c <= 14); }

/* StreamTokenizer.nextToken with resetSyntax, word chars 0-255, comment
 * char '%', whitespace ' ' and 9-14, lowerCaseMode(true).  Returns 0 at
 * end of input, else the word in tok (truncated to cap-1 characters). */
static int next_word(reader *r, char *tok, int cap) {
    for (;;) {
        while (r->pos < r->n && ws((unsigned char)r->s[r->pos])) r->pos++;
        if (r->pos >= r->n) return 0;
        if (r->s[r->pos] == '%') {
            while (r->pos < r->n && r->s[r->pos] != '\n' && r->s[r->pos] != '\r') r->pos++;
            continue;
        }
        int k = 0;
        while (r->pos < r->n && !ws((unsigned char)r->s[r->pos]) && r->s[r->pos] != '%') {
            if (k < cap - 1) tok[k++] = (char)tolower((unsigned char)r->s[r->pos]);
            r->pos++;
        }
        tok[k] = 0;
        return 1;
    }
}

static regex_t re_dec, re_hex;
static int re_ready = 0;

static void re_init(void) {
    if (re_ready) return;
    regcomp(&re_dec, "^[+-]?([0-9]+[.]?[0-9]*|[.][0-9]+)(e[+-]?[0-9]+)?[fd]?$", REG_EXTENDED | REG_NOSUB);
    regcomp(&re_hex, "^[+-]?0x([0-9a-f]+[.]?|[0-9a-f]*[.][0-9a-f]+)p[+-]?[0-9]+[fd]?$", REG_EXTENDED | REG_NOSUB);
    re_ready = 1;
}

/* Double.parseDouble (String.trim first).  Returns 0 on a format error. */
static int parse_double(const char *tok, double *out) {
    char buf[600];
    size_t b = 0, e = strlen(tok);
    while (b < e && (unsigned char)tok[b] <= ' ') b++;
    while (e > b && (unsigned char)tok[e - 1] <= ' ') e--;
    if (e - b >= sizeof buf) return 0;
    memcpy(buf, tok + b, e - b);
    buf[e - b] = 0;
    re_init();
    if (regexec(&re_dec, buf, 0, NULL, 0) != 0 && regexec(&re_hex, buf, 0, NULL, 0) != 0) return 0;
    size_t L = strlen(buf);
    if (L && (buf[L - 1] == 'f' || buf[L - 1] == 'd')) buf[L - 1] = 0;
    *out = strtod(buf, NULL);
    return 1;
}

/* Double.valueOf(s).intValue() */
static int parse_int(const char *tok, int32_t *out) {
    double d;
    if (!parse_double(tok, &d)) return 0;
    if (d != d) *out = 0;
    else if (d >= 2147483647.0) *out = 2147483647;
    else if (d <= -2147483648.0) *out = (int32_t)(-2147483647 - 1);
    else *out = (int32_t)d;
    return 1;
}

static int get_int(reader *r, int32_t *v) {
    char tok[600];
    if (!next_word(r, tok, sizeof tok)) return OC_IO;
    return parse_int(tok, v) ? OC_OK : OC_BAD_ARG;
}

static int get_double(reader *r, double *v) {
    char tok[600];
    if (!next_word(r, tok, sizeof tok)) return OC_IO;
    return parse_double(tok, v) ? OC_OK : OC_BAD_ARG;
}

static int ieq(const char *a, long alen, const char *b) {
    long n = (long)strlen(b);
    if (alen != n) return 0;
    for (long k = 0; k < n; ++k)
        if (tolower((unsigned char)a[k]) != tolower((unsigned char)b[k])) return 0;
    return 1;
}

/* Splits the trimmed line [b,e) on runs of ' ' (String.split(" +")). */
static int split_words(const char *s, long b, long e, long *wb, long *we, int maxw) {
    while (b < e && (unsigned char)s[b] <= ' ') b++;
    while (e > b && (unsigned char)s[e - 1] <= ' ') e--;
    int n = 0;
    long c = b;
    if (b == e) {
        wb[0] = we[0] = b;
        return 1;
    }
    while (c < e) {
        long s0 = c;
        while (c < e && s[c] != ' ') c++;
        if (n < maxw) {
            wb[n] = s0;
            we[n] = c;
        }
        n++;
        while (c < e && s[c] == ' ') c++;
    }
    return n;
}

/* hasInfo + read*Info.  field: 0 real, 1 integer, 2 complex, 3 pattern;
 * sym: 0 general, 1 symmetric, 2 skew, 3 hermitian. */
static int read_info(reader *r, int vector, int *sparse, int *field, int *sym) {
    *sparse = 1;
    *field = 0;
    *sym = 0;
    long b, e;
    reader probe = *r;
    if (!read_line(&probe, &b, &e)) return OC_IO; /* readTrimmedLine: EOF */
    long wb[8] = {0}, we[8] = {0};
    int nw = split_words(r->s, b, e, wb, we, 8);
    if (!ieq(r->s + wb[0], we[0] - wb[0], "%%MatrixMarket")) return OC_OK;
    *r = probe; /* the info line is consumed */
    if (nw != (vector ? 4 : 5)) return OC_IO;
    if (!ieq(r->s + wb[1], we[1] - wb[1], vector ? "vector" : "matrix")) return OC_IO;
    const char *s = r->s;
    if (ieq(s + wb[2], we[2] - wb[2], "coordinate")) *sparse = 1;
    else if (ieq(s + wb[2], we[2] - wb[2], "array")) *sparse = 0;
    else return OC_IO;
    if (ieq(s + wb[3], we[3] - wb[3], "real")) *field = 0;
    else if (ieq(s + wb[3], we[3] - wb[3], "integer")) *field = 1;
    else if (ieq(s + wb[3], we[3] - wb[3], "complex")) *field = 2;
    else if (ieq(s + wb[3], we[3] - wb[3], "pattern")) *field = 3;
    else return OC_IO;
    if (vector) {
        if (!*sparse && *field == 3) return OC_BAD_ARG;
        return OC_OK;
    }
    if (ieq(s + wb[4], we[4] - wb[4], "general")) *sym = 0;
    else if (ieq(s + wb[4], we[4] - wb[4], "symmetric")) *sym = 1;
    else if (ieq(s + wb[4], we[4] - wb[4], "skew-symmetric")) *sym = 2;
    else if (ieq(s + wb[4], we[4] - wb[4], "hermitian")) *sym = 3;
    else return OC_IO;
    if (!*sparse && *field == 3) return OC_BAD_ARG;
    if (*field != 2 && *sym == 3) return OC_BAD_ARG;
    if (*field == 3 && *sym == 2) return OC_BAD_ARG;
    return OC_OK;
}

typedef struct {
    int32_t r, c;
} pair;

static int pair_cmp(const void *a, const void *b) {
    const pair *x = a, *y = b;
    if (x->r != y->r) return x->r < y->r ? -1 : 1;
    if (x->c != y->c) return x->c < y->c ? -1 : 1;
    return 0;
}

/* CompRowMatrix.set -> getIndex: binary search of column c in row r. */
static long entry_index(const oc_mm *m, int32_t r, int32_t c) {
    long lo = m->rp[r], hi = m->rp[r + 1] - 1;
    while (lo <= hi) {
        long mid = (lo + hi) / 2;
        if (m->ci[mid] == c) return mid;
        if (m->ci[mid] < c) lo = mid + 1;
        else hi = mid - 1;
    }
    return -1;
}

int oc_mm_matrix(const char *text, long len, oc_mm *out) {
    memset(out, 0, sizeof *out);
    reader rd = {text, len, 0};
    int sparse, field, sym;
    int rc = read_info(&rd, 0, &sparse, &field, &sym);
    if (rc) return rc;
    if (field == 3 || !sparse || field == 2) return OC_UNSUPPORTED;
    int32_t rows, cols, ne;
    if ((rc = get_int(&rd, &rows)) || (rc = get_int(&rd, &cols)) || (rc = get_int(&rd, &ne))) return rc;
    if (rows < 0 || cols < 0 || ne < 0) return OC_BAD_ARG;
    int32_t *row = malloc(sizeof(int32_t) * (ne ? ne : 1));
    int32_t *col = malloc(sizeof(int32_t) * (ne ? ne : 1));
    double *ent = malloc(sizeof(double) * (ne ? ne : 1));
    pair *pp = malloc(sizeof(pair) * (2 * (size_t)ne + 1));
    rc = OC_OK;
    for (int32_t i = 0; i < ne && !rc; ++i) {
        if ((rc = get_int(&rd, &row[i]))) break;
        if ((rc = get_int(&rd, &col[i]))) break;
        rc = get_double(&rd, &ent[i]);
    }
    long np = 0;
    if (!rc) {
        for (int32_t i = 0; i < ne; ++i) {
            row[i] -= 1;
            col[i] -= 1;
        }
        /* rnz.get(row[i]).add(column[i]) */
        for (int32_t i = 0; i < ne && !rc; ++i) {
            if (row[i] < 0 || row[i] >= rows) rc = OC_INDEX;
            else pp[np++] = (pair){row[i], col[i]};
        }
        if (!rc && (sym == 1 || sym == 2))
            for (int32_t i = 0; i < ne && !rc; ++i)
                if (row[i] != col[i]) {
                    if (col[i] < 0 || col[i] >= rows) rc = OC_INDEX;
                    else pp[np++] = (pair){col[i], row[i]};
                }
    }
    if (!rc) {
        /* construct(nz): per-row sets, sorted; column range check */
        qsort(pp, (size_t)np, sizeof(pair), pair_cmp);
        long u = 0;
        for (long k = 0; k < np; ++k)
            if (u == 0 || pair_cmp(&pp[u - 1], &pp[k]) != 0) pp[u++] = pp[k];
        for (long k = 0; k < u && !rc; ++k)
            if (pp[k].c < 0 || pp[k].c >= cols) rc = OC_BAD_ARG;
        if (!rc) {
            out->rows = rows;
            out->cols = cols;
            out->symmetry = sym;
            out->nnz = u;
            out->rp = calloc((size_t)rows + 1, sizeof(int32_t));
            out->ci = malloc(sizeof(int32_t) * (u ? u : 1));
            out->val = calloc(u ? u : 1, sizeof(double));
            for (long k = 0; k < u; ++k) {
                out->rp[pp[k].r + 1]++;
                out->ci[k] = pp[k].c;
            }
            for (int32_t i = 0; i < rows; ++i) out->rp[i + 1] += out->rp[i];
            /* set(row, column, entry) in file order, then the mirror */
            for (int32_t i = 0; i < ne; ++i) out->val[entry_index(out, row[i], col[i])] = ent[i];
            if (sym == 1)
                for (int32_t i = 0; i < ne; ++i) {
                    if (row[i] != col[i]) out->val[entry_index(out, col[i], row[i])] = ent[i];
                }
            else if (sym == 2)
                for (int32_t i = 0; i < ne; ++i) {
                    if (row[i] != col[i]) out->val[entry_index(out, col[i], row[i])] = -ent[i];
                }
        }
    }
    free(row);
    free(col);
    free(ent);
    free(pp);
    return rc;
}

int oc_mm_vector(const char *text, long len, oc_mm *out) {
    memset(out, 0, sizeof *out);
    reader rd = {text, len, 0};
    int sparse, field, sym;
    int rc = read_info(&rd, 1, &sparse, &field, &sym);
    if (rc) return rc;
    int32_t size, ne = 0;
    if ((rc = get_int(&rd, &size))) return rc;
    if (sparse && (rc = get_int(&rd, &ne))) return rc;
    if (size < 0 || ne < 0) return OC_BAD_ARG;
    if (field == 3 || field == 2) return OC_UNSUPPORTED;
    out->rows = size;
    out->cols = 1;
    out->val = calloc(size ? size : 1, sizeof(double));
    if (sparse) {
        int32_t *idx = malloc(sizeof(int32_t) * (ne ? ne : 1));
        double *ent = malloc(sizeof(double) * (ne ? ne : 1));
        for (int32_t i = 0; i < ne && !rc; ++i) {
            if ((rc = get_int(&rd, &idx[i]))) break;
            rc = get_double(&rd, &ent[i]);
        }
        for (int32_t i = 0; i < ne && !rc; ++i) {
            int32_t k = idx[i] - 1;
            if (k < 0 || k >= size) rc = OC_INDEX;
            else out->val[k] = ent[i];
        }
        out->nnz = ne;
        free(idx);
        free(ent);
    } else {
        for (int32_t i = 0; i < size && !rc; ++i) rc = get_double(&rd, &out->val[i]);
        out->nnz = size;
    }
    if (rc) oc_mm_free(out);
    return rc;
}

void oc_mm_free(oc_mm *m) {
    free(m->rp);
    free(m->ci);
    free(m->val);
    memset(m, 0, sizeof *m);
}
