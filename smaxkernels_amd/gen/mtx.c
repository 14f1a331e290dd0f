/*
 * Matrix Market (coordinate) reader producing the CRS arrays the reference
 * benchmarks build from a .mtx file: COOMatrix::read_from_mtx
 * (examples/matrix_structs.hpp:52-180, the default path without
 * USE_FAST_MMIO) followed by CRSMatrix::convert_coo_to_crs
 * (examples/matrix_structs.hpp:338-388).  That path
 *   - accepts "matrix coordinate" files with real, integer or pattern values
 *     and general or symmetric storage (anything else is rejected);
 *   - converts indices from 1-based to 0-based; pattern entries get 0.01;
 *   - for symmetric storage appends the mirrored entry (j, i, v) right after
 *     every off-diagonal entry (i, j, v);
 *   - stable-sorts the entries by row only, so within a row they keep file
 *     order (columns are not sorted);
 *   - builds row_ptr from the per-row counts.
 * The CRS here follows the same steps (a stable counting sort by row), so
 * a CRS SpMV over it runs the reference's chains in the reference's order.
 * The format itself is the public NIST Matrix Market exchange format.
 */
#include <ctype.h>
#include <errno.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <strings.h>

enum {
    SMAX_MTX_OK = 0,
    SMAX_MTX_OPEN = -1,      /* file cannot be opened            */
    SMAX_MTX_BANNER = -2,    /* missing or malformed banner      */
    SMAX_MTX_UNSUPPORTED = -3, /* not coordinate real/int/pattern general/symmetric */
    SMAX_MTX_SIZE = -4,      /* malformed size line              */
    SMAX_MTX_ENTRY = -5,     /* malformed or out-of-range entry  */
    SMAX_MTX_CAPACITY = -6,  /* caller's arrays too small        */
    SMAX_MTX_OVERFLOW = -7,  /* does not fit int32 CRS           */
    SMAX_MTX_NOMEM = -8
};

#define SMAX_MTX_SYMMETRIC 1
#define SMAX_MTX_PATTERN 2

typedef struct {
    FILE *f;
    int64_t n_rows, n_cols, n_entries; /* entries as listed in the file */
    int flags;
} mtx_file;

static int next_content_line(FILE *f, char *buf, size_t cap) {
    while (fgets(buf, (int)cap, f)) {
        const char *p = buf;
        while (*p == ' ' || *p == '\t') ++p;
        if (*p == '%' || *p == '\n' || *p == '\r' || *p == '\0') continue;
        return 1;
    }
    return 0;
}

static int open_mtx(const char *path, mtx_file *m) {
    char line[1024], banner[64], object[64], format[64], field[64], symmetry[64];
    memset(m, 0, sizeof(*m));
    m->f = fopen(path, "r");
    if (!m->f) return SMAX_MTX_OPEN;
    if (!fgets(line, sizeof line, m->f)) return SMAX_MTX_BANNER;
    if (sscanf(line, "%63s %63s %63s %63s %63s", banner, object, format, field,
               symmetry) != 5 ||
        strcmp(banner, "%%MatrixMarket") != 0)
        return SMAX_MTX_BANNER;
    if (strcasecmp(object, "matrix") != 0 || strcasecmp(format, "coordinate") != 0)
        return SMAX_MTX_UNSUPPORTED;
    if (strcasecmp(field, "pattern") == 0)
        m->flags |= SMAX_MTX_PATTERN;
    else if (strcasecmp(field, "real") != 0 && strcasecmp(field, "integer") != 0)
        return SMAX_MTX_UNSUPPORTED;
    if (strcasecmp(symmetry, "symmetric") == 0)
        m->flags |= SMAX_MTX_SYMMETRIC;
    else if (strcasecmp(symmetry, "general") != 0)
        return SMAX_MTX_UNSUPPORTED;
    if (!next_content_line(m->f, line, sizeof line)) return SMAX_MTX_SIZE;
    long long r, c, n;
    if (sscanf(line, "%lld %lld %lld", &r, &c, &n) != 3 || r < 0 || c < 0 || n < 0)
        return SMAX_MTX_SIZE;
    m->n_rows = r;
    m->n_cols = c;
    m->n_entries = n;
    return SMAX_MTX_OK;
}

/* Shape, entries listed in the file, flags (bit 0 symmetric, bit 1
 * pattern).  The CRS holds at most 2 * entries nonzeros (symmetric). */
int smax_mtx_header(const char *path, int64_t *n_rows, int64_t *n_cols,
                    int64_t *n_entries, int *flags) {
    mtx_file m;
    int rc = open_mtx(path, &m);
    if (m.f) fclose(m.f);
    if (rc) return rc;
    *n_rows = m.n_rows;
    *n_cols = m.n_cols;
    *n_entries = m.n_entries;
    *flags = m.flags;
    return SMAX_MTX_OK;
}

/* Reads the file into CRS arrays row_ptr[n_rows + 1], col[cap], val[cap];
 * returns the number of nonzeros (>= 0) or an SMAX_MTX_* error. */
int64_t smax_mtx_read(const char *path, int64_t cap, int32_t *row_ptr, int32_t *col,
                      double *val) {
    mtx_file m;
    int rc = open_mtx(path, &m);
    if (rc) {
        if (m.f) fclose(m.f);
        return rc;
    }
    const int sym = (m.flags & SMAX_MTX_SYMMETRIC) != 0;
    const int pat = (m.flags & SMAX_MTX_PATTERN) != 0;
    if (m.n_rows > INT32_MAX - 1 || m.n_cols > INT32_MAX ||
        (sym ? 2 * m.n_entries : m.n_entries) > INT32_MAX) {
        fclose(m.f);
        return SMAX_MTX_OVERFLOW;
    }
    const int64_t max_nnz = sym ? 2 * m.n_entries : m.n_entries;
    int32_t *ci = malloc(sizeof(int32_t) * (size_t)(max_nnz > 0 ? max_nnz : 1));
    int32_t *cj = malloc(sizeof(int32_t) * (size_t)(max_nnz > 0 ? max_nnz : 1));
    double *cv = malloc(sizeof(double) * (size_t)(max_nnz > 0 ? max_nnz : 1));
    if (!ci || !cj || !cv) {
        free(ci); free(cj); free(cv);
        fclose(m.f);
        return SMAX_MTX_NOMEM;
    }
    char line[1024];
    int64_t nnz = 0;
    int64_t err = 0;
    for (int64_t e = 0; e < m.n_entries && !err; ++e) {
        if (!next_content_line(m.f, line, sizeof line)) {
            err = SMAX_MTX_ENTRY;
            break;
        }
        char *p = line, *q;
        errno = 0;
        const long long i = strtoll(p, &q, 10);
        if (q == p) { err = SMAX_MTX_ENTRY; break; }
        p = q;
        const long long j = strtoll(p, &q, 10);
        if (q == p) { err = SMAX_MTX_ENTRY; break; }
        p = q;
        double v = 0.01;
        if (!pat) {
            v = strtod(p, &q);
            if (q == p) { err = SMAX_MTX_ENTRY; break; }
        }
        if (errno || i < 1 || j < 1 || i > m.n_rows || j > m.n_cols) {
            err = SMAX_MTX_ENTRY;
            break;
        }
        ci[nnz] = (int32_t)(i - 1);
        cj[nnz] = (int32_t)(j - 1);
        cv[nnz] = v;
        ++nnz;
        if (sym && i != j) {
            ci[nnz] = (int32_t)(j - 1);
            cj[nnz] = (int32_t)(i - 1);
            cv[nnz] = v;
            ++nnz;
        }
    }
    fclose(m.f);
    if (!err && nnz > cap) err = SMAX_MTX_CAPACITY;
    if (!err) {
        /* stable counting sort by row */
        for (int64_t r = 0; r <= m.n_rows; ++r) row_ptr[r] = 0;
        for (int64_t e = 0; e < nnz; ++e) ++row_ptr[ci[e] + 1];
        for (int64_t r = 0; r < m.n_rows; ++r) row_ptr[r + 1] += row_ptr[r];
        int32_t *next = malloc(sizeof(int32_t) * (size_t)(m.n_rows > 0 ? m.n_rows : 1));
        if (!next) {
            err = SMAX_MTX_NOMEM;
        } else {
            for (int64_t r = 0; r < m.n_rows; ++r) next[r] = row_ptr[r];
            for (int64_t e = 0; e < nnz; ++e) {
                const int32_t d = next[ci[e]]++;
                col[d] = cj[e];
                val[d] = cv[e];
            }
            free(next);
        }
    }
    free(ci); free(cj); free(cv);
    return err ? err : nnz;
}
