/*
 * ORACLE — test infrastructure only.
 *
 * A plain-C restatement of the reference SmaxKernels CPU algorithms on the
 * SpMV path.  It is the checker for the HIP library: tests, the smoke check
 * and the benchmark's CPU-baseline leg call it; the product never does.
 *
 * Summation order follows the reference exactly.  The reference is built with
 * `-O3 -march=native` (CMakeLists.txt:221,227), under which GCC contracts
 * `sum += val*x` into fused multiply-adds, so each output here is a
 * sequential fma() chain in the reference's loop order.
 *
 * Functions and the reference code each restates:
 *   oracle_crs_spmv_*      src/kernels/spmv/spmv_cpu/spmv_cpu_crs_impl.hpp:9-38
 *   oracle_scs_spmv_*      src/kernels/spmv/spmv_cpu/spmv_cpu_scs_impl.hpp:9-49
 *   oracle_bcrs_spmv_*     src/kernels/spmv/spmv_cpu/spmv_cpu_bcrs_impl.hpp:19-85
 *   oracle_bcrs_spmm_*     the above applied to each of k right-hand sides
 *   oracle_crs_spmm_*      src/kernels/spmm/spmm_cpu/spmm_cpu_crs_impl.hpp:9-89
 *   oracle_sell_sort       std::sort of libstdc++ 11.4 (bits/stl_algo.h:
 *                          1855-1960, bits/stl_heap.h:130-425) on the
 *                          SellCSigmaPair type with the comparator of
 *                          src/utils/convert_formats.hpp:51-54
 *   oracle_crs_to_scs*     src/utils/convert_formats.hpp:16-132
 *   oracle_crs_to_bcrs*    src/utils/convert_formats.hpp:134-247
 *   oracle_crs_transpose   explicit A^T (stable counting sort by column)
 *   oracle_sddmm_f64       dvals_j = sum_v G[row_j,v] * X[col_j,v] (FP64)
 *
 * Pinning (tests/test_oracle.py): the reference's own known-answer tests
 * (examples/tests), libstdc++'s real std::sort compiled here, closed forms
 * (Laplacian row sums, tie-independent SELL metadata, A^T via dense
 * transposition) and brute force on small inputs.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef int64_t i64;
typedef int32_t i32;

static inline float bf16_to_f32(uint16_t h) {
    uint32_t u = (uint32_t)h << 16;
    float f;
    memcpy(&f, &u, 4);
    return f;
}

/* ============================================================ CRS SpMV */

void oracle_crs_spmv_f64(i64 n_rows, const i32 *row_ptr, const i32 *col,
                         const double *val, const double *x, double *y) {
#pragma omp parallel for schedule(static)
    for (i64 r = 0; r < n_rows; ++r) {
        double sum = 0.0;
        for (i32 j = row_ptr[r]; j < row_ptr[r + 1]; ++j)
            sum = fma(val[j], x[col[j]], sum);
        y[r] = sum;
    }
}

void oracle_crs_spmv_f32(i64 n_rows, const i32 *row_ptr, const i32 *col,
                         const float *val, const float *x, float *y) {
#pragma omp parallel for schedule(static)
    for (i64 r = 0; r < n_rows; ++r) {
        float sum = 0.0f;
        for (i32 j = row_ptr[r]; j < row_ptr[r + 1]; ++j)
            sum = fmaf(val[j], x[col[j]], sum);
        y[r] = sum;
    }
}

/* BF16-stored values, FP32 vectors and accumulation (no reference kernel:
 * the FP32 loop above with each value widened exactly to FP32). */
void oracle_crs_spmv_bf16(i64 n_rows, const i32 *row_ptr, const i32 *col,
                          const uint16_t *val, const float *x, float *y) {
#pragma omp parallel for schedule(static)
    for (i64 r = 0; r < n_rows; ++r) {
        float sum = 0.0f;
        for (i32 j = row_ptr[r]; j < row_ptr[r + 1]; ++j)
            sum = fmaf(bf16_to_f32(val[j]), x[col[j]], sum);
        y[r] = sum;
    }
}

/* ============================================================ SELL SpMV */

void oracle_scs_spmv_f64(i64 C, i64 n_chunks, const i32 *chunk_ptr,
                         const i32 *chunk_len, const i32 *col,
                         const double *val, const double *x, double *y) {
#pragma omp parallel for schedule(static)
    for (i64 c = 0; c < n_chunks; ++c) {
        for (i64 i = 0; i < C; ++i) {
            double t = 0.0;
            for (i64 j = 0; j < chunk_len[c]; ++j) {
                i64 o = (i64)chunk_ptr[c] + j * C + i;
                t = fma(val[o], x[col[o]], t);
            }
            y[c * C + i] = t;
        }
    }
}

void oracle_scs_spmv_f32(i64 C, i64 n_chunks, const i32 *chunk_ptr,
                         const i32 *chunk_len, const i32 *col,
                         const float *val, const float *x, float *y) {
#pragma omp parallel for schedule(static)
    for (i64 c = 0; c < n_chunks; ++c) {
        for (i64 i = 0; i < C; ++i) {
            float t = 0.0f;
            for (i64 j = 0; j < chunk_len[c]; ++j) {
                i64 o = (i64)chunk_ptr[c] + j * C + i;
                t = fmaf(val[o], x[col[o]], t);
            }
            y[c * C + i] = t;
        }
    }
}

void oracle_scs_spmv_bf16(i64 C, i64 n_chunks, const i32 *chunk_ptr,
                          const i32 *chunk_len, const i32 *col,
                          const uint16_t *val, const float *x, float *y) {
#pragma omp parallel for schedule(static)
    for (i64 c = 0; c < n_chunks; ++c) {
        for (i64 i = 0; i < C; ++i) {
            float t = 0.0f;
            for (i64 j = 0; j < chunk_len[c]; ++j) {
                i64 o = (i64)chunk_ptr[c] + j * C + i;
                t = fmaf(bf16_to_f32(val[o]), x[col[o]], t);
            }
            y[c * C + i] = t;
        }
    }
}

/* ============================================================ BCRS SpMV */
/*
 * Per block row, y_buf[ix] accumulates over blocks j (CRS order) and then the
 * block's columns jx in ascending order; both block layouts give the same
 * per-output order (spmv_cpu_bcrs_impl.hpp:52-67).
 */
#define BCRS_BODY(T, LOAD, FMA)                                               \
    _Pragma("omp parallel for schedule(static)")                               \
    for (i64 r = 0; r < n_rows; ++r) {                                         \
        for (i64 ix = 0; ix < b_h; ++ix) {                                     \
            T acc = 0;                                                         \
            for (i32 j = row_ptr[r]; j < row_ptr[r + 1]; ++j) {                \
                for (i64 jx = 0; jx < b_w; ++jx) {                             \
                    i64 o = (i64)j * h_pad * w_pad +                           \
                            (col_major ? jx * h_pad + ix : ix * w_pad + jx);   \
                    acc = FMA(LOAD(val[o]), x[(i64)col[j] * w_pad + jx], acc); \
                }                                                              \
            }                                                                  \
            y[r * h_pad + ix] = acc;                                           \
        }                                                                      \
    }

#define ID(v) (v)

void oracle_bcrs_spmv_f64(i64 n_rows, i64 b_h, i64 b_w, i64 h_pad, i64 w_pad,
                          const i32 *col, const i32 *row_ptr,
                          const double *val, int col_major, const double *x,
                          double *y) {
    BCRS_BODY(double, ID, fma)
}

void oracle_bcrs_spmv_f32(i64 n_rows, i64 b_h, i64 b_w, i64 h_pad, i64 w_pad,
                          const i32 *col, const i32 *row_ptr, const float *val,
                          int col_major, const float *x, float *y) {
    BCRS_BODY(float, ID, fmaf)
}

void oracle_bcrs_spmv_bf16(i64 n_rows, i64 b_h, i64 b_w, i64 h_pad,
                           i64 w_pad, const i32 *col, const i32 *row_ptr,
                           const uint16_t *val, int col_major, const float *x,
                           float *y) {
    BCRS_BODY(float, bf16_to_f32, fmaf)
}

/* BCRS times k right-hand sides, X and Y row-major (x index * k + v). */
#define BCRS_SPMM_BODY(T, LOAD, FMA)                                          \
    _Pragma("omp parallel for schedule(static)")                               \
    for (i64 r = 0; r < n_rows; ++r) {                                         \
        for (i64 ix = 0; ix < b_h; ++ix) {                                     \
            for (i64 v = 0; v < k; ++v) {                                      \
                T acc = 0;                                                     \
                for (i32 j = row_ptr[r]; j < row_ptr[r + 1]; ++j) {            \
                    for (i64 jx = 0; jx < b_w; ++jx) {                         \
                        i64 o = (i64)j * h_pad * w_pad +                       \
                                (col_major ? jx * h_pad + ix                   \
                                           : ix * w_pad + jx);                 \
                        acc = FMA(LOAD(val[o]),                                \
                                  X[((i64)col[j] * w_pad + jx) * k + v], acc); \
                    }                                                          \
                }                                                              \
                Y[(r * h_pad + ix) * k + v] = acc;                             \
            }                                                                  \
        }                                                                      \
    }

void oracle_bcrs_spmm_f32(i64 n_rows, i64 b_h, i64 b_w, i64 h_pad, i64 w_pad,
                          const i32 *col, const i32 *row_ptr, const float *val,
                          int col_major, i64 k, const float *X, float *Y) {
    BCRS_SPMM_BODY(float, ID, fmaf)
}

void oracle_bcrs_spmm_bf16(i64 n_rows, i64 b_h, i64 b_w, i64 h_pad,
                           i64 w_pad, const i32 *col, const i32 *row_ptr,
                           const uint16_t *val, int col_major, i64 k,
                           const float *X, float *Y) {
    BCRS_SPMM_BODY(float, bf16_to_f32, fmaf)
}

void oracle_bcrs_spmm_f64(i64 n_rows, i64 b_h, i64 b_w, i64 h_pad, i64 w_pad,
                          const i32 *col, const i32 *row_ptr,
                          const double *val, int col_major, i64 k,
                          const double *X, double *Y) {
    BCRS_SPMM_BODY(double, ID, fma)
}

/* ============================================================ CRS SpMM */
/*
 * Column-major: X[n_rows*v + col], Y[row + v*n_rows] (the reference indexes
 * X with n_rows as leading dimension, spmm_cpu_crs_impl.hpp:29); row-major:
 * X[col*k + v], Y[row*k + v].  The reference's loop nest
 * (spmm_cpu_crs_impl.hpp:18-45, 58-87): per row a tmp[k] zeroed, then for
 * each stored entry j a SIMD pass over the k columns, tmp[v] += val*X; so each
 * output is a sequential chain over j (an fma under -march=native).
 */
#define SPMM_BODY(T, LOAD, FMA)                                               \
    _Pragma("omp parallel for schedule(static)")                               \
    for (i64 r = 0; r < n_rows; ++r) {                                         \
        T tmp[k];                                                              \
        for (i64 v = 0; v < k; ++v) tmp[v] = 0;                                \
        for (i32 j = row_ptr[r]; j < row_ptr[r + 1]; ++j) {                    \
            const T a = LOAD(val[j]);                                          \
            const T *xr = row_major ? X + (i64)col[j] * k : X + col[j];        \
            const i64 xs = row_major ? 1 : n_rows;                             \
            _Pragma("omp simd")                                                \
            for (i64 v = 0; v < k; ++v) tmp[v] = FMA(a, xr[v * xs], tmp[v]);   \
        }                                                                      \
        for (i64 v = 0; v < k; ++v) {                                          \
            if (row_major) Y[r * k + v] = tmp[v]; else Y[r + v * n_rows] = tmp[v]; \
        }                                                                      \
    }

void oracle_crs_spmm_f64(i64 n_rows, const i32 *row_ptr, const i32 *col,
                         const double *val, i64 k, int row_major,
                         const double *X, double *Y) {
    SPMM_BODY(double, ID, fma)
}

void oracle_crs_spmm_f32(i64 n_rows, const i32 *row_ptr, const i32 *col,
                         const float *val, i64 k, int row_major,
                         const float *X, float *Y) {
    SPMM_BODY(float, ID, fmaf)
}

void oracle_crs_spmm_bf16(i64 n_rows, const i32 *row_ptr, const i32 *col,
                          const uint16_t *val, i64 k, int row_major,
                          const float *X, float *Y) {
    SPMM_BODY(float, bf16_to_f32, fmaf)
}

/* ===================================================== libstdc++ std::sort */
/*
 * Restatement of libstdc++ 11 std::sort (introsort) specialised to
 * SellCSigmaPair {int index; int count;} with comp(a,b) = a.count > b.count
 * (convert_formats.hpp:10-13,51-54).  Function names follow bits/stl_algo.h
 * and bits/stl_heap.h so each step can be checked against that source.
 */
typedef struct {
    i32 index;
    i32 count;
} sell_pair;

#define COMP(a, b) ((a).count > (b).count)

static inline void pswap(sell_pair *a, sell_pair *b) {
    sell_pair t = *a;
    *a = *b;
    *b = t;
}

/* stl_algo.h:79-97 */
static void move_median_to_first(sell_pair *result, sell_pair *a,
                                 sell_pair *b, sell_pair *c) {
    if (COMP(*a, *b)) {
        if (COMP(*b, *c)) pswap(result, b);
        else if (COMP(*a, *c)) pswap(result, c);
        else pswap(result, a);
    } else if (COMP(*a, *c)) pswap(result, a);
    else if (COMP(*b, *c)) pswap(result, c);
    else pswap(result, b);
}

/* stl_algo.h:1878-1895 */
static sell_pair *unguarded_partition(sell_pair *first, sell_pair *last,
                                      const sell_pair *pivot) {
    for (;;) {
        while (COMP(*first, *pivot)) ++first;
        --last;
        while (COMP(*pivot, *last)) --last;
        if (!(first < last)) return first;
        pswap(first, last);
        ++first;
    }
}

/* stl_heap.h:130-145 */
static void push_heap(sell_pair *first, i64 hole, i64 top, sell_pair value) {
    i64 parent = (hole - 1) / 2;
    while (hole > top && COMP(first[parent], value)) {
        first[hole] = first[parent];
        hole = parent;
        parent = (hole - 1) / 2;
    }
    first[hole] = value;
}

/* stl_heap.h:223-248 */
static void adjust_heap(sell_pair *first, i64 hole, i64 len,
                        sell_pair value) {
    const i64 top = hole;
    i64 second = hole;
    while (second < (len - 1) / 2) {
        second = 2 * (second + 1);
        if (COMP(first[second], first[second - 1])) second--;
        first[hole] = first[second];
        hole = second;
    }
    if ((len & 1) == 0 && second == (len - 2) / 2) {
        second = 2 * (second + 1);
        first[hole] = first[second - 1];
        hole = second - 1;
    }
    push_heap(first, hole, top, value);
}

/* stl_heap.h:339-360 */
static void make_heap(sell_pair *first, sell_pair *last) {
    i64 len = last - first;
    if (len < 2) return;
    i64 parent = (len - 2) / 2;
    for (;;) {
        sell_pair value = first[parent];
        adjust_heap(first, parent, len, value);
        if (parent == 0) return;
        parent--;
    }
}

/* stl_heap.h:253-265 (__pop_heap) and 415-425 (__sort_heap) */
static void sort_heap(sell_pair *first, sell_pair *last) {
    while (last - first > 1) {
        --last;
        sell_pair value = *last;
        *last = *first;
        adjust_heap(first, 0, last - first, value);
    }
}

/* stl_algo.h:1799-1813 */
static void unguarded_linear_insert(sell_pair *last) {
    sell_pair val = *last;
    sell_pair *next = last - 1;
    while (COMP(val, *next)) {
        *last = *next;
        last = next;
        --next;
    }
    *last = val;
}

/* stl_algo.h:1819-1838 */
static void insertion_sort(sell_pair *first, sell_pair *last) {
    if (first == last) return;
    for (sell_pair *i = first + 1; i != last; ++i) {
        if (COMP(*i, *first)) {
            sell_pair val = *i;
            memmove(first + 1, first, sizeof(sell_pair) * (size_t)(i - first));
            *first = val;
        } else {
            unguarded_linear_insert(i);
        }
    }
}

/* stl_algo.h:1861-1872 */
static void final_insertion_sort(sell_pair *first, sell_pair *last) {
    if (last - first > 16) {
        insertion_sort(first, first + 16);
        for (sell_pair *i = first + 16; i != last; ++i)
            unguarded_linear_insert(i);
    } else {
        insertion_sort(first, last);
    }
}

/* Number of segments sorted by the heap fallback (lets tests confirm that
 * an input exercises it). */
long long oracle_heap_segments = 0;

/* stl_algo.h:1925-1941; heap fallback is __partial_sort(first,last,last):
 * __heap_select with middle == last is make_heap, then __sort_heap. */
static void introsort_loop(sell_pair *first, sell_pair *last, i64 depth) {
    while (last - first > 16) {
        if (depth == 0) {
            ++oracle_heap_segments;
            make_heap(first, last);
            sort_heap(first, last);
            return;
        }
        --depth;
        sell_pair *mid = first + (last - first) / 2;
        move_median_to_first(first, first + 1, mid, last - 1);
        sell_pair *cut = unguarded_partition(first + 1, last, first);
        introsort_loop(cut, last, depth);
        last = cut;
    }
}

static i64 lg64(i64 n) { return 63 - __builtin_clzll((unsigned long long)n); }

/* stl_algo.h:1945-1959 */
static void std_sort(sell_pair *first, sell_pair *last) {
    if (first != last) {
        introsort_loop(first, last, lg64(last - first) * 2);
        final_insertion_sort(first, last);
    }
}

/* Sorts (index[i], count[i]) pairs in place exactly as std::sort would. */
void oracle_sell_sort(i64 n, i32 *index, i32 *count) {
    sell_pair *p = (sell_pair *)malloc(sizeof(sell_pair) * (size_t)(n + 1));
    for (i64 i = 0; i < n; ++i) { p[i].index = index[i]; p[i].count = count[i]; }
    std_sort(p, p + n);
    for (i64 i = 0; i < n; ++i) { index[i] = p[i].index; count[i] = p[i].count; }
    free(p);
}

/* ================================================ CRS -> SELL-C-sigma */

/* Builds the sigma-sorted (index,count) list of convert_formats.hpp:33-55. */
static sell_pair *sorted_rows(i64 n_rows, const i32 *row_ptr, i64 C,
                              i64 sigma, i64 *n_chunks, i64 *n_rows_padded) {
    *n_chunks = (n_rows + C - 1) / C;
    *n_rows_padded = *n_chunks * C;
    sell_pair *e = (sell_pair *)malloc(sizeof(sell_pair) *
                                       (size_t)(*n_rows_padded + 1));
    for (i64 i = 0; i < *n_rows_padded; ++i) {
        e[i].index = (i32)i;
        e[i].count = i < n_rows ? row_ptr[i + 1] - row_ptr[i] : 0;
    }
    for (i64 i = 0; i < *n_rows_padded; i += sigma) {
        i64 end = i + sigma < *n_rows_padded ? i + sigma : *n_rows_padded;
        std_sort(e + i, e + end);
    }
    return e;
}

/* Metadata of the conversion: returns 0, or 1 on invalid arguments. */
int oracle_crs_to_scs_meta(i64 n_rows, const i32 *row_ptr, i64 C, i64 sigma,
                           i64 *n_chunks, i64 *n_rows_padded,
                           i64 *n_elements) {
    if (n_rows <= 0 || C <= 0 || sigma <= 0) return 1;
    sell_pair *e = sorted_rows(n_rows, row_ptr, C, sigma, n_chunks,
                               n_rows_padded);
    i64 acc = 0;
    for (i64 c = 0; c < *n_chunks; ++c) {
        i32 mx = e[c * C].count;
        for (i64 i = 1; i < C; ++i)
            if (e[c * C + i].count > mx) mx = e[c * C + i].count;
        acc += (i64)mx * C;
    }
    *n_elements = acc;
    free(e);
    return 0;
}

/* Fills caller-sized arrays (sizes from oracle_crs_to_scs_meta). */
int oracle_crs_to_scs(i64 n_rows, const i32 *row_ptr, const i32 *col,
                      const void *val, i64 val_bytes, i64 C, i64 sigma,
                      i32 *chunk_ptr, i32 *chunk_len, i32 *scs_col,
                      void *scs_val, i32 *perm) {
    if (n_rows <= 0 || C <= 0 || sigma <= 0) return 1;
    i64 n_chunks, n_rows_padded;
    sell_pair *e = sorted_rows(n_rows, row_ptr, C, sigma, &n_chunks,
                               &n_rows_padded);
    i64 acc = 0;
    for (i64 c = 0; c < n_chunks; ++c) {
        i32 mx = e[c * C].count;
        for (i64 i = 1; i < C; ++i)
            if (e[c * C + i].count > mx) mx = e[c * C + i].count;
        chunk_len[c] = mx;
        chunk_ptr[c] = (i32)acc;
        acc += (i64)mx * C;
    }
    chunk_ptr[n_chunks] = (i32)acc;
    for (i64 i = 0; i < n_rows_padded; ++i)
        if (e[i].index < n_rows) perm[e[i].index] = (i32)i;
    memset(scs_col, 0, sizeof(i32) * (size_t)acc);
    memset(scs_val, 0, (size_t)(val_bytes * acc));
    for (i64 r = 0; r < n_rows; ++r) {
        i64 nr = perm[r];
        i64 base = (i64)chunk_ptr[nr / C] + nr % C;
        for (i32 j = row_ptr[r]; j < row_ptr[r + 1]; ++j) {
            i64 idx = base + (i64)(j - row_ptr[r]) * C;
            scs_col[idx] = col[j];
            memcpy((char *)scs_val + idx * val_bytes,
                   (const char *)val + (i64)j * val_bytes, (size_t)val_bytes);
        }
    }
    free(e);
    return 0;
}

/* ==================================================== CRS -> BCRS */
/*
 * Block pattern per block row: sorted unique col/b_w over its b_h rows
 * (convert_formats.hpp:178-194); values placed at
 * col-major h_pad*(c%b_w)+l_row or row-major w_pad*l_row+(c%b_w) (238).
 * Returns n_blocks, or -1 when b_h/b_w do not divide the shape or the
 * padding is smaller than the block.
 */
static int cmp_i32(const void *a, const void *b) {
    i32 x = *(const i32 *)a, y = *(const i32 *)b;
    return (x > y) - (x < y);
}

static i64 block_row_cols(const i32 *row_ptr, const i32 *col, i64 br, i64 b_h,
                          i64 b_w, i32 *buf) {
    i64 n = 0;
    for (i64 r = br * b_h; r < (br + 1) * b_h; ++r)
        for (i32 j = row_ptr[r]; j < row_ptr[r + 1]; ++j)
            buf[n++] = (i32)(col[j] / b_w);
    qsort(buf, (size_t)n, sizeof(i32), cmp_i32);
    i64 u = 0;
    for (i64 i = 0; i < n; ++i)
        if (u == 0 || buf[u - 1] != buf[i]) buf[u++] = buf[i];
    return u;
}

i64 oracle_crs_to_bcrs_nblocks(i64 n_rows, i64 n_cols, const i32 *row_ptr,
                               const i32 *col, i64 b_h, i64 b_w, i64 h_pad,
                               i64 w_pad) {
    if (b_h <= 0 || b_w <= 0 || n_rows % b_h || n_cols % b_w || h_pad < b_h ||
        w_pad < b_w)
        return -1;
    i64 nbr = n_rows / b_h, total = 0;
    i32 *buf = (i32 *)malloc(sizeof(i32) * (size_t)(row_ptr[n_rows] + 1));
    for (i64 br = 0; br < nbr; ++br)
        total += block_row_cols(row_ptr, col, br, b_h, b_w, buf);
    free(buf);
    return total;
}

int oracle_crs_to_bcrs(i64 n_rows, i64 n_cols, const i32 *row_ptr,
                       const i32 *col, const void *val, i64 val_bytes, i64 b_h,
                       i64 b_w, i64 h_pad, i64 w_pad, int col_major,
                       i32 *bcol, i32 *brow_ptr, void *bval) {
    if (b_h <= 0 || b_w <= 0 || n_rows % b_h || n_cols % b_w || h_pad < b_h ||
        w_pad < b_w)
        return 1;
    i64 nbr = n_rows / b_h, nb = 0;
    i32 *buf = (i32 *)malloc(sizeof(i32) * (size_t)(row_ptr[n_rows] + 1));
    brow_ptr[0] = 0;
    for (i64 br = 0; br < nbr; ++br) {
        i64 u = block_row_cols(row_ptr, col, br, b_h, b_w, buf);
        memcpy(bcol + nb, buf, sizeof(i32) * (size_t)u);
        nb += u;
        brow_ptr[br + 1] = (i32)nb;
    }
    free(buf);
    i64 bs = h_pad * w_pad;
    memset(bval, 0, (size_t)(nb * bs * val_bytes));
    for (i64 br = 0; br < nbr; ++br) {
        for (i64 lr = 0; lr < b_h; ++lr) {
            i64 r = br * b_h + lr, bi = brow_ptr[br];
            for (i32 j = row_ptr[r]; j < row_ptr[r + 1]; ++j) {
                i32 bc = col[j] / (i32)b_w;
                while (bcol[bi] < bc) ++bi;
                /* convert_formats.hpp:230-234 throws here (columns of a row
                 * must be ascending) */
                if (bcol[bi] != bc) return 2;
                i64 inc = col_major ? h_pad * (col[j] % b_w) + lr
                                    : w_pad * lr + (col[j] % b_w);
                memcpy((char *)bval + (bi * bs + inc) * val_bytes,
                       (const char *)val + (i64)j * val_bytes,
                       (size_t)val_bytes);
            }
        }
    }
    return 0;
}

/* ================================================== adjoint helpers */

/* Explicit A^T in CRS (rows of A^T ascending in the original row index). */
void oracle_crs_transpose(i64 n_rows, i64 n_cols, const i32 *row_ptr,
                          const i32 *col, const void *val, i64 val_bytes,
                          i32 *t_row_ptr, i32 *t_col, void *t_val) {
    memset(t_row_ptr, 0, sizeof(i32) * (size_t)(n_cols + 1));
    i64 nnz = row_ptr[n_rows];
    for (i64 j = 0; j < nnz; ++j) t_row_ptr[col[j] + 1]++;
    for (i64 c = 0; c < n_cols; ++c) t_row_ptr[c + 1] += t_row_ptr[c];
    i32 *fill = (i32 *)malloc(sizeof(i32) * (size_t)(n_cols + 1));
    memcpy(fill, t_row_ptr, sizeof(i32) * (size_t)(n_cols + 1));
    for (i64 r = 0; r < n_rows; ++r) {
        for (i32 j = row_ptr[r]; j < row_ptr[r + 1]; ++j) {
            i32 dst = fill[col[j]]++;
            t_col[dst] = (i32)r;
            memcpy((char *)t_val + (i64)dst * val_bytes,
                   (const char *)val + (i64)j * val_bytes, (size_t)val_bytes);
        }
    }
    free(fill);
}

/* dvals[j] = sum_v G[row_j*k+v] * X[col_j*k+v], sequential FP64. */
void oracle_sddmm_f64(i64 n_rows, const i32 *row_ptr, const i32 *col, i64 k,
                      const double *G, const double *X, double *dvals) {
#pragma omp parallel for schedule(static)
    for (i64 r = 0; r < n_rows; ++r)
        for (i32 j = row_ptr[r]; j < row_ptr[r + 1]; ++j) {
            double acc = 0.0;
            for (i64 v = 0; v < k; ++v)
                acc = fma(G[r * k + v], X[(i64)col[j] * k + v], acc);
            dvals[j] = acc;
        }
}
