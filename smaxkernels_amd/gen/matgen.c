/*
 * Seeded synthetic workload generators for the SpMV path.
 *
 * Every generator is a pure function of (shape, seed): integer hashing
 * (splitmix64 finaliser, Lemire bounded integers, Floyd sampling) plus
 * correctly rounded IEEE-754 +, *, / and sqrt only, so a given seed gives the
 * same matrix on any x86-64 host and in any thread count.
 *
 * Matrices:
 *   poisson2d   2-D 5-point Laplacian, the exact row order of the reference's
 *               create2DPoissonMatrixCRS (examples/applications/
 *               applications_common.hpp:61-111): left, bottom, diag, top, right.
 *   lomax_band  rows of length L = min(cap, 1 + floor(lambda*(U^(-2/3) - 1)))
 *               (Lomax, shape 1.5), columns L distinct uniform integers in a
 *               band [i - half, i + half] clipped to [0, N), sorted ascending.
 *   stencil27   3-D 27-point stencil on an n^3 grid, columns ascending.
 *   block7      scalar CRS of a 7-point block adjacency on an n^3 block grid
 *               with b x b dense blocks, diagonal blocks shifted by +shift*I.
 *
 * This file is input synthesis for benchmarks and tests, not a kernel.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>

#define GOLDEN 0x9E3779B97F4A7C15ull

static inline uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

/* i-th output of a splitmix64 stream started at `seed` (counter-based). */
static inline uint64_t hash_at(uint64_t seed, uint64_t i) {
    return mix64(seed + (i + 1) * GOLDEN);
}

static inline uint64_t next64(uint64_t *s) {
    *s += GOLDEN;
    return mix64(*s);
}

/* Uniform integer in [0, range) (Lemire's nearly-divisionless method). */
static inline uint64_t bounded(uint64_t *s, uint64_t range) {
    uint64_t x = next64(s);
    __uint128_t m = (__uint128_t)x * range;
    uint64_t l = (uint64_t)m;
    if (l < range) {
        uint64_t t = (0 - range) % range;
        while (l < t) {
            x = next64(s);
            m = (__uint128_t)x * range;
            l = (uint64_t)m;
        }
    }
    return (uint64_t)(m >> 64);
}

/* Uniform double in [-1, 1) from the top 53 bits: exact, no rounding. */
static inline double unit_f64(uint64_t h) {
    return (double)(h >> 11) * (2.0 / 9007199254740992.0) - 1.0;
}

static inline uint64_t row_stream(uint64_t seed, uint64_t row) {
    return mix64(seed ^ mix64(row + 0x632BE59BD9B4E019ull));
}

/* ---------------------------------------------------------------- vectors */

void smax_gen_uniform_f64(uint64_t seed, int64_t n, double *out) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i)
        out[i] = unit_f64(hash_at(seed, (uint64_t)i));
}

/* Uniform float in [-1, 1) from the top 24 bits (exact in binary32). */
void smax_gen_uniform_f32(uint64_t seed, int64_t n, float *out) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i)
        out[i] = (float)((double)(hash_at(seed, (uint64_t)i) >> 40) *
                             (2.0 / 16777216.0) -
                         1.0);
}

/* ------------------------------------------------------------- poisson2d */

int64_t smax_gen_poisson2d_nnz(int64_t n) { return 5 * n * n - 4 * n; }

void smax_gen_poisson2d(int64_t n, int32_t *row_ptr, int32_t *col,
                        double *val) {
    int64_t k = 0;
    row_ptr[0] = 0;
    for (int64_t i = 0; i < n; ++i) {
        for (int64_t j = 0; j < n; ++j) {
            int64_t row = i * n + j;
            if (j > 0) { val[k] = -1.0; col[k++] = (int32_t)(row - 1); }
            if (i > 0) { val[k] = -1.0; col[k++] = (int32_t)(row - n); }
            val[k] = 4.0; col[k++] = (int32_t)row;
            if (i < n - 1) { val[k] = -1.0; col[k++] = (int32_t)(row + n); }
            if (j < n - 1) { val[k] = -1.0; col[k++] = (int32_t)(row + 1); }
            row_ptr[row + 1] = (int32_t)k;
        }
    }
}

/* ------------------------------------------------------------ lomax_band */

static void band_of(int64_t i, int64_t N, int64_t half, int64_t *lo,
                    int64_t *hi) {
    *lo = i - half < 0 ? 0 : i - half;
    *hi = i + half > N - 1 ? N - 1 : i + half;
}

/*
 * Row lengths.  L >= k  <=>  u <= t_k^(-3/2) with t_k = 1 + (k-1)/lambda and
 * u in (0, 1]; the thresholds use only +, /, *, sqrt (correctly rounded).
 * Fills row_ptr[0..N] and returns nnz (or -1 if nnz overflows int32).
 */
int64_t smax_gen_lomax_rowptr(int64_t N, uint64_t seed, double lambda,
                              int32_t cap, int64_t half, int32_t *row_ptr) {
    double *thr = (double *)malloc(sizeof(double) * ((size_t)cap + 1));
    thr[0] = 2.0;
    for (int32_t k = 1; k <= cap; ++k) {
        double t = 1.0 + (double)(k - 1) / lambda;
        thr[k] = 1.0 / (t * sqrt(t));
    }
    int64_t *len = (int64_t *)malloc(sizeof(int64_t) * (size_t)N);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < N; ++i) {
        uint64_t s = row_stream(seed, (uint64_t)i);
        double u = (double)((next64(&s) >> 11) + 1) / 9007199254740992.0;
        int32_t a = 1, b = cap; /* largest k in [1,cap] with u <= thr[k] */
        while (a < b) {
            int32_t m = a + (b - a + 1) / 2;
            if (u <= thr[m]) a = m; else b = m - 1;
        }
        int64_t lo, hi;
        band_of(i, N, half, &lo, &hi);
        int64_t L = a;
        if (L > hi - lo + 1) L = hi - lo + 1;
        len[i] = L;
    }
    int64_t acc = 0;
    row_ptr[0] = 0;
    for (int64_t i = 0; i < N; ++i) {
        acc += len[i];
        if (acc > 2147483647LL) { free(len); free(thr); return -1; }
        row_ptr[i + 1] = (int32_t)acc;
    }
    free(len);
    free(thr);
    return acc;
}

/* Columns: Floyd's sampling of L distinct offsets in the row's band. */
void smax_gen_lomax_cols(int64_t N, uint64_t seed, int64_t half,
                         const int32_t *row_ptr, int32_t *col) {
#pragma omp parallel
    {
        int64_t capacity = 0;
        int64_t *set = NULL;
#pragma omp for schedule(dynamic, 1024)
        for (int64_t i = 0; i < N; ++i) {
            int64_t L = row_ptr[i + 1] - row_ptr[i];
            if (L > capacity) {
                free(set);
                capacity = L;
                set = (int64_t *)malloc(sizeof(int64_t) * (size_t)capacity);
            }
            uint64_t s = row_stream(seed ^ 0xC0FFEEull, (uint64_t)i);
            int64_t lo, hi;
            band_of(i, N, half, &lo, &hi);
            int64_t n = hi - lo + 1, cnt = 0;
            for (int64_t j = n - L; j < n; ++j) {
                int64_t t = (int64_t)bounded(&s, (uint64_t)(j + 1));
                /* insert t (or j if t is taken) into the sorted set */
                int64_t v = t;
                for (int pass = 0; pass < 2; ++pass) {
                    int64_t a = 0, b = cnt;
                    while (a < b) {
                        int64_t m = (a + b) / 2;
                        if (set[m] < v) a = m + 1; else b = m;
                    }
                    if (a < cnt && set[a] == v) { v = j; continue; }
                    memmove(set + a + 1, set + a,
                            sizeof(int64_t) * (size_t)(cnt - a));
                    set[a] = v;
                    ++cnt;
                    break;
                }
            }
            int32_t *out = col + row_ptr[i];
            for (int64_t k = 0; k < L; ++k) out[k] = (int32_t)(lo + set[k]);
        }
        free(set);
    }
}

/* Values for nnz index j: uniform [-1,1) keyed by j. */
void smax_gen_values_f64(uint64_t seed, int64_t nnz, double *val) {
    smax_gen_uniform_f64(seed, nnz, val);
}

/* -------------------------------------------------------------- stencil27 */

int64_t smax_gen_stencil27_nnz(int64_t n) {
    int64_t m = 3 * n - 2; /* neighbours summed over one axis */
    return m * m * m;
}

void smax_gen_stencil27(int64_t n, double diag, int32_t *row_ptr,
                        int32_t *col, double *val) {
    int64_t N = n * n * n;
    /* row lengths are a product of per-axis counts: build row_ptr first */
    row_ptr[0] = 0;
    for (int64_t r = 0; r < N; ++r) {
        int64_t z = r / (n * n), y = (r / n) % n, x = r % n;
        int64_t cz = 3 - (z == 0) - (z == n - 1);
        int64_t cy = 3 - (y == 0) - (y == n - 1);
        int64_t cx = 3 - (x == 0) - (x == n - 1);
        row_ptr[r + 1] = (int32_t)(row_ptr[r] + cz * cy * cx);
    }
#pragma omp parallel for schedule(static)
    for (int64_t r = 0; r < N; ++r) {
        int64_t z = r / (n * n), y = (r / n) % n, x = r % n;
        int64_t k = row_ptr[r];
        for (int64_t dz = -1; dz <= 1; ++dz) {
            if (z + dz < 0 || z + dz >= n) continue;
            for (int64_t dy = -1; dy <= 1; ++dy) {
                if (y + dy < 0 || y + dy >= n) continue;
                for (int64_t dx = -1; dx <= 1; ++dx) {
                    if (x + dx < 0 || x + dx >= n) continue;
                    col[k] = (int32_t)((z + dz) * n * n + (y + dy) * n + x + dx);
                    val[k] = (dz == 0 && dy == 0 && dx == 0) ? diag : -1.0;
                    ++k;
                }
            }
        }
    }
}

/* ----------------------------------------------------------------- block7 */

static int64_t block7_count(int64_t n, int64_t br) {
    int64_t z = br / (n * n), y = (br / n) % n, x = br % n;
    return 1 + (z > 0) + (y > 0) + (x > 0) + (x < n - 1) + (y < n - 1) +
           (z < n - 1);
}

int64_t smax_gen_block7_nnz(int64_t n, int64_t b) {
    int64_t nb = 7 * n * n * n - 6 * n * n;
    return nb * b * b;
}

/*
 * Scalar CRS (N = b*n^3 rows) whose b x b block pattern is the 7-point
 * adjacency of an n^3 block grid.  Entries uniform [-1,1) keyed by nnz index,
 * +shift on the diagonal of diagonal blocks.  Columns ascending.
 */
void smax_gen_block7(int64_t n, int64_t b, uint64_t seed, double shift,
                     int32_t *row_ptr, int32_t *col, double *val) {
    int64_t NB = n * n * n;
    row_ptr[0] = 0;
    for (int64_t br = 0; br < NB; ++br) {
        int64_t len = block7_count(n, br) * b;
        for (int64_t lr = 0; lr < b; ++lr)
            row_ptr[br * b + lr + 1] = (int32_t)(row_ptr[br * b + lr] + len);
    }
#pragma omp parallel for schedule(static)
    for (int64_t br = 0; br < NB; ++br) {
        int64_t z = br / (n * n), y = (br / n) % n, x = br % n;
        int64_t nbr[7], c = 0;
        if (z > 0) nbr[c++] = br - n * n;
        if (y > 0) nbr[c++] = br - n;
        if (x > 0) nbr[c++] = br - 1;
        nbr[c++] = br;
        if (x < n - 1) nbr[c++] = br + 1;
        if (y < n - 1) nbr[c++] = br + n;
        if (z < n - 1) nbr[c++] = br + n * n;
        for (int64_t lr = 0; lr < b; ++lr) {
            int64_t k = row_ptr[br * b + lr];
            for (int64_t t = 0; t < c; ++t) {
                for (int64_t lc = 0; lc < b; ++lc) {
                    col[k] = (int32_t)(nbr[t] * b + lc);
                    double v = unit_f64(hash_at(seed, (uint64_t)k));
                    if (nbr[t] == br && lr == lc) v += shift;
                    val[k] = v;
                    ++k;
                }
            }
        }
    }
}

/* ------------------------------------------------------------ conversions */

/* Round-to-nearest-even float -> bfloat16 bits (NaN kept quiet). */
void smax_gen_f64_to_bf16(int64_t n, const double *in, uint16_t *out) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        float f = (float)in[i];
        uint32_t u;
        memcpy(&u, &f, 4);
        if ((u & 0x7fffffffu) > 0x7f800000u) {
            out[i] = (uint16_t)((u >> 16) | 0x40u);
        } else {
            uint32_t r = 0x7fffu + ((u >> 16) & 1u);
            out[i] = (uint16_t)((u + r) >> 16);
        }
    }
}

void smax_gen_f64_to_f32(int64_t n, const double *in, float *out) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) out[i] = (float)in[i];
}
