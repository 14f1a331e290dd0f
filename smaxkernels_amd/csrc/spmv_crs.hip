// CRS SpMV and SpMM for gfx950.
//
// Every output is the reference's sequential row sum
// (spmv_cpu_crs_impl.hpp:20-33, spmm_cpu_crs_impl.hpp:22-44) computed as one
// fma chain in CRS order, so results are bit-identical to the reference CPU
// kernels built with FMA contraction.  Two SpMV schedules:
//
//  * sequential: one lane per row reading straight from HBM (the reference's
//    CUDA schedule, spmv_cuda_crs_impl.cu:5-22);
//  * blocked: a plan cuts the rows into row blocks of at most T rows whose
//    first nonzeros fall in one W-entry window; rows of 512 or more nonzeros
//    form a block of their own.  Each workgroup of T threads streams its
//    block's values and column indices through LDS in coalesced W-entry tiles
//    (gathering x on the way), and each row's lane then runs its fma chain out
//    of LDS.  Short rows are thereby reduced lane-per-row while the loads stay
//    fully coalesced; a long row is streamed by its whole workgroup.  Row
//    blocks are dealt to the XCDs in contiguous ranges (xcd_block), so the x
//    band of neighbouring blocks is shared in one L2.  Two geometries:
//    T = 256, W = 1024 for rows of even length, and single-wave blocks,
//    T = 64, W = 512, when the longest row exceeds 4x the mean: a wave then
//    waits at its own tiles' barriers only (4M-row band matrix, FP64:
//    0.692 ms against 0.766; BF16 values 0.512 against 0.692; the 512^2
//    Laplacian and a 27-point stencil run 6.0 / 25.5 us with the wide blocks
//    against 6.5 / 28.2 us).
#include <algorithm>
#include <cmath>
#include <mutex>
#include <vector>

#include "common.hpp"
#include "scan.hpp"

struct smax_hip_crs_plan_s {
    int64_t n_rows = 0, nnz = 0, n_blocks = 0;
    int32_t block_threads = 256;     // row-block geometry: 256 (wide) or 64 (narrow)
    int32_t *block_start = nullptr;  // n_blocks + 1 row indices
    int32_t *row_order = nullptr;    // row-major SpMM execution order (n_rows)
    int64_t spmm_lds = 0;            // dynamic LDS per SpMM workgroup (occupancy cap)
};

namespace smaxhip {

constexpr int kCrsThreads = 256;
constexpr int kCrsLongRow = 512;     // rows at least this long stream alone
// row-block geometries: threads per block (= rows at most) and the tile and
// block-start window in nonzeros
constexpr int kBlkWideThreads = 256, kBlkWideTile = 1024;
constexpr int kBlkNarrowThreads = 64, kBlkNarrowTile = 512;

// ------------------------------------------------------------ sequential
template <int VT>
__global__ __launch_bounds__(kCrsThreads) void crs_spmv_lane(
    int64_t n_rows, const int32_t *__restrict__ row_ptr,
    const int32_t *__restrict__ col,
    const typename vt<VT>::val_t *__restrict__ val,
    const typename vt<VT>::vec_t *__restrict__ x,
    typename vt<VT>::vec_t *__restrict__ y) {
    using T = vt<VT>;
    const int64_t r = static_cast<int64_t>(blockIdx.x) * kCrsThreads + threadIdx.x;
    if (r >= n_rows) return;
    const int64_t s = row_ptr[r], e = row_ptr[r + 1];
    typename T::vec_t acc = 0;
    int64_t j = s;
    constexpr int U = 4;
    for (; j + U <= e; j += U) {
        int32_t c[U];
        typename T::val_t v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            c[u] = col[j + u];
            v[u] = val[j + u];
        }
        typename T::vec_t xv[U];
#pragma unroll
        for (int u = 0; u < U; ++u) xv[u] = x[c[u]];
#pragma unroll
        for (int u = 0; u < U; ++u) acc = T::fma_(T::widen(v[u]), xv[u], acc);
    }
    for (; j < e; ++j) acc = T::fma_(T::widen(val[j]), x[col[j]], acc);
    y[r] = acc;
}

// ------------------------------------------------------------ blocked
template <int VT, int NT, int TILE>
__global__ __launch_bounds__(NT) void crs_spmv_blocked(
    const int32_t *__restrict__ block_start,
    const int32_t *__restrict__ row_ptr, const int32_t *__restrict__ col,
    const typename vt<VT>::val_t *__restrict__ val,
    const typename vt<VT>::vec_t *__restrict__ x,
    typename vt<VT>::vec_t *__restrict__ y) {
    using T = vt<VT>;
    using A = typename T::vec_t;
    __shared__ A sv[TILE];
    __shared__ A sx[TILE];
    const int tid = threadIdx.x;
    const int64_t b = xcd_block(blockIdx.x, gridDim.x);
    const int64_t r0 = block_start[b], r1 = block_start[b + 1];
    const int64_t s = row_ptr[r0], e = row_ptr[r1];
    const int64_t row = r0 + tid;
    const bool has = row < r1;
    const int64_t rs = has ? row_ptr[row] : 0, re = has ? row_ptr[row + 1] : 0;
    A acc = 0;
    constexpr int PER = TILE / NT;
    for (int64_t t0 = s; t0 < e; t0 += TILE) {
        const int64_t nt = e - t0 < TILE ? e - t0 : TILE;
        int32_t c[PER];
        typename T::val_t v[PER];
#pragma unroll
        for (int u = 0; u < PER; ++u) {
            const int64_t k = tid + u * NT;
            if (k < nt) {
                c[u] = ld_stream(col + t0 + k);
                v[u] = ld_stream(val + t0 + k);
            }
        }
#pragma unroll
        for (int u = 0; u < PER; ++u) {
            const int64_t k = tid + u * NT;
            if (k < nt) {
                sx[k] = x[c[u]];
                sv[k] = T::widen(v[u]);
            }
        }
        __syncthreads();
        if (has) {
            const int64_t lo = rs > t0 ? rs : t0;
            const int64_t hi = re < t0 + nt ? re : t0 + nt;
            for (int64_t j = lo; j < hi; ++j)
                acc = T::fma_(sv[j - t0], sx[j - t0], acc);
        }
        __syncthreads();
    }
    if (has) y[row] = acc;
}

// A new block starts at row r when r is a multiple of nt (the block's
// threads), when r's first nonzero enters a new window of `window` entries,
// or when r or r-1 is a long row.
__global__ void crs_block_flags(int64_t n_rows, const int32_t *__restrict__ row_ptr,
                                int32_t *__restrict__ flag, int32_t nt, int32_t window) {
    const int64_t r = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (r >= n_rows) return;
    int32_t f = 1;
    if (r % nt != 0) {
        const int32_t a = row_ptr[r - 1], b = row_ptr[r], c = row_ptr[r + 1];
        f = (a / window != b / window) || (b - a >= kCrsLongRow) || (c - b >= kCrsLongRow);
    }
    flag[r] = f;
}

__global__ void crs_block_starts(int64_t n_rows,
                                 const int32_t *__restrict__ flag,
                                 const int32_t *__restrict__ pos,
                                 int32_t *__restrict__ block_start) {
    const int64_t r = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (r >= n_rows) return;
    if (flag[r]) block_start[pos[r] - 1] = static_cast<int32_t>(r);
    if (r == n_rows - 1) block_start[pos[r]] = static_cast<int32_t>(n_rows);
}

// ------------------------------------------------------------ SpMM
// Thread (row, v) with v fastest for row-major operands (k lanes share one
// broadcast val/col load, X rows are read contiguously), row fastest for
// column-major ones (Y columns are written contiguously).
template <int VT, bool ROW_MAJOR>
__global__ __launch_bounds__(kCrsThreads) void crs_spmm_kernel(
    int64_t n_rows, int64_t k, const int32_t *__restrict__ order,
    const int32_t *__restrict__ row_ptr, const int32_t *__restrict__ col,
    const typename vt<VT>::val_t *__restrict__ val,
    const typename vt<VT>::vec_t *__restrict__ X, int64_t ldx,
    typename vt<VT>::vec_t *__restrict__ Y, int64_t ldy) {
    using T = vt<VT>;
    const int64_t t = static_cast<int64_t>(blockIdx.x) * kCrsThreads + threadIdx.x;
    int64_t r, v;
    if (ROW_MAJOR) {
        r = t / k;
        v = t - r * k;
    } else {
        v = t / n_rows;
        r = t - v * n_rows;
    }
    if (r >= n_rows || v >= k) return;
    if (ROW_MAJOR && order) r = order[r];
    const int64_t s = row_ptr[r], e = row_ptr[r + 1];
    typename T::vec_t acc = 0;
    int64_t j = s;
    constexpr int U = 4;
    for (; j + U <= e; j += U) {
        int32_t c[U];
        typename T::val_t w[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            c[u] = col[j + u];
            w[u] = val[j + u];
        }
        typename T::vec_t xv[U];
#pragma unroll
        for (int u = 0; u < U; ++u)
            xv[u] = ROW_MAJOR ? X[static_cast<int64_t>(c[u]) * ldx + v]
                              : X[v * ldx + c[u]];
#pragma unroll
        for (int u = 0; u < U; ++u) acc = T::fma_(T::widen(w[u]), xv[u], acc);
    }
    for (; j < e; ++j) {
        const typename T::vec_t xv =
            ROW_MAJOR ? X[static_cast<int64_t>(col[j]) * ldx + v] : X[v * ldx + col[j]];
        acc = T::fma_(T::widen(val[j]), xv, acc);
    }
    if (ROW_MAJOR) Y[r * ldy + v] = acc;
    else Y[v * ldy + r] = acc;
}

// Row-major SpMM with k = K a power of two and 16-byte aligned X/Y rows: a
// lane owns Q = 16 / sizeof(vec) consecutive columns v (K / Q lanes per row),
// reads its part of each X row with one 16-byte load and keeps Q fma chains;
// the chain is software-pipelined U elements deep.  Every Y[r, v] is the
// same sequential chain over the row as in crs_spmm_kernel.
template <int VT, int K>
__global__ __launch_bounds__(kCrsThreads) void crs_spmm_rowvec(
    int64_t n_rows, const int32_t *__restrict__ order,
    const int32_t *__restrict__ row_ptr, const int32_t *__restrict__ col,
    const typename vt<VT>::val_t *__restrict__ val,
    const typename vt<VT>::vec_t *__restrict__ X, int64_t ldx,
    typename vt<VT>::vec_t *__restrict__ Y, int64_t ldy) {
    using T = vt<VT>;
    using A = typename T::vec_t;
    using V = typename T::val_t;
    constexpr int Q = 16 / sizeof(A);
    constexpr int GL = K / Q;
    constexpr int U = 8;
    static_assert(GL * Q == K, "K must be a multiple of the run length");
    const int64_t t = static_cast<int64_t>(blockIdx.x) * kCrsThreads + threadIdx.x;
    int64_t r = t / GL;
    const int l = static_cast<int>(t % GL);
    if (r >= n_rows) return;
    if (order) r = order[r];
    const int64_t s = row_ptr[r], e = row_ptr[r + 1];
    A acc[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q) acc[q] = 0;
    int64_t j = s;
    if constexpr ((GL == U || GL == 2 * U) && sizeof(V) <= 4) {
        // 8 or 16 lanes per row (k = 32 or 64 with 4-byte X entries): lane l
        // loads entry l % 8 of each group of 8 and a swizzle within the 8
        // lanes (a DPP row broadcast within the 16) hands every lane all of
        // them, instead of every lane loading each; the next group's are
        // requested before this group's X rows
        if (e - s >= U) {
            const int le = l % U;  // the entry of each group this lane loads
            int cm = col[j + le];
            int wm = val_bits<V>(val[j + le]);
            for (;;) {
                const bool more = j + 2 * U <= e;
                int cn = 0, wn = 0;
                if (more) {
                    cn = col[j + U + le];
                    wn = val_bits<V>(val[j + U + le]);
                }
                int cs[U], ws[U];
                if constexpr (GL == U)
                    grp8_bcast8(cm, wm, cs, ws, std::make_integer_sequence<int, U>{});
                else
                    row16_bcast8(cm, wm, cs, ws, std::make_integer_sequence<int, U>{});
                A xv[U][Q];
#pragma unroll
                for (int u = 0; u < U; ++u)
                    load_run<A, Q>(X + static_cast<int64_t>(cs[u]) * ldx + l * Q, xv[u]);
#pragma unroll
                for (int u = 0; u < U; ++u)
#pragma unroll
                    for (int q = 0; q < Q; ++q)
                        acc[q] = T::fma_(T::widen(bits_val<V>(ws[u])), xv[u][q], acc[q]);
                j += U;
                if (!more) break;
                cm = cn;
                wm = wn;
            }
        }
    } else if (e - s >= U) {
        int32_t c0[U];
        V w0[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            c0[u] = col[j + u];
            w0[u] = val[j + u];
        }
        for (;;) {
            const bool more = j + 2 * U <= e;
            int32_t c1[U];
            V w1[U];
            if (more) {
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    c1[u] = col[j + U + u];
                    w1[u] = val[j + U + u];
                }
            }
            A xv[U][Q];
#pragma unroll
            for (int u = 0; u < U; ++u)
                load_run<A, Q>(X + static_cast<int64_t>(c0[u]) * ldx + l * Q, xv[u]);
#pragma unroll
            for (int u = 0; u < U; ++u)
#pragma unroll
                for (int q = 0; q < Q; ++q) acc[q] = T::fma_(T::widen(w0[u]), xv[u][q], acc[q]);
            j += U;
            if (!more) break;
#pragma unroll
            for (int u = 0; u < U; ++u) {
                c0[u] = c1[u];
                w0[u] = w1[u];
            }
        }
    }
    for (; j < e; ++j) {
        A xv[Q];
        load_run<A, Q>(X + static_cast<int64_t>(col[j]) * ldx + l * Q, xv);
        const A w = T::widen(val[j]);
#pragma unroll
        for (int q = 0; q < Q; ++q) acc[q] = T::fma_(w, xv[q], acc[q]);
    }
    store_run<A, Q>(Y + r * ldy + l * Q, acc);
}

// Launches crs_spmm_rowvec when k is a power of two in [4Q, 64] and X, Y and
// their row strides are 16-byte aligned; false otherwise.
template <int VT>
static bool launch_crs_spmm_rowvec(smax_hip_ctx ctx, int64_t n_rows, int64_t k,
                                   const int32_t *order, unsigned lds, const int32_t *row_ptr,
                                   const int32_t *col,
                                   const void *val, const void *X, int64_t ldx, void *Y,
                                   int64_t ldy) {
    using T = vt<VT>;
    using A = typename T::vec_t;
    constexpr int64_t Q = 16 / sizeof(A);
    // at least 4 lanes per row (as for the BCRS SpMM, where 2 lost time)
    if (k < 4 * Q || k > 64 || (k & (k - 1)) != 0 || !aligned16(X, Y) ||
        (ldx * sizeof(A)) % 16 != 0 || (ldy * sizeof(A)) % 16 != 0)
        return false;
    const dim3 g = grid_for(n_rows * (k / Q), kCrsThreads);
    auto v = static_cast<const typename T::val_t *>(val);
    auto xv = static_cast<const A *>(X);
    auto yv = static_cast<A *>(Y);
#define SMAX_CRS_ROWVEC(KK)                                                                  \
    case KK:                                                                                  \
        if constexpr (KK >= Q)                                                                \
            hipLaunchKernelGGL((crs_spmm_rowvec<VT, KK>), g, dim3(kCrsThreads), lds,            \
                               ctx->stream, n_rows, order, row_ptr, col, v, xv, ldx, yv, ldy); \
        return true;
    switch (k) {
        SMAX_CRS_ROWVEC(2)
        SMAX_CRS_ROWVEC(4)
        SMAX_CRS_ROWVEC(8)
        SMAX_CRS_ROWVEC(16)
        SMAX_CRS_ROWVEC(32)
        SMAX_CRS_ROWVEC(64)
    default:
        return false;
    }
#undef SMAX_CRS_ROWVEC
}

// Row-major SpMM execution order: every thread group runs one row's chains,
// so a wave lasts as long as its longest row.  Rows of at least
// max(128, 8 x mean) entries start first, longest first (they set the
// critical path); the rest follow window by window (kSpmmWindow rows, so a
// window's X band stays together), longest first inside each window so that
// the rows sharing a wave have similar lengths.  Results do not depend on
// the order.
constexpr int64_t kSpmmWindow = 65536;
constexpr int64_t kSpmmCapLds = 53 * 1024;

static int build_row_order(smax_hip_ctx ctx, int64_t n_rows, int64_t nnz,
                           const std::vector<int32_t> &rp, int32_t **out, int64_t *n_long) {
    *out = nullptr;
    *n_long = 0;
    const double mean = static_cast<double>(nnz) / static_cast<double>(n_rows);
    const int64_t thr =
        std::max<int64_t>(128, static_cast<int64_t>(std::ceil(8 * mean)));
    auto len = [&](int32_t r) { return rp[r + 1] - rp[r]; };
    std::vector<int32_t> ord;
    ord.reserve(static_cast<size_t>(n_rows));
    for (int64_t r = 0; r < n_rows; ++r)
        if (len(static_cast<int32_t>(r)) >= thr) ord.push_back(static_cast<int32_t>(r));
    *n_long = static_cast<int64_t>(ord.size());
    auto longer = [&](int32_t a, int32_t b) { return len(a) > len(b); };
    std::stable_sort(ord.begin(), ord.end(), longer);
    for (int64_t w = 0; w < n_rows; w += kSpmmWindow) {
        const size_t first = ord.size();
        for (int64_t r = w; r < std::min(n_rows, w + kSpmmWindow); ++r)
            if (len(static_cast<int32_t>(r)) < thr) ord.push_back(static_cast<int32_t>(r));
        std::stable_sort(ord.begin() + static_cast<std::ptrdiff_t>(first), ord.end(), longer);
    }
    int rc = dev_alloc(ctx, reinterpret_cast<void **>(out), sizeof(int32_t) * n_rows,
                       "crs plan row order");
    if (rc) return rc;
    SMAX_HIP_CALL(ctx, hipMemcpyAsync(*out, ord.data(), sizeof(int32_t) * n_rows,
                                      hipMemcpyHostToDevice, ctx->stream));
    SMAX_HIP_CALL(ctx, hipStreamSynchronize(ctx->stream));
    return SMAX_HIP_SUCCESS;
}

}  // namespace smaxhip

using namespace smaxhip;

namespace smaxhip {

// Dense rows x k block between column-major (element (r, v) at r + v*ld) and
// row-major (r*k + v) layouts, through a 32 x 33 LDS tile so both the read
// and the write are coalesced.  COL_IN: column-major in, row-major out.
template <typename T, bool COL_IN>
__global__ __launch_bounds__(256) void k_relayout(int64_t rows, int64_t k,
                                                  const T *__restrict__ in,
                                                  int64_t ld, T *__restrict__ out) {
    __shared__ T tile[32][33];
    const int64_t r0 = static_cast<int64_t>(blockIdx.x) * 32;
    const int64_t v0 = static_cast<int64_t>(blockIdx.y) * 32;
    const int tx = static_cast<int>(threadIdx.x) & 31, ty = static_cast<int>(threadIdx.x) >> 5;
#pragma unroll
    for (int j = 0; j < 32; j += 8) {
        // the read walks the unit-stride index of `in` with tx
        const int64_t r = COL_IN ? r0 + tx : r0 + ty + j;
        const int64_t v = COL_IN ? v0 + ty + j : v0 + tx;
        if (r < rows && v < k) {
            const T e = COL_IN ? in[r + v * ld] : in[r * k + v];
            if (COL_IN) tile[ty + j][tx] = e;   // tile[v - v0][r - r0]
            else tile[tx][ty + j] = e;          // tile[v - v0][r - r0]
        }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < 32; j += 8) {
        // the write walks the unit-stride index of `out` with tx
        const int64_t r = COL_IN ? r0 + ty + j : r0 + tx;
        const int64_t v = COL_IN ? v0 + tx : v0 + ty + j;
        if (r < rows && v < k) {
            if (COL_IN) out[r * k + v] = tile[tx][ty + j];
            else out[r + v * ld] = tile[ty + j][tx];
        }
    }
}

// The same relayout with 16-byte accesses on both sides, for operands whose
// unit-stride extents (rows and k) and leading dimension are multiples of
// E = 16 / sizeof(T) and whose pointers are 16-byte aligned: a 64-row x
// 32-column tile, the column-major side read or written as runs of E rows of
// one column, the row-major side as runs of E columns of one row.  The LDS
// tile is kept as [column][row] with a pitch of 65 elements, so the E-row
// runs of the column-major side and the E-column gathers of the row-major
// side both spread over the banks.
template <typename T, bool COL_IN>
__global__ __launch_bounds__(256) void k_relayout_vec(int64_t rows, int64_t k,
                                                      const T *__restrict__ in,
                                                      int64_t ld, T *__restrict__ out) {
    constexpr int E = 16 / static_cast<int>(sizeof(T));
    constexpr int TR = 64, TV = 32, P = TR + 1;
    constexpr int VPC = TR / E;   // runs per tile column (column-major side)
    constexpr int CPP = 256 / VPC;
    constexpr int VPR = TV / E;   // runs per tile row (row-major side)
    constexpr int RPP = 256 / VPR;
    using V = T __attribute__((ext_vector_type(E)));
    __shared__ T tile[TV * P];   // tile[(v - v0) * P + (r - r0)]
    const int64_t r0 = static_cast<int64_t>(blockIdx.x) * TR;
    const int64_t v0 = static_cast<int64_t>(blockIdx.y) * TV;
    const int t = static_cast<int>(threadIdx.x);
    if (COL_IN) {
#pragma unroll
        for (int p = 0; p < TV; p += CPP) {
            const int vc = p + t / VPC, rc = (t % VPC) * E;
            const int64_t r = r0 + rc, v = v0 + vc;
            if (r < rows && v < k) {
                const V e = *reinterpret_cast<const V *>(in + r + v * ld);
#pragma unroll
                for (int i = 0; i < E; ++i) tile[vc * P + rc + i] = e[i];
            }
        }
        __syncthreads();
#pragma unroll
        for (int p = 0; p < TR; p += RPP) {
            const int rr = p + t / VPR, vv = (t % VPR) * E;
            const int64_t r = r0 + rr, v = v0 + vv;
            if (r < rows && v < k) {
                V e;
#pragma unroll
                for (int i = 0; i < E; ++i) e[i] = tile[(vv + i) * P + rr];
                *reinterpret_cast<V *>(out + r * k + v) = e;
            }
        }
    } else {
#pragma unroll
        for (int p = 0; p < TR; p += RPP) {
            const int rr = p + t / VPR, vv = (t % VPR) * E;
            const int64_t r = r0 + rr, v = v0 + vv;
            if (r < rows && v < k) {
                const V e = *reinterpret_cast<const V *>(in + r * k + v);
#pragma unroll
                for (int i = 0; i < E; ++i) tile[(vv + i) * P + rr] = e[i];
            }
        }
        __syncthreads();
#pragma unroll
        for (int p = 0; p < TV; p += CPP) {
            const int vc = p + t / VPC, rc = (t % VPC) * E;
            const int64_t r = r0 + rc, v = v0 + vc;
            if (r < rows && v < k) {
                V e;
#pragma unroll
                for (int i = 0; i < E; ++i) e[i] = tile[vc * P + rc + i];
                *reinterpret_cast<V *>(out + r + v * ld) = e;
            }
        }
    }
}

// Column-major <-> row-major copy of a rows x k block (COL_IN: column-major
// `in` with leading dimension ld to row-major `out`; else the reverse), with
// 16-byte accesses where the extents and alignment allow.
template <typename T, bool COL_IN>
static void launch_relayout(smax_hip_ctx ctx, int64_t rows, int64_t k, const T *in, int64_t ld,
                            T *out) {
    constexpr int64_t E = 16 / sizeof(T);
    const bool vec = rows % E == 0 && k % E == 0 && ld % E == 0 &&
                     ((reinterpret_cast<uintptr_t>(in) | reinterpret_cast<uintptr_t>(out)) & 15) == 0;
    const unsigned gy = static_cast<unsigned>(ceil_div(k, 32));
    if (vec)
        hipLaunchKernelGGL((k_relayout_vec<T, COL_IN>),
                           dim3(static_cast<unsigned>(ceil_div(rows, 64)), gy), dim3(256), 0,
                           ctx->stream, rows, k, in, ld, out);
    else
        hipLaunchKernelGGL((k_relayout<T, COL_IN>),
                           dim3(static_cast<unsigned>(ceil_div(rows, 32)), gy), dim3(256), 0,
                           ctx->stream, rows, k, in, ld, out);
}

}  // namespace smaxhip

namespace smaxhip {

// Fixed staging area of smax_hip_crs_spmm_noalloc (row-major copies of a
// column block of X and of Y): a file-scope device array, so that entry
// point allocates nothing.  Each device holds its own copy for the life of
// the process, shared by every call on that device.  Host code reaches it
// only through the address hipGetSymbolAddress returns for the current
// device; the name itself is a host-side stand-in.
constexpr size_t kNoAllocScratchBytes = size_t(32) << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void *g_scratch[kMaxDevices];

// The scratch array's device address on ctx's device (current on entry),
// looked up once per device.
static int noalloc_scratch(smax_hip_ctx ctx, void **out) {
    *out = nullptr;
    if (ctx->device < 0 || ctx->device >= kMaxDevices)
        return fail(ctx, SMAX_HIP_ERROR_NO_DEVICE,
                    "smax_hip_crs_spmm_noalloc: device %d beyond %d", ctx->device, kMaxDevices);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[ctx->device]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return hip_fail(ctx, e != hipSuccess ? e : hipErrorInvalidSymbol,
                            "hipGetSymbolAddress(scratch)");
        g_scratch[ctx->device] = p;
    }
    *out = g_scratch[ctx->device];
    return SMAX_HIP_SUCCESS;
}

}  // namespace smaxhip

extern "C" {

int smax_hip_crs_plan_create(smax_hip_ctx ctx, int64_t n_rows, int64_t nnz,
                             const int32_t *row_ptr, smax_hip_crs_plan *out) {
    SMAX_HIP_CHECK_CTX(ctx);
    SMAX_HIP_REQUIRE(ctx, out, "smax_hip_crs_plan_create: null output");
    *out = nullptr;
    SMAX_HIP_REQUIRE(ctx, n_rows > 0 && n_rows < INT32_MAX && nnz >= 0 &&
                              nnz <= INT32_MAX && row_ptr,
                     "smax_hip_crs_plan_create: bad arguments");
    // the row lengths decide the block geometry and the SpMM row order
    std::vector<int32_t> rp(static_cast<size_t>(n_rows + 1));
    SMAX_HIP_CALL(ctx, hipMemcpyAsync(rp.data(), row_ptr, sizeof(int32_t) * (n_rows + 1),
                                      hipMemcpyDeviceToHost, ctx->stream));
    SMAX_HIP_CALL(ctx, hipStreamSynchronize(ctx->stream));
    int64_t longest = 0;
    for (int64_t r = 0; r < n_rows; ++r)
        longest = std::max<int64_t>(longest, static_cast<int64_t>(rp[r + 1]) - rp[r]);
    const bool narrow = static_cast<double>(longest) >
                        4.0 * static_cast<double>(nnz) / static_cast<double>(n_rows);
    const int32_t nt = narrow ? kBlkNarrowThreads : kBlkWideThreads;
    const int32_t window = narrow ? kBlkNarrowTile : kBlkWideTile;
    int32_t *flag = nullptr, *tmp = nullptr, *starts = nullptr;
    int rc = dev_alloc(ctx, reinterpret_cast<void **>(&flag),
                       sizeof(int32_t) * n_rows, "crs plan");
    if (!rc)
        rc = dev_alloc(ctx, reinterpret_cast<void **>(&tmp),
                       sizeof(int32_t) * scan_tmp_elems(n_rows), "crs plan");
    if (!rc)
        rc = dev_alloc(ctx, reinterpret_cast<void **>(&starts),
                       sizeof(int32_t) * (n_rows + 1), "crs plan");
    int64_t n_blocks = 0;
    if (!rc) {
        hipLaunchKernelGGL(crs_block_flags, grid_for(n_rows, 256), dim3(256), 0,
                           ctx->stream, n_rows, row_ptr, flag, nt, window);
        rc = check_launch(ctx, "crs_block_flags");
    }
    // flag -> inclusive positions (in place); tmp is the scan scratch.
    if (!rc) rc = scan_inclusive<int32_t, false>(ctx, flag, flag, n_rows, tmp);
    int32_t last = 0;
    if (!rc) {
        hipError_t e = hipMemcpyAsync(&last, flag + n_rows - 1, sizeof(int32_t),
                                      hipMemcpyDeviceToHost, ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) rc = hip_fail(ctx, e, "crs plan read-back");
        n_blocks = last;
    }
    if (!rc) {
        // Recompute the flags (the scan overwrote them) into tmp-sized reuse.
        int32_t *flag2 = nullptr;
        rc = dev_alloc(ctx, reinterpret_cast<void **>(&flag2),
                       sizeof(int32_t) * n_rows, "crs plan");
        if (!rc) {
            hipLaunchKernelGGL(crs_block_flags, grid_for(n_rows, 256), dim3(256),
                               0, ctx->stream, n_rows, row_ptr, flag2, nt, window);
            rc = check_launch(ctx, "crs_block_flags");
        }
        if (!rc) {
            hipLaunchKernelGGL(crs_block_starts, grid_for(n_rows, 256), dim3(256),
                               0, ctx->stream, n_rows, flag2, flag, starts);
            rc = check_launch(ctx, "crs_block_starts");
        }
        (void)hipStreamSynchronize(ctx->stream);
        dev_free(flag2);
    }
    dev_free(flag);
    dev_free(tmp);
    int32_t *order = nullptr;
    int64_t n_long = 0;
    if (!rc) rc = build_row_order(ctx, n_rows, nnz, rp, &order, &n_long);
    if (rc) {
        dev_free(starts);
        dev_free(order);
        return rc;
    }
    smax_hip_crs_plan p = new smax_hip_crs_plan_s();
    p->n_rows = n_rows;
    p->nnz = nnz;
    p->n_blocks = n_blocks;
    p->block_threads = nt;
    p->block_start = starts;
    p->row_order = order;
    // With long rows leading, three workgroups per CU (53 KB of LDS each out
    // of 160 KB) run the row-major SpMM faster than full occupancy: 0.46 ->
    // 0.40 ms on the 1M-row band matrix, k = 32.
    p->spmm_lds = n_long > 0 ? kSpmmCapLds : 0;
    *out = p;
    return SMAX_HIP_SUCCESS;
}

int smax_hip_crs_plan_destroy(smax_hip_ctx ctx, smax_hip_crs_plan plan) {
    if (!plan) return SMAX_HIP_SUCCESS;
    if (ctx) {
        (void)bind(ctx);
        (void)hipStreamSynchronize(ctx->stream);
    }
    dev_free(plan->block_start);
    dev_free(plan->row_order);
    delete plan;
    return SMAX_HIP_SUCCESS;
}

int smax_hip_crs_spmv(smax_hip_ctx ctx, int value_type, int64_t n_rows,
                      int64_t n_cols, int64_t nnz, const int32_t *row_ptr,
                      const int32_t *col, const void *val, const void *x,
                      void *y, smax_hip_crs_plan plan) {
    SMAX_HIP_CHECK_CTX(ctx);
    if (!valid_value_type(value_type))
        return fail(ctx, SMAX_HIP_ERROR_UNSUPPORTED,
                    "smax_hip_crs_spmv: value type %d", value_type);
    SMAX_HIP_REQUIRE(ctx, n_rows > 0 && n_cols > 0 && nnz >= 0 &&
                              nnz <= INT32_MAX,
                     "smax_hip_crs_spmv: bad shape");
    if (n_rows > INT32_MAX || n_cols > INT32_MAX)
        return fail(ctx, SMAX_HIP_ERROR_OVERFLOW, "smax_hip_crs_spmv: %lld x %lld exceeds int32",
                    (long long)n_rows, (long long)n_cols);
    SMAX_HIP_REQUIRE(ctx, row_ptr && x && y && (nnz == 0 || (col && val)),
                     "smax_hip_crs_spmv: null pointer");
    if (plan)
        SMAX_HIP_REQUIRE(ctx, plan->n_rows == n_rows && plan->nnz == nnz,
                         "smax_hip_crs_spmv: plan built for another pattern");
    SMAX_HIP_DISPATCH_VT(value_type, {
        using T = vt<VT>;
        auto v = static_cast<const typename T::val_t *>(val);
        auto xv = static_cast<const typename T::vec_t *>(x);
        auto yv = static_cast<typename T::vec_t *>(y);
        if (plan) {
            const dim3 g(static_cast<unsigned>(plan->n_blocks));
            if (plan->block_threads == kBlkNarrowThreads)
                hipLaunchKernelGGL((crs_spmv_blocked<VT, kBlkNarrowThreads, kBlkNarrowTile>), g,
                                   dim3(kBlkNarrowThreads), 0, ctx->stream, plan->block_start,
                                   row_ptr, col, v, xv, yv);
            else
                hipLaunchKernelGGL((crs_spmv_blocked<VT, kBlkWideThreads, kBlkWideTile>), g,
                                   dim3(kBlkWideThreads), 0, ctx->stream, plan->block_start,
                                   row_ptr, col, v, xv, yv);
            SMAX_HIP_LAUNCHED(ctx, "crs_spmv_blocked");
        } else {
            hipLaunchKernelGGL((crs_spmv_lane<VT>), grid_for(n_rows, kCrsThreads),
                               dim3(kCrsThreads), 0, ctx->stream, n_rows,
                               row_ptr, col, v, xv, yv);
            SMAX_HIP_LAUNCHED(ctx, "crs_spmv_lane");
        }
        return SMAX_HIP_SUCCESS;
    });
    return SMAX_HIP_ERROR_UNSUPPORTED;
}

int smax_hip_crs_spmm(smax_hip_ctx ctx, int value_type, int64_t n_rows,
                      int64_t n_cols, int64_t nnz, const int32_t *row_ptr,
                      const int32_t *col, const void *val, int64_t k,
                      int row_major, const void *X, int64_t ldx, void *Y,
                      int64_t ldy, smax_hip_crs_plan plan) {
    SMAX_HIP_CHECK_CTX(ctx);
    if (!valid_value_type(value_type))
        return fail(ctx, SMAX_HIP_ERROR_UNSUPPORTED,
                    "smax_hip_crs_spmm: value type %d", value_type);
    SMAX_HIP_REQUIRE(ctx, n_rows > 0 && n_cols > 0 && nnz >= 0 && k > 0,
                     "smax_hip_crs_spmm: bad shape");
    if (!extent_ok(n_rows, k) || n_cols > INT32_MAX)
        return fail(ctx, SMAX_HIP_ERROR_OVERFLOW, "%s: %lld rows x %lld columns exceed the "
                    "launch limits", __func__, (long long)n_rows, (long long)k);
    SMAX_HIP_REQUIRE(ctx, row_ptr && X && Y && (nnz == 0 || (col && val)),
                     "smax_hip_crs_spmm: null pointer");
    if (row_major)
        SMAX_HIP_REQUIRE(ctx, ldx >= k && ldy >= k,
                         "smax_hip_crs_spmm: row-major ld must be >= k");
    else
        SMAX_HIP_REQUIRE(ctx, (ldx >= n_cols || ldx >= n_rows) && (k == 1 || ldy >= n_rows),
                         "smax_hip_crs_spmm: column-major ld too small (need ldx >= "
                         "min(n_rows, n_cols) and ldy >= n_rows)");
    if (plan)
        SMAX_HIP_REQUIRE(ctx, plan->n_rows == n_rows && plan->nnz == nnz,
                         "smax_hip_crs_spmm: plan built for another pattern");
    if (!row_major && k >= 2) {
        // Column-major operands run through row-major copies in the context's
        // workspace: each X row (the k values a nonzero multiplies) becomes
        // contiguous, so the row-major kernels and the plan's row order apply.
        // Every output is the same chain, so Y is unchanged bit for bit.
        SMAX_HIP_DISPATCH_VT(value_type, {
            using A = typename vt<VT>::vec_t;
            const size_t xb = (static_cast<size_t>(n_cols) * k * sizeof(A) + 255) / 256 * 256;
            const size_t yb = static_cast<size_t>(n_rows) * k * sizeof(A);
            void *w = nullptr;
            int rc = ctx_workspace(ctx, xb + yb, &w);
            if (rc) return rc;
            A *Xr = static_cast<A *>(w);
            A *Yr = reinterpret_cast<A *>(static_cast<char *>(w) + xb);
            launch_relayout<A, true>(ctx, n_cols, k, static_cast<const A *>(X), ldx, Xr);
            SMAX_HIP_LAUNCHED(ctx, "k_relayout");
            rc = smax_hip_crs_spmm(ctx, value_type, n_rows, n_cols, nnz, row_ptr, col, val,
                                   k, 1, Xr, k, Yr, k, plan);
            if (rc) return rc;
            launch_relayout<A, false>(ctx, n_rows, k, Yr, ldy, static_cast<A *>(Y));
            SMAX_HIP_LAUNCHED(ctx, "k_relayout");
            return SMAX_HIP_SUCCESS;
        });
        return SMAX_HIP_ERROR_UNSUPPORTED;
    }
    const int32_t *order = plan ? plan->row_order : nullptr;
    const int64_t spmm_lds = plan ? plan->spmm_lds : 0;
    const int64_t threads = n_rows * k;
    SMAX_HIP_DISPATCH_VT(value_type, {
        using T = vt<VT>;
        auto v = static_cast<const typename T::val_t *>(val);
        auto xv = static_cast<const typename T::vec_t *>(X);
        auto yv = static_cast<typename T::vec_t *>(Y);
        if (row_major && launch_crs_spmm_rowvec<VT>(ctx, n_rows, k, order,
                                                    static_cast<unsigned>(spmm_lds), row_ptr,
                                                    col, val, X, ldx, Y, ldy)) {
            SMAX_HIP_LAUNCHED(ctx, "crs_spmm_rowvec");
            return SMAX_HIP_SUCCESS;
        }
        if (row_major)
            hipLaunchKernelGGL((crs_spmm_kernel<VT, true>),
                               grid_for(threads, kCrsThreads), dim3(kCrsThreads),
                               0, ctx->stream, n_rows, k, order, row_ptr, col, v, xv,
                               ldx, yv, ldy);
        else
            hipLaunchKernelGGL((crs_spmm_kernel<VT, false>),
                               grid_for(threads, kCrsThreads), dim3(kCrsThreads),
                               0, ctx->stream, n_rows, k, nullptr, row_ptr, col, v, xv,
                               ldx, yv, ldy);
        SMAX_HIP_LAUNCHED(ctx, "crs_spmm_kernel");
        return SMAX_HIP_SUCCESS;
    });
    return SMAX_HIP_ERROR_UNSUPPORTED;
}

int smax_hip_crs_spmm_noalloc(smax_hip_ctx ctx, int value_type, int64_t n_rows,
                              int64_t n_cols, int64_t nnz, const int32_t *row_ptr,
                              const int32_t *col, const void *val, int64_t k,
                              int row_major, const void *X, int64_t ldx, void *Y,
                              int64_t ldy, smax_hip_crs_plan plan) {
    SMAX_HIP_CHECK_CTX(ctx);
    // row-major products and k = 1 stage nothing: the same kernels, no scratch
    if (row_major || k == 1)
        return smax_hip_crs_spmm(ctx, value_type, n_rows, n_cols, nnz, row_ptr, col, val, k,
                                 row_major, X, ldx, Y, ldy, plan);
    if (!valid_value_type(value_type))
        return fail(ctx, SMAX_HIP_ERROR_UNSUPPORTED,
                    "smax_hip_crs_spmm_noalloc: value type %d", value_type);
    SMAX_HIP_REQUIRE(ctx, n_rows > 0 && n_cols > 0 && nnz >= 0 && k > 0,
                     "smax_hip_crs_spmm_noalloc: bad shape");
    if (!extent_ok(n_rows, k) || n_cols > INT32_MAX)
        return fail(ctx, SMAX_HIP_ERROR_OVERFLOW, "%s: %lld rows x %lld columns exceed the "
                    "launch limits", __func__, (long long)n_rows, (long long)k);
    SMAX_HIP_REQUIRE(ctx, row_ptr && X && Y && (nnz == 0 || (col && val)),
                     "smax_hip_crs_spmm_noalloc: null pointer");
    SMAX_HIP_REQUIRE(ctx, (ldx >= n_cols || ldx >= n_rows) && ldy >= n_rows,
                     "smax_hip_crs_spmm_noalloc: column-major ld too small (need ldx >= "
                     "min(n_rows, n_cols) and ldy >= n_rows)");
    if (plan)
        SMAX_HIP_REQUIRE(ctx, plan->n_rows == n_rows && plan->nnz == nnz,
                         "smax_hip_crs_spmm_noalloc: plan built for another pattern");
    SMAX_HIP_DISPATCH_VT(value_type, {
        using A = typename vt<VT>::vec_t;
        const int64_t s = static_cast<int64_t>(sizeof(A));
        // columns per pass: the X block (rounded up to 256 bytes) and the Y
        // block must fit the scratch together
        const int64_t per_col = (n_cols + n_rows) * s;
        const int64_t kb_max = (static_cast<int64_t>(kNoAllocScratchBytes) - 256) / per_col;
        if (kb_max < 1)
            return fail(ctx, SMAX_HIP_ERROR_INVALID_VALUE,
                        "smax_hip_crs_spmm_noalloc: one column of X and Y takes %lld bytes, "
                        "more than the %zu-byte scratch (n_rows + n_cols <= %lld for this "
                        "type); use smax_hip_crs_spmm",
                        static_cast<long long>(per_col), kNoAllocScratchBytes,
                        static_cast<long long>((static_cast<int64_t>(kNoAllocScratchBytes) - 256) / s));
        int64_t kb = std::min(k, kb_max);
        if (kb < k && kb >= 16) {
            // a power of two in [16, 64] takes the 16-byte-run kernel
            int64_t p = 16;
            while (p * 2 <= std::min<int64_t>(kb, 64)) p *= 2;
            kb = p;
        }
        void *base = nullptr;
        int rc = noalloc_scratch(ctx, &base);
        if (rc) return rc;
        A *Xr = static_cast<A *>(base);
        const int64_t xbytes = (n_cols * kb * s + 255) / 256 * 256;
        A *Yr = reinterpret_cast<A *>(static_cast<char *>(base) + xbytes);
        for (int64_t v0 = 0; v0 < k; v0 += kb) {
            const int64_t kc = std::min(kb, k - v0);
            launch_relayout<A, true>(ctx, n_cols, kc, static_cast<const A *>(X) + v0 * ldx, ldx,
                                     Xr);
            SMAX_HIP_LAUNCHED(ctx, "k_relayout");
            rc = smax_hip_crs_spmm(ctx, value_type, n_rows, n_cols, nnz, row_ptr, col, val,
                                   kc, 1, Xr, kc, Yr, kc, plan);
            if (rc) return rc;
            launch_relayout<A, false>(ctx, n_rows, kc, Yr, ldy, static_cast<A *>(Y) + v0 * ldy);
            SMAX_HIP_LAUNCHED(ctx, "k_relayout");
        }
        return SMAX_HIP_SUCCESS;
    });
    return SMAX_HIP_ERROR_UNSUPPORTED;
}

}  // extern "C"
