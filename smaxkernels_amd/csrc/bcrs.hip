// BCRS SpMV / SpMM and CRS -> BCRS conversion for gfx950.
//
// Block row r, block j (CRS order over blocks), block column jx:
//   y[r*h_pad + ix] = sum_j sum_jx val[j*h_pad*w_pad + off(ix, jx)] *
//                     x[col[j]*w_pad + jx]
// with off = jx*h_pad + ix (column-major blocks) or ix*w_pad + jx
// (row-major), accumulated per output in (j, jx) order exactly as
// naive_bcrs_spmv does (spmv_cpu_bcrs_impl.hpp:45-74): bit-identical.
//
// SpMV stages the blocks of a range of block rows through LDS: the values
// of 256/b_h block rows are contiguous in memory, so the workgroup streams
// them (and the x entries their columns select) with coalesced loads, then
// lane (r, ix) runs its fma chain out of LDS.
#include "common.hpp"
#include "scan.hpp"

#include <type_traits>
#include <utility>

struct smax_hip_bcrs_s {
    smax_hip_bcrs_info info;
};

namespace smaxhip {

constexpr int kBThreads = 256;
constexpr int kBTileVals = 2048;  // block values staged per LDS tile

template <int VT>
__global__ __launch_bounds__(kBThreads) void bcrs_spmv_lds(
    int64_t n_block_rows, int b_h, int b_w, int h_pad, int w_pad, int col_major,
    const int32_t *__restrict__ col, const int32_t *__restrict__ row_ptr,
    const typename vt<VT>::val_t *__restrict__ val,
    const typename vt<VT>::vec_t *__restrict__ x,
    typename vt<VT>::vec_t *__restrict__ y) {
    using T = vt<VT>;
    using A = typename T::vec_t;
    __shared__ A sv[kBTileVals];
    __shared__ A sx[kBTileVals];
    const int rows_per_wg = kBThreads / b_h;
    const int64_t r0 = static_cast<int64_t>(blockIdx.x) * rows_per_wg;
    const int64_t r1 = min<int64_t>(r0 + rows_per_wg, n_block_rows);
    const int tid = threadIdx.x;
    const int64_t my_r = r0 + tid / b_h;
    const int ix = tid % b_h;
    const bool has = tid < rows_per_wg * b_h && my_r < r1;
    const int64_t bs = static_cast<int64_t>(h_pad) * w_pad;
    const int64_t s = row_ptr[r0], e = row_ptr[r1];
    const int64_t js = has ? row_ptr[my_r] : 0, je = has ? row_ptr[my_r + 1] : 0;
    // blocks per tile: values and gathered x entries both fit
    const int64_t tb = min<int64_t>(kBTileVals / bs, kBTileVals / w_pad);
    A acc = 0;
    for (int64_t t0 = s; t0 < e; t0 += tb) {
        const int64_t nb = min<int64_t>(tb, e - t0);
        const int64_t nv = nb * bs;
        const typename T::val_t *vsrc = val + t0 * bs;
        for (int64_t k = tid; k < nv; k += kBThreads) sv[k] = T::widen(ld_stream(vsrc + k));
        const int64_t nx = nb * w_pad;
        for (int64_t k = tid; k < nx; k += kBThreads) {
            const int64_t b = k / w_pad, jx = k - b * w_pad;
            sx[k] = jx < b_w ? x[static_cast<int64_t>(col[t0 + b]) * w_pad + jx] : A(0);
        }
        __syncthreads();
        if (has) {
            const int64_t lo = max(js, t0), hi = min(je, t0 + nb);
            for (int64_t j = lo; j < hi; ++j) {
                const int64_t b = j - t0;
                for (int jx = 0; jx < b_w; ++jx) {
                    const int64_t off = col_major ? jx * h_pad + ix : ix * w_pad + jx;
                    acc = T::fma_(sv[b * bs + off], sx[b * w_pad + jx], acc);
                }
            }
        }
        __syncthreads();
    }
    if (has) y[my_r * h_pad + ix] = acc;
}

// Thread (r, ix, v), v fastest: the k lanes of one output row share each
// block value (broadcast) and read one contiguous row of X per (j, jx).
// BW > 0 fixes b_w at compile time: a block's b_w values and X entries are
// then loaded together, and the next block's while this one's fmas run, so
// the chain waits about one memory round trip per block.  The fma order is
// the reference's (j, then jx) either way.
template <int VT, int BW>
__global__ __launch_bounds__(kBThreads) void bcrs_spmm_kernel(
    int64_t n_block_rows, int b_h, int b_w_rt, int h_pad, int w_pad, int col_major,
    int64_t k, const int32_t *__restrict__ col,
    const int32_t *__restrict__ row_ptr,
    const typename vt<VT>::val_t *__restrict__ val,
    const typename vt<VT>::vec_t *__restrict__ X,
    typename vt<VT>::vec_t *__restrict__ Y) {
    using T = vt<VT>;
    using A = typename T::vec_t;
    using V = typename T::val_t;
    const int64_t t = static_cast<int64_t>(blockIdx.x) * kBThreads + threadIdx.x;
    const int64_t v = t % k, rix = t / k;
    const int64_t r = rix / b_h;
    const int ix = static_cast<int>(rix - r * b_h);
    if (r >= n_block_rows) return;
    const int64_t bs = static_cast<int64_t>(h_pad) * w_pad;
    // value (ix, jx) of a block sits at ix*vs_i + jx*vs_j
    const int64_t vs_i = col_major ? 1 : w_pad, vs_j = col_major ? h_pad : 1;
    const V *vrow = val + ix * vs_i;
    A acc = 0;
    int64_t j = row_ptr[r];
    const int64_t je = row_ptr[r + 1];
    if (BW == 0) {
        for (; j < je; ++j) {
            const V *blk = vrow + j * bs;
            const A *xr = X + static_cast<int64_t>(col[j]) * w_pad * k + v;
            for (int jx = 0; jx < b_w_rt; ++jx)
                acc = T::fma_(T::widen(blk[jx * vs_j]), xr[jx * k], acc);
        }
    } else if (j < je) {
        V w0[BW > 0 ? BW : 1];
        A x0[BW > 0 ? BW : 1];
        {
            const V *blk = vrow + j * bs;
            const A *xr = X + static_cast<int64_t>(col[j]) * w_pad * k + v;
#pragma unroll
            for (int jx = 0; jx < BW; ++jx) {
                w0[jx] = blk[jx * vs_j];
                x0[jx] = xr[jx * k];
            }
        }
        for (; j + 1 < je; ++j) {
            V w1[BW > 0 ? BW : 1];
            A x1[BW > 0 ? BW : 1];
            const V *blk = vrow + (j + 1) * bs;
            const A *xr = X + static_cast<int64_t>(col[j + 1]) * w_pad * k + v;
#pragma unroll
            for (int jx = 0; jx < BW; ++jx) {
                w1[jx] = blk[jx * vs_j];
                x1[jx] = xr[jx * k];
            }
#pragma unroll
            for (int jx = 0; jx < BW; ++jx) acc = T::fma_(T::widen(w0[jx]), x0[jx], acc);
#pragma unroll
            for (int jx = 0; jx < BW; ++jx) {
                w0[jx] = w1[jx];
                x0[jx] = x1[jx];
            }
        }
#pragma unroll
        for (int jx = 0; jx < BW; ++jx) acc = T::fma_(T::widen(w0[jx]), x0[jx], acc);
    }
    Y[(r * h_pad + ix) * k + v] = acc;
}

// One block's values, widened to the accumulation type.  DENSE: h_pad == BH
// and w_pad == BW, so the block is BH*BW contiguous values (row- or
// column-major), read as 16-byte loads when that size is a multiple of 16
// bytes (val is then 16-byte aligned, checked by the host); BF16 blocks are
// read as 32-bit words holding two values each.
template <int VT, int BH, int BW, bool DENSE, bool CM>
__device__ __forceinline__ void load_block(const typename vt<VT>::val_t *__restrict__ blk,
                                           int64_t vs_i, int64_t vs_j,
                                           typename vt<VT>::vec_t (&w)[BH][BW]) {
    using T = vt<VT>;
    using A = typename T::vec_t;
    if constexpr (DENSE) {
        A flat[BH * BW];
        if constexpr (VT == SMAX_HIP_BF16 && (BH * BW) % 2 == 0) {
            uint32_t words[BH * BW / 2];
            load_run<uint32_t, BH * BW / 2>(reinterpret_cast<const uint32_t *>(blk), words);
#pragma unroll
            for (int i = 0; i < BH * BW / 2; ++i) {
                flat[2 * i] = __uint_as_float(words[i] << 16);
                flat[2 * i + 1] = __uint_as_float(words[i] & 0xffff0000u);
            }
        } else {
            typename T::val_t raw[BH * BW];
            load_run<typename T::val_t, BH * BW>(blk, raw);
#pragma unroll
            for (int i = 0; i < BH * BW; ++i) flat[i] = T::widen(raw[i]);
        }
#pragma unroll
        for (int ix = 0; ix < BH; ++ix)
#pragma unroll
            for (int jx = 0; jx < BW; ++jx)
                w[ix][jx] = CM ? flat[jx * BH + ix] : flat[ix * BW + jx];
    } else {
#pragma unroll
        for (int jx = 0; jx < BW; ++jx)
#pragma unroll
            for (int ix = 0; ix < BH; ++ix) w[ix][jx] = T::widen(blk[ix * vs_i + jx * vs_j]);
    }
}

// Register-blocked SpMM for compile-time block shapes: thread (r, l) keeps
// the BH outputs of block row r in columns v = l*Q .. l*Q+Q-1, so each X row
// of a block is loaded once for all BH rows (instead of once per row), and
// the BH*Q chains give independent fmas.  Q = 16 bytes' worth (16-byte runs:
// k a multiple of Q, X and Y 16-byte aligned) reads each lane's part of an X
// row with one 16-byte load; Q = 2 with 4-byte entries takes column pairs
// (no alignment needed).  Blocks go in groups of G: the G column indices,
// then all G blocks' values and X entries are requested before any fma, so a
// row of G blocks costs two memory round trips.  Per output the fma order is
// still the reference's (block j, then jx).
template <int VT, int BH, int BW, int Q, bool DENSE, bool CM, int G>
__global__ __launch_bounds__(kBThreads) void bcrs_spmm_reg(
    int64_t n_block_rows, int h_pad, int w_pad, int col_major, int64_t k,
    const int32_t *__restrict__ col, const int32_t *__restrict__ row_ptr,
    const typename vt<VT>::val_t *__restrict__ val,
    const typename vt<VT>::vec_t *__restrict__ X,
    typename vt<VT>::vec_t *__restrict__ Y) {
    using T = vt<VT>;
    using A = typename T::vec_t;
    const int64_t gl = k / Q;  // lanes per block row
    const int64_t t = static_cast<int64_t>(blockIdx.x) * kBThreads + threadIdx.x;
    const int64_t r = t / gl, v = (t - r * gl) * Q;
    if (r >= n_block_rows) return;
    const int64_t bs = static_cast<int64_t>(h_pad) * w_pad;
    const int64_t vs_i = col_major ? 1 : w_pad, vs_j = col_major ? h_pad : 1;
    A acc[BH][Q];
#pragma unroll
    for (int ix = 0; ix < BH; ++ix)
#pragma unroll
        for (int q = 0; q < Q; ++q) acc[ix][q] = 0;
    const int64_t je = row_ptr[r + 1];
    for (int64_t j = row_ptr[r]; j < je; j += G) {
        int32_t c[G];
#pragma unroll
        for (int g = 0; g < G; ++g)
            if (j + g < je) c[g] = col[j + g];
        A w[G][BH][BW];
        A xv[G][BW][Q];
#pragma unroll
        for (int g = 0; g < G; ++g)
            if (j + g < je) {
                const A *xr = X + static_cast<int64_t>(c[g]) * w_pad * k + v;
#pragma unroll
                for (int jx = 0; jx < BW; ++jx) load_run<A, Q>(xr + jx * k, xv[g][jx]);
                load_block<VT, BH, BW, DENSE, CM>(val + (j + g) * bs, vs_i, vs_j, w[g]);
            }
#pragma unroll
        for (int g = 0; g < G; ++g)
            if (j + g < je) {
#pragma unroll
                for (int ix = 0; ix < BH; ++ix)
#pragma unroll
                    for (int jx = 0; jx < BW; ++jx)
#pragma unroll
                        for (int q = 0; q < Q; ++q)
                            acc[ix][q] = T::fma_(w[g][ix][jx], xv[g][jx][q], acc[ix][q]);
            }
    }
#pragma unroll
    for (int ix = 0; ix < BH; ++ix) store_run<A, Q>(Y + (r * h_pad + ix) * k + v, acc[ix]);
}

// Lane N of each 16-lane DPP row, handed to every lane of that row
// (row_newbcast).
template <int N>
__device__ __forceinline__ float row_bcast(float v) {
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x150 + N, 0xf, 0xf, false));
}

template <int... N>
__device__ __forceinline__ void row_bcast16(float mine, float (&out)[16],
                                            std::integer_sequence<int, N...>) {
    ((out[N] = row_bcast<N>(mine)), ...);
}

// bcrs_spmm_reg for dense 4x4 blocks and k = 64 four-byte columns, where a
// block row's 16 lanes (four columns each) are exactly one DPP row.  There,
// every lane loading the whole 16-value block makes a third of the bytes the
// vector memory path (texture address unit busy ~88 % of the kernel's cycles
// on the 96^3 block matrix) hands back.  Here each lane loads one value of
// the block and 16 row broadcasts give every lane all of them; the next
// group's column indices are requested while this group's X rows are in
// flight.  Per output the fma order is bcrs_spmm_reg's (block j,
// then jx), so Y is bit-identical.
template <int VT, bool CM, int G>
__global__ __launch_bounds__(kBThreads) void bcrs_spmm_bcast(
    int64_t n_block_rows, int64_t k, const int32_t *__restrict__ col,
    const int32_t *__restrict__ row_ptr, const typename vt<VT>::val_t *__restrict__ val,
    const float *__restrict__ X, float *__restrict__ Y) {
    using T = vt<VT>;
    constexpr int BH = 4, BW = 4, Q = 4, kRow = 16;
    const int64_t t = static_cast<int64_t>(blockIdx.x) * kBThreads + threadIdx.x;
    const int64_t r = t / kRow;  // whole DPP rows leave together
    const int l = static_cast<int>(threadIdx.x) % kRow;
    const int64_t v = static_cast<int64_t>(l) * Q;
    if (r >= n_block_rows) return;
    float acc[BH][Q];
#pragma unroll
    for (int ix = 0; ix < BH; ++ix)
#pragma unroll
        for (int q = 0; q < Q; ++q) acc[ix][q] = 0;
    const int64_t jb = row_ptr[r], je = row_ptr[r + 1];
    int32_t cn[G];
#pragma unroll
    for (int g = 0; g < G; ++g) cn[g] = jb + g < je ? col[jb + g] : 0;
    for (int64_t j = jb; j < je; j += G) {
        int32_t c[G];
        float mine[G];
        float xv[G][BW][Q];
#pragma unroll
        for (int g = 0; g < G; ++g) c[g] = cn[g];
#pragma unroll
        for (int g = 0; g < G; ++g)
            if (j + g < je) {
                const float *xr = X + static_cast<int64_t>(c[g]) * BW * k + v;
#pragma unroll
                for (int jx = 0; jx < BW; ++jx) load_run<float, Q>(xr + jx * k, xv[g][jx]);
                mine[g] = T::widen(val[(j + g) * (BH * BW) + l]);
            }
#pragma unroll
        for (int g = 0; g < G; ++g)
            if (j + G + g < je) cn[g] = col[j + G + g];
#pragma unroll
        for (int g = 0; g < G; ++g)
            if (j + g < je) {
                float w[BH * BW];
                row_bcast16(mine[g], w, std::make_integer_sequence<int, BH * BW>{});
#pragma unroll
                for (int ix = 0; ix < BH; ++ix)
#pragma unroll
                    for (int jx = 0; jx < BW; ++jx)
#pragma unroll
                        for (int q = 0; q < Q; ++q)
                            acc[ix][q] = T::fma_(w[CM ? jx * BH + ix : ix * BW + jx],
                                                 xv[g][jx][q], acc[ix][q]);
            }
    }
#pragma unroll
    for (int ix = 0; ix < BH; ++ix) store_run<float, Q>(Y + (r * BH + ix) * k + v, acc[ix]);
}

// Dword v of lane S of each DPP quad, handed to the quad's four lanes
// (quad_perm [S,S,S,S]).
template <int S>
__device__ __forceinline__ uint32_t quad_bcast(uint32_t v) {
    return static_cast<uint32_t>(
        __builtin_amdgcn_mov_dpp(static_cast<int>(v), S * 0x55, 0xf, 0xf, false));
}

// Dword v of lane S of each aligned group of NL lanes (4: a DPP quad, 8: a
// ds_swizzle group), handed to the group's lanes.
template <int NL, int S>
__device__ __forceinline__ uint32_t group_bcast(uint32_t v) {
    if constexpr (NL == 4)
        return quad_bcast<S>(v);
    else
        return static_cast<uint32_t>(grp8_bcast<S>(static_cast<int>(v)));
}

template <int NL, int M, int... S>
__device__ __forceinline__ void group_bcast_all(const uint32_t (&mine)[M],
                                                uint32_t (&words)[NL * M],
                                                std::integer_sequence<int, S...>) {
#pragma unroll
    for (int d = 0; d < M; ++d) ((words[S * M + d] = group_bcast<NL, S>(mine[d])), ...);
}

// bcrs_spmm_reg for dense 4x4 blocks with four-byte columns and NL = 4 or 8
// lanes per block row (k = 8 in column pairs, Q = 2; k = 16 or 32 in
// 16-byte runs, Q = 4), so that a block row's lanes are one DPP quad or one
// swizzle group of 8: each lane loads 1/NL of the block and broadcasts
// within the group give every lane the whole block, instead of every lane
// loading all of it.  Blocks go in groups of 4 (the quarters take few registers),
// with the next group's column indices requested while this group's X rows
// are in flight.  Per output the fma order is bcrs_spmm_reg's (block j, then
// jx), so Y is bit-identical.
template <int VT, int Q, int NL, bool CM, int G>
__global__ __launch_bounds__(kBThreads) void bcrs_spmm_quad(
    int64_t n_block_rows, int64_t k, const int32_t *__restrict__ col,
    const int32_t *__restrict__ row_ptr, const typename vt<VT>::val_t *__restrict__ val,
    const float *__restrict__ X, float *__restrict__ Y) {
    using T = vt<VT>;
    using V = typename T::val_t;
    constexpr int BH = 4, BW = 4;
    constexpr int kWords = BH * BW * static_cast<int>(sizeof(V)) / 4;  // dwords per block
    constexpr int kMine = kWords / NL;                                 // loaded per lane
    static_assert(kMine * NL == kWords, "a block splits evenly over the group");
    const int64_t t = static_cast<int64_t>(blockIdx.x) * kBThreads + threadIdx.x;
    const int64_t r = t / NL;  // whole groups leave together
    const int l = static_cast<int>(threadIdx.x) % NL;
    const int64_t v = static_cast<int64_t>(l) * Q;
    if (r >= n_block_rows) return;
    float acc[BH][Q];
#pragma unroll
    for (int ix = 0; ix < BH; ++ix)
#pragma unroll
        for (int q = 0; q < Q; ++q) acc[ix][q] = 0;
    const int64_t jb = row_ptr[r], je = row_ptr[r + 1];
    int32_t cn[G];
#pragma unroll
    for (int g = 0; g < G; ++g) cn[g] = jb + g < je ? col[jb + g] : 0;
    const uint32_t *vw = reinterpret_cast<const uint32_t *>(val);
    for (int64_t j = jb; j < je; j += G) {
        int32_t c[G];
        uint32_t mine[G][kMine];
        float xv[G][BW][Q];
#pragma unroll
        for (int g = 0; g < G; ++g) c[g] = cn[g];
#pragma unroll
        for (int g = 0; g < G; ++g)
            if (j + g < je) {
                const float *xr = X + static_cast<int64_t>(c[g]) * BW * k + v;
#pragma unroll
                for (int jx = 0; jx < BW; ++jx) load_run<float, Q>(xr + jx * k, xv[g][jx]);
                load_run<uint32_t, kMine>(vw + (j + g) * kWords + l * kMine, mine[g]);
            }
#pragma unroll
        for (int g = 0; g < G; ++g)
            if (j + G + g < je) cn[g] = col[j + G + g];
#pragma unroll
        for (int g = 0; g < G; ++g)
            if (j + g < je) {
                uint32_t words[kWords];
                group_bcast_all<NL>(mine[g], words, std::make_integer_sequence<int, NL>{});
                float w[BH * BW];
#pragma unroll
                for (int i = 0; i < BH * BW; ++i) {
                    if constexpr (sizeof(V) == 2)  // BF16: two values per dword, low first
                        w[i] = __uint_as_float((i & 1) ? (words[i >> 1] & 0xffff0000u)
                                                       : (words[i >> 1] << 16));
                    else
                        w[i] = __uint_as_float(words[i]);
                }
#pragma unroll
                for (int ix = 0; ix < BH; ++ix)
#pragma unroll
                    for (int jx = 0; jx < BW; ++jx)
#pragma unroll
                        for (int q = 0; q < Q; ++q)
                            acc[ix][q] = T::fma_(w[CM ? jx * BH + ix : ix * BW + jx],
                                                 xv[g][jx][q], acc[ix][q]);
            }
    }
#pragma unroll
    for (int ix = 0; ix < BH; ++ix) store_run<float, Q>(Y + (r * BH + ix) * k + v, acc[ix]);
}

template <int VT, int BH, int BW>
static void launch_bcrs_spmm_reg(smax_hip_ctx ctx, int64_t n_block_rows, int64_t h_pad,
                                 int64_t w_pad, bool cm, int64_t k, const int32_t *col,
                                 const int32_t *row_ptr, const void *val, const void *X,
                                 void *Y) {
    using T = vt<VT>;
    using A = typename T::vec_t;
    constexpr int Q = 16 / sizeof(A);
    auto v = static_cast<const typename T::val_t *>(val);
    auto xv = static_cast<const A *>(X);
    auto yv = static_cast<A *>(Y);
    // 16-byte runs need at least 4 lanes per block row: with fewer (k = 8 in
    // FP32) the per-lane chains run longer than the saved instructions pay for
    const bool runs = k % Q == 0 && k >= 4 * Q && aligned16(X, Y);
    // unpadded blocks load as one contiguous run (16-byte loads if the block
    // size is a multiple of 16 bytes and val is aligned)
    constexpr bool kBlock16 = (BH * BW * sizeof(typename T::val_t)) % 16 == 0;
    const bool dense = h_pad == BH && w_pad == BW && (!kBlock16 || aligned16(val, val));
    // otherwise, with 4-byte X entries and an even k >= 8, each lane takes two
    // columns: every lane loads its block row's blocks whole, so half as many
    // lanes load half as many copies (96^3 block matrix, k = 8: BF16 values
    // 0.155 -> 0.151 ms, FP32 0.198 -> 0.183 ms)
    const bool pairs = !runs && dense && sizeof(A) == 4 && k % 2 == 0 && k >= 8;
    const int64_t lanes = runs ? k / Q : pairs ? k / 2 : k;  // per block row
    const dim3 g = grid_for(n_block_rows * lanes, kBThreads);
    // 4x4 dense blocks at k = 64 (16 lanes per block row): one value per lane
    // and DPP row broadcasts, in groups of 4 blocks (96^3 block matrix, BF16
    // values: 0.646 -> 0.615 ms forward and through the transpose; groups of
    // 2: 0.621 ms)
    if constexpr (BH == 4 && BW == 4 && sizeof(A) == 4) {
        if (runs && dense && lanes == 16) {
            if (cm)
                hipLaunchKernelGGL((bcrs_spmm_bcast<VT, true, 4>), g, dim3(kBThreads), 0,
                                   ctx->stream, n_block_rows, k, col, row_ptr, v, xv, yv);
            else
                hipLaunchKernelGGL((bcrs_spmm_bcast<VT, false, 4>), g, dim3(kBThreads), 0,
                                   ctx->stream, n_block_rows, k, col, row_ptr, v, xv, yv);
            return;
        }
        // 4 lanes per block row (k = 8 in pairs, k = 16 in runs): a quarter of
        // each block per lane and DPP quad broadcasts, in groups of 4 blocks
        // (k = 8: BF16 values 0.152 -> 0.139 ms forward and through the
        // transpose, FP32 0.183 -> 0.160 ms; groups of 2: 0.150 / 0.171 ms;
        // k = 16, FP32 values: 0.240 -> 0.223 ms).  8 lanes (k = 32 in runs):
        // an eighth per lane and swizzles within the 8 (BF16 values 0.331 ->
        // 0.318 ms, FP32 0.368 -> 0.345 ms)
#define SMAX_BCRS_QUAD(QQ, NN)                                                             \
    if (cm)                                                                                \
        hipLaunchKernelGGL((bcrs_spmm_quad<VT, QQ, NN, true, 4>), g, dim3(kBThreads), 0,   \
                           ctx->stream, n_block_rows, k, col, row_ptr, v, xv, yv);         \
    else                                                                                   \
        hipLaunchKernelGGL((bcrs_spmm_quad<VT, QQ, NN, false, 4>), g, dim3(kBThreads), 0,  \
                           ctx->stream, n_block_rows, k, col, row_ptr, v, xv, yv)
        if (pairs && lanes == 4) {
            SMAX_BCRS_QUAD(2, 4);
            return;
        }
        if (runs && dense && lanes == 4) {
            SMAX_BCRS_QUAD(4, 4);
            return;
        }
        if (runs && dense && lanes == 8) {
            SMAX_BCRS_QUAD(4, 8);
            return;
        }
#undef SMAX_BCRS_QUAD
    }
    // block groups: 8 when a single thread holds the whole block row (k = 1),
    // else 2 (measured on the 96^3 block matrix: k = 1 0.108 -> 0.058 ms,
    // k = 8 0.173 -> 0.160 ms, k = 64 0.663 -> 0.649 ms in BF16).  Padded
    // blocks load value by value, with an address each: groups of 8 would
    // spill registers, so a single thread takes kPaddedGroup blocks at once.
    constexpr int kPaddedGroup = 4;
    const bool single = !runs && k == 1;
#define SMAX_BCRS_REG_LAUNCH(QQ, DD, CC, GG)                                                \
    hipLaunchKernelGGL((bcrs_spmm_reg<VT, BH, BW, QQ, DD, CC, GG>), g, dim3(kBThreads), 0,  \
                       ctx->stream, n_block_rows, static_cast<int>(h_pad),                 \
                       static_cast<int>(w_pad), cm ? 1 : 0, k, col, row_ptr, v, xv, yv)
#define SMAX_BCRS_REG_G(DD, CC, GS)                                                         \
    if (single) SMAX_BCRS_REG_LAUNCH(1, DD, CC, GS);                                        \
    else SMAX_BCRS_REG_LAUNCH(1, DD, CC, 2)
    // runs imply k >= 4Q lanes' worth of columns, so never a single thread
    if (runs && dense && cm) { SMAX_BCRS_REG_LAUNCH(Q, true, true, 2); }
    else if (runs && dense) { SMAX_BCRS_REG_LAUNCH(Q, true, false, 2); }
    else if (runs) { SMAX_BCRS_REG_LAUNCH(Q, false, false, 2); }
    else if (pairs && cm) { SMAX_BCRS_REG_LAUNCH(2, true, true, 2); }
    else if (pairs) { SMAX_BCRS_REG_LAUNCH(2, true, false, 2); }
    else if (dense && cm) { SMAX_BCRS_REG_G(true, true, 8); }
    else if (dense) { SMAX_BCRS_REG_G(true, false, 8); }
    else { SMAX_BCRS_REG_G(false, false, kPaddedGroup); }
#undef SMAX_BCRS_REG_G
#undef SMAX_BCRS_REG_LAUNCH
}

template <int VT, int BW>
static void launch_bcrs_spmm(smax_hip_ctx ctx, int64_t threads, int64_t n_block_rows,
                             int64_t b_h, int64_t b_w, int64_t h_pad, int64_t w_pad,
                             bool cm, int64_t k, const int32_t *col,
                             const int32_t *row_ptr, const void *val, const void *X,
                             void *Y) {
    using T = vt<VT>;
    hipLaunchKernelGGL((bcrs_spmm_kernel<VT, BW>), grid_for(threads, kBThreads),
                       dim3(kBThreads), 0, ctx->stream, n_block_rows,
                       static_cast<int>(b_h), static_cast<int>(b_w),
                       static_cast<int>(h_pad), static_cast<int>(w_pad), cm ? 1 : 0, k,
                       col, row_ptr, static_cast<const typename T::val_t *>(val),
                       static_cast<const typename T::vec_t *>(X),
                       static_cast<typename T::vec_t *>(Y));
}

// ------------------------------------------------------------ conversion
// Sorted unique block columns of each block row: a merge of its b_h rows,
// whose block-column sequences must be non-decreasing (the reference's fill
// loop, convert_formats.hpp:222-234, throws otherwise).
//
// A workgroup takes 256/GS consecutive block rows (GS the power of two >=
// b_h), whose CRS entries are one contiguous range; when the range fits the
// LDS tile it is loaded once, coalesced (columns, and values in the fill
// pass), and the merge reads LDS.  A group of GS lanes takes one block row,
// lane l its CRS row br*b_h + l, read front to back.  Each step the group's
// smallest head block column m is found with an xor-butterfly; every lane
// then steps past its entries in block column m, and m is the block row's
// next block column.  In the fill pass a lane places each value it steps
// past, in row order, so of two entries with the same column the later one
// stays, as in the reference's fill loop.

// c / b_w for 0 <= c < 2^31 as a multiply and shift: with l = ceil(log2
// b_w) and m = ceil(2^(31+l) / b_w), m*b_w - 2^(31+l) < b_w <= 2^l, which
// makes floor(c*m / 2^(31+l)) exact for every 31-bit c (Granlund and
// Montgomery, "Division by invariant integers using multiplication", 1994,
// Theorem 4.2).
struct BlockDiv {
    uint32_t m;
    int shift;
    __device__ int32_t operator()(int32_t c) const {
        return static_cast<int32_t>(
            (static_cast<uint64_t>(static_cast<uint32_t>(c)) * m) >> shift);
    }
};

inline BlockDiv block_div(int64_t b_w) {
    int l = 0;
    while ((int64_t{1} << l) < b_w) ++l;
    const uint64_t m = ((uint64_t{1} << (31 + l)) + b_w - 1) / b_w;
    return {static_cast<uint32_t>(m), 31 + l};
}

template <typename V>
struct BcrsBuild {
    int64_t nbr, n_cols;
    int b_h, b_w, h_pad, w_pad, col_major;
    int cap;  // entries the LDS tile holds
    BlockDiv div;
    const int32_t *row_ptr, *col;
    const V *val;
    int64_t *cnt;        // count pass: blocks per block row
    int64_t *bad;        // count pass: [0] descending rows, [1] columns out of range
    const int64_t *incl; // fill pass: inclusive scan of cnt
    int32_t *brow_ptr, *bcol;
    V *bval;
};

// Entry j of the CRS arrays, from the LDS tile (which starts at entry
// base) or from memory.
template <bool STAGED, typename T>
struct Entries {
    const T *p;
    int32_t base;
    __device__ T operator[](int32_t j) const { return STAGED ? p[j - base] : p[j]; }
};

template <int GS, bool FILL, bool STAGED, typename V>
__device__ void merge_block_row(const BcrsBuild<V> &a, int64_t br, bool row_valid,
                                Entries<STAGED, int32_t> cs, Entries<STAGED, V> vs) {
    const int lane = static_cast<int>(threadIdx.x) & (GS - 1);
    const bool mine = row_valid && lane < a.b_h;
    int32_t pos = 0, end = 0;
    if (mine) {
        const int64_t r = br * a.b_h + lane;
        pos = a.row_ptr[r];
        end = a.row_ptr[r + 1];
    }
    int desc = 0, range = 0;
    if (!FILL && mine) {
        int32_t prev = -1;
        for (int32_t j = pos; j < end; ++j) {
            const int32_t c = cs[j];
            range |= static_cast<uint32_t>(c) >= static_cast<uint64_t>(a.n_cols);
            const int32_t bc = a.div(c);
            desc |= bc < prev;
            prev = bc;
        }
    }
    if (!FILL) {
#pragma unroll
        for (int off = GS / 2; off >= 1; off >>= 1) {
            desc |= __shfl_xor(desc, off, GS);
            range |= __shfl_xor(range, off, GS);
        }
        if (desc || range) {
            if (row_valid && lane == 0) {
                a.cnt[br] = 0;
                if (desc) atomicAdd(reinterpret_cast<unsigned long long *>(a.bad), 1ull);
                if (range) atomicAdd(reinterpret_cast<unsigned long long *>(a.bad + 1), 1ull);
            }
            return;
        }
    }
    const int64_t start = FILL && row_valid ? (br == 0 ? 0 : a.incl[br - 1]) : 0;
    const int64_t bs = static_cast<int64_t>(a.h_pad) * a.w_pad;
    int64_t u = 0;
    for (;;) {
        int32_t m = pos < end ? a.div(cs[pos]) : INT32_MAX;
#pragma unroll
        for (int off = GS / 2; off >= 1; off >>= 1) m = min(m, __shfl_xor(m, off, GS));
        if (m == INT32_MAX) break;
        const int64_t blk = start + u;
        while (pos < end) {
            const int32_t c = cs[pos];
            if (a.div(c) != m) break;
            if (FILL) {
                const int64_t in = c - static_cast<int64_t>(m) * a.b_w;
                const int64_t inc = a.col_major ? static_cast<int64_t>(a.h_pad) * in + lane
                                                : static_cast<int64_t>(a.w_pad) * lane + in;
                a.bval[blk * bs + inc] = vs[pos];
            }
            ++pos;
        }
        if (FILL && lane == 0) a.bcol[blk] = m;
        ++u;
    }
    if (row_valid && lane == 0) {
        if (FILL) {
            a.brow_ptr[br + 1] = static_cast<int32_t>(a.incl[br]);
            if (br == 0) a.brow_ptr[0] = 0;
        } else {
            a.cnt[br] = u;
        }
    }
}

// Count pass (FILL false): blocks per block row, descending rows and
// out-of-range columns.  Fill pass: block row pointers, block columns and
// values (into zeroed blocks).  Dynamic LDS: cap columns, then (fill pass)
// cap values.
template <int GS, bool FILL, typename V>
__global__ __launch_bounds__(256) void bcrs_build(BcrsBuild<V> a) {
    extern __shared__ int32_t tile_col[];
    V *tile_val = reinterpret_cast<V *>(tile_col + a.cap);
    constexpr int kRows = 256 / GS;
    const int64_t br0 = static_cast<int64_t>(blockIdx.x) * kRows;
    const int64_t br1 = min<int64_t>(br0 + kRows, a.nbr);
    const int64_t br = br0 + static_cast<int>(threadIdx.x) / GS;
    const bool row_valid = br < br1;
    const int32_t base = a.row_ptr[br0 * a.b_h];
    const int32_t len = a.row_ptr[br1 * a.b_h] - base;
    if (len >= 0 && len <= a.cap) {
        for (int32_t t = threadIdx.x; t < len; t += 256) {
            tile_col[t] = a.col[base + t];
            if (FILL) tile_val[t] = a.val[base + t];
        }
        __syncthreads();
        merge_block_row<GS, FILL, true>(a, br, row_valid, Entries<true, int32_t>{tile_col, base},
                                        Entries<true, V>{tile_val, base});
    } else {
        merge_block_row<GS, FILL, false>(a, br, row_valid, Entries<false, int32_t>{a.col, 0},
                                         Entries<false, V>{a.val, 0});
    }
}

// Lanes per block row of the conversion: the power of two >= b_h (<= 64).
inline int bcrs_group(int64_t b_h) {
    int g = 1;
    while (g < b_h) g <<= 1;
    return g;
}

// LDS tile of a pass: room for 1.25x a workgroup's average entries (plus
// slack), at most 64 KB.
inline int bcrs_tile_entries(int64_t nnz, int64_t n_wg, int64_t entry_bytes) {
    const int64_t avg = nnz / (n_wg > 0 ? n_wg : 1);
    const int64_t most = (65536 / entry_bytes) / 256 * 256;
    const int64_t want = (avg + avg / 4 + 256 + 255) / 256 * 256;
    return static_cast<int>(want < most ? want : most);
}

template <bool FILL, typename V>
static int launch_bcrs_build(smax_hip_ctx ctx, BcrsBuild<V> a, int64_t nnz) {
    const int gs = bcrs_group(a.b_h);
    const int64_t n_wg = (a.nbr + 256 / gs - 1) / (256 / gs);
    const int64_t entry = 4 + (FILL ? sizeof(V) : 0);
    a.cap = bcrs_tile_entries(nnz, n_wg, entry);
    const size_t lds = static_cast<size_t>(a.cap) * entry;
    switch (gs) {
#define SMAX_BCRS_BUILD(G)                                                                \
    case G:                                                                               \
        hipLaunchKernelGGL((bcrs_build<G, FILL, V>), dim3(static_cast<unsigned>(n_wg)),   \
                           dim3(256), lds, ctx->stream, a);                               \
        break;
        SMAX_BCRS_BUILD(1)
        SMAX_BCRS_BUILD(2)
        SMAX_BCRS_BUILD(4)
        SMAX_BCRS_BUILD(8)
        SMAX_BCRS_BUILD(16)
        SMAX_BCRS_BUILD(32)
        SMAX_BCRS_BUILD(64)
#undef SMAX_BCRS_BUILD
    }
    return check_launch(ctx, FILL ? "bcrs_build(fill)" : "bcrs_build(count)");
}

static int check_bcrs(smax_hip_ctx ctx, const char *fn, int value_type,
                      int64_t nbr, int64_t b_h, int64_t b_w, int64_t h_pad,
                      int64_t w_pad) {
    if (!valid_value_type(value_type))
        return fail(ctx, SMAX_HIP_ERROR_UNSUPPORTED, "%s: value type %d", fn,
                    value_type);
    SMAX_HIP_REQUIRE(ctx, nbr > 0 && b_h > 0 && b_w > 0 && h_pad >= b_h &&
                              w_pad >= b_w,
                     "%s: bad block shape (%lld rows, %lldx%lld pad %lldx%lld)",
                     fn, (long long)nbr, (long long)b_h, (long long)b_w,
                     (long long)h_pad, (long long)w_pad);
    SMAX_HIP_REQUIRE(ctx, b_h <= kBThreads && h_pad * w_pad <= kBTileVals,
                     "%s: blocks larger than %d values are not supported", fn,
                     kBTileVals);
    if (nbr > INT32_MAX)
        return fail(ctx, SMAX_HIP_ERROR_OVERFLOW, "%s: %lld block rows exceed int32", fn,
                    (long long)nbr);
    return SMAX_HIP_SUCCESS;
}

}  // namespace smaxhip

using namespace smaxhip;

extern "C" {

int smax_hip_bcrs_spmv(smax_hip_ctx ctx, int value_type,
                       int64_t n_block_rows, int64_t b_h, int64_t b_w,
                       int64_t h_pad, int64_t w_pad, const int32_t *col,
                       const int32_t *row_ptr, const void *val,
                       int block_col_major, const void *x, void *y) {
    SMAX_HIP_CHECK_CTX(ctx);
    int rc = check_bcrs(ctx, "smax_hip_bcrs_spmv", value_type, n_block_rows,
                        b_h, b_w, h_pad, w_pad);
    if (rc) return rc;
    SMAX_HIP_REQUIRE(ctx, row_ptr && x && y, "smax_hip_bcrs_spmv: null pointer");
    // b_h, b_w in {1, 2, 4}: the register-blocked product with one right-hand
    // side (a thread per block row, the same chains; twice as fast as staging
    // the blocks in LDS on the 96^3 block matrix)
    if ((b_h == 1 || b_h == 2 || b_h == 4) && (b_w == 1 || b_w == 2 || b_w == 4))
        return smax_hip_bcrs_spmm(ctx, value_type, n_block_rows, b_h, b_w, h_pad, w_pad, col,
                                  row_ptr, val, block_col_major, 1, x, y);
    const int64_t rows_per_wg = kBThreads / b_h;
    SMAX_HIP_DISPATCH_VT(value_type, {
        using T = vt<VT>;
        hipLaunchKernelGGL((bcrs_spmv_lds<VT>), grid_for(n_block_rows, rows_per_wg),
                           dim3(kBThreads), 0, ctx->stream, n_block_rows,
                           static_cast<int>(b_h), static_cast<int>(b_w),
                           static_cast<int>(h_pad), static_cast<int>(w_pad),
                           block_col_major ? 1 : 0, col, row_ptr,
                           static_cast<const typename T::val_t *>(val),
                           static_cast<const typename T::vec_t *>(x),
                           static_cast<typename T::vec_t *>(y));
        SMAX_HIP_LAUNCHED(ctx, "bcrs_spmv_lds");
        return SMAX_HIP_SUCCESS;
    });
    return SMAX_HIP_ERROR_UNSUPPORTED;
}

int smax_hip_bcrs_spmm(smax_hip_ctx ctx, int value_type,
                       int64_t n_block_rows, int64_t b_h, int64_t b_w,
                       int64_t h_pad, int64_t w_pad, const int32_t *col,
                       const int32_t *row_ptr, const void *val,
                       int block_col_major, int64_t k, const void *X, void *Y) {
    SMAX_HIP_CHECK_CTX(ctx);
    int rc = check_bcrs(ctx, "smax_hip_bcrs_spmm", value_type, n_block_rows,
                        b_h, b_w, h_pad, w_pad);
    if (rc) return rc;
    SMAX_HIP_REQUIRE(ctx, k > 0 && row_ptr && X && Y,
                     "smax_hip_bcrs_spmm: bad k or null pointer");
    if (!extent_ok(n_block_rows * h_pad, k))
        return fail(ctx, SMAX_HIP_ERROR_OVERFLOW,
                    "smax_hip_bcrs_spmm: %lld rows x %lld columns exceed the launch limits",
                    (long long)(n_block_rows * h_pad), (long long)k);
    const int64_t threads = n_block_rows * b_h * k;
    const bool cm = block_col_major != 0;
    // register-blocked shapes: every (b_h, b_w) in {1, 2, 4} x {1, 2, 4}
    const bool reg = (b_h == 1 || b_h == 2 || b_h == 4) && (b_w == 1 || b_w == 2 || b_w == 4);
#define SMAX_BCRS_REG(H, W)                                                            \
    if (b_h == H && b_w == W) {                                                        \
        launch_bcrs_spmm_reg<VT, H, W>(ctx, n_block_rows, h_pad, w_pad, cm, k, col,    \
                                       row_ptr, val, X, Y);                            \
        SMAX_HIP_LAUNCHED(ctx, "bcrs_spmm_reg");                                       \
        return SMAX_HIP_SUCCESS;                                                       \
    }
    if (reg) {
        SMAX_HIP_DISPATCH_VT(value_type, {
            SMAX_BCRS_REG(1, 1) SMAX_BCRS_REG(1, 2) SMAX_BCRS_REG(1, 4)
            SMAX_BCRS_REG(2, 1) SMAX_BCRS_REG(2, 2) SMAX_BCRS_REG(2, 4)
            SMAX_BCRS_REG(4, 1) SMAX_BCRS_REG(4, 2) SMAX_BCRS_REG(4, 4)
        });
    }
#undef SMAX_BCRS_REG
    SMAX_HIP_DISPATCH_VT(value_type, {
        switch (b_w) {
        case 1:
            launch_bcrs_spmm<VT, 1>(ctx, threads, n_block_rows, b_h, b_w, h_pad, w_pad,
                                    cm, k, col, row_ptr, val, X, Y);
            break;
        case 2:
            launch_bcrs_spmm<VT, 2>(ctx, threads, n_block_rows, b_h, b_w, h_pad, w_pad,
                                    cm, k, col, row_ptr, val, X, Y);
            break;
        case 4:
            launch_bcrs_spmm<VT, 4>(ctx, threads, n_block_rows, b_h, b_w, h_pad, w_pad,
                                    cm, k, col, row_ptr, val, X, Y);
            break;
        case 8:
            launch_bcrs_spmm<VT, 8>(ctx, threads, n_block_rows, b_h, b_w, h_pad, w_pad,
                                    cm, k, col, row_ptr, val, X, Y);
            break;
        default:
            launch_bcrs_spmm<VT, 0>(ctx, threads, n_block_rows, b_h, b_w, h_pad, w_pad,
                                    cm, k, col, row_ptr, val, X, Y);
            break;
        }
        SMAX_HIP_LAUNCHED(ctx, "bcrs_spmm_kernel");
        return SMAX_HIP_SUCCESS;
    });
    return SMAX_HIP_ERROR_UNSUPPORTED;
}

int smax_hip_bcrs_from_crs(smax_hip_ctx ctx, int64_t n_rows, int64_t n_cols,
                           int64_t nnz, const int32_t *row_ptr,
                           const int32_t *col, const void *val,
                           int64_t val_bytes, int64_t b_h, int64_t b_w,
                           int64_t h_pad, int64_t w_pad, int block_col_major,
                           smax_hip_bcrs *out) {
    SMAX_HIP_CHECK_CTX(ctx);
    SMAX_HIP_REQUIRE(ctx, out, "smax_hip_bcrs_from_crs: null output");
    *out = nullptr;
    SMAX_HIP_REQUIRE(ctx, n_rows > 0 && n_cols > 0 && nnz >= 0 && row_ptr &&
                              (nnz == 0 || (col && val)),
                     "smax_hip_bcrs_from_crs: bad input");
    if (n_rows > INT32_MAX || n_cols > INT32_MAX || nnz > INT32_MAX)
        return fail(ctx, SMAX_HIP_ERROR_OVERFLOW,
                    "smax_hip_bcrs_from_crs: %lld x %lld with %lld entries exceeds int32",
                    (long long)n_rows, (long long)n_cols, (long long)nnz);
    SMAX_HIP_REQUIRE(ctx, val_bytes == 2 || val_bytes == 4 || val_bytes == 8,
                     "smax_hip_bcrs_from_crs: val_bytes must be 2, 4 or 8");
    // convert_formats.hpp:142-157
    SMAX_HIP_REQUIRE(ctx, b_h > 0 && b_w > 0 && n_rows % b_h == 0,
                     "Target block height does not divide number of input rows");
    SMAX_HIP_REQUIRE(ctx, n_cols % b_w == 0,
                     "Target block width does not divide number of input columns");
    SMAX_HIP_REQUIRE(ctx, h_pad >= b_h,
                     "Height padding has to be at least target block height");
    SMAX_HIP_REQUIRE(ctx, w_pad >= b_w,
                     "Width padding has to be at least target block width");
    SMAX_HIP_REQUIRE(ctx, b_h <= 64, "smax_hip_bcrs_from_crs: b_h > 64");
    const int64_t nbr = n_rows / b_h;
    int64_t *cnt = nullptr, *incl = nullptr, *tmp = nullptr, *bad = nullptr;
    smax_hip_bcrs_info info{};
    int rc = dev_alloc(ctx, reinterpret_cast<void **>(&cnt), 8 * nbr, "bcrs");
    if (!rc) rc = dev_alloc(ctx, reinterpret_cast<void **>(&incl), 8 * nbr, "bcrs");
    if (!rc)
        rc = dev_alloc(ctx, reinterpret_cast<void **>(&tmp),
                       8 * scan_tmp_elems(nbr), "bcrs");
    if (!rc) rc = dev_alloc(ctx, reinterpret_cast<void **>(&bad), 16, "bcrs");
    if (!rc) rc = smax_hip_zero(ctx, bad, 16);
    BcrsBuild<uint32_t> args{};
    args.nbr = nbr;
    args.n_cols = n_cols;
    args.b_h = static_cast<int>(b_h);
    args.b_w = static_cast<int>(b_w);
    args.h_pad = static_cast<int>(h_pad);
    args.w_pad = static_cast<int>(w_pad);
    args.col_major = block_col_major ? 1 : 0;
    args.div = block_div(b_w);
    args.row_ptr = row_ptr;
    args.col = col;
    args.cnt = cnt;
    args.bad = bad;
    int64_t n_blocks = 0, nbad[2] = {0, 0};
    if (!rc) rc = launch_bcrs_build<false>(ctx, args, nnz);
    if (!rc) rc = scan_inclusive<int64_t, false>(ctx, cnt, incl, nbr, tmp);
    if (!rc) rc = read_i64(ctx, bad, &nbad[0]);
    if (!rc) rc = read_i64(ctx, bad + 1, &nbad[1]);
    if (!rc) rc = read_i64(ctx, incl + nbr - 1, &n_blocks);
    if (!rc && nbad[1])
        rc = fail(ctx, SMAX_HIP_ERROR_INVALID_VALUE,
                  "smax_hip_bcrs_from_crs: %lld block rows hold column indices outside "
                  "[0, %lld)", (long long)nbad[1], (long long)n_cols);
    if (!rc && nbad[0])
        rc = fail(ctx, SMAX_HIP_ERROR_INVALID_VALUE,
                  "Blocked Column array does not fit with crs columns "
                  "(%lld block rows with descending columns)", (long long)nbad[0]);
    if (!rc && n_blocks * h_pad * w_pad > INT32_MAX * 4LL)
        rc = fail(ctx, SMAX_HIP_ERROR_OVERFLOW, "bcrs: too many blocks");
    if (!rc)
        rc = dev_alloc(ctx, reinterpret_cast<void **>(&info.row_ptr),
                       4 * (nbr + 1), "bcrs row_ptr");
    if (!rc)
        rc = dev_alloc(ctx, reinterpret_cast<void **>(&info.col), 4 * n_blocks,
                       "bcrs col");
    if (!rc)
        rc = dev_alloc(ctx, &info.val, val_bytes * n_blocks * h_pad * w_pad,
                       "bcrs val");
    if (!rc) rc = smax_hip_zero(ctx, info.val, val_bytes * n_blocks * h_pad * w_pad);
    if (!rc) {
        auto fill = [&](auto *vals) {
            using V = std::remove_const_t<std::remove_pointer_t<decltype(vals)>>;
            BcrsBuild<V> f{};
            f.nbr = args.nbr;
            f.n_cols = args.n_cols;
            f.b_h = args.b_h;
            f.b_w = args.b_w;
            f.h_pad = args.h_pad;
            f.w_pad = args.w_pad;
            f.col_major = args.col_major;
            f.div = args.div;
            f.row_ptr = row_ptr;
            f.col = col;
            f.val = vals;
            f.incl = incl;
            f.brow_ptr = info.row_ptr;
            f.bcol = info.col;
            f.bval = static_cast<V *>(info.val);
            return launch_bcrs_build<true>(ctx, f, nnz);
        };
        if (val_bytes == 8)
            rc = fill(static_cast<const uint64_t *>(val));
        else if (val_bytes == 4)
            rc = fill(static_cast<const uint32_t *>(val));
        else
            rc = fill(static_cast<const uint16_t *>(val));
    }
    (void)hipStreamSynchronize(ctx->stream);
    dev_free(cnt);
    dev_free(incl);
    dev_free(tmp);
    dev_free(bad);
    if (rc) {
        dev_free(info.row_ptr);
        dev_free(info.col);
        dev_free(info.val);
        return rc;
    }
    info.n_block_rows = nbr;
    info.n_block_cols = n_cols / b_w;
    info.n_blocks = n_blocks;
    info.b_h = b_h;
    info.b_w = b_w;
    info.h_pad = h_pad;
    info.w_pad = w_pad;
    info.val_bytes = val_bytes;
    info.block_col_major = block_col_major ? 1 : 0;
    smax_hip_bcrs m = new smax_hip_bcrs_s();
    m->info = info;
    *out = m;
    return SMAX_HIP_SUCCESS;
}

int smax_hip_bcrs_get_info(smax_hip_bcrs bcrs, smax_hip_bcrs_info *info) {
    if (!bcrs || !info)
        return fail(nullptr, SMAX_HIP_ERROR_INVALID_VALUE,
                    "smax_hip_bcrs_get_info: null argument");
    *info = bcrs->info;
    return SMAX_HIP_SUCCESS;
}

int smax_hip_bcrs_destroy(smax_hip_ctx ctx, smax_hip_bcrs bcrs) {
    if (!bcrs) return SMAX_HIP_SUCCESS;
    if (ctx) {
        (void)bind(ctx);
        (void)hipStreamSynchronize(ctx->stream);
    }
    dev_free(bcrs->info.row_ptr);
    dev_free(bcrs->info.col);
    dev_free(bcrs->info.val);
    dev_free(bcrs->info.src_block);
    delete bcrs;
    return SMAX_HIP_SUCCESS;
}

}  // extern "C"

// ------------------------------------------------------------ adjoint
#include "radix.hpp"

namespace smaxhip {

template <typename V>
__global__ void bcrs_gather_blocks(int64_t total, int64_t bs,
                                   const int32_t *__restrict__ src_block,
                                   const V *__restrict__ src, V *__restrict__ out) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= total) return;
    const int64_t p = i / bs, t = i - p * bs;
    out[i] = src[static_cast<int64_t>(src_block[p]) * bs + t];
}

static int gather_blocks(smax_hip_ctx ctx, const smax_hip_bcrs_info &T,
                         const void *src) {
    const int64_t bs = T.h_pad * T.w_pad, total = T.n_blocks * bs;
    if (total == 0) return SMAX_HIP_SUCCESS;
    const dim3 g = grid_for(total, 256);
    if (T.val_bytes == 8)
        hipLaunchKernelGGL(bcrs_gather_blocks<uint64_t>, g, dim3(256), 0, ctx->stream,
                           total, bs, T.src_block, static_cast<const uint64_t *>(src),
                           static_cast<uint64_t *>(T.val));
    else if (T.val_bytes == 4)
        hipLaunchKernelGGL(bcrs_gather_blocks<uint32_t>, g, dim3(256), 0, ctx->stream,
                           total, bs, T.src_block, static_cast<const uint32_t *>(src),
                           static_cast<uint32_t *>(T.val));
    else
        hipLaunchKernelGGL(bcrs_gather_blocks<uint16_t>, g, dim3(256), 0, ctx->stream,
                           total, bs, T.src_block, static_cast<const uint16_t *>(src),
                           static_cast<uint16_t *>(T.val));
    SMAX_HIP_LAUNCHED(ctx, "bcrs_gather_blocks");
    return SMAX_HIP_SUCCESS;
}

// Group of G lanes per block position; fixed xor-butterfly over the group.
template <int VT, int G>
__global__ __launch_bounds__(256) void bcrs_sddmm_kernel(
    int64_t n_positions, int64_t n_block_rows, int b_h, int b_w, int h_pad, int w_pad,
    int col_major, int64_t k, const int32_t *__restrict__ row_ptr,
    const int32_t *__restrict__ col, const typename vt<VT>::vec_t *__restrict__ Gm,
    const typename vt<VT>::vec_t *__restrict__ X,
    typename vt<VT>::vec_t *__restrict__ dvals) {
    using T = vt<VT>;
    using A = typename T::vec_t;
    const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    const int64_t e = t / G;
    const int lane = static_cast<int>(t % G);
    const int64_t bs = static_cast<int64_t>(h_pad) * w_pad;
    const int64_t j = e / bs, off = e - j * bs;
    const int ix = static_cast<int>(col_major ? off % h_pad : off / w_pad);
    const int jx = static_cast<int>(col_major ? off / h_pad : off % w_pad);
    const bool real = e < n_positions && ix < b_h && jx < b_w;
    A acc = 0;
    if (real) {
        int64_t lo = 0, hi = n_block_rows;  // block row of block j
        while (hi - lo > 1) {
            const int64_t mid = (lo + hi) >> 1;
            if (row_ptr[mid] <= j) lo = mid; else hi = mid;
        }
        const A *gr = Gm + (lo * h_pad + ix) * k;
        const A *xr = X + (static_cast<int64_t>(col[j]) * w_pad + jx) * k;
        for (int64_t v = lane; v < k; v += G) acc = T::fma_(gr[v], xr[v], acc);
    }
#pragma unroll
    for (int o = G / 2; o >= 1; o >>= 1) acc += __shfl_xor(acc, o, G);
    if (e < n_positions && lane == 0) dvals[e] = real ? acc : A(0);
}

template <int VT, int G>
static void launch_bcrs_sddmm_g(smax_hip_ctx ctx, const smax_hip_bcrs_info &in,
                                int64_t k, const void *Gm, const void *X, void *dv) {
    using T = vt<VT>;
    const int64_t n = in.n_blocks * in.h_pad * in.w_pad;
    hipLaunchKernelGGL((bcrs_sddmm_kernel<VT, G>), grid_for(n * G, 256), dim3(256), 0,
                       ctx->stream, n, in.n_block_rows, static_cast<int>(in.b_h),
                       static_cast<int>(in.b_w), static_cast<int>(in.h_pad),
                       static_cast<int>(in.w_pad), static_cast<int>(in.block_col_major),
                       k, in.row_ptr, in.col,
                       static_cast<const typename T::vec_t *>(Gm),
                       static_cast<const typename T::vec_t *>(X),
                       static_cast<typename T::vec_t *>(dv));
}

template <int VT>
static int launch_bcrs_sddmm(smax_hip_ctx ctx, const smax_hip_bcrs_info &in,
                             int64_t k, int g, const void *Gm, const void *X,
                             void *dv) {
    switch (g) {
    case 1: launch_bcrs_sddmm_g<VT, 1>(ctx, in, k, Gm, X, dv); break;
    case 2: launch_bcrs_sddmm_g<VT, 2>(ctx, in, k, Gm, X, dv); break;
    case 4: launch_bcrs_sddmm_g<VT, 4>(ctx, in, k, Gm, X, dv); break;
    case 8: launch_bcrs_sddmm_g<VT, 8>(ctx, in, k, Gm, X, dv); break;
    case 16: launch_bcrs_sddmm_g<VT, 16>(ctx, in, k, Gm, X, dv); break;
    case 32: launch_bcrs_sddmm_g<VT, 32>(ctx, in, k, Gm, X, dv); break;
    default: launch_bcrs_sddmm_g<VT, 64>(ctx, in, k, Gm, X, dv); break;
    }
    SMAX_HIP_LAUNCHED(ctx, "bcrs_sddmm_kernel");
    return SMAX_HIP_SUCCESS;
}

}  // namespace smaxhip

extern "C" {

int smax_hip_bcrs_transpose(smax_hip_ctx ctx, smax_hip_bcrs A, smax_hip_bcrs *out) {
    SMAX_HIP_CHECK_CTX(ctx);
    SMAX_HIP_REQUIRE(ctx, A && out, "smax_hip_bcrs_transpose: null argument");
    *out = nullptr;
    const smax_hip_bcrs_info &a = A->info;
    const int64_t nb = a.n_blocks, m = nb > 0 ? nb : 1;
    smax_hip_bcrs_info t{};
    t.n_block_rows = a.n_block_cols;
    t.n_block_cols = a.n_block_rows;
    t.n_blocks = nb;
    t.b_h = a.b_w;
    t.b_w = a.b_h;
    t.h_pad = a.w_pad;
    t.w_pad = a.h_pad;
    t.val_bytes = a.val_bytes;
    t.block_col_major = a.block_col_major ? 0 : 1;
    RadixSort sorter(ctx);
    int rc = sorter.reserve(m);
    if (!rc)
        rc = dev_alloc(ctx, reinterpret_cast<void **>(&t.row_ptr),
                       4 * (t.n_block_rows + 1), "bcrs transpose");
    if (!rc) rc = dev_alloc(ctx, reinterpret_cast<void **>(&t.col), 4 * m, "bcrs transpose");
    if (!rc)
        rc = dev_alloc(ctx, reinterpret_cast<void **>(&t.src_block), 4 * m,
                       "bcrs transpose");
    if (!rc) rc = dev_alloc(ctx, &t.val, a.val_bytes * m * a.h_pad * a.w_pad, "bcrs transpose");
    if (!rc && nb > 0) {
        hipLaunchKernelGGL(iota_keys, grid_for(nb, 256), dim3(256), 0, ctx->stream, nb,
                           a.col, sorter.keys(), sorter.vals());
        rc = check_launch(ctx, "iota_keys");
        if (!rc) rc = sorter.sort(nb, a.n_block_cols);
    }
    if (!rc) {
        hipLaunchKernelGGL(transpose_row_ptr, grid_for(t.n_block_rows + 1, 256),
                           dim3(256), 0, ctx->stream, t.n_block_rows, nb, sorter.keys(),
                           t.row_ptr);
        rc = check_launch(ctx, "transpose_row_ptr");
    }
    if (!rc && nb > 0) {
        hipLaunchKernelGGL(transpose_cols, grid_for(nb, 256), dim3(256), 0, ctx->stream,
                           nb, a.n_block_rows, a.row_ptr, sorter.vals(), t.col);
        rc = check_launch(ctx, "transpose_cols");
        if (!rc) {
            const hipError_t e = hipMemcpyAsync(t.src_block, sorter.vals(), 4 * nb,
                                                hipMemcpyDeviceToDevice, ctx->stream);
            if (e != hipSuccess) rc = hip_fail(ctx, e, "src_block copy");
        }
        if (!rc) rc = gather_blocks(ctx, t, a.val);
    }
    (void)hipStreamSynchronize(ctx->stream);
    if (rc) {
        dev_free(t.row_ptr);
        dev_free(t.col);
        dev_free(t.src_block);
        dev_free(t.val);
        return rc;
    }
    smax_hip_bcrs h = new smax_hip_bcrs_s();
    h->info = t;
    *out = h;
    return SMAX_HIP_SUCCESS;
}

int smax_hip_bcrs_refresh_values(smax_hip_ctx ctx, smax_hip_bcrs T,
                                 const void *src_val) {
    SMAX_HIP_CHECK_CTX(ctx);
    SMAX_HIP_REQUIRE(ctx, T && src_val && T->info.src_block,
                     "smax_hip_bcrs_refresh_values: not a transpose or null input");
    return gather_blocks(ctx, T->info, src_val);
}

int smax_hip_bcrs_sddmm(smax_hip_ctx ctx, int value_type, smax_hip_bcrs A, int64_t k,
                        const void *G, const void *X, void *dvals) {
    SMAX_HIP_CHECK_CTX(ctx);
    if (!valid_value_type(value_type))
        return fail(ctx, SMAX_HIP_ERROR_UNSUPPORTED, "smax_hip_bcrs_sddmm: value type %d",
                    value_type);
    SMAX_HIP_REQUIRE(ctx, A && k > 0 && G && X && dvals,
                     "smax_hip_bcrs_sddmm: bad arguments");
    if (!extent_ok(A->info.n_block_rows * A->info.h_pad, k) ||
        !extent_ok(A->info.n_block_cols * A->info.w_pad, k))
        return fail(ctx, SMAX_HIP_ERROR_OVERFLOW,
                    "smax_hip_bcrs_sddmm: k = %lld exceeds the launch limits", (long long)k);
    if (A->info.n_blocks == 0) return SMAX_HIP_SUCCESS;
    int g = 1;
    while (g < k && g < 64) g <<= 1;
    SMAX_HIP_DISPATCH_VT(value_type,
                         return launch_bcrs_sddmm<VT>(ctx, A->info, k, g, G, X, dvals));
    return SMAX_HIP_ERROR_UNSUPPORTED;
}

}  // extern "C"
