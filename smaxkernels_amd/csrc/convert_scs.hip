// CRS -> SELL-C-sigma on the GPU, bit-identical to
// SMAX::Utils::convert_crs_to_scs (src/utils/convert_formats.hpp:16-132).
//
// The only non-trivial step is the per-sigma-window std::sort of
// (index, count) pairs by count descending (convert_formats.hpp:46-55).
// std::sort is unstable, so tied rows end up in the order libstdc++'s
// introsort leaves them; that order is reproduced exactly here by running the
// introsort recursion level-synchronously over every window at once:
//
//  * a segment [h, e) with e - h > 16 and depth > 0 is partitioned exactly as
//    __unguarded_partition_pivot does (bits/stl_algo.h:1878-1907): the
//    median of (h+1, h+(e-h)/2, e-1) is swapped to h, then the Hoare scan
//    swaps the k-th "left stopper" (count <= pivot, scanning up from h+1)
//    with the k-th "right stopper" (count >= pivot, scanning down from e-1)
//    for every k with l_k < r_k.  With L(q) / R(q) the number of left / right
//    stoppers in (h, q] / [q, e), the swap count is s = max_q min(L(q),
//    R(q+1)) and the returned cut is min(l_{s+1}, r_s) — both read off
//    prefix sums, so every element's destination is known in parallel;
//  * both halves inherit depth - 1 (bits/stl_algo.h:1925-1941);
//  * a segment still longer than 16 at depth 0 is heap-sorted sequentially
//    (make_heap + sort_heap, bits/stl_heap.h);
//  * the closing __final_insertion_sort never moves an element across a
//    partition boundary (left parts hold counts >= pivot, right parts
//    <= pivot, and insertion only passes strictly smaller counts), so it is
//    a stable insertion sort of every segment of <= 16 elements.
//
// Everything else (chunk lengths, chunk_ptr, perm, zero-padded col/val and
// the CRS-order scatter) is data-parallel.
#include <algorithm>
#include <vector>

#include "common.hpp"
#include "convert_internal.hpp"
#include "scan.hpp"

namespace smaxhip {

enum : int32_t { SEG_ACTIVE = 0, SEG_LEAF = 1, SEG_HEAP = 2, SEG_NONE = 3 };
constexpr int kThreshold = 16;  // _S_threshold, bits/stl_algo.h:1855
constexpr int kThreads = 256;

__device__ __forceinline__ int32_t lg2(int64_t n) {
    return 63 - __clzll(static_cast<unsigned long long>(n));
}

// ------------------------------------------------------------ init
__global__ void k_init(int64_t n, int64_t n_rows, int64_t sigma,
                       const int32_t *__restrict__ row_ptr,
                       int32_t *__restrict__ idx, int32_t *__restrict__ cnt,
                       int32_t *__restrict__ head, int32_t *__restrict__ end,
                       int32_t *__restrict__ depth, int32_t *__restrict__ state,
                       int32_t *__restrict__ list, int64_t *__restrict__ nlist) {
    int64_t q = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (q >= n) return;
    idx[q] = static_cast<int32_t>(q);
    cnt[q] = q < n_rows ? row_ptr[q + 1] - row_ptr[q] : 0;
    int64_t h = (q / sigma) * sigma;
    head[q] = static_cast<int32_t>(h);
    if (h == q) {
        int64_t e = h + sigma < n ? h + sigma : n;
        int64_t len = e - h;
        end[q] = static_cast<int32_t>(e);
        depth[q] = 2 * lg2(len);  // std::__lg(last - first) * 2
        if (len > kThreshold) {
            state[q] = SEG_ACTIVE;
            list[atomicAdd(reinterpret_cast<unsigned long long *>(nlist), 1ull)] =
                static_cast<int32_t>(q);
        } else {
            state[q] = SEG_LEAF;
        }
    } else {
        state[q] = SEG_NONE;
    }
}

// ------------------------------------------------------------ level kernels
__device__ __forceinline__ void swap_pair(int32_t *idx, int32_t *cnt,
                                          int64_t a, int64_t b) {
    int32_t ti = idx[a], tc = cnt[a];
    idx[a] = idx[b];
    cnt[a] = cnt[b];
    idx[b] = ti;
    cnt[b] = tc;
}

// __move_median_to_first (bits/stl_algo.h:79-97) with comp(a,b) = a > b.
__global__ void k_pivot(int64_t nl, const int32_t *__restrict__ list,
                        const int32_t *__restrict__ end,
                        const int32_t *__restrict__ depth,
                        int32_t *__restrict__ state, int32_t *__restrict__ idx,
                        int32_t *__restrict__ cnt, int32_t *__restrict__ pv,
                        int32_t *__restrict__ sw) {
    int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= nl) return;
    const int64_t h = list[t];
    if (depth[h] == 0) {  // __partial_sort fallback, bits/stl_algo.h:1931-1935
        state[h] = SEG_HEAP;
        return;
    }
    const int64_t e = end[h];
    const int64_t a = h + 1, b = h + (e - h) / 2, c = e - 1;
    const int32_t ca = cnt[a], cb = cnt[b], cc = cnt[c];
    int64_t m;
    if (ca > cb) {
        if (cb > cc) m = b;
        else if (ca > cc) m = c;
        else m = a;
    } else if (ca > cc) m = a;
    else if (cb > cc) m = c;
    else m = b;
    swap_pair(idx, cnt, h, m);
    pv[h] = cnt[h];
    sw[h] = 0;
}

// Left / right stopper flags of every element inside an active segment.
__global__ void k_flags(int64_t n, const int32_t *__restrict__ head,
                        const int32_t *__restrict__ state,
                        const int32_t *__restrict__ cnt,
                        const int32_t *__restrict__ pv,
                        int32_t *__restrict__ lf, int32_t *__restrict__ rf) {
    int64_t q = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const int64_t h = head[q];
    int32_t l = 0, r = 0;
    if (state[h] == SEG_ACTIVE && q > h) {
        const int32_t c = cnt[q], p = pv[h];
        l = c <= p;  // !comp(*first, *pivot): left scan stops
        r = c >= p;  // !comp(*pivot, *last):  right scan stops
    }
    lf[q] = l;
    rf[q] = r;
}

__device__ __forceinline__ int32_t suffix_at(const int32_t *SR, int64_t q,
                                             int64_t n) {
    return q < n ? SR[q] : 0;
}

// Stopper ranks, position lists and the swap count s of every segment.
// s is a maximum over the segment's positions: a wavefront whose lanes all
// lie in one segment takes its maximum first and issues one atomic (one
// per element would serialise millions of atomics on one address when a
// sigma window is one large segment).
__global__ __launch_bounds__(256) void k_rank(
    int64_t n, const int32_t *__restrict__ head, const int32_t *__restrict__ state,
    const int32_t *__restrict__ end, const int32_t *__restrict__ cnt,
    const int32_t *__restrict__ pv, const int32_t *__restrict__ PL,
    const int32_t *__restrict__ SR, int32_t *__restrict__ posL,
    int32_t *__restrict__ posR, int32_t *__restrict__ sw) {
    const int64_t q = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    int64_t h = -1;
    int32_t cand = 0;
    if (q < n) {
        h = head[q];
        if (state[h] == SEG_ACTIVE) {
            const int64_t e = end[h];
            const int32_t se = suffix_at(SR, e, n);
            const int32_t Lq = PL[q] - PL[h];
            if (q > h) {
                const int32_t c = cnt[q], p = pv[h];
                if (c <= p) posL[h + Lq] = static_cast<int32_t>(q);
                if (c >= p) posR[h + (SR[q] - se)] = static_cast<int32_t>(q);
            }
            const int32_t Rq1 = q + 1 < e ? SR[q + 1] - se : 0;
            cand = Lq < Rq1 ? Lq : Rq1;
        }
    }
    // every lane of the wavefront reaches this point
    if (__all(h == __shfl(h, 0))) {
        int32_t m = cand;
#pragma unroll
        for (int off = kWave / 2; off >= 1; off >>= 1) m = max(m, __shfl_xor(m, off));
        if ((threadIdx.x & (kWave - 1)) == 0 && m > 0) atomicMax(sw + h, m);
    } else if (cand > 0) {
        atomicMax(sw + h, cand);
    }
}

__global__ void k_cut(int64_t nl, const int32_t *__restrict__ list,
                      const int32_t *__restrict__ state,
                      const int32_t *__restrict__ end,
                      const int32_t *__restrict__ PL,
                      const int32_t *__restrict__ posL,
                      const int32_t *__restrict__ posR,
                      const int32_t *__restrict__ sw, int32_t *__restrict__ cut,
                      int64_t *__restrict__ errflag) {
    int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= nl) return;
    const int64_t h = list[t];
    if (state[h] != SEG_ACTIVE) return;
    const int64_t e = end[h];
    const int32_t s = sw[h];
    const int32_t ltot = PL[e - 1] - PL[h];
    const int64_t big = e;  // never returned: one of the two always exists
    const int64_t lnext = s + 1 <= ltot ? posL[h + s + 1] : big;
    const int64_t rs = s >= 1 ? posR[h + s] : big;
    int64_t c = lnext < rs ? lnext : rs;
    if (c <= h || c >= e) {  // cannot happen for a valid partition
        atomicAdd(reinterpret_cast<unsigned long long *>(errflag), 1ull);
        c = h + 1;
    }
    cut[h] = static_cast<int32_t>(c);
}

__global__ void k_move(int64_t n, const int32_t *__restrict__ head,
                       const int32_t *__restrict__ state,
                       const int32_t *__restrict__ end,
                       const int32_t *__restrict__ pv,
                       const int32_t *__restrict__ sw,
                       const int32_t *__restrict__ PL,
                       const int32_t *__restrict__ SR,
                       const int32_t *__restrict__ posL,
                       const int32_t *__restrict__ posR,
                       const int32_t *__restrict__ idx,
                       const int32_t *__restrict__ cnt,
                       int32_t *__restrict__ idx2, int32_t *__restrict__ cnt2) {
    int64_t q = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const int64_t h = head[q];
    int64_t dest = q;
    const int32_t c = cnt[q];
    if (state[h] == SEG_ACTIVE && q > h) {
        const int32_t p = pv[h], s = sw[h];
        if (c <= p) {
            const int32_t Lq = PL[q] - PL[h];
            if (Lq <= s) dest = posR[h + Lq];
        }
        if (dest == q && c >= p) {
            const int32_t Rq = SR[q] - suffix_at(SR, end[h], n);
            if (Rq <= s) dest = posL[h + Rq];
        }
    }
    idx2[dest] = idx[q];
    cnt2[dest] = c;
}

__global__ void k_split(int64_t n, int32_t *__restrict__ head,
                        const int32_t *__restrict__ state,
                        const int32_t *__restrict__ cut) {
    int64_t q = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const int64_t h = head[q];
    if (state[h] == SEG_ACTIVE && q >= cut[h]) head[q] = cut[h];
}

__global__ void k_newsegs(int64_t nl, const int32_t *__restrict__ list,
                          int32_t *__restrict__ state, int32_t *__restrict__ end,
                          int32_t *__restrict__ depth,
                          const int32_t *__restrict__ cut,
                          int32_t *__restrict__ next,
                          int64_t *__restrict__ nnext) {
    int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= nl) return;
    const int32_t h = list[t];
    if (state[h] != SEG_ACTIVE) return;
    const int32_t e = end[h], c = cut[h], d = depth[h] - 1;
    end[c] = e;
    depth[c] = d;
    end[h] = c;
    depth[h] = d;
    const bool right = e - c > kThreshold, left = c - h > kThreshold;
    state[c] = right ? SEG_ACTIVE : SEG_LEAF;
    state[h] = left ? SEG_ACTIVE : SEG_LEAF;
    unsigned long long *cntp = reinterpret_cast<unsigned long long *>(nnext);
    if (right) next[atomicAdd(cntp, 1ull)] = c;
    if (left) next[atomicAdd(cntp, 1ull)] = h;
}

// ------------------------------------------------------------ heap fallback
// make_heap + sort_heap of bits/stl_heap.h on (idx, cnt) pairs,
// comp(a, b) = a.count > b.count.
__device__ void push_heap_d(int32_t *I, int32_t *K, int64_t hole, int64_t top,
                            int32_t vi, int32_t vk) {
    int64_t parent = (hole - 1) / 2;
    while (hole > top && K[parent] > vk) {
        I[hole] = I[parent];
        K[hole] = K[parent];
        hole = parent;
        parent = (hole - 1) / 2;
    }
    I[hole] = vi;
    K[hole] = vk;
}

__device__ void adjust_heap_d(int32_t *I, int32_t *K, int64_t hole,
                              int64_t len, int32_t vi, int32_t vk) {
    const int64_t top = hole;
    int64_t second = hole;
    while (second < (len - 1) / 2) {
        second = 2 * (second + 1);
        if (K[second] > K[second - 1]) second--;
        I[hole] = I[second];
        K[hole] = K[second];
        hole = second;
    }
    if ((len & 1) == 0 && second == (len - 2) / 2) {
        second = 2 * (second + 1);
        I[hole] = I[second - 1];
        K[hole] = K[second - 1];
        hole = second - 1;
    }
    push_heap_d(I, K, hole, top, vi, vk);
}

__global__ void k_heapsort(int64_t n, const int32_t *__restrict__ head,
                           const int32_t *__restrict__ state,
                           const int32_t *__restrict__ end, int32_t *idx,
                           int32_t *cnt) {
    int64_t q = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (q >= n || head[q] != q || state[q] != SEG_HEAP) return;
    int32_t *I = idx + q;
    int32_t *K = cnt + q;
    const int64_t len = end[q] - q;
    if (len >= 2) {  // __make_heap
        int64_t parent = (len - 2) / 2;
        for (;;) {
            adjust_heap_d(I, K, parent, len, I[parent], K[parent]);
            if (parent == 0) break;
            parent--;
        }
    }
    for (int64_t last = len; last > 1;) {  // __sort_heap / __pop_heap
        --last;
        const int32_t vi = I[last], vk = K[last];
        I[last] = I[0];
        K[last] = K[0];
        adjust_heap_d(I, K, 0, last, vi, vk);
    }
}

// Stable insertion sort (count descending) of every leaf segment.
__global__ void k_leaf(int64_t n, const int32_t *__restrict__ head,
                       const int32_t *__restrict__ state,
                       const int32_t *__restrict__ end, int32_t *idx,
                       int32_t *cnt) {
    int64_t q = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (q >= n || head[q] != q || state[q] != SEG_LEAF) return;
    const int len = static_cast<int>(end[q] - q);
    int32_t I[kThreshold], K[kThreshold];
    for (int i = 0; i < len; ++i) {
        I[i] = idx[q + i];
        K[i] = cnt[q + i];
    }
    for (int i = 1; i < len; ++i) {
        const int32_t vi = I[i], vk = K[i];
        int j = i;
        while (j > 0 && vk > K[j - 1]) {
            I[j] = I[j - 1];
            K[j] = K[j - 1];
            --j;
        }
        I[j] = vi;
        K[j] = vk;
    }
    for (int i = 0; i < len; ++i) {
        idx[q + i] = I[i];
        cnt[q + i] = K[i];
    }
}

// ------------------------------------------------------------ chunk layout
__global__ void k_chunk_len(int64_t n_chunks, int64_t C,
                            const int32_t *__restrict__ cnt,
                            int32_t *__restrict__ chunk_len,
                            int64_t *__restrict__ chunk_elems) {
    int64_t c = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (c >= n_chunks) return;
    int32_t m = cnt[c * C];
    for (int64_t i = 1; i < C; ++i) m = max(m, cnt[c * C + i]);
    chunk_len[c] = m;
    chunk_elems[c] = static_cast<int64_t>(m) * C;
}

__global__ void k_chunk_ptr(int64_t n_chunks,
                            const int64_t *__restrict__ incl,
                            int32_t *__restrict__ chunk_ptr) {
    int64_t c = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (c > n_chunks) return;
    chunk_ptr[c] = c == 0 ? 0 : static_cast<int32_t>(incl[c - 1]);
}

__global__ void k_perm(int64_t n, int64_t n_rows,
                       const int32_t *__restrict__ idx,
                       int32_t *__restrict__ perm,
                       int32_t *__restrict__ slot_row) {
    int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int32_t old = idx[i];
    slot_row[i] = old;
    if (old < n_rows) perm[old] = static_cast<int32_t>(i);
}

__global__ void k_fill_i32(int64_t n, int32_t v, int32_t *__restrict__ p) {
    int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// Row of CRS position j: the last r with row_ptr[r] <= j.
__device__ __forceinline__ int64_t row_of(const int32_t *__restrict__ row_ptr,
                                          int64_t n_rows, int64_t j) {
    int64_t lo = 0, hi = n_rows;  // row_ptr[lo] <= j < row_ptr[hi]
    while (hi - lo > 1) {
        int64_t mid = (lo + hi) >> 1;
        if (row_ptr[mid] <= j) lo = mid; else hi = mid;
    }
    return lo;
}

// Scatter in CRS order (convert_formats.hpp:101-128): element k of old row r
// goes to chunk_ptr[p/C] + k*C + p%C with p = perm[r].
template <typename V>
__global__ void k_scatter(int64_t nnz, int64_t n_rows, int64_t C,
                          const int32_t *__restrict__ row_ptr,
                          const int32_t *__restrict__ col,
                          const V *__restrict__ val,
                          const int32_t *__restrict__ src_map,
                          const int32_t *__restrict__ perm,
                          const int32_t *__restrict__ chunk_ptr,
                          int32_t *__restrict__ out_col, V *__restrict__ out_val,
                          int32_t *__restrict__ out_src) {
    int64_t j = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (j >= nnz) return;
    const int64_t r = row_of(row_ptr, n_rows, j);
    const int64_t k = j - row_ptr[r];
    const int64_t p = perm[r];
    const int64_t dst = static_cast<int64_t>(chunk_ptr[p / C]) + k * C + p % C;
    const int32_t s = src_map ? src_map[j] : static_cast<int32_t>(j);
    out_col[dst] = col[j];
    out_val[dst] = val[s];
    out_src[dst] = s;
}

template <typename V>
__global__ void k_refresh(int64_t n_elem, const int32_t *__restrict__ src,
                          const V *__restrict__ crs_val, V *__restrict__ val) {
    int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= n_elem) return;
    const int32_t s = src[e];
    val[e] = s >= 0 ? crs_val[s] : V(0);
}

// ------------------------------------------------------- caller SELL arrays
// Structure checks of SELL arrays handed in by a caller: every chunk's range
// is C * chunk_lengths[c] entries from chunk_ptr[c] (chunk_ptr[0] = 0,
// chunk_ptr[n_chunks] = n_elements) and every column lies in [0, n_cols).
// Any failure sets *bad (every writer stores the same 1).
__global__ void k_check_sell_chunks(int64_t n_chunks, int64_t C, int64_t n_elements,
                                    const int32_t *__restrict__ chunk_ptr,
                                    const int32_t *__restrict__ chunk_len,
                                    int32_t *__restrict__ bad) {
    const int64_t c = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (c > n_chunks) return;
    if (c == 0 && chunk_ptr[0] != 0) *bad = 1;
    if (c == n_chunks) {
        if (chunk_ptr[c] != n_elements) *bad = 1;
        return;
    }
    const int64_t len = chunk_len[c];
    if (len < 0 || static_cast<int64_t>(chunk_ptr[c + 1]) - chunk_ptr[c] != C * len) *bad = 1;
}

__global__ void k_check_sell_cols(int64_t n_elements, int64_t n_cols,
                                  const int32_t *__restrict__ col, int32_t *__restrict__ bad) {
    const int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e < n_elements && (col[e] < 0 || col[e] >= n_cols)) *bad = 1;
}

// slot_row[perm[r]] = r (slot_row prefilled with n_rows, which marks padding);
// perm entries outside [0, n_slots) set *bad and are skipped.
__global__ void k_slot_row_from_perm(int64_t n_rows, int64_t n_slots,
                                     const int32_t *__restrict__ perm,
                                     int32_t *__restrict__ slot_row, int32_t *__restrict__ bad) {
    const int64_t r = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (r >= n_rows) return;
    const int32_t s = perm[r];
    if (s < 0 || s >= n_slots) {
        *bad = 1;
        return;
    }
    slot_row[s] = static_cast<int32_t>(r);
}

// After the scatter: a perm that maps two rows to one slot leaves one of them
// unmatched.
__global__ void k_check_perm(int64_t n_rows, const int32_t *__restrict__ perm,
                             const int32_t *__restrict__ slot_row, int32_t *__restrict__ bad) {
    const int64_t r = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (r < n_rows && slot_row[perm[r]] != r) *bad = 1;
}

// Stored length of each slot of caller SELL arrays: the slot's elements up to
// its last one that is not +0.0 (all value bits zero) at column 0.  Trailing
// +0.0 entries at column 0 are exactly the reference's padding elements: each
// adds fma(+0, x[0], acc), the padding term, which a lane that stops early
// applies once (it is idempotent), so the product is unchanged bit for bit.
// A trailing -0.0 stays: fma(-0, x, -0) = -0 where fma(+0, x, -0) = +0.
template <typename E>
__global__ void k_sell_slot_len(int64_t n_slots, int64_t C,
                                const int32_t *__restrict__ chunk_ptr,
                                const int32_t *__restrict__ chunk_len,
                                const int32_t *__restrict__ col, const E *__restrict__ val,
                                int32_t *__restrict__ slot_len) {
    const int64_t s = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (s >= n_slots) return;
    const int64_t c = s / C, base = static_cast<int64_t>(chunk_ptr[c]) + s % C;
    int j = chunk_len[c];
    while (j > 0) {
        const int64_t e = base + static_cast<int64_t>(j - 1) * C;
        if (val[e] != E(0) || col[e] != 0) break;
        --j;
    }
    slot_len[s] = j;
}

// ------------------------------------------------------------ host side
// Temporary device buffers, released (after the stream drains) on scope exit.
struct Scratch {
    smax_hip_ctx ctx;
    std::vector<void *> ptrs;
    explicit Scratch(smax_hip_ctx c) : ctx(c) {}
    ~Scratch() {
        if (!ptrs.empty()) (void)hipStreamSynchronize(ctx->stream);
        for (void *p : ptrs) dev_free(p);
    }
    template <typename T>
    int get(smax_hip_ctx ctx, T **p, int64_t n, const char *what) {
        void *q = nullptr;
        int rc = dev_alloc(ctx, &q, sizeof(T) * static_cast<size_t>(n > 0 ? n : 1),
                           what);
        if (rc) return rc;
        ptrs.push_back(q);
        *p = static_cast<T *>(q);
        return SMAX_HIP_SUCCESS;
    }
};

void scs_free_arrays(smax_hip_scs_info &in) {
    dev_free(in.chunk_ptr);
    dev_free(in.chunk_lengths);
    dev_free(in.col);
    dev_free(in.val);
    dev_free(in.perm);
    dev_free(in.slot_row);
    dev_free(in.src_index);
    dev_free(in.slot_len);
    in.chunk_ptr = in.chunk_lengths = in.col = in.perm = in.slot_row =
        in.src_index = in.slot_len = nullptr;
    in.val = nullptr;
}

#define LAUNCH(kernel, n, ...)                                                \
    do {                                                                      \
        hipLaunchKernelGGL(kernel, grid_for((n), kThreads), dim3(kThreads), 0, \
                           ctx->stream, __VA_ARGS__);                         \
        SMAX_HIP_LAUNCHED(ctx, #kernel);                                      \
    } while (0)

// Sorted (index, count) pairs of every sigma window, exactly as std::sort.
int sell_sort_windows(smax_hip_ctx ctx, int64_t n, int64_t n_rows,
                      int64_t sigma, const int32_t *row_ptr, int32_t **idx_out,
                      int32_t **cnt_out, Scratch &S) {
    int32_t *idx, *cnt, *idx2, *cnt2, *head, *end, *depth, *state;
    int32_t *pv, *sw, *cut, *PL, *SR, *posL, *posR, *listA, *listB, *tmp;
    int64_t *counters;
    const int64_t nlist_cap = n / (kThreshold + 1) + 2;
    int rc = 0;
    rc |= S.get(ctx, &idx, n, "sell sort");
    rc |= S.get(ctx, &cnt, n, "sell sort");
    rc |= S.get(ctx, &idx2, n, "sell sort");
    rc |= S.get(ctx, &cnt2, n, "sell sort");
    rc |= S.get(ctx, &head, n, "sell sort");
    rc |= S.get(ctx, &end, n, "sell sort");
    rc |= S.get(ctx, &depth, n, "sell sort");
    rc |= S.get(ctx, &state, n, "sell sort");
    rc |= S.get(ctx, &listA, nlist_cap, "sell sort");
    rc |= S.get(ctx, &listB, nlist_cap, "sell sort");
    rc |= S.get(ctx, &counters, 4, "sell sort");
    if (rc) return rc ? SMAX_HIP_ERROR_OUT_OF_MEMORY : 0;
    rc = smax_hip_zero(ctx, counters, 4 * sizeof(int64_t));
    if (rc) return rc;
    const int64_t sig = sigma < n ? sigma : n;
    LAUNCH(k_init, n, n, n_rows, sig, row_ptr, idx, cnt, head, end, depth,
           state, listA, counters + 0);
    int64_t nl = 0;
    rc = read_i64(ctx, counters + 0, &nl);
    if (rc) return rc;
    if (nl > 0) {
        rc |= S.get(ctx, &pv, n, "sell sort");
        rc |= S.get(ctx, &sw, n, "sell sort");
        rc |= S.get(ctx, &cut, n, "sell sort");
        rc |= S.get(ctx, &PL, n, "sell sort");
        rc |= S.get(ctx, &SR, n, "sell sort");
        rc |= S.get(ctx, &posL, n, "sell sort");
        rc |= S.get(ctx, &posR, n, "sell sort");
        rc |= S.get(ctx, &tmp, scan_tmp_elems(n), "sell sort");
        if (rc) return SMAX_HIP_ERROR_OUT_OF_MEMORY;
    }
    int level = 0;
    while (nl > 0) {
        if (++level > 200)
            return fail(ctx, SMAX_HIP_ERROR_RUNTIME,
                        "sell sort: recursion did not terminate");
        LAUNCH(k_pivot, nl, nl, listA, end, depth, state, idx, cnt, pv, sw);
        LAUNCH(k_flags, n, n, head, state, cnt, pv, PL, SR);
        rc = scan_inclusive<int32_t, false>(ctx, PL, PL, n, tmp);
        if (rc) return rc;
        rc = scan_inclusive<int32_t, true>(ctx, SR, SR, n, tmp);
        if (rc) return rc;
        LAUNCH(k_rank, n, n, head, state, end, cnt, pv, PL, SR, posL, posR, sw);
        LAUNCH(k_cut, nl, nl, listA, state, end, PL, posL, posR, sw, cut,
               counters + 2);
        LAUNCH(k_move, n, n, head, state, end, pv, sw, PL, SR, posL, posR, idx,
               cnt, idx2, cnt2);
        std::swap(idx, idx2);
        std::swap(cnt, cnt2);
        LAUNCH(k_split, n, n, head, state, cut);
        rc = smax_hip_zero(ctx, counters + 1, sizeof(int64_t));
        if (rc) return rc;
        LAUNCH(k_newsegs, nl, nl, listA, state, end, depth, cut, listB,
               counters + 1);
        rc = read_i64(ctx, counters + 1, &nl);
        if (rc) return rc;
        std::swap(listA, listB);
    }
    int64_t errs = 0;
    rc = read_i64(ctx, counters + 2, &errs);
    if (rc) return rc;
    if (errs)
        return fail(ctx, SMAX_HIP_ERROR_RUNTIME,
                    "sell sort: %lld partitions produced an invalid cut",
                    static_cast<long long>(errs));
    LAUNCH(k_heapsort, n, n, head, state, end, idx, cnt);
    LAUNCH(k_leaf, n, n, head, state, end, idx, cnt);
    *idx_out = idx;
    *cnt_out = cnt;
    return SMAX_HIP_SUCCESS;
}

int build_scs(smax_hip_ctx ctx, int64_t n_rows, int64_t n_cols, int64_t nnz,
              const int32_t *row_ptr, const int32_t *col, const void *val,
              int64_t val_bytes, const int32_t *src_map, int64_t C,
              int64_t sigma, smax_hip_scs *out) {
    SMAX_HIP_REQUIRE(ctx, out, "scs conversion: null output");
    *out = nullptr;
    SMAX_HIP_REQUIRE(ctx, n_rows > 0 && n_cols > 0 && nnz >= 0,
                     "scs conversion: bad shape (%lld x %lld, nnz %lld)",
                     (long long)n_rows, (long long)n_cols, (long long)nnz);
    SMAX_HIP_REQUIRE(ctx, C > 0 && sigma > 0,
                     "scs conversion: C and sigma must be positive");
    SMAX_HIP_REQUIRE(ctx, val_bytes == 2 || val_bytes == 4 || val_bytes == 8,
                     "scs conversion: val_bytes must be 2, 4 or 8");
    SMAX_HIP_REQUIRE(ctx, row_ptr && (nnz == 0 || (col && val)),
                     "scs conversion: null input array");
    if (n_rows > INT32_MAX || n_cols > INT32_MAX || C > INT32_MAX || nnz > INT32_MAX)
        return fail(ctx, SMAX_HIP_ERROR_OVERFLOW,
                    "scs conversion: %lld x %lld, C = %lld or %lld entries exceed int32",
                    (long long)n_rows, (long long)n_cols, (long long)C, (long long)nnz);
    const int64_t n_chunks = ceil_div(n_rows, C);
    const int64_t n = n_chunks * C;
    if (n > INT32_MAX)
        return fail(ctx, SMAX_HIP_ERROR_OVERFLOW,
                    "scs conversion: %lld padded rows exceed int32",
                    (long long)n);
    Scratch S(ctx);
    int32_t *idx = nullptr, *cnt = nullptr;
    int rc = sell_sort_windows(ctx, n, n_rows, sigma, row_ptr, &idx, &cnt, S);
    if (rc) return rc;

    smax_hip_scs_info info{};
    info.C = C;
    info.sigma = sigma;
    info.n_rows = n_rows;
    info.n_rows_padded = n;
    info.n_cols = n_cols;
    info.n_chunks = n_chunks;
    info.nnz = nnz;
    info.val_bytes = val_bytes;
    struct Guard {
        smax_hip_ctx ctx;
        smax_hip_scs_info *p;
        bool keep = false;
        ~Guard() {
            if (keep) return;
            (void)hipStreamSynchronize(ctx->stream);
            scs_free_arrays(*p);
        }
    } guard{ctx, &info};

    int64_t *elems, *incl, *tmp64;
    rc |= S.get(ctx, &elems, n_chunks, "scs conversion");
    rc |= S.get(ctx, &incl, n_chunks, "scs conversion");
    rc |= S.get(ctx, &tmp64, scan_tmp_elems(n_chunks), "scs conversion");
    rc |= dev_alloc(ctx, reinterpret_cast<void **>(&info.chunk_lengths),
                    sizeof(int32_t) * n_chunks, "scs chunk_lengths");
    rc |= dev_alloc(ctx, reinterpret_cast<void **>(&info.chunk_ptr),
                    sizeof(int32_t) * (n_chunks + 1), "scs chunk_ptr");
    rc |= dev_alloc(ctx, reinterpret_cast<void **>(&info.perm),
                    sizeof(int32_t) * n_rows, "scs perm");
    rc |= dev_alloc(ctx, reinterpret_cast<void **>(&info.slot_row),
                    sizeof(int32_t) * n, "scs slot_row");
    rc |= dev_alloc(ctx, reinterpret_cast<void **>(&info.slot_len),
                    sizeof(int32_t) * n, "scs slot_len");
    if (rc) return SMAX_HIP_ERROR_OUT_OF_MEMORY;

    LAUNCH(k_chunk_len, n_chunks, n_chunks, C, cnt, info.chunk_lengths, elems);
    rc = scan_inclusive<int64_t, false>(ctx, elems, incl, n_chunks, tmp64);
    if (rc) return rc;
    int64_t n_elements = 0;
    rc = read_i64(ctx, incl + (n_chunks - 1), &n_elements);
    if (rc) return rc;
    if (n_elements > INT32_MAX)
        return fail(ctx, SMAX_HIP_ERROR_OVERFLOW,
                    "scs conversion: %lld SELL elements exceed int32 chunk_ptr",
                    (long long)n_elements);
    info.n_elements = n_elements;
    LAUNCH(k_chunk_ptr, n_chunks + 1, n_chunks, incl, info.chunk_ptr);
    LAUNCH(k_perm, n, n, n_rows, idx, info.perm, info.slot_row);
    SMAX_HIP_CALL(ctx, hipMemcpyAsync(info.slot_len, cnt, sizeof(int32_t) * n,
                                      hipMemcpyDeviceToDevice, ctx->stream));

    rc |= dev_alloc(ctx, reinterpret_cast<void **>(&info.col),
                    sizeof(int32_t) * n_elements, "scs col");
    rc |= dev_alloc(ctx, &info.val, val_bytes * n_elements, "scs val");
    rc |= dev_alloc(ctx, reinterpret_cast<void **>(&info.src_index),
                    sizeof(int32_t) * n_elements, "scs src_index");
    if (rc) return SMAX_HIP_ERROR_OUT_OF_MEMORY;
    rc = smax_hip_zero(ctx, info.col, sizeof(int32_t) * n_elements);
    if (!rc) rc = smax_hip_zero(ctx, info.val, val_bytes * n_elements);
    if (rc) return rc;
    LAUNCH(k_fill_i32, n_elements, n_elements, -1, info.src_index);
    if (nnz > 0) {
        switch (val_bytes) {
        case 8:
            LAUNCH(k_scatter<uint64_t>, nnz, nnz, n_rows, C, row_ptr, col,
                   static_cast<const uint64_t *>(val), src_map, info.perm,
                   info.chunk_ptr, info.col, static_cast<uint64_t *>(info.val),
                   info.src_index);
            break;
        case 4:
            LAUNCH(k_scatter<uint32_t>, nnz, nnz, n_rows, C, row_ptr, col,
                   static_cast<const uint32_t *>(val), src_map, info.perm,
                   info.chunk_ptr, info.col, static_cast<uint32_t *>(info.val),
                   info.src_index);
            break;
        default:
            LAUNCH(k_scatter<uint16_t>, nnz, nnz, n_rows, C, row_ptr, col,
                   static_cast<const uint16_t *>(val), src_map, info.perm,
                   info.chunk_ptr, info.col, static_cast<uint16_t *>(info.val),
                   info.src_index);
            break;
        }
    }
    // Scratch is released on return; wait so it is not freed under the GPU.
    SMAX_HIP_CALL(ctx, hipStreamSynchronize(ctx->stream));
    smax_hip_scs scs = new smax_hip_scs_s();
    scs->info = info;
    scs->src_consecutive = src_map == nullptr;
    guard.keep = true;
    rc = build_schedule(ctx, C, sigma, n, n_chunks, n_elements, nnz, info.chunk_lengths,
                        &scs->sched, info.slot_row);
    if (!rc) rc = build_tails(ctx, scs);
    if (rc) {
        smax_hip_scs_destroy(ctx, scs);
        return rc;
    }
    *out = scs;
    return SMAX_HIP_SUCCESS;
}

int build_scs_from_sell(smax_hip_ctx ctx, int64_t C, int64_t sigma, int64_t n_rows,
                        int64_t n_rows_padded, int64_t n_cols, int64_t n_chunks,
                        int64_t n_elements, int64_t nnz, const int32_t *chunk_ptr,
                        const int32_t *chunk_lengths, const int32_t *col, const void *val,
                        int64_t val_bytes, const int32_t *perm, smax_hip_scs *out) {
    SMAX_HIP_REQUIRE(ctx, out, "smax_hip_scs_from_sell: null output");
    *out = nullptr;
    // n_rows_padded == n_chunks * C, tested without forming the product
    SMAX_HIP_REQUIRE(ctx, C > 0 && sigma > 0 && n_chunks > 0 && n_rows_padded % C == 0 &&
                              n_rows_padded / C == n_chunks && n_rows > 0 &&
                              n_rows <= n_rows_padded && n_cols > 0 && n_elements >= 0 &&
                              nnz >= 0 && nnz <= n_elements,
                     "smax_hip_scs_from_sell: inconsistent sizes");
    SMAX_HIP_REQUIRE(ctx, val_bytes == 2 || val_bytes == 4 || val_bytes == 8,
                     "smax_hip_scs_from_sell: val_bytes must be 2, 4 or 8");
    SMAX_HIP_REQUIRE(ctx, chunk_ptr && chunk_lengths && perm && (n_elements == 0 || (col && val)),
                     "smax_hip_scs_from_sell: null array");
    if (n_rows_padded > INT32_MAX || n_elements > INT32_MAX)
        return fail(ctx, SMAX_HIP_ERROR_OVERFLOW, "smax_hip_scs_from_sell: sizes exceed int32");
    smax_hip_scs_info info{};
    info.C = C;
    info.sigma = sigma;
    info.n_rows = n_rows;
    info.n_rows_padded = n_rows_padded;
    info.n_cols = n_cols;
    info.n_chunks = n_chunks;
    info.n_elements = n_elements;
    info.nnz = nnz;
    info.val_bytes = val_bytes;
    struct Guard {
        smax_hip_ctx ctx;
        smax_hip_scs_info *p;
        bool keep = false;
        ~Guard() {
            if (keep) return;
            (void)hipStreamSynchronize(ctx->stream);
            scs_free_arrays(*p);
        }
    } guard{ctx, &info};
    Scratch S(ctx);
    int32_t *bad = nullptr;
    int rc = S.get(ctx, &bad, 1, "scs adoption");
    rc |= dev_alloc(ctx, reinterpret_cast<void **>(&info.chunk_ptr),
                    sizeof(int32_t) * (n_chunks + 1), "scs chunk_ptr");
    rc |= dev_alloc(ctx, reinterpret_cast<void **>(&info.chunk_lengths),
                    sizeof(int32_t) * n_chunks, "scs chunk_lengths");
    rc |= dev_alloc(ctx, reinterpret_cast<void **>(&info.col),
                    sizeof(int32_t) * (n_elements > 0 ? n_elements : 1), "scs col");
    rc |= dev_alloc(ctx, &info.val, val_bytes * (n_elements > 0 ? n_elements : 1), "scs val");
    rc |= dev_alloc(ctx, reinterpret_cast<void **>(&info.perm), sizeof(int32_t) * n_rows,
                    "scs perm");
    rc |= dev_alloc(ctx, reinterpret_cast<void **>(&info.slot_row),
                    sizeof(int32_t) * n_rows_padded, "scs slot_row");
    rc |= dev_alloc(ctx, reinterpret_cast<void **>(&info.slot_len),
                    sizeof(int32_t) * n_rows_padded, "scs slot_len");
    if (rc) return SMAX_HIP_ERROR_OUT_OF_MEMORY;
    const auto d2d = [&](void *dst, const void *src, size_t bytes) {
        return hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, ctx->stream);
    };
    SMAX_HIP_CALL(ctx, d2d(info.chunk_ptr, chunk_ptr, sizeof(int32_t) * (n_chunks + 1)));
    SMAX_HIP_CALL(ctx, d2d(info.chunk_lengths, chunk_lengths, sizeof(int32_t) * n_chunks));
    if (n_elements > 0) {
        SMAX_HIP_CALL(ctx, d2d(info.col, col, sizeof(int32_t) * n_elements));
        SMAX_HIP_CALL(ctx, d2d(info.val, val, val_bytes * n_elements));
    }
    SMAX_HIP_CALL(ctx, d2d(info.perm, perm, sizeof(int32_t) * n_rows));
    LAUNCH(k_fill_i32, 1, 1, 0, bad);
    LAUNCH(k_check_sell_chunks, n_chunks + 1, n_chunks, C, n_elements, info.chunk_ptr,
           info.chunk_lengths, bad);
    int32_t h_bad = 0;
    SMAX_HIP_CALL(ctx, hipMemcpyAsync(&h_bad, bad, sizeof(int32_t), hipMemcpyDeviceToHost,
                                      ctx->stream));
    SMAX_HIP_CALL(ctx, hipStreamSynchronize(ctx->stream));
    if (h_bad)
        return fail(ctx, SMAX_HIP_ERROR_INVALID_VALUE,
                    "smax_hip_scs_from_sell: chunk_ptr does not match C * chunk_lengths "
                    "(or chunk_ptr[n_chunks] != n_elements)");
    if (n_elements > 0)
        LAUNCH(k_check_sell_cols, n_elements, n_elements, n_cols, info.col, bad);
    LAUNCH(k_fill_i32, n_rows_padded, n_rows_padded, static_cast<int32_t>(n_rows),
           info.slot_row);
    LAUNCH(k_slot_row_from_perm, n_rows, n_rows, n_rows_padded, info.perm, info.slot_row, bad);
    SMAX_HIP_CALL(ctx, hipMemcpyAsync(&h_bad, bad, sizeof(int32_t), hipMemcpyDeviceToHost,
                                      ctx->stream));
    SMAX_HIP_CALL(ctx, hipStreamSynchronize(ctx->stream));
    if (h_bad)
        return fail(ctx, SMAX_HIP_ERROR_INVALID_VALUE,
                    "smax_hip_scs_from_sell: a column index lies outside [0, n_cols) or a perm "
                    "entry outside [0, n_rows_padded)");
    LAUNCH(k_check_perm, n_rows, n_rows, info.perm, info.slot_row, bad);
    switch (val_bytes) {
    case 8:
        LAUNCH(k_sell_slot_len<uint64_t>, n_rows_padded, n_rows_padded, C, info.chunk_ptr,
               info.chunk_lengths, info.col, static_cast<const uint64_t *>(info.val),
               info.slot_len);
        break;
    case 4:
        LAUNCH(k_sell_slot_len<uint32_t>, n_rows_padded, n_rows_padded, C, info.chunk_ptr,
               info.chunk_lengths, info.col, static_cast<const uint32_t *>(info.val),
               info.slot_len);
        break;
    default:
        LAUNCH(k_sell_slot_len<uint16_t>, n_rows_padded, n_rows_padded, C, info.chunk_ptr,
               info.chunk_lengths, info.col, static_cast<const uint16_t *>(info.val),
               info.slot_len);
        break;
    }
    SMAX_HIP_CALL(ctx, hipMemcpyAsync(&h_bad, bad, sizeof(int32_t), hipMemcpyDeviceToHost,
                                      ctx->stream));
    SMAX_HIP_CALL(ctx, hipStreamSynchronize(ctx->stream));
    if (h_bad)
        return fail(ctx, SMAX_HIP_ERROR_INVALID_VALUE,
                    "smax_hip_scs_from_sell: perm maps two rows to one slot");
    smax_hip_scs scs = new smax_hip_scs_s();
    scs->info = info;
    scs->from_sell = true;
    guard.keep = true;
    rc = build_schedule(ctx, C, sigma, n_rows_padded, n_chunks, n_elements, nnz,
                        info.chunk_lengths, &scs->sched, info.slot_row);
    if (!rc) rc = build_tails(ctx, scs);
    if (rc) {
        smax_hip_scs_destroy(ctx, scs);
        return rc;
    }
    *out = scs;
    return SMAX_HIP_SUCCESS;
}

}  // namespace smaxhip

using namespace smaxhip;

extern "C" {

int smax_hip_scs_from_sell(smax_hip_ctx ctx, int64_t C, int64_t sigma, int64_t n_rows,
                           int64_t n_rows_padded, int64_t n_cols, int64_t n_chunks,
                           int64_t n_elements, int64_t nnz, const int32_t *chunk_ptr,
                           const int32_t *chunk_lengths, const int32_t *col, const void *val,
                           int64_t val_bytes, const int32_t *perm, smax_hip_scs *out) {
    SMAX_HIP_CHECK_CTX(ctx);
    return build_scs_from_sell(ctx, C, sigma, n_rows, n_rows_padded, n_cols, n_chunks,
                               n_elements, nnz, chunk_ptr, chunk_lengths, col, val, val_bytes,
                               perm, out);
}

int smax_hip_scs_from_crs(smax_hip_ctx ctx, int64_t n_rows, int64_t n_cols,
                          int64_t nnz, const int32_t *row_ptr,
                          const int32_t *col, const void *val,
                          int64_t val_bytes, int64_t C, int64_t sigma,
                          smax_hip_scs *out) {
    SMAX_HIP_CHECK_CTX(ctx);
    return build_scs(ctx, n_rows, n_cols, nnz, row_ptr, col, val, val_bytes,
                     nullptr, C, sigma, out);
}

int smax_hip_scs_get_info(smax_hip_scs scs, smax_hip_scs_info *info) {
    if (!scs || !info)
        return fail(nullptr, SMAX_HIP_ERROR_INVALID_VALUE,
                    "smax_hip_scs_get_info: null argument");
    *info = scs->info;
    return SMAX_HIP_SUCCESS;
}

int smax_hip_scs_destroy(smax_hip_ctx ctx, smax_hip_scs scs) {
    if (!scs) return SMAX_HIP_SUCCESS;
    if (ctx) {
        (void)bind(ctx);
        (void)hipStreamSynchronize(ctx->stream);
    }
    scs_free_arrays(scs->info);
    dev_free(scs->col_slot);
    dev_free(scs->col_rel);
    dev_free(scs->col_idx);
    dev_free(scs->col_idx_row);
    dev_free(scs->col_tab);
    dev_free(scs->val_f32);
    dev_free(scs->val_bf16);
    dev_free(scs->diag_f32);
    free_schedule(scs->sched);
    free_tails(scs->tails);
    dev_free(scs->elem_row);
    delete scs;
    return SMAX_HIP_SUCCESS;
}

int smax_hip_scs_refresh_values(smax_hip_ctx ctx, smax_hip_scs scs,
                                const void *crs_val) {
    SMAX_HIP_CHECK_CTX(ctx);
    SMAX_HIP_REQUIRE(ctx, scs && crs_val, "smax_hip_scs_refresh_values: null");
    SMAX_HIP_REQUIRE(ctx, !scs->from_sell,
                     "smax_hip_scs_refresh_values: the handle was built from SELL arrays and "
                     "has no CRS positions; build it with smax_hip_scs_from_crs");
    const smax_hip_scs_info &in = scs->info;
    const int64_t m = in.n_elements;
    // the solvers' rounded copies follow the old values: drop them (rebuilt
    // on the next solve); without copies the refresh stays asynchronous
    if (scs->val_f32 || scs->val_bf16 || scs->diag_f32) {
        (void)hipStreamSynchronize(ctx->stream);
        for (void **p : {reinterpret_cast<void **>(&scs->val_f32),
                         reinterpret_cast<void **>(&scs->val_bf16),
                         reinterpret_cast<void **>(&scs->diag_f32)}) {
            dev_free(*p);
            *p = nullptr;
        }
        scs->bf16_split = -1;
    }
    if (m == 0) return SMAX_HIP_SUCCESS;
    switch (in.val_bytes) {
    case 8:
        LAUNCH(k_refresh<uint64_t>, m, m, in.src_index,
               static_cast<const uint64_t *>(crs_val),
               static_cast<uint64_t *>(in.val));
        break;
    case 4:
        LAUNCH(k_refresh<uint32_t>, m, m, in.src_index,
               static_cast<const uint32_t *>(crs_val),
               static_cast<uint32_t *>(in.val));
        break;
    default:
        LAUNCH(k_refresh<uint16_t>, m, m, in.src_index,
               static_cast<const uint16_t *>(crs_val),
               static_cast<uint16_t *>(in.val));
        break;
    }
    // the tail waves hold copies of their slots' values
    return refresh_tail_values(ctx, scs);
}

}  // extern "C"
