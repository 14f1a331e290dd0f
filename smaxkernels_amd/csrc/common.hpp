// Internal helpers shared by the translation units of libsmax_hip.so.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <string>
#include <utility>
#include <vector>

#include "smax_hip.h"

struct smax_hip_ctx_s {
    int device = 0;
    hipStream_t own_stream = nullptr;
    hipStream_t stream = nullptr;
    hipEvent_t ev_start = nullptr;
    hipEvent_t ev_stop = nullptr;
    std::string last_error;
    // Small pinned host word used to read back device counters.
    int64_t *host_word = nullptr;
    // Device workspace of entry points that stage operands (column-major
    // CRS SpMM), grown on demand.  A buffer that is outgrown is retired, not
    // freed: work already queued, and HIP graphs captured with it, keep a
    // valid address until the context is destroyed.
    void *work = nullptr;
    size_t work_bytes = 0;
    std::vector<void *> retired_work;
};

namespace smaxhip {

// Records a failure on ctx (if any) and in the thread-local slot.
int fail(smax_hip_ctx ctx, int status, const char *fmt, ...);

// Maps a HIP error to a status and records it.
int hip_fail(smax_hip_ctx ctx, hipError_t err, const char *what);

// Makes ctx's device current for the calling thread.
int bind(smax_hip_ctx ctx);

// Checks the launch that was just enqueued.
int check_launch(smax_hip_ctx ctx, const char *kernel);

// Allocation helpers that record failures.
int dev_alloc(smax_hip_ctx ctx, void **p, size_t bytes, const char *what);
void dev_free(void *p);

// ctx's workspace of at least `bytes` (256-byte aligned).  Growing it
// allocates a new buffer and retires the old one (freed with the context).
int ctx_workspace(smax_hip_ctx ctx, size_t bytes, void **out);

// Device counter read-back (int64 at a device address), synchronous.
int read_i64(smax_hip_ctx ctx, const int64_t *dptr, int64_t *out);

constexpr int kWave = 64;

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

// Grid for a 1-D launch with `threads` per block over n items.
inline dim3 grid_for(int64_t n, int threads) {
    int64_t g = ceil_div(n, threads);
    return dim3(static_cast<unsigned>(g > 0 ? g : 1));
}

// ---------------------------------------------------------------- types
// BF16 storage is the upper half of an IEEE binary32; widening is exact.
__device__ __forceinline__ float bf16_to_f32(uint16_t h) {
    return __uint_as_float(static_cast<uint32_t>(h) << 16);
}

template <int VT> struct vt;
template <> struct vt<SMAX_HIP_F64> {
    using val_t = double;
    using vec_t = double;
    __device__ __forceinline__ static double widen(double v) { return v; }
    __device__ __forceinline__ static double fma_(double a, double b,
                                                  double c) {
        return __builtin_fma(a, b, c);
    }
};
template <> struct vt<SMAX_HIP_F32> {
    using val_t = float;
    using vec_t = float;
    __device__ __forceinline__ static float widen(float v) { return v; }
    __device__ __forceinline__ static float fma_(float a, float b, float c) {
        return __builtin_fmaf(a, b, c);
    }
};
template <> struct vt<SMAX_HIP_BF16> {
    using val_t = uint16_t;
    using vec_t = float;
    __device__ __forceinline__ static float widen(uint16_t v) {
        return bf16_to_f32(v);
    }
    __device__ __forceinline__ static float fma_(float a, float b, float c) {
        return __builtin_fmaf(a, b, c);
    }
};

inline int64_t vt_val_bytes(int v) {
    return v == SMAX_HIP_F64 ? 8 : v == SMAX_HIP_F32 ? 4 : 2;
}

inline bool valid_value_type(int v) {
    return v == SMAX_HIP_F64 || v == SMAX_HIP_F32 || v == SMAX_HIP_BF16;
}

// Workgroup order remapped so each XCD (workgroups are dealt round-robin to
// the 8 XCDs) walks one contiguous range of work items: neighbouring items,
// which gather overlapping parts of x, then share that XCD's L2.
__device__ __forceinline__ int64_t xcd_block(int64_t bid, int64_t nblocks) {
    constexpr int64_t kXcd = 8;
    const int64_t xcd = bid % kXcd, local = bid / kXcd;
    const int64_t q = nblocks / kXcd, r = nblocks % kXcd;
    return xcd < r ? xcd * (q + 1) + local : r * (q + 1) + (xcd - r) * q + local;
}

// Lane U of each aligned group of 8 lanes, handed to the group (ds_swizzle
// in bit mode: and-mask 0x18 keeps the group, or-mask U picks the lane).
// Used where 8 lanes share one row or slot and would otherwise each load
// the same column index and value.
template <int U>
__device__ __forceinline__ int grp8_bcast(int v) {
    return __builtin_amdgcn_ds_swizzle(v, 0x18 | (U << 5));
}
template <int... U>
__device__ __forceinline__ void grp8_bcast8(int c, int w, int (&cs)[8], int (&ws)[8],
                                            std::integer_sequence<int, U...>) {
    ((cs[U] = grp8_bcast<U>(c)), ...);
    ((ws[U] = grp8_bcast<U>(w)), ...);
}
// The same for 16-lane groups (a DPP row): lane U of the row, U < 8, handed
// to all 16 lanes (row_newbcast).
template <int U>
__device__ __forceinline__ int row16_bcast(int v) {
    return __builtin_amdgcn_mov_dpp(v, 0x150 + U, 0xf, 0xf, false);
}
template <int... U>
__device__ __forceinline__ void row16_bcast8(int c, int w, int (&cs)[8], int (&ws)[8],
                                             std::integer_sequence<int, U...>) {
    ((cs[U] = row16_bcast<U>(c)), ...);
    ((ws[U] = row16_bcast<U>(w)), ...);
}
template <typename V>
__device__ __forceinline__ int val_bits(V v) {
    if constexpr (sizeof(V) == 4) return __float_as_int(v);
    else return static_cast<int>(v);
}
template <typename V>
__device__ __forceinline__ V bits_val(int b) {
    if constexpr (sizeof(V) == 4) return __int_as_float(b);
    else return static_cast<V>(b);
}

// Q consecutive entries at p.  When Q * sizeof(A) is a multiple of 16 the
// caller guarantees p is 16-byte aligned, and they move as 16-byte accesses.
template <typename A, int Q>
__device__ __forceinline__ void load_run(const A *__restrict__ p, A (&out)[Q]) {
    if constexpr ((Q * sizeof(A)) % 16 == 0) {
        constexpr int W = 16 / sizeof(A);
        typedef A V __attribute__((ext_vector_type(W)));
        const V *q = reinterpret_cast<const V *>(p);
#pragma unroll
        for (int h = 0; h < Q / W; ++h) {
            const V w = q[h];
#pragma unroll
            for (int i = 0; i < W; ++i) out[h * W + i] = w[i];
        }
    } else {
#pragma unroll
        for (int i = 0; i < Q; ++i) out[i] = p[i];
    }
}

template <typename A, int Q>
__device__ __forceinline__ void store_run(A *__restrict__ p, const A (&in)[Q]) {
    if constexpr ((Q * sizeof(A)) % 16 == 0) {
        constexpr int W = 16 / sizeof(A);
        typedef A V __attribute__((ext_vector_type(W)));
        V *q = reinterpret_cast<V *>(p);
#pragma unroll
        for (int h = 0; h < Q / W; ++h) {
            V w;
#pragma unroll
            for (int i = 0; i < W; ++i) w[i] = in[h * W + i];
            q[h] = w;
        }
    } else {
#pragma unroll
        for (int i = 0; i < Q; ++i) p[i] = in[i];
    }
}

// Whether `rows` output rows of `k` columns fit the launch arithmetic: both
// counts within int32, and one thread per (row, column) within a grid of
// 2^32 workgroups of 256 (the products are formed only after the bounds).
inline bool extent_ok(int64_t rows, int64_t k) {
    return rows > 0 && k > 0 && rows <= INT32_MAX && k <= INT32_MAX &&
           rows * k <= (static_cast<int64_t>(1) << 40);
}

// Whether every pointer is 16-byte aligned (the run helpers' precondition).
inline bool aligned16(const void *a, const void *b) {
    return ((reinterpret_cast<uintptr_t>(a) | reinterpret_cast<uintptr_t>(b)) & 15) == 0;
}

// Streamed-once loads bypass cache allocation.
template <typename T> __device__ __forceinline__ T ld_stream(const T *p) {
    return __builtin_nontemporal_load(p);
}

}  // namespace smaxhip

// Dispatch a templated launcher on the runtime value type.
#define SMAX_HIP_DISPATCH_VT(vtype, ...)                                      \
    switch (vtype) {                                                          \
    case SMAX_HIP_F64: { constexpr int VT = SMAX_HIP_F64; __VA_ARGS__; break; } \
    case SMAX_HIP_F32: { constexpr int VT = SMAX_HIP_F32; __VA_ARGS__; break; } \
    case SMAX_HIP_BF16: { constexpr int VT = SMAX_HIP_BF16; __VA_ARGS__; break; } \
    default: break;                                                           \
    }

#define SMAX_HIP_CHECK_CTX(ctx)                                               \
    do {                                                                      \
        if (!(ctx))                                                           \
            return smaxhip::fail(nullptr, SMAX_HIP_ERROR_INVALID_VALUE,       \
                                 "%s: null context", __func__);               \
        int bind_rc_ = smaxhip::bind(ctx);                                    \
        if (bind_rc_) return bind_rc_;                                        \
    } while (0)

#define SMAX_HIP_REQUIRE(ctx, cond, ...)                                      \
    do {                                                                      \
        if (!(cond))                                                          \
            return smaxhip::fail((ctx), SMAX_HIP_ERROR_INVALID_VALUE,         \
                                 __VA_ARGS__);                                \
    } while (0)

#define SMAX_HIP_CALL(ctx, expr)                                              \
    do {                                                                      \
        hipError_t e_ = (expr);                                               \
        if (e_ != hipSuccess) return smaxhip::hip_fail((ctx), e_, #expr);     \
    } while (0)

#define SMAX_HIP_LAUNCHED(ctx, name)                                          \
    do {                                                                      \
        int rc_ = smaxhip::check_launch((ctx), (name));                       \
        if (rc_) return rc_;                                                  \
    } while (0)
