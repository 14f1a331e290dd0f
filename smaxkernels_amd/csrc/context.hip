// Context, memory, timing and stream-bandwidth entry points.
#include "common.hpp"

#include <cstring>

namespace smaxhip {

static thread_local std::string g_thread_error;

int fail(smax_hip_ctx ctx, int status, const char *fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_thread_error = buf;
    if (ctx) ctx->last_error = buf;
    return status;
}

int hip_fail(smax_hip_ctx ctx, hipError_t err, const char *what) {
    int status = SMAX_HIP_ERROR_RUNTIME;
    if (err == hipErrorOutOfMemory || err == hipErrorMemoryAllocation)
        status = SMAX_HIP_ERROR_OUT_OF_MEMORY;
    else if (err == hipErrorNoDevice || err == hipErrorInvalidDevice)
        status = SMAX_HIP_ERROR_NO_DEVICE;
    else if (err == hipErrorLaunchFailure || err == hipErrorInvalidConfiguration ||
             err == hipErrorLaunchOutOfResources)
        status = SMAX_HIP_ERROR_LAUNCH;
    return fail(ctx, status, "%s: %s", what, hipGetErrorString(err));
}

int bind(smax_hip_ctx ctx) {
    int cur = -1;
    hipError_t e = hipGetDevice(&cur);
    if (e != hipSuccess) return hip_fail(ctx, e, "hipGetDevice");
    if (cur != ctx->device) {
        e = hipSetDevice(ctx->device);
        if (e != hipSuccess) return hip_fail(ctx, e, "hipSetDevice");
    }
    return SMAX_HIP_SUCCESS;
}

int check_launch(smax_hip_ctx ctx, const char *kernel) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        int st = hip_fail(ctx, e, kernel);
        return st == SMAX_HIP_ERROR_RUNTIME ? SMAX_HIP_ERROR_LAUNCH : st;
    }
    return SMAX_HIP_SUCCESS;
}

int dev_alloc(smax_hip_ctx ctx, void **p, size_t bytes, const char *what) {
    *p = nullptr;
    if (bytes == 0) bytes = 1;
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
        *p = nullptr;
        return fail(ctx, SMAX_HIP_ERROR_OUT_OF_MEMORY,
                    "%s: hipMalloc(%zu bytes) failed: %s", what, bytes,
                    hipGetErrorString(e));
    }
    return SMAX_HIP_SUCCESS;
}

void dev_free(void *p) {
    if (p) (void)hipFree(p);
}

int ctx_workspace(smax_hip_ctx ctx, size_t bytes, void **out) {
    if (ctx->work_bytes < bytes) {
        void *grown = nullptr;
        const int rc = dev_alloc(ctx, &grown, bytes, "workspace");
        if (rc) return rc;
        if (ctx->work) ctx->retired_work.push_back(ctx->work);
        ctx->work = grown;
        ctx->work_bytes = bytes;
    }
    *out = ctx->work;
    return SMAX_HIP_SUCCESS;
}

int read_i64(smax_hip_ctx ctx, const int64_t *dptr, int64_t *out) {
    SMAX_HIP_CALL(ctx, hipMemcpyAsync(ctx->host_word, dptr, sizeof(int64_t),
                                      hipMemcpyDeviceToHost, ctx->stream));
    SMAX_HIP_CALL(ctx, hipStreamSynchronize(ctx->stream));
    *out = *ctx->host_word;
    return SMAX_HIP_SUCCESS;
}

// Byte-wise zero fill, 16 bytes per lane where aligned.
__global__ void zero_kernel(uint8_t *p, size_t bytes) {
    size_t tid = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
    size_t head = (16 - (reinterpret_cast<uintptr_t>(p) & 15)) & 15;
    if (head > bytes) head = bytes;
    for (size_t i = tid; i < head; i += stride) p[i] = 0;
    uint4 *q = reinterpret_cast<uint4 *>(p + head);
    size_t n16 = (bytes - head) / 16;
    for (size_t i = tid; i < n16; i += stride) q[i] = make_uint4(0, 0, 0, 0);
    for (size_t i = head + n16 * 16 + tid; i < bytes; i += stride) p[i] = 0;
}

typedef double f64x2 __attribute__((ext_vector_type(2)));

__global__ void copy_kernel(const double *__restrict__ src,
                            double *__restrict__ dst, int64_t n) {
    int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
    const int64_t n2 = n / 2;
    const f64x2 *s2 = reinterpret_cast<const f64x2 *>(src);
    f64x2 *d2 = reinterpret_cast<f64x2 *>(dst);
    for (int64_t k = i; k < n2; k += stride)
        __builtin_nontemporal_store(__builtin_nontemporal_load(s2 + k), d2 + k);
    if (i == 0 && (n & 1)) dst[n - 1] = src[n - 1];
}

// The copy with L independent 16-byte loads in flight per lane: a workgroup
// iteration covers 256*L consecutive pairs, lane t takes pairs t + 256*l, so
// every load instruction of a wavefront is one contiguous 1 KiB run.
template <int L>
__global__ __launch_bounds__(256) void copy_wide_kernel(const double *__restrict__ src,
                                                        double *__restrict__ dst, int64_t n) {
    const int64_t n2 = n / 2;
    const f64x2 *s2 = reinterpret_cast<const f64x2 *>(src);
    f64x2 *d2 = reinterpret_cast<f64x2 *>(dst);
    const int64_t span = static_cast<int64_t>(256) * L;
    for (int64_t base = static_cast<int64_t>(blockIdx.x) * span; base < n2;
         base += static_cast<int64_t>(gridDim.x) * span) {
        f64x2 v[L];
#pragma unroll
        for (int l = 0; l < L; ++l) {
            const int64_t k = base + threadIdx.x + 256 * l;
            if (k < n2) v[l] = __builtin_nontemporal_load(s2 + k);
        }
#pragma unroll
        for (int l = 0; l < L; ++l) {
            const int64_t k = base + threadIdx.x + 256 * l;
            if (k < n2) __builtin_nontemporal_store(v[l], d2 + k);
        }
    }
    if (blockIdx.x == 0 && threadIdx.x == 0 && (n & 1)) dst[n - 1] = src[n - 1];
}

__global__ void triad_kernel(double *__restrict__ a,
                             const double *__restrict__ b,
                             const double *__restrict__ c, double s,
                             int64_t n) {
    int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
    const int64_t n2 = n / 2;
    const f64x2 *b2 = reinterpret_cast<const f64x2 *>(b);
    const f64x2 *c2 = reinterpret_cast<const f64x2 *>(c);
    f64x2 *a2 = reinterpret_cast<f64x2 *>(a);
    for (int64_t k = i; k < n2; k += stride) {
        f64x2 x = __builtin_nontemporal_load(b2 + k);
        f64x2 y = __builtin_nontemporal_load(c2 + k);
        f64x2 r;
        r.x = __builtin_fma(s, y.x, x.x);
        r.y = __builtin_fma(s, y.y, x.y);
        __builtin_nontemporal_store(r, a2 + k);
    }
    if (i == 0 && (n & 1)) a[n - 1] = __builtin_fma(s, c[n - 1], b[n - 1]);
}

// Random-gather rate: lane i reads K entries of x at pseudo-random positions
// of the band [i - window/2, i - window/2 + window) (shifted inside [0, n)),
// all K requested before any is used, as an SpMV lane gathers its row's x
// entries.  No index stream is read, so the time is the gather path alone.
__device__ __forceinline__ uint32_t gather_mix(uint32_t a) {
    a ^= a >> 16;
    a *= 0x7feb352dU;
    a ^= a >> 15;
    a *= 0x846ca68bU;
    a ^= a >> 16;
    return a;
}

template <int K>
__global__ __launch_bounds__(256) void gather_kernel(const double *__restrict__ x,
                                                     double *__restrict__ y, int64_t n,
                                                     int64_t window) {
    const int64_t i = smaxhip::xcd_block(blockIdx.x, gridDim.x) * 256 + threadIdx.x;
    if (i >= n) return;
    int64_t lo = i - window / 2;
    lo = lo < 0 ? 0 : (lo > n - window ? n - window : lo);
    uint32_t h = gather_mix(static_cast<uint32_t>(i) * 2654435761U + 12345U);
    double v[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        h = gather_mix(h + static_cast<uint32_t>(k));
        v[k] = x[lo + static_cast<int64_t>(h % static_cast<uint32_t>(window))];
    }
    double acc = 0;
#pragma unroll
    for (int k = 0; k < K; ++k) acc += v[k];
    y[i] = acc;
}

// Persistent-style grid: enough blocks to fill 256 CUs several times over.
static dim3 stream_grid(int64_t n2) {
    int64_t blocks = ceil_div(n2, 256);
    if (blocks > 256 * 32) blocks = 256 * 32;
    if (blocks < 1) blocks = 1;
    return dim3(static_cast<unsigned>(blocks));
}

}  // namespace smaxhip

using namespace smaxhip;

extern "C" {

int smax_hip_abi_version(void) { return SMAX_HIP_ABI_VERSION; }

int smax_hip_check_types(int index_type, int value_type) {
    if (index_type != SMAX_HIP_INT32)
        return fail(nullptr, SMAX_HIP_ERROR_UNSUPPORTED,
                    "index type %d not supported on HIP (int32 only)",
                    index_type);
    if (!valid_value_type(value_type))
        return fail(nullptr, SMAX_HIP_ERROR_UNSUPPORTED,
                    "value type %d not supported on HIP", value_type);
    return SMAX_HIP_SUCCESS;
}

int smax_hip_ctx_create(int device, smax_hip_ctx *out) {
    if (!out)
        return fail(nullptr, SMAX_HIP_ERROR_INVALID_VALUE,
                    "smax_hip_ctx_create: null output");
    *out = nullptr;
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count == 0)
        return fail(nullptr, SMAX_HIP_ERROR_NO_DEVICE,
                    "smax_hip_ctx_create: no HIP device (%s)",
                    e != hipSuccess ? hipGetErrorString(e) : "count 0");
    if (device < 0 || device >= count)
        return fail(nullptr, SMAX_HIP_ERROR_NO_DEVICE,
                    "smax_hip_ctx_create: device %d of %d", device, count);
    smax_hip_ctx ctx = new smax_hip_ctx_s();
    ctx->device = device;
    e = hipSetDevice(device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&ctx->own_stream,
                                                      hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&ctx->ev_start);
    if (e == hipSuccess) e = hipEventCreate(&ctx->ev_stop);
    if (e == hipSuccess)
        e = hipHostMalloc(reinterpret_cast<void **>(&ctx->host_word),
                          sizeof(int64_t) * 8, hipHostMallocDefault);
    if (e != hipSuccess) {
        int st = hip_fail(nullptr, e, "smax_hip_ctx_create");
        smax_hip_ctx_destroy(ctx);
        return st;
    }
    ctx->stream = ctx->own_stream;
    *out = ctx;
    return SMAX_HIP_SUCCESS;
}

int smax_hip_ctx_destroy(smax_hip_ctx ctx) {
    if (!ctx) return SMAX_HIP_SUCCESS;
    (void)hipSetDevice(ctx->device);
    if (ctx->own_stream) (void)hipStreamSynchronize(ctx->own_stream);
    if (ctx->stream && ctx->stream != ctx->own_stream)
        (void)hipStreamSynchronize(ctx->stream);
    if (ctx->ev_start) (void)hipEventDestroy(ctx->ev_start);
    if (ctx->ev_stop) (void)hipEventDestroy(ctx->ev_stop);
    if (ctx->own_stream) (void)hipStreamDestroy(ctx->own_stream);
    if (ctx->host_word) (void)hipHostFree(ctx->host_word);
    if (ctx->work) (void)hipFree(ctx->work);
    for (void *p : ctx->retired_work) (void)hipFree(p);
    delete ctx;
    return SMAX_HIP_SUCCESS;
}

const char *smax_hip_last_error(smax_hip_ctx ctx) {
    if (ctx) return ctx->last_error.c_str();
    return g_thread_error.c_str();
}

int smax_hip_set_stream(smax_hip_ctx ctx, void *stream) {
    SMAX_HIP_CHECK_CTX(ctx);
    ctx->stream = stream ? static_cast<hipStream_t>(stream) : ctx->own_stream;
    return SMAX_HIP_SUCCESS;
}

void *smax_hip_get_stream(smax_hip_ctx ctx) {
    return ctx ? static_cast<void *>(ctx->stream) : nullptr;
}

int smax_hip_sync(smax_hip_ctx ctx) {
    SMAX_HIP_CHECK_CTX(ctx);
    SMAX_HIP_CALL(ctx, hipStreamSynchronize(ctx->stream));
    return SMAX_HIP_SUCCESS;
}

int smax_hip_malloc(smax_hip_ctx ctx, void **dptr, size_t bytes) {
    SMAX_HIP_CHECK_CTX(ctx);
    SMAX_HIP_REQUIRE(ctx, dptr, "smax_hip_malloc: null output");
    return dev_alloc(ctx, dptr, bytes, "smax_hip_malloc");
}

int smax_hip_free(smax_hip_ctx ctx, void *dptr) {
    SMAX_HIP_CHECK_CTX(ctx);
    if (dptr) SMAX_HIP_CALL(ctx, hipFree(dptr));
    return SMAX_HIP_SUCCESS;
}

static int copy(smax_hip_ctx ctx, void *dst, const void *src, size_t bytes,
                hipMemcpyKind kind, const char *what) {
    SMAX_HIP_CHECK_CTX(ctx);
    if (bytes == 0) return SMAX_HIP_SUCCESS;
    SMAX_HIP_REQUIRE(ctx, dst && src, "%s: null pointer", what);
    SMAX_HIP_CALL(ctx, hipMemcpyAsync(dst, src, bytes, kind, ctx->stream));
    SMAX_HIP_CALL(ctx, hipStreamSynchronize(ctx->stream));
    return SMAX_HIP_SUCCESS;
}

int smax_hip_memcpy_h2d(smax_hip_ctx ctx, void *dst, const void *src,
                        size_t bytes) {
    return copy(ctx, dst, src, bytes, hipMemcpyHostToDevice,
                "smax_hip_memcpy_h2d");
}

int smax_hip_memcpy_d2h(smax_hip_ctx ctx, void *dst, const void *src,
                        size_t bytes) {
    return copy(ctx, dst, src, bytes, hipMemcpyDeviceToHost,
                "smax_hip_memcpy_d2h");
}

int smax_hip_memcpy_d2d(smax_hip_ctx ctx, void *dst, const void *src,
                        size_t bytes) {
    return copy(ctx, dst, src, bytes, hipMemcpyDeviceToDevice,
                "smax_hip_memcpy_d2d");
}

int smax_hip_zero(smax_hip_ctx ctx, void *dptr, size_t bytes) {
    SMAX_HIP_CHECK_CTX(ctx);
    if (bytes == 0) return SMAX_HIP_SUCCESS;
    SMAX_HIP_REQUIRE(ctx, dptr, "smax_hip_zero: null pointer");
    int64_t blocks = ceil_div(static_cast<int64_t>(bytes / 16 + 1), 256);
    if (blocks > 8192) blocks = 8192;
    hipLaunchKernelGGL(zero_kernel, dim3(static_cast<unsigned>(blocks)),
                       dim3(256), 0, ctx->stream,
                       static_cast<uint8_t *>(dptr), bytes);
    SMAX_HIP_LAUNCHED(ctx, "zero_kernel");
    return SMAX_HIP_SUCCESS;
}

int smax_hip_timer_start(smax_hip_ctx ctx) {
    SMAX_HIP_CHECK_CTX(ctx);
    SMAX_HIP_CALL(ctx, hipEventRecord(ctx->ev_start, ctx->stream));
    return SMAX_HIP_SUCCESS;
}

int smax_hip_timer_stop(smax_hip_ctx ctx, float *elapsed_ms) {
    SMAX_HIP_CHECK_CTX(ctx);
    SMAX_HIP_REQUIRE(ctx, elapsed_ms, "smax_hip_timer_stop: null output");
    SMAX_HIP_CALL(ctx, hipEventRecord(ctx->ev_stop, ctx->stream));
    SMAX_HIP_CALL(ctx, hipEventSynchronize(ctx->ev_stop));
    SMAX_HIP_CALL(ctx,
                  hipEventElapsedTime(elapsed_ms, ctx->ev_start, ctx->ev_stop));
    return SMAX_HIP_SUCCESS;
}

int smax_hip_stream_copy(smax_hip_ctx ctx, const double *src, double *dst,
                         int64_t n) {
    SMAX_HIP_CHECK_CTX(ctx);
    SMAX_HIP_REQUIRE(ctx, src && dst && n > 0, "smax_hip_stream_copy: bad args");
    SMAX_HIP_REQUIRE(ctx,
                     ((reinterpret_cast<uintptr_t>(src) |
                       reinterpret_cast<uintptr_t>(dst)) & 15) == 0,
                     "smax_hip_stream_copy: pointers must be 16-byte aligned");
    hipLaunchKernelGGL(copy_kernel, stream_grid(n / 2), dim3(256), 0,
                       ctx->stream, src, dst, n);
    SMAX_HIP_LAUNCHED(ctx, "copy_kernel");
    return SMAX_HIP_SUCCESS;
}

int smax_hip_stream_copy_wide(smax_hip_ctx ctx, const double *src, double *dst,
                              int64_t n, int loads_per_lane) {
    SMAX_HIP_CHECK_CTX(ctx);
    SMAX_HIP_REQUIRE(ctx, src && dst && n > 0, "smax_hip_stream_copy_wide: bad args");
    SMAX_HIP_REQUIRE(ctx, aligned16(src, dst),
                     "smax_hip_stream_copy_wide: pointers must be 16-byte aligned");
    const dim3 g = stream_grid(ceil_div(n / 2, loads_per_lane));
    switch (loads_per_lane) {
    case 1: hipLaunchKernelGGL(copy_wide_kernel<1>, g, dim3(256), 0, ctx->stream, src, dst, n); break;
    case 2: hipLaunchKernelGGL(copy_wide_kernel<2>, g, dim3(256), 0, ctx->stream, src, dst, n); break;
    case 4: hipLaunchKernelGGL(copy_wide_kernel<4>, g, dim3(256), 0, ctx->stream, src, dst, n); break;
    case 8: hipLaunchKernelGGL(copy_wide_kernel<8>, g, dim3(256), 0, ctx->stream, src, dst, n); break;
    default:
        return fail(ctx, SMAX_HIP_ERROR_INVALID_VALUE,
                    "smax_hip_stream_copy_wide: loads_per_lane must be 1, 2, 4 or 8");
    }
    SMAX_HIP_LAUNCHED(ctx, "copy_wide_kernel");
    return SMAX_HIP_SUCCESS;
}

int smax_hip_stream_triad(smax_hip_ctx ctx, double *a, const double *b,
                          const double *c, double s, int64_t n) {
    SMAX_HIP_CHECK_CTX(ctx);
    SMAX_HIP_REQUIRE(ctx, a && b && c && n > 0, "smax_hip_stream_triad: bad args");
    SMAX_HIP_REQUIRE(ctx,
                     ((reinterpret_cast<uintptr_t>(a) |
                       reinterpret_cast<uintptr_t>(b) |
                       reinterpret_cast<uintptr_t>(c)) & 15) == 0,
                     "smax_hip_stream_triad: pointers must be 16-byte aligned");
    hipLaunchKernelGGL(triad_kernel, stream_grid(n / 2), dim3(256), 0,
                       ctx->stream, a, b, c, s, n);
    SMAX_HIP_LAUNCHED(ctx, "triad_kernel");
    return SMAX_HIP_SUCCESS;
}

int smax_hip_gather_rate(smax_hip_ctx ctx, const double *x, double *y, int64_t n,
                         int64_t window, int per_lane) {
    SMAX_HIP_CHECK_CTX(ctx);
    SMAX_HIP_REQUIRE(ctx, x && y && n > 0 && n <= INT32_MAX && window > 0 && window <= n,
                     "smax_hip_gather_rate: need 0 < window <= n <= 2^31-1");
    SMAX_HIP_REQUIRE(ctx, per_lane == 8 || per_lane == 16 || per_lane == 32,
                     "smax_hip_gather_rate: per_lane must be 8, 16 or 32");
    const dim3 g = grid_for(n, 256);
    switch (per_lane) {
    case 8:
        hipLaunchKernelGGL(gather_kernel<8>, g, dim3(256), 0, ctx->stream, x, y, n, window);
        break;
    case 16:
        hipLaunchKernelGGL(gather_kernel<16>, g, dim3(256), 0, ctx->stream, x, y, n, window);
        break;
    default:
        hipLaunchKernelGGL(gather_kernel<32>, g, dim3(256), 0, ctx->stream, x, y, n, window);
        break;
    }
    SMAX_HIP_LAUNCHED(ctx, "gather_kernel");
    return SMAX_HIP_SUCCESS;
}

}  // extern "C"
