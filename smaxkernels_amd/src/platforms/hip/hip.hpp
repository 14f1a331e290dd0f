// HIP platform memory: device buffers owned (and freed) by the platform,
// replacing transfer_HtoD / transfer_DtoH of src/platforms/cuda/cuda.cuh:12-82
// (which allocate on every call and never free).
#pragma once

#include <cstddef>
#include <string>
#include <utility>

#include "../../error_handler.hpp"
#include "smax_hip.h"

namespace SMAX::PLATFORM::HIP {

// Logs a failed C-ABI call as a warning and returns 1 (the reference's
// error convention for kernel phases); 0 on success.
inline int report(smax_hip_ctx ctx, int status, const char *what) {
    if (status == SMAX_HIP_SUCCESS) return 0;
    ErrorHandler::warning(std::string("[HIP] ") + what + ": " +
                          smax_hip_last_error(ctx) + " (status " +
                          std::to_string(status) + ")");
    return 1;
}

// One context (device + stream) per kernel, created on first use so that a
// HIP kernel can be registered on a host without a GPU.
class Stream {
    smax_hip_ctx ctx_ = nullptr;

  public:
    Stream() = default;
    Stream(const Stream &) = delete;
    Stream &operator=(const Stream &) = delete;
    ~Stream() { smax_hip_ctx_destroy(ctx_); }
    int open(int device = 0) {
        if (ctx_) return 0;
        return report(nullptr, smax_hip_ctx_create(device, &ctx_), "context");
    }
    smax_hip_ctx get() const { return ctx_; }
};

class DeviceBuffer {
    smax_hip_ctx ctx_ = nullptr;
    void *ptr_ = nullptr;
    size_t bytes_ = 0;

  public:
    DeviceBuffer() = default;
    DeviceBuffer(const DeviceBuffer &) = delete;
    DeviceBuffer &operator=(const DeviceBuffer &) = delete;
    ~DeviceBuffer() { release(); }

    void release() {
        if (ptr_) smax_hip_free(ctx_, ptr_);
        ptr_ = nullptr;
        bytes_ = 0;
    }
    int reserve(smax_hip_ctx ctx, size_t bytes) {
        if (ptr_ && bytes_ >= bytes && ctx_ == ctx) return 0;
        release();
        ctx_ = ctx;
        int rc = report(ctx, smax_hip_malloc(ctx, &ptr_, bytes ? bytes : 1), "malloc");
        if (!rc) bytes_ = bytes;
        return rc;
    }
    int upload(smax_hip_ctx ctx, const void *host, size_t bytes) {
        if (reserve(ctx, bytes)) return 1;
        if (!bytes) return 0;
        return report(ctx, smax_hip_memcpy_h2d(ctx, ptr_, host, bytes), "copy to device");
    }
    int download(void *host, size_t bytes) const {
        if (!bytes) return 0;
        return report(ctx_, smax_hip_memcpy_d2h(ctx_, host, ptr_, bytes), "copy to host");
    }
    void *get() const { return ptr_; }
    template <typename T> T *as() const { return static_cast<T *>(ptr_); }
    void swap(DeviceBuffer &o) {
        std::swap(ctx_, o.ctx_);
        std::swap(ptr_, o.ptr_);
        std::swap(bytes_, o.bytes_);
    }
};

} // namespace SMAX::PLATFORM::HIP
