#include "gpu_state.hpp"

#include <string>

namespace vdx::gpu {

void throw_hip(hipError_t err, const char* what) {
    const std::string msg = std::string(what) + ": " + hipGetErrorName(err) + " (" +
                            hipGetErrorString(err) + ")";
    if (err == hipErrorOutOfMemory) {
        throw std::bad_alloc();
    }
    throw device_error(msg);
}

void DeviceBuffer::resize(std::size_t bytes) {
    release();
    if (bytes == 0) {
        return;
    }
    check(hipMalloc(&ptr_, bytes), "hipMalloc");
    bytes_ = bytes;
}

void DeviceBuffer::release() noexcept {
    if (ptr_ != nullptr) {
        (void)hipFree(ptr_);
        ptr_ = nullptr;
        bytes_ = 0;
    }
}

void PinnedBuffer::reserve(std::size_t bytes) {
    if (bytes <= bytes_) {
        return;
    }
    release();
    check(hipHostMalloc(&ptr_, bytes, hipHostMallocDefault), "hipHostMalloc");
    bytes_ = bytes;
    check(hipHostGetDevicePointer(&dev_, ptr_, 0), "hipHostGetDevicePointer");
}

void PinnedBuffer::release() noexcept {
    if (ptr_ != nullptr) {
        (void)hipHostFree(ptr_);
        ptr_ = nullptr;
        dev_ = nullptr;
        bytes_ = 0;
    }
}

Workspace::~Workspace() {
    if (stream != nullptr) {
        (void)hipStreamSynchronize(stream);
        (void)hipStreamDestroy(stream);
    }
}

WorkspaceLease::WorkspaceLease(vdx_gpu_ctx& ctx) : ctx_(ctx) {
    {
        const std::lock_guard<std::mutex> lock(ctx.pool_mutex);
        if (!ctx.free_workspaces.empty()) {
            ws_ = std::move(ctx.free_workspaces.back());
            ctx.free_workspaces.pop_back();
        }
    }
    if (!ws_) {
        ws_ = std::make_unique<Workspace>();
        check(hipStreamCreateWithFlags(&ws_->stream, hipStreamNonBlocking), "hipStreamCreate");
    }
}

WorkspaceLease::~WorkspaceLease() {
    const std::lock_guard<std::mutex> lock(ctx_.pool_mutex);
    try {
        ctx_.free_workspaces.push_back(std::move(ws_));
    } catch (...) {
        // the workspace is simply freed
    }
}

DeviceGuard::DeviceGuard(int device) {
    check(hipGetDevice(&previous_), "hipGetDevice");
    if (previous_ != device) {
        check(hipSetDevice(device), "hipSetDevice");
    }
}

DeviceGuard::~DeviceGuard() { (void)hipSetDevice(previous_); }

DimSplit split_dim(std::uint64_t dim) {
    DimSplit s{};
    s.nb = static_cast<std::uint32_t>(dim / 32);
    const std::uint64_t rem = dim - 32ull * s.nb;
    s.n8 = static_cast<std::uint32_t>(rem / 8);
    s.tail = static_cast<std::uint32_t>(rem - 8ull * s.n8);
    s.first_tail = 32 * s.nb + 8 * s.n8;
    s.ld = static_cast<std::uint32_t>((dim + 3) / 4 * 4);
    return s;
}

}  // namespace vdx::gpu

vdx_gpu_ctx::~vdx_gpu_ctx() {
    host_pool.reset();
    free_workspaces.clear();
    for (hipEvent_t& e : events) {
        if (e != nullptr) {
            (void)hipEventDestroy(e);
        }
    }
    if (stream != nullptr) {
        (void)hipStreamSynchronize(stream);
        (void)hipStreamDestroy(stream);
    }
}
