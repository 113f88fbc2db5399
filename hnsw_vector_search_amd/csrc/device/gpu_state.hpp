#pragma once

// Host-side state behind the GPU half of the C ABI: the device context, device buffers, the
// uploaded HNSW graph and brute-force base, and per-call workspaces.

#include <hip/hip_runtime.h>

#include <atomic>
#include <cstddef>
#include <cstdint>
#include <memory>
#include <mutex>
#include <utility>
#include <vector>

#include "../host/common.hpp"
#include "../host/host_pool.hpp"

namespace vdx::gpu {

[[noreturn]] void throw_hip(hipError_t err, const char* what);

inline void check(hipError_t err, const char* what) {
    if (err != hipSuccess) {
        throw_hip(err, what);
    }
}

// An owned device allocation.
class DeviceBuffer {
public:
    DeviceBuffer() = default;
    explicit DeviceBuffer(std::size_t bytes) { resize(bytes); }
    ~DeviceBuffer() { release(); }
    DeviceBuffer(const DeviceBuffer&) = delete;
    DeviceBuffer& operator=(const DeviceBuffer&) = delete;
    DeviceBuffer(DeviceBuffer&& o) noexcept : ptr_(o.ptr_), bytes_(o.bytes_) {
        o.ptr_ = nullptr;
        o.bytes_ = 0;
    }
    // Grows to at least `bytes` (contents not kept).
    void reserve(std::size_t bytes) {
        if (bytes > bytes_) {
            resize(bytes);
        }
    }
    void resize(std::size_t bytes);
    void release() noexcept;
    void swap(DeviceBuffer& o) noexcept {
        std::swap(ptr_, o.ptr_);
        std::swap(bytes_, o.bytes_);
    }
    template <typename T>
    T* as() const noexcept {
        return static_cast<T*>(ptr_);
    }
    std::size_t bytes() const noexcept { return bytes_; }

private:
    void* ptr_ = nullptr;
    std::size_t bytes_ = 0;
};

// An owned pinned host allocation (hipHostMalloc).
class PinnedBuffer {
public:
    PinnedBuffer() = default;
    ~PinnedBuffer() { release(); }
    PinnedBuffer(const PinnedBuffer&) = delete;
    PinnedBuffer& operator=(const PinnedBuffer&) = delete;
    void reserve(std::size_t bytes);
    void release() noexcept;
    template <typename T>
    T* as() const noexcept {
        return static_cast<T*>(ptr_);
    }
    // The allocation's address for kernels (mapped: a kernel reads and writes it over PCIe).
    template <typename T>
    T* dev() const noexcept {
        return static_cast<T*>(dev_);
    }

private:
    void* ptr_ = nullptr;
    void* dev_ = nullptr;
    std::size_t bytes_ = 0;
};

// Device buffers and a stream for one in-flight search call; pooled per context so that
// concurrent calls never share one (the GPU analogue of vecdex's ScratchPool,
// src/search_scratch.hpp:61-105).
struct Workspace {
    hipStream_t stream = nullptr;
    DeviceBuffer queries, ks, efs, out, count, counters, aux0, aux1, aux2, aux3;
    DeviceBuffer extra[20];  // the MFMA brute force's per-call buffers
    DeviceBuffer bitmap;     // the HNSW search's visited bitmaps (visited_form 2)
    DeviceBuffer fp64[4];    // the brute force's float64 pass: window, counts, distances, counters
    PinnedBuffer h_queries, h_out, h_count;
    PinnedBuffer h_qbytes;  // byte-valued queries staged as 128 bytes each (the integer kernel)
    ~Workspace();
};

}  // namespace vdx::gpu

struct vdx_gpu_ctx {
    int device = 0;
    int cu_count = 0;
    std::size_t lds_per_block = 65536;
    hipStream_t stream = nullptr;
    hipEvent_t events[16] = {};
    std::mutex pool_mutex;
    std::vector<std::unique_ptr<vdx::gpu::Workspace>> free_workspaces;
    // host threads for the search entry points' validation, staging and unpacking
    std::unique_ptr<vdx::HostPool> host_pool;
    ~vdx_gpu_ctx();
};

namespace vdx::gpu {

// RAII loan of a workspace from the context's pool.
class WorkspaceLease {
public:
    explicit WorkspaceLease(vdx_gpu_ctx& ctx);
    ~WorkspaceLease();
    WorkspaceLease(const WorkspaceLease&) = delete;
    WorkspaceLease& operator=(const WorkspaceLease&) = delete;
    Workspace& operator*() const noexcept { return *ws_; }
    Workspace* operator->() const noexcept { return ws_.get(); }

private:
    vdx_gpu_ctx& ctx_;
    std::unique_ptr<Workspace> ws_;
};

// Makes `device` current for the calling thread while alive.
class DeviceGuard {
public:
    explicit DeviceGuard(int device);
    ~DeviceGuard();

private:
    int previous_ = 0;
};

// The split of a dimension in the distance summation order.
struct DimSplit {
    std::uint32_t nb, n8, tail, first_tail, ld;
};
DimSplit split_dim(std::uint64_t dim);

}  // namespace vdx::gpu
