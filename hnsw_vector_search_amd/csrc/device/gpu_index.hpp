#pragma once

// The HBM copies behind the C ABI's opaque handles.

#include <atomic>
#include <cstdint>
#include <mutex>
#include <vector>

#include "gpu_state.hpp"

struct vdx_gpu_hnsw {
    vdx_gpu_ctx* ctx = nullptr;
    std::uint32_t metric = 0;
    std::uint64_t dim = 0, n = 0, M = 0, max_level = 0, entry_point = 0;
    std::uint64_t upper_blocks = 0;  // U of the uploaded graph (its upper section's blocks)
    // Counts vdx_gpu_hnsw_update calls: launches captured before one name buffers and sizes it
    // may have replaced.
    std::uint64_t generation = 0;
    // Host copies of what updates check against: each node's level and first upper block, and
    // the link checker's per-node marks (kept so an update costs what it changes, not n).
    std::vector<std::uint8_t> levels;
    std::vector<std::uint32_t> offsets;
    std::vector<std::uint32_t> link_marks;
    std::uint32_t link_epoch = 0;
    // staging of an update's changed blocks (pinned) and their device copy, kept between updates
    vdx::gpu::PinnedBuffer staging;
    vdx::gpu::DeviceBuffer changes;
    vdx::gpu::DimSplit split{};
    vdx::gpu::DeviceBuffer vectors, layer0, upper, upper_offset;
    // Dimension 128 with every stored value an integer in [0, 255]: the rows again as bytes, 128
    // per row in the beam's lane order (a quarter of the gather traffic, the same floats).
    vdx::gpu::DeviceBuffer rows_u8, u8_flag;
    bool u8 = false;
    // Queue counters and stats of the device entry point: one 256-byte slot per call, taken in
    // turn. A call holds its slot's lock while it waits for the slot's previous launch (its
    // event) and enqueues its own, so calls in flight on different streams never share one.
    static constexpr std::uint32_t kRingSlots = 64;
    // stamps of the byte-query flag, one per launch that checks its queries (never 0)
    std::atomic<std::uint32_t> query_epoch{0};
    vdx::gpu::DeviceBuffer ring;
    std::atomic<std::uint32_t> next_slot{0};
    std::mutex ring_mutex[kRingSlots];
    hipEvent_t ring_event[kRingSlots] = {};
    // the device entry point's visited bitmaps (visited_form 2): one buffer, so launches that
    // use it follow one another (each waits for the previous one's event)
    std::mutex bitmap_mutex;
    vdx::gpu::DeviceBuffer bitmap;
    hipEvent_t bitmap_done = nullptr;

    vdx_gpu_hnsw() = default;
    vdx_gpu_hnsw(const vdx_gpu_hnsw&) = delete;
    vdx_gpu_hnsw& operator=(const vdx_gpu_hnsw&) = delete;
    ~vdx_gpu_hnsw() {
        if (bitmap_done != nullptr) {
            (void)hipEventSynchronize(bitmap_done);
            (void)hipEventDestroy(bitmap_done);
        }
        for (hipEvent_t& e : ring_event) {
            if (e != nullptr) {
                (void)hipEventSynchronize(e);
                (void)hipEventDestroy(e);
            }
        }
    }
};

namespace vdx::gpu {

// A base quantized for one MFMA precision, made on first use and kept.
struct BfOperands {
    bool ready = false;
    std::uint32_t kp = 0;      // operand row length (elements)
    std::uint64_t rows = 0;    // padded row count (multiple of 64)
    DeviceBuffer xq, norm2, nhalf, scale, qerr, qnorm, integral;
    bool all_integral = false;
    float max_qerr = 0.f, max_norm = 0.f, max_norm2 = 0.f, max_qnorm = 0.f;
};

}  // namespace vdx::gpu

struct vdx_gpu_bf {
    vdx_gpu_ctx* ctx = nullptr;
    std::uint32_t metric = 0;
    std::uint64_t dim = 0, n = 0;
    vdx::gpu::DimSplit split{};
    vdx::gpu::DeviceBuffer vectors;
    std::mutex prep_mutex;
    vdx::gpu::BfOperands prep[3];
    // the BF16 filter's GEMM shape (launch_bf_gemm's `shape`), set by vdx_gpu_bf_tune
    std::atomic<int> gemm_shape{0};
};

struct vdx_gpu_mlp {
    vdx_gpu_ctx* ctx = nullptr;
    vdx::gpu::DeviceBuffer master, m, v, grad, weights, loss, perm;
    std::uint64_t step = 0;
};
