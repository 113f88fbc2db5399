#pragma once

// Host-side planning and launch of the HNSW beam kernels, shared by the search entry points
// and the batch planner.

#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../../include/vdxgpu.h"
#include "beam_args.hpp"
#include "gpu_index.hpp"

namespace vdx::gpu {

// Launch shape choices; 0 keeps the default.
struct BeamConfig {
    std::uint32_t visited_log2 = 0;   // visited-cache slots per wave, log2
    std::uint32_t waves = 0;          // wavefronts per workgroup (1..4)
    std::uint32_t blocks_per_cu = 0;  // resident workgroups per CU to fill (queries in flight)
    bool f32_rows = false;            // never the 8-bit row copy
    std::uint32_t visited_form = 0;   // 0: the LDS cache; 2: the HBM bitmap (the LDS kernel only)
};

struct BeamPlan {
    std::uint32_t ef_cap, visited_log2, waves, grid;
    std::uint32_t grid_int;  // D128U8: the grid of the integer kernel (hnsw_beam_reg_u8int_kernel)
    int chunks;  // > 0: the register-resident kernel with R in `chunks` x 64 keys; 0: LDS kernel
    std::size_t lds_per_wave, lds_bytes;
    vdx::dev::BeamShape shape;  // D128U8: the plain register kernel over the 8-bit row copy
    // the visited set as an HBM bitmap: words per wave and the bytes the launch needs (grid x
    // waves rows), the caller's buffer; 0 with the LDS cache
    std::uint32_t bitmap_words;
    std::uint64_t bitmap_bytes;
};

vdx::dev::BeamShape beam_shape(std::uint64_t dim);

// The kernel mode an options struct asks for (validating a learned model's settings).
int beam_mode(const vdx_search_opts* opts, bool trace);

// R capacity and kernel for a launch whose largest max(ef, k) is max_efp: the register kernel
// with the smallest power-of-two chunk count that holds it, else the LDS kernel.
BeamPlan plan_beam(const vdx_gpu_hnsw& h, std::uint32_t max_efp, std::uint64_t nq, const BeamConfig& cfg,
                   int mode);

// Kernel arguments for one launch over prepared device queries.
vdx::dev::BeamArgs beam_args(const vdx_gpu_hnsw& h, const BeamPlan& plan, int mode,
                             const vdx_search_opts* opts, const float* d_queries, std::uint64_t nq,
                             const std::uint32_t* d_ks, std::uint32_t k, const std::uint32_t* d_efs,
                             std::uint32_t ef, vdx::dev::Neighbor* d_out, std::uint64_t out_stride,
                             std::uint32_t* d_count, std::uint32_t* d_queue, unsigned long long* d_stats);

// Enqueues the plan's kernels. Over the 8-bit rows with a.qflag set, the integer kernel and the
// float kernel both go (the flag lets one work), or the integer kernel alone with `int_only`
// (the caller knows every query is byte-valued and the flag says so).
void launch_beam(const vdx_gpu_hnsw& h, const BeamPlan& plan, int mode, const vdx::dev::BeamArgs& a,
                 hipStream_t stream, bool int_only = false);

BeamConfig config_from(const vdx_search_opts* opts);

// Locates the no-allocation entry point's scratch array on `device` (which must be current),
// once per device; vdx_gpu_open calls it.
void noalloc_scratch_attach(int device);

}  // namespace vdx::gpu
