#pragma once

// Host orchestration of the brute-force kernels.

#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../../include/vdxgpu.h"
#include "beam_args.hpp"
#include "gpu_index.hpp"

namespace vdx::gpu {

// Exact scan of every row (bf_exact.hip).
void bf_exact(vdx_gpu_bf& b, const float* d_queries, std::uint64_t nq, std::uint32_t k,
              vdx::dev::Neighbor* d_out, std::uint64_t out_stride, std::uint32_t* d_count,
              Workspace& ws, hipStream_t s);

// MFMA shortlist + exact re-rank + certification, exact fallback per uncertified query
// (bf_gemm.hip). prec: VDX_PREC_FP32, VDX_PREC_BF16 or VDX_PREC_FP8.
void bf_mfma(vdx_gpu_bf& b, int prec, const float* d_queries, std::uint64_t nq, std::uint32_t k,
             vdx::dev::Neighbor* d_out, std::uint64_t out_stride, std::uint32_t* d_count,
             Workspace& ws, hipStream_t s, vdx_bf_stats* stats);

}  // namespace vdx::gpu
