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

// MFMA shortlist + exact re-rank + certification (bf_gemm.hip). prec: VDX_PREC_FP32,
// VDX_PREC_BF16 or VDX_PREC_FP8. The shortlist holds `widen` times the precision's multiple of
// k; the queries it does not certify are redone once with an 8x wider shortlist, then by the
// exact scan.
void bf_mfma(vdx_gpu_bf& b, int prec, const float* d_queries, std::uint64_t nq, std::uint32_t k,
             vdx::dev::Neighbor* d_out, std::uint64_t out_stride, std::uint32_t* d_count,
             Workspace& ws, hipStream_t s, vdx_bf_stats* stats, std::uint32_t widen = 1);

}  // namespace vdx::gpu
