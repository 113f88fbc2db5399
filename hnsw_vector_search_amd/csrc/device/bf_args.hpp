#pragma once

// Arguments of the brute-force kernels, shared with their host-side launchers.

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "beam_args.hpp"

namespace vdx::dev {

struct BfScanArgs {
    const float* vectors;  // n rows of ld floats
    std::uint32_t ld, dim, n, l2;
    std::uint32_t nb, n8, tail, first_tail;
    const float* queries;  // nq rows of qld floats, prepared
    std::uint32_t qld, nq;
    std::uint32_t k;     // results wanted per query
    std::uint32_t kcap;  // list capacity per (slab, query) in `partial`, >= min(k, n)
    std::uint64_t* partial;        // [slabs][nq][kcap] sorted keys
    std::uint32_t* partial_count;  // [slabs][nq]
    Neighbor* out;
    std::uint64_t out_stride;
    std::uint32_t* out_count;
};

// Shape ids: 0 for dim 128, 1 for dim 100, 2 for any dim below 288.
hipError_t launch_bf_exact_scan(const BfScanArgs& args, unsigned qgroups, unsigned slabs,
                                unsigned waves_per_group, std::size_t lds_bytes,
                                hipStream_t stream, int shape);

}  // namespace vdx::dev
