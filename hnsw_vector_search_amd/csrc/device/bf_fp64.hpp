#pragma once

// Arguments of the brute force's float64 pass (bf_fp64.hip), shared with its host launcher.

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "beam_args.hpp"

namespace vdx::dev {

struct BfFp64Args {
    const float* vectors;  // n stored rows of ld floats (unit length under cosine)
    std::uint32_t ld, dim, metric, n;  // metric: FORMAT code (0 L2, 1 inner product, 2 cosine)
    const float* queries;  // nq prepared rows of qld floats
    std::uint32_t qld, nq;
    const Neighbor* cand;             // [nq][m] each query's first results in `closer` order
    const std::uint32_t* cand_count;  // [nq]
    std::uint32_t m;                  // window length
    double* d64;                      // [nq][m] float64 distances of the window
    Neighbor* out;                    // [nq][out_stride] the first `keep` after the re-rank (or null)
    std::uint64_t out_stride;
    std::uint32_t* out_count;         // [nq] (or null)
    std::uint32_t keep;
    std::uint32_t ties;               // 1: re-rank one-ulp groups by float64; 0: measure only
    // [0] queries whose last group reached the window's end, [1] candidates in groups of two or
    // more, [2] queries whose first `keep` changed, [3] max |float32 - float64| and [4] the max of that
    // over 1 + |float64| (bit patterns of non-negative doubles), over the returned candidates
    unsigned long long* stats;
};

constexpr std::uint32_t kFp64MaxWindow = 1024;  // 16 KB of LDS per query in the re-rank

inline std::size_t bf_fp64_lds_bytes(std::uint32_t m) { return 16u * static_cast<std::size_t>(m); }

hipError_t launch_bf_fp64(const BfFp64Args& args, hipStream_t stream);

}  // namespace vdx::dev
