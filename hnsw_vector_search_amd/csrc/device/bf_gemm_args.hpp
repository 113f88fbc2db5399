#pragma once

// Arguments of the MFMA brute-force kernels (bf_gemm.hip), shared with the host launchers.

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "beam_args.hpp"

namespace vdx::dev {

enum : int { kPrecF32 = 0, kPrecBF16 = 1, kPrecFP8 = 2 };

struct PrepArgs {
    const float* src;
    std::uint32_t ld_src, rows, dim, kp;
    int prec;
    void* dst;      // rows x kp operands
    float* norm2;   // ||v||^2
    float* nhalf;   // -||v||^2 / 2 (the same float halved), or null
    float* scale;   // FP8 per-vector scale (1 otherwise)
    float* qerr;    // ||v~ - v||
    float* qnorm;   // ||v~||
    std::uint32_t* integral;  // 1 if every element is an integer and v~ == v exactly
};

struct GemmArgs {
    const void* xq;     // base operands, rows of kp
    const float* xn;    // ||x||^2
    const float* xh;    // -||x||^2 / 2: the accumulator start of the BF16 LDS-fed kernel (L2)
    const float* xs;    // FP8 scales, or null
    const void* qq;     // query operands, rows of kp (a multiple of 128 rows, zero padded)
    const float* qn;    // ||q||^2
    const float* qs;    // FP8 scales, or null
    const float* tau;   // per-query threshold (filter mode), or null
    float xn_max;       // >= every ||x||^2 (the filter's pre-test margin)
    std::uint32_t nq, kp, ks, l2;
    std::uint32_t rows;        // rows covered: row j is base row j * row_stride
    std::uint32_t row_stride;  // 1 in the filter (mode 1)
    std::uint32_t mode;        // 0: dense store, 1: threshold filter
    float* dense;              // [nq][dense_ld]
    std::uint64_t dense_ld;
    std::uint64_t* buf;        // [nq][cap] candidate keys
    std::uint32_t cap;
    std::uint32_t* cnt;        // [nq] candidates seen (may exceed cap)
};

struct ThresholdArgs {
    const float* dense;
    std::uint64_t dense_ld;
    std::uint32_t count, rank;
    float* tau;
};

struct ShortlistArgs {
    const std::uint64_t* buf;
    const std::uint32_t* cnt;
    std::uint32_t cap, S;
    const float* tau;   // null when every row was kept
    std::uint32_t* list;        // [nq][S] ids
    float* approx;              // [nq][S] GEMM distances
    std::uint32_t* list_count;  // [nq]
    float* left;                // [nq] smallest GEMM distance left out (see strict)
    std::uint32_t* strict;      // [nq] 1: left-out distances >= left; 0: > left
};

struct RerankArgs {
    const float* vectors;  // exact base rows (ld floats)
    std::uint32_t ld, l2, n, nq, k;
    std::uint32_t nb, n8, tail, first_tail;
    const float* queries;  // prepared, rows of qld
    std::uint32_t qld, S, sort_cap, lds_per_wave;
    const std::uint32_t* list;
    const float* approx;
    const std::uint32_t* list_count;
    const float* left;
    const std::uint32_t* strict;
    const float* q_norm2;  // ||q||^2
    const float* q_qerr;   // ||q~ - q||
    const float* q_qnorm;  // ||q~||
    const std::uint32_t* q_integral;
    std::uint32_t x_integral;  // every base row integral
    float x_max_qerr, x_max_norm, x_max_norm2, x_max_qnorm, gamma_k, gamma_d;
    Neighbor* out;
    std::uint64_t out_stride;
    std::uint32_t* out_count;
    std::uint32_t* fail;
    float* max_err;
};

hipError_t launch_bf_prepare(const PrepArgs& args, hipStream_t stream);
hipError_t launch_bf_gather_rows(const float* src, const std::uint32_t* idx, std::uint32_t count,
                                 std::uint32_t dim, float* dst, hipStream_t stream);
hipError_t launch_bf_scatter_results(const Neighbor* part, const std::uint32_t* part_count,
                                     const std::uint32_t* idx, std::uint32_t count,
                                     std::uint32_t width, Neighbor* out, std::uint64_t out_stride,
                                     std::uint32_t* out_count, hipStream_t stream);
// One workgroup per 256 queries (waves = 8) and slab of base rows. shape (BF16 with 128-element
// operand rows): 0, eight waves of one 32-query block each (bf_gemm_bf16_dma_kernel); 1, four
// waves of two blocks each (bf_gemm_bf16_dma2_kernel). Other precisions and widths ignore it.
hipError_t launch_bf_gemm(const GemmArgs& args, unsigned qtiles, unsigned slabs, int prec, int waves,
                          hipStream_t stream, int shape = 0);
hipError_t launch_bf_threshold(const ThresholdArgs& args, unsigned nq, hipStream_t stream);
hipError_t launch_bf_shortlist(const ShortlistArgs& args, unsigned nq, hipStream_t stream);
hipError_t launch_bf_rerank(const RerankArgs& args, unsigned nq, std::size_t lds_per_wave,
                            hipStream_t stream, int shape);

}  // namespace vdx::dev
