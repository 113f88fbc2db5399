#pragma once

// Host distance arithmetic. Every distance this library computes, on the host or on the GPU,
// follows one summation order: the order of vecdex's AVX2+FMA kernels
// (src/kernels_avx2.cpp:17-80), which is what vecdex runs on any x86-64 host with AVX2 and FMA.
// On such a host the intrinsics below compute it directly; elsewhere a scalar emulation
// computes the same bits. Fixing one order makes CPU-built graphs, host-normalized vectors and
// GPU distances agree bit for bit.
//
// The order, for dim = 32*nb + 8*n8 + t (0 <= n8 < 4, t < 8):
//   32 slots s = 8j + l (accumulator j = 0..3, lane l = 0..7), each a chain of fused
//   multiply-adds over the nb full 32-float blocks: acc[s] = fma(x, y, acc[s]) with
//   x = y = a - b for L2 and x = a, y = b for the inner product;
//   then the n8 8-float blocks into accumulator 0 (slots 0..7);
//   v[l] = (acc[l] + acc[8+l]) + (acc[16+l] + acc[24+l]);
//   h[i] = v[i] + v[i+4];  sum = (h[0] + h[2]) + (h[1] + h[3]);
//   then the t tail elements left to right, multiply and add rounded separately (ISO C++
//   builds of vecdex, -std=c++20, do not contract them).

#include <cstddef>

namespace vdx::cpu {

using DistanceFn = float (*)(const float* a, const float* b, std::size_t dim) noexcept;

float l2_squared(const float* a, const float* b, std::size_t dim) noexcept;
float inner_product(const float* a, const float* b, std::size_t dim) noexcept;

// The ranking distance under a FORMAT metric code: l2_squared, or 1 - inner_product.
float metric_distance(unsigned metric, const float* a, const float* b, std::size_t dim) noexcept;

// The scalar emulation of the same order, exposed for tests; equal to the above bit for bit.
float l2_squared_emulated(const float* a, const float* b, std::size_t dim) noexcept;
float inner_product_emulated(const float* a, const float* b, std::size_t dim) noexcept;

// True when the AVX2+FMA intrinsics are in use.
bool using_avx2() noexcept;

}  // namespace vdx::cpu
