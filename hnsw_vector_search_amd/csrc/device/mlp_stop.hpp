#pragma once

// The learned stop model: an MLP 16 -> 64 -> 64 -> 1 with ReLU and a sigmoid output, run in
// BF16 with FP32 accumulation. Inputs, weights and hidden activations are BF16 values; every
// sum is accumulated in FP32. The same device functions serve inference inside the beam kernel
// and the forward pass of training, so both see exactly the same arithmetic.
//
// Parameter vector (FP32, the C ABI's order; kMlpParams values):
//   mean[16], inv_std[16]              feature standardisation (not trained)
//   W1[16][64], b1[64], W2[64][64], b2[64], w3[64], b3

#include <hip/hip_runtime.h>

#include <cstdint>

namespace vdx::dev {

constexpr int kFeatures = 16;
constexpr int kHidden = 64;
constexpr int kMlpParams = 16 + 16 + 16 * 64 + 64 + 64 * 64 + 64 + 64 + 1;  // 5345

// Offsets into the FP32 parameter vector.
constexpr int kOffMean = 0, kOffInvStd = 16, kOffW1 = 32, kOffB1 = kOffW1 + 16 * 64,
              kOffW2 = kOffB1 + 64, kOffB2 = kOffW2 + 64 * 64, kOffW3 = kOffB2 + 64,
              kOffB3 = kOffW3 + 64;

// The device copy used for inference and the forward pass: BF16 weights, FP32 biases.
struct MlpWeights {
    float mean[16];
    float inv_std[16];
    float b1[64];
    float b2[64];
    float b3;
    float pad[3];
    std::uint16_t w1[16 * 64];  // bf16 bits, [i][u]
    std::uint16_t w2[64 * 64];  // bf16 bits, [j][u]
    std::uint16_t w3[64];       // bf16 bits
};

__device__ __forceinline__ float bf16_bits_to_float(std::uint16_t b) {
    return __uint_as_float(static_cast<std::uint32_t>(b) << 16);
}

// Round to the nearest BF16 value, ties to even (finite inputs).
__device__ __forceinline__ float bf16_round(float x) {
    const std::uint32_t u = __float_as_uint(x);
    const std::uint32_t r = u + 0x7FFFu + ((u >> 16) & 1u);
    return __uint_as_float(r & 0xFFFF0000u);
}

__device__ __forceinline__ std::uint16_t float_to_bf16_bits(float x) {
    return static_cast<std::uint16_t>(__float_as_uint(bf16_round(x)) >> 16);
}

// Standardised, BF16-rounded input i from raw feature value v.
__device__ __forceinline__ float mlp_input(const MlpWeights& w, int i, float v) {
    return bf16_round((v - w.mean[i]) * w.inv_std[i]);
}

// Forward pass for one sample across a wavefront: lane u is hidden unit u. `x` holds the 16
// standardised inputs (same in every lane). Returns the logit (same in every lane) and the
// lane's BF16 activations h1, h2.
__device__ __forceinline__ float mlp_forward(const MlpWeights& w, const float (&x)[kFeatures],
                                             int lane, float& h1, float& h2) {
    float a = w.b1[lane];
#pragma unroll
    for (int i = 0; i < kFeatures; ++i) {
        a = __builtin_fmaf(x[i], bf16_bits_to_float(w.w1[i * 64 + lane]), a);
    }
    h1 = bf16_round(fmaxf(a, 0.f));
    float b = w.b2[lane];
#pragma unroll 8
    for (int j = 0; j < kHidden; ++j) {
        b = __builtin_fmaf(__shfl(h1, j, 64), bf16_bits_to_float(w.w2[j * 64 + lane]), b);
    }
    h2 = bf16_round(fmaxf(b, 0.f));
    float z = h2 * bf16_bits_to_float(w.w3[lane]);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        z += __shfl_xor(z, off, 64);
    }
    // The butterfly leaves each lane a differently associated sum: take lane 0's, so every
    // lane (and the stop decision) sees one value.
    return __shfl(z, 0, 64) + w.b3;
}

}  // namespace vdx::dev
