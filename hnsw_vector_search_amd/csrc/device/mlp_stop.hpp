#pragma once

// The learned stop model: an MLP 16 -> 64 -> 64 -> 1 with ReLU and a sigmoid output, run in
// BF16 with FP32 accumulation. Inputs, weights and hidden activations are BF16 values; the
// layer sums are chains of two-term BF16 dot products accumulated in FP32 (v_dot2_f32_bf16). The same device functions serve inference inside the beam kernel
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

// The device copy used for inference and the forward pass: BF16 weights, FP32 biases. W1 and
// W2 are stored as BF16 pairs along the input index, w1p[p * 64 + u] = (W[2p][u] low half,
// W[2p + 1][u] high half), so one 32-bit word feeds one two-term BF16 dot product.
struct MlpWeights {
    float mean[16];
    float inv_std[16];
    float b1[64];
    float b2[64];
    float b3;
    float pad[3];
    std::uint32_t w1p[8 * 64];
    std::uint32_t w2p[32 * 64];
    std::uint16_t w3[64];  // bf16 bits
};

// Position of W[i][u] (input i, unit u) in the BF16-pair storage, as a 16-bit index.
__host__ __device__ constexpr int pair_index(int i, int u) { return ((i >> 1) * 64 + u) * 2 + (i & 1); }

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

typedef __bf16 mlp_bf16x2 __attribute__((ext_vector_type(2)));

// acc + lo(a) * lo(b) + hi(a) * hi(b) over BF16 pairs, FP32 accumulation.
__device__ __forceinline__ float dot2_bf16(std::uint32_t a, std::uint32_t b, float acc) {
    return __builtin_amdgcn_fdot2_f32_bf16(__builtin_bit_cast(mlp_bf16x2, a),
                                           __builtin_bit_cast(mlp_bf16x2, b), acc, false);
}

// Forward pass for one sample across a wavefront: lane u is hidden unit u. `input_pair(p)`
// gives the standardised BF16 inputs 2p and 2p + 1 packed as a pair (same in every lane);
// `hbuf` is 128 bytes of the wave's LDS (16-byte aligned) through which the first layer's
// activations reach every lane. Returns the logit (same in every lane) and the lane's BF16
// activations h1, h2.
template <typename InputPair>
__device__ __forceinline__ float mlp_forward_core(const MlpWeights& w, InputPair input_pair, int lane,
                                                  float& h1, float& h2, std::uint16_t* hbuf) {
    float a = w.b1[lane];
#pragma unroll
    for (int p = 0; p < kFeatures / 2; ++p) {
        a = dot2_bf16(input_pair(p), w.w1p[p * 64 + lane], a);
    }
    h1 = bf16_round(fmaxf(a, 0.f));
    hbuf[lane] = static_cast<std::uint16_t>(__float_as_uint(h1) >> 16);
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    const uint4* hp = reinterpret_cast<const uint4*>(hbuf);
    float b = w.b2[lane];
#pragma unroll
    for (int q = 0; q < kHidden / 8; ++q) {
        const uint4 v = hp[q];
        b = dot2_bf16(v.x, w.w2p[(4 * q + 0) * 64 + lane], b);
        b = dot2_bf16(v.y, w.w2p[(4 * q + 1) * 64 + lane], b);
        b = dot2_bf16(v.z, w.w2p[(4 * q + 2) * 64 + lane], b);
        b = dot2_bf16(v.w, w.w2p[(4 * q + 3) * 64 + lane], b);
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
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

__device__ __forceinline__ std::uint32_t bf16_pair(float lo, float hi) {
    return (__float_as_uint(lo) >> 16) | (__float_as_uint(hi) & 0xFFFF0000u);
}

// The forward pass from standardised inputs held in registers (training).
__device__ __forceinline__ float mlp_forward(const MlpWeights& w, const float (&x)[kFeatures],
                                             int lane, float& h1, float& h2, std::uint16_t* hbuf) {
    return mlp_forward_core(w, [&](int p) { return bf16_pair(x[2 * p], x[2 * p + 1]); }, lane, h1, h2,
                            hbuf);
}

// The forward pass from raw features in LDS (inference inside the search): each input pair is
// standardised as it is consumed, the same values mlp_input gives.
__device__ __forceinline__ float mlp_forward_raw(const MlpWeights& w, const float* raw, int lane,
                                                 float& h1, float& h2, std::uint16_t* hbuf) {
    return mlp_forward_core(
        w,
        [&](int p) {
            return bf16_pair(mlp_input(w, 2 * p, raw[2 * p]), mlp_input(w, 2 * p + 1, raw[2 * p + 1]));
        },
        lane, h1, h2, hbuf);
}

// Lane i's search-state feature i after an expansion (lanes 16 and up: 0). Every argument is
// wave-uniform, so every lane computes all 16 values and keeps its own (no branch per lane); the
// quotients multiply by reciprocals of the per-query scales (inv_scale: of the entry distance;
// inv_efp: of ef'; inv_m2: of 2M) and of the k-th distance. The trace mode records exactly these
// values, so the model is trained on the arithmetic it later sees.
//   0 best / entry   1 k-th / entry   2 expanded / k-th   3 best / k-th   4 steps / ef'
//   5 log2(1 + steps) / 8   6 |R| / ef'   7 improvements among the last 8 expansions / 8
//   8 expanded position in R / ef'   9 inserted / 2M   10 fresh neighbours / 2M
//   11 min(expansions since the top-k improved, 64) / 16   12 ef' / 512   13 k / 100
//   14 worst / k-th   15 1 for squared L2
__device__ __forceinline__ float step_feature(int lane, float d_best, float d_k, float d_worst, float d_cur,
                                              float inv_scale, std::uint32_t steps, std::uint32_t size,
                                              std::uint32_t hist, std::uint32_t pos, std::uint32_t inserted,
                                              std::uint32_t nf, std::uint32_t since, float inv_efp,
                                              std::uint32_t efp, float inv_m2, std::uint32_t k, bool l2) {
    const float inv_sk = __builtin_amdgcn_rcpf(fmaxf(fabsf(d_k), 1e-6f));
    const float f[kFeatures] = {d_best * inv_scale,
                                d_k * inv_scale,
                                d_cur * inv_sk,
                                d_best * inv_sk,
                                static_cast<float>(steps) * inv_efp,
                                __log2f(1.0f + static_cast<float>(steps)) * 0.125f,
                                static_cast<float>(size) * inv_efp,
                                static_cast<float>(__popc(hist)) * 0.125f,
                                static_cast<float>(pos) * inv_efp,
                                static_cast<float>(inserted) * inv_m2,
                                static_cast<float>(nf) * inv_m2,
                                static_cast<float>(min(since, 64u)) * 0.0625f,
                                static_cast<float>(efp) * (1.0f / 512.0f),
                                static_cast<float>(k) * 0.01f,
                                d_worst * inv_sk,
                                l2 ? 1.0f : 0.0f};
    float v = 0.0f;
#pragma unroll
    for (int i = 0; i < kFeatures; ++i) {
        v = lane == i ? f[i] : v;
    }
    return v;
}

// The stop model's logit from the feature each lane holds (step_feature): lane i standardises
// and rounds its own input (mlp_input), the 16 BF16 inputs meet in `xbuf` (32 bytes of the
// wave's LDS, 16-byte aligned), and mlp_forward_core runs on them — the values
// mlp_forward_raw computes from the raw features, with one LDS round trip instead of 16 reads.
__device__ __forceinline__ float mlp_logit_lanes(const MlpWeights& w, float feature, int lane, std::uint16_t* xbuf,
                                                 std::uint16_t* hbuf) {
    if (lane < kFeatures) {
        xbuf[lane] = float_to_bf16_bits(mlp_input(w, lane, feature));
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    const uint4 x0 = reinterpret_cast<const uint4*>(xbuf)[0];
    const uint4 x1 = reinterpret_cast<const uint4*>(xbuf)[1];
    const std::uint32_t words[kFeatures / 2] = {x0.x, x0.y, x0.z, x0.w, x1.x, x1.y, x1.z, x1.w};
    float h1, h2;
    return mlp_forward_core(w, [&](int p) { return words[p]; }, lane, h1, h2, hbuf);
}

}  // namespace vdx::dev
