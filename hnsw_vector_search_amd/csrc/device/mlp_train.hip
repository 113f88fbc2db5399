// Training of the learned stop model (mlp_stop.hpp): mini-batch Adam on binary cross-entropy.
//
// Forward and backward run in BF16 with FP32 accumulation: the forward pass is the inference
// function itself (BF16 inputs, weights and activations), gradients are accumulated in FP32,
// and Adam keeps FP32 master weights and moments, writing BF16 copies after every step.
//
// mlp_grad_kernel: one wavefront per group of samples, lane u = hidden unit u; the model is
// staged in LDS per workgroup (with a transposed W2 for the backward pass); each wave sums its
// samples' gradients in registers and adds them to the batch gradient once.
// mlp_adam_kernel: one thread per parameter; it also clears the gradient for the next batch.

#include <hip/hip_runtime.h>

#include "mlp_stop.hpp"
#include "mlp_train.hpp"

namespace vdx::dev {

__global__ __launch_bounds__(256) void mlp_grad_kernel(MlpGradArgs a) {
    __shared__ MlpWeights w;
    __shared__ std::uint16_t w2t[64 * 64];  // w2t[j * 64 + u] = W2[u][j]
    __shared__ __align__(16) std::uint16_t hbuf[4][64];
    for (std::uint32_t i = threadIdx.x; i < sizeof(MlpWeights) / 4; i += blockDim.x) {
        reinterpret_cast<std::uint32_t*>(&w)[i] = reinterpret_cast<const std::uint32_t*>(a.weights)[i];
    }
    __syncthreads();
    const std::uint16_t* w2h = reinterpret_cast<const std::uint16_t*>(w.w2p);
    for (std::uint32_t i = threadIdx.x; i < 64 * 64; i += blockDim.x) {
        const int j = static_cast<int>(i / 64), u = static_cast<int>(i % 64);
        w2t[j * 64 + u] = w2h[pair_index(u, j)];
    }
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const std::uint32_t wave = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const std::uint32_t first = wave * a.per_wave;
    float gw1[kFeatures], gw2[kHidden];
#pragma unroll
    for (int i = 0; i < kFeatures; ++i) gw1[i] = 0.f;
#pragma unroll
    for (int j = 0; j < kHidden; ++j) gw2[j] = 0.f;
    float gb1 = 0.f, gb2 = 0.f, gw3 = 0.f, gb3 = 0.f, loss = 0.f;
    const float w3u = bf16_bits_to_float(w.w3[lane]);
    for (std::uint32_t t = 0; t < a.per_wave; ++t) {
        const std::uint32_t slot = first + t;
        if (slot >= a.count) break;
        const std::uint64_t sample = a.perm[a.start + slot];
        const float raw = lane < kFeatures ? a.features[sample * kFeatures + lane] : 0.f;
        const float y = a.labels[sample];
        float x[kFeatures];
#pragma unroll
        for (int i = 0; i < kFeatures; ++i) {
            x[i] = mlp_input(w, i, __shfl(raw, i, 64));
        }
        float h1, h2;
        const float z = mlp_forward(w, x, lane, h1, h2, hbuf[threadIdx.x >> 6]);
        const float p = 1.0f / (1.0f + __expf(-z));
        // binary cross-entropy, from the logit for stability
        loss += fmaxf(z, 0.f) - z * y + __logf(1.0f + __expf(-fabsf(z)));
        const float dz = p - y;
        gb3 += dz;
        gw3 = __builtin_fmaf(dz, h2, gw3);
        const float dh2 = h2 > 0.f ? dz * w3u : 0.f;
        gb2 += dh2;
        float back = 0.f;
#pragma unroll 8
        for (int j = 0; j < kHidden; ++j) {
            gw2[j] = __builtin_fmaf(__shfl(h1, j, 64), dh2, gw2[j]);
            back = __builtin_fmaf(bf16_bits_to_float(w2t[j * 64 + lane]), __shfl(dh2, j, 64), back);
        }
        const float dh1 = h1 > 0.f ? back : 0.f;
        gb1 += dh1;
#pragma unroll
        for (int i = 0; i < kFeatures; ++i) {
            gw1[i] = __builtin_fmaf(x[i], dh1, gw1[i]);
        }
    }
#pragma unroll
    for (int i = 0; i < kFeatures; ++i) atomicAdd(a.grad + kOffW1 + i * 64 + lane, gw1[i]);
    atomicAdd(a.grad + kOffB1 + lane, gb1);
#pragma unroll 8
    for (int j = 0; j < kHidden; ++j) atomicAdd(a.grad + kOffW2 + j * 64 + lane, gw2[j]);
    atomicAdd(a.grad + kOffB2 + lane, gb2);
    atomicAdd(a.grad + kOffW3 + lane, gw3);
    if (lane == 0) {
        atomicAdd(a.grad + kOffB3, gb3);
        atomicAdd(a.loss, loss);
    }
}

__global__ void mlp_adam_kernel(MlpAdamArgs a) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < kOffW1 || p >= kMlpParams) return;  // standardisation constants are not trained
    const float g = a.grad[p] * a.inv_batch;
    a.grad[p] = 0.f;
    const float m = a.beta1 * a.m[p] + (1.f - a.beta1) * g;
    const float v = a.beta2 * a.v[p] + (1.f - a.beta2) * g * g;
    a.m[p] = m;
    a.v[p] = v;
    const float mhat = m * a.bc1;
    const float vhat = v * a.bc2;
    const float value = a.master[p] - a.lr * mhat / (sqrtf(vhat) + 1e-8f);
    a.master[p] = value;
    MlpWeights* w = a.weights;
    if (p < kOffB1) {
        const int e = p - kOffW1;
        reinterpret_cast<std::uint16_t*>(w->w1p)[pair_index(e / 64, e % 64)] = float_to_bf16_bits(value);
    } else if (p < kOffW2) {
        w->b1[p - kOffB1] = value;
    } else if (p < kOffB2) {
        const int e = p - kOffW2;
        reinterpret_cast<std::uint16_t*>(w->w2p)[pair_index(e / 64, e % 64)] = float_to_bf16_bits(value);
    } else if (p < kOffW3) {
        w->b2[p - kOffB2] = value;
    } else if (p < kOffB3) {
        w->w3[p - kOffW3] = float_to_bf16_bits(value);
    } else {
        w->b3 = value;
    }
}

// Writes the device copy (BF16 weights) from FP32 parameters.
__global__ void mlp_pack_kernel(const float* params, MlpWeights* w) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= kMlpParams) return;
    const float value = params[p];
    if (p < kOffInvStd) {
        w->mean[p] = value;
    } else if (p < kOffW1) {
        w->inv_std[p - kOffInvStd] = value;
    } else if (p < kOffB1) {
        const int e = p - kOffW1;
        reinterpret_cast<std::uint16_t*>(w->w1p)[pair_index(e / 64, e % 64)] = float_to_bf16_bits(value);
    } else if (p < kOffW2) {
        w->b1[p - kOffB1] = value;
    } else if (p < kOffB2) {
        const int e = p - kOffW2;
        reinterpret_cast<std::uint16_t*>(w->w2p)[pair_index(e / 64, e % 64)] = float_to_bf16_bits(value);
    } else if (p < kOffW3) {
        w->b2[p - kOffB2] = value;
    } else if (p < kOffB3) {
        w->w3[p - kOffW3] = float_to_bf16_bits(value);
    } else {
        w->b3 = value;
        w->pad[0] = w->pad[1] = w->pad[2] = 0.f;
    }
}

// Inference over a batch of feature rows (for checks): one wave per row, logit out.
__global__ __launch_bounds__(256) void mlp_infer_kernel(const MlpWeights* weights, const float* features,
                                                       std::uint64_t n, float* logits) {
    __shared__ MlpWeights w;
    __shared__ __align__(16) std::uint16_t hbuf[4][64];
    for (std::uint32_t i = threadIdx.x; i < sizeof(MlpWeights) / 4; i += blockDim.x) {
        reinterpret_cast<std::uint32_t*>(&w)[i] = reinterpret_cast<const std::uint32_t*>(weights)[i];
    }
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const std::uint64_t row = static_cast<std::uint64_t>(blockIdx.x) * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (row >= n) return;
    const float raw = lane < kFeatures ? features[row * kFeatures + lane] : 0.f;
    float x[kFeatures];
#pragma unroll
    for (int i = 0; i < kFeatures; ++i) x[i] = mlp_input(w, i, __shfl(raw, i, 64));
    float h1, h2;
    const float z = mlp_forward(w, x, lane, h1, h2, hbuf[threadIdx.x >> 6]);
    if (lane == 0) logits[row] = z;
}

hipError_t launch_mlp_grad(const MlpGradArgs& a, hipStream_t s) {
    const std::uint32_t waves = (a.count + a.per_wave - 1) / a.per_wave;
    hipLaunchKernelGGL(mlp_grad_kernel, dim3((waves + 3) / 4), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_mlp_adam(const MlpAdamArgs& a, hipStream_t s) {
    hipLaunchKernelGGL(mlp_adam_kernel, dim3((kMlpParams + 255) / 256), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_mlp_pack(const float* params, MlpWeights* w, hipStream_t s) {
    hipLaunchKernelGGL(mlp_pack_kernel, dim3((kMlpParams + 255) / 256), dim3(256), 0, s, params, w);
    return hipGetLastError();
}

hipError_t launch_mlp_infer(const MlpWeights* w, const float* features, std::uint64_t n, float* logits,
                            hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(mlp_infer_kernel, dim3(static_cast<unsigned>((n + 3) / 4)), dim3(256), 0, s, w,
                       features, n, logits);
    return hipGetLastError();
}

}  // namespace vdx::dev
