#pragma once

// Arguments of the stop model's training and inference kernels (mlp_train.hip).

#include <hip/hip_runtime.h>

#include <cstdint>

#include "mlp_stop.hpp"

namespace vdx::dev {

struct MlpGradArgs {
    const MlpWeights* weights;
    const float* features;     // [n][16] raw features
    const float* labels;       // [n] 0 or 1
    const std::uint32_t* perm; // sample order of this epoch
    std::uint64_t start;       // first position of this batch in perm
    std::uint32_t count;       // samples in this batch
    std::uint32_t per_wave;    // samples per wavefront
    float* grad;               // [kMlpParams] summed gradient
    float* loss;               // summed loss
};

struct MlpAdamArgs {
    float* master;
    float* m;
    float* v;
    float* grad;
    MlpWeights* weights;
    float lr, beta1, beta2, bc1, bc2, inv_batch;
};

hipError_t launch_mlp_grad(const MlpGradArgs& a, hipStream_t s);
hipError_t launch_mlp_adam(const MlpAdamArgs& a, hipStream_t s);
hipError_t launch_mlp_pack(const float* params, MlpWeights* w, hipStream_t s);
hipError_t launch_mlp_infer(const MlpWeights* w, const float* features, std::uint64_t n, float* logits,
                            hipStream_t s);

}  // namespace vdx::dev
