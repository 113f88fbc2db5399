// The C ABI of the learned stop model: create, read back, train (mlp_train.hip), infer.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <numeric>
#include <random>
#include <vector>

#include "../host/capi_common.hpp"
#include "gpu_index.hpp"
#include "gpu_state.hpp"
#include "mlp_train.hpp"

using vdx::guarded;
using vdx::require;
using vdx::gpu::check;
using vdx::gpu::DeviceGuard;

extern "C" {

int vdx_gpu_mlp_create(vdx_gpu_ctx* ctx, const float* params, uint64_t count, vdx_gpu_mlp** out) {
    return guarded([&] {
        require(ctx != nullptr && params != nullptr && out != nullptr, "vdx_gpu_mlp_create: null argument");
        require(count == static_cast<uint64_t>(vdx::dev::kMlpParams), "vdx_gpu_mlp_create: wrong parameter count");
        for (uint64_t i = 0; i < count; ++i) {
            require(std::isfinite(params[i]), "vdx_gpu_mlp_create: parameters must be finite");
        }
        DeviceGuard guard(ctx->device);
        auto m = std::make_unique<vdx_gpu_mlp>();
        m->ctx = ctx;
        const std::size_t bytes = count * sizeof(float);
        m->master.resize(bytes);
        m->m.resize(bytes);
        m->v.resize(bytes);
        m->grad.resize(bytes);
        m->weights.resize(sizeof(vdx::dev::MlpWeights));
        m->loss.resize(256);
        hipStream_t s = ctx->stream;
        check(hipMemcpyAsync(m->master.as<void>(), params, bytes, hipMemcpyHostToDevice, s), "copy");
        for (vdx::gpu::DeviceBuffer* b : {&m->m, &m->v, &m->grad}) {
            check(hipMemsetAsync(b->as<void>(), 0, bytes, s), "hipMemsetAsync");
        }
        check(hipMemsetAsync(m->weights.as<void>(), 0, sizeof(vdx::dev::MlpWeights), s), "hipMemsetAsync");
        check(vdx::dev::launch_mlp_pack(m->master.as<float>(), m->weights.as<vdx::dev::MlpWeights>(), s),
              "mlp_pack launch");
        check(hipStreamSynchronize(s), "mlp create");
        *out = m.release();
    });
}

int vdx_gpu_mlp_params(vdx_gpu_mlp* m, float* params, uint64_t count) {
    return guarded([&] {
        require(m != nullptr && params != nullptr, "vdx_gpu_mlp_params: null argument");
        require(count == static_cast<uint64_t>(vdx::dev::kMlpParams), "vdx_gpu_mlp_params: wrong parameter count");
        DeviceGuard guard(m->ctx->device);
        check(hipMemcpy(params, m->master.as<void>(), count * sizeof(float), hipMemcpyDeviceToHost), "copy");
    });
}

void vdx_gpu_mlp_free(vdx_gpu_mlp* m) {
    if (m != nullptr) {
        DeviceGuard guard(m->ctx->device);
        delete m;
    }
}

int vdx_gpu_mlp_train(vdx_gpu_mlp* m, const float* d_features, const float* d_labels, uint64_t n,
                      uint32_t epochs, uint32_t batch, float lr, uint64_t seed, float* epoch_loss) {
    return guarded([&] {
        require(m != nullptr && d_features != nullptr && d_labels != nullptr, "vdx_gpu_mlp_train: null argument");
        require(n >= 1 && n < 0xFFFFFFFFull, "vdx_gpu_mlp_train: sample count out of range");
        require(batch >= 1 && epochs >= 1 && lr > 0.f, "vdx_gpu_mlp_train: bad hyper-parameters");
        DeviceGuard guard(m->ctx->device);
        hipStream_t s = m->ctx->stream;
        m->perm.reserve(n * 4);
        std::vector<std::uint32_t> order(n);
        std::iota(order.begin(), order.end(), 0u);
        std::mt19937_64 rng(seed);
        const float beta1 = 0.9f, beta2 = 0.999f;
        for (uint32_t e = 0; e < epochs; ++e) {
            std::shuffle(order.begin(), order.end(), rng);
            check(hipMemcpyAsync(m->perm.as<void>(), order.data(), n * 4, hipMemcpyHostToDevice, s), "copy");
            check(hipMemsetAsync(m->loss.as<void>(), 0, 4, s), "hipMemsetAsync");
            for (uint64_t b0 = 0; b0 < n; b0 += batch) {
                const std::uint32_t cnt = static_cast<std::uint32_t>(std::min<uint64_t>(batch, n - b0));
                vdx::dev::MlpGradArgs g{};
                g.weights = m->weights.as<vdx::dev::MlpWeights>();
                g.features = d_features;
                g.labels = d_labels;
                g.perm = m->perm.as<std::uint32_t>();
                g.start = b0;
                g.count = cnt;
                g.per_wave = 16;
                g.grad = m->grad.as<float>();
                g.loss = m->loss.as<float>();
                check(vdx::dev::launch_mlp_grad(g, s), "mlp_grad launch");
                ++m->step;
                vdx::dev::MlpAdamArgs ad{};
                ad.master = m->master.as<float>();
                ad.m = m->m.as<float>();
                ad.v = m->v.as<float>();
                ad.grad = m->grad.as<float>();
                ad.weights = m->weights.as<vdx::dev::MlpWeights>();
                ad.lr = lr;
                ad.beta1 = beta1;
                ad.beta2 = beta2;
                ad.bc1 = 1.0f / (1.0f - std::pow(beta1, static_cast<float>(m->step)));
                ad.bc2 = 1.0f / (1.0f - std::pow(beta2, static_cast<float>(m->step)));
                ad.inv_batch = 1.0f / static_cast<float>(cnt);
                check(vdx::dev::launch_mlp_adam(ad, s), "mlp_adam launch");
            }
            float loss = 0.f;
            check(hipMemcpyAsync(&loss, m->loss.as<void>(), 4, hipMemcpyDeviceToHost, s), "copy");
            check(hipStreamSynchronize(s), "mlp train epoch");
            if (epoch_loss != nullptr) {
                epoch_loss[e] = loss / static_cast<float>(n);
            }
        }
    });
}

int vdx_gpu_mlp_infer(vdx_gpu_mlp* m, const float* d_features, uint64_t n, float* d_logits) {
    return guarded([&] {
        require(m != nullptr && d_features != nullptr && d_logits != nullptr, "vdx_gpu_mlp_infer: null argument");
        DeviceGuard guard(m->ctx->device);
        hipStream_t s = m->ctx->stream;
        check(vdx::dev::launch_mlp_infer(m->weights.as<vdx::dev::MlpWeights>(), d_features, n, d_logits, s),
              "mlp_infer launch");
        check(hipStreamSynchronize(s), "mlp infer");
    });
}

}  // extern "C"
