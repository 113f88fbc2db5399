#include "cpu_distance.hpp"

#include <cmath>

#include <immintrin.h>

namespace vdx::cpu {
namespace {

// ---- scalar emulation of the 32-slot order --------------------------------------------------

template <bool kL2>
float emulate(const float* a, const float* b, std::size_t dim) noexcept {
    float acc[32] = {};
    std::size_t i = 0;
    for (; dim - i >= 32; i += 32) {
        for (int s = 0; s < 32; ++s) {
            if constexpr (kL2) {
                const float d = a[i + s] - b[i + s];
                acc[s] = std::fma(d, d, acc[s]);
            } else {
                acc[s] = std::fma(a[i + s], b[i + s], acc[s]);
            }
        }
    }
    for (; dim - i >= 8; i += 8) {
        for (int l = 0; l < 8; ++l) {
            if constexpr (kL2) {
                const float d = a[i + l] - b[i + l];
                acc[l] = std::fma(d, d, acc[l]);
            } else {
                acc[l] = std::fma(a[i + l], b[i + l], acc[l]);
            }
        }
    }
    float v[8];
    for (int l = 0; l < 8; ++l) {
        v[l] = (acc[l] + acc[8 + l]) + (acc[16 + l] + acc[24 + l]);
    }
    float h[4];
    for (int l = 0; l < 4; ++l) {
        h[l] = v[l] + v[l + 4];
    }
    float sum = (h[0] + h[2]) + (h[1] + h[3]);
    for (; i < dim; ++i) {
        if constexpr (kL2) {
            const float d = a[i] - b[i];
            const float sq = d * d;
            sum = sum + sq;
        } else {
            const float p = a[i] * b[i];
            sum = sum + p;
        }
    }
    return sum;
}

// ---- AVX2 + FMA -------------------------------------------------------------------------------
// Only these functions carry the avx2/fma target; they run only after the CPU check below.

[[gnu::target("avx2,fma")]] float reduce8(__m256 acc0, __m256 acc1, __m256 acc2,
                                          __m256 acc3) noexcept {
    const __m256 v = _mm256_add_ps(_mm256_add_ps(acc0, acc1), _mm256_add_ps(acc2, acc3));
    const __m128 h = _mm_add_ps(_mm256_castps256_ps128(v), _mm256_extractf128_ps(v, 1));
    // p0 = h0 + h2, p1 = h1 + h3, sum = p0 + p1.
    const __m128 p = _mm_add_ps(h, _mm_movehl_ps(h, h));
    return _mm_cvtss_f32(_mm_add_ss(p, _mm_shuffle_ps(p, p, 0x55)));
}

[[gnu::target("avx2,fma")]] float l2_avx2(const float* a, const float* b,
                                          std::size_t dim) noexcept {
    __m256 acc[4] = {_mm256_setzero_ps(), _mm256_setzero_ps(), _mm256_setzero_ps(),
                     _mm256_setzero_ps()};
    std::size_t i = 0;
    for (; dim - i >= 32; i += 32) {
        for (int j = 0; j < 4; ++j) {
            const __m256 d =
                _mm256_sub_ps(_mm256_loadu_ps(a + i + 8 * j), _mm256_loadu_ps(b + i + 8 * j));
            acc[j] = _mm256_fmadd_ps(d, d, acc[j]);
        }
    }
    for (; dim - i >= 8; i += 8) {
        const __m256 d = _mm256_sub_ps(_mm256_loadu_ps(a + i), _mm256_loadu_ps(b + i));
        acc[0] = _mm256_fmadd_ps(d, d, acc[0]);
    }
    float sum = reduce8(acc[0], acc[1], acc[2], acc[3]);
    for (; i < dim; ++i) {
        const float d = a[i] - b[i];
        const float sq = d * d;
        sum = sum + sq;
    }
    return sum;
}

[[gnu::target("avx2,fma")]] float ip_avx2(const float* a, const float* b,
                                          std::size_t dim) noexcept {
    __m256 acc[4] = {_mm256_setzero_ps(), _mm256_setzero_ps(), _mm256_setzero_ps(),
                     _mm256_setzero_ps()};
    std::size_t i = 0;
    for (; dim - i >= 32; i += 32) {
        for (int j = 0; j < 4; ++j) {
            acc[j] = _mm256_fmadd_ps(_mm256_loadu_ps(a + i + 8 * j), _mm256_loadu_ps(b + i + 8 * j),
                                     acc[j]);
        }
    }
    for (; dim - i >= 8; i += 8) {
        acc[0] = _mm256_fmadd_ps(_mm256_loadu_ps(a + i), _mm256_loadu_ps(b + i), acc[0]);
    }
    float sum = reduce8(acc[0], acc[1], acc[2], acc[3]);
    for (; i < dim; ++i) {
        const float p = a[i] * b[i];
        sum = sum + p;
    }
    return sum;
}

bool detect_avx2() noexcept {
    __builtin_cpu_init();
    return __builtin_cpu_supports("avx2") && __builtin_cpu_supports("fma");
}

struct Active {
    DistanceFn l2;
    DistanceFn ip;
    bool avx2;
};

const Active& active() noexcept {
    static const Active set = detect_avx2()
                                  ? Active{&l2_avx2, &ip_avx2, true}
                                  : Active{&emulate<true>, &emulate<false>, false};
    return set;
}

}  // namespace

float l2_squared(const float* a, const float* b, std::size_t dim) noexcept {
    return active().l2(a, b, dim);
}

float inner_product(const float* a, const float* b, std::size_t dim) noexcept {
    return active().ip(a, b, dim);
}

float metric_distance(unsigned metric, const float* a, const float* b, std::size_t dim) noexcept {
    return metric == 0 ? l2_squared(a, b, dim) : 1.0F - inner_product(a, b, dim);
}

float l2_squared_emulated(const float* a, const float* b, std::size_t dim) noexcept {
    return emulate<true>(a, b, dim);
}

float inner_product_emulated(const float* a, const float* b, std::size_t dim) noexcept {
    return emulate<false>(a, b, dim);
}

bool using_avx2() noexcept { return active().avx2; }

}  // namespace vdx::cpu
