#pragma once

// Exact-order distances from one query to a batch of gathered rows, 8 lanes per row.

#include "device_common.hpp"

namespace vdx::dev {

// Where the rows live and how to compare: l2 != 0 for squared L2, else 1 - <a, b>.
struct RowSource {
    const float* vectors;
    std::uint32_t ld;
    std::uint32_t l2;
};

// Distances of cnt <= 8 * R candidates (ids in LDS) to the query, into LDS `out`. Group g
// (lanes 8g..8g+7) computes candidate 8r + g of round r; all rounds' loads are issued before
// any arithmetic, so up to 8R rows are in flight together. NB/N8/TAIL are the maximum block
// counts this instantiation holds in registers; the runtime shape `s` may use fewer.
template <int NB, int N8, int TAIL, int R>
__device__ __forceinline__ void distances(const RowSource& a, const std::uint32_t* ids,
                                          std::uint32_t cnt, const float4 (&qb)[NB > 0 ? NB : 1],
                                          const float4 (&q8)[N8 > 0 ? N8 : 1], const float (&qt)[8],
                                          int nb, int n8, int tail, int first_tail, float* out) {
    const int lane = lane_id();
    const int g = lane >> 3;
    const int m = lane & 7;
    const std::uint32_t rounds = (cnt + 7) / 8;
    float4 x[R][NB > 0 ? NB : 1];
    float4 x8[R][N8 > 0 ? N8 : 1];
    float xt[R][8];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const std::uint32_t c = r * 8 + g;
        const bool valid = static_cast<std::uint32_t>(r) < rounds && c < cnt;
        const float* row = a.vectors + static_cast<std::size_t>(valid ? ids[c] : 0) * a.ld;
#pragma unroll
        for (int i = 0; i < NB; ++i) {
            x[r][i] = make_float4(0.f, 0.f, 0.f, 0.f);
            if (valid && i < nb) {
                x[r][i] = *reinterpret_cast<const float4*>(row + 32 * i + 4 * m);
            }
        }
#pragma unroll
        for (int t = 0; t < N8; ++t) {
            x8[r][t] = make_float4(0.f, 0.f, 0.f, 0.f);
            if (valid && t < n8 && m < 2) {
                x8[r][t] = *reinterpret_cast<const float4*>(row + 32 * nb + 8 * t + 4 * m);
            }
        }
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            xt[r][e] = 0.f;
            if (e < TAIL && valid && e < tail && m == 0) {
                xt[r][e] = row[first_tail + e];
            }
        }
    }
    const bool l2 = a.l2 != 0;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        if (static_cast<std::uint32_t>(r) < rounds) {
            float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
            for (int i = 0; i < NB; ++i) {
                if (i < nb) {
                    if (l2) {
                        const float4 d = f4_sub(qb[i], x[r][i]);
                        acc = f4_fma(d, d, acc);
                    } else {
                        acc = f4_fma(qb[i], x[r][i], acc);
                    }
                }
            }
            // 8-float blocks add into accumulator 0 (slots 0..7, lanes 0 and 1); the other
            // lanes hold zeros there and fma(0, 0, acc) leaves acc unchanged.
#pragma unroll
            for (int t = 0; t < N8; ++t) {
                if (t < n8) {
                    if (l2) {
                        const float4 d = f4_sub(q8[t], x8[r][t]);
                        acc = f4_fma(d, d, acc);
                    } else {
                        acc = f4_fma(q8[t], x8[r][t], acc);
                    }
                }
            }
            const float d = finish_distance(l2, acc, qt, xt[r], TAIL > 0 ? tail : 0);
            const std::uint32_t c = r * 8 + g;
            if (m == 0 && c < cnt) {
                out[c] = d;
            }
        }
    }
}

}  // namespace vdx::dev
