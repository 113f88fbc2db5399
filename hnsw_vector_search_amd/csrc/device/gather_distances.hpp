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

namespace vdx::dev {

// Cross-lane moves inside a row of 16 lanes (DPP): the value of lane i ^ 1, lane i ^ 2 and
// lane i + 4 (lanes 12..15 of a row read 0; no caller uses them).
__device__ __forceinline__ float dpp_xor1(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0xB1, 0xF, 0xF, true));
}
__device__ __forceinline__ float dpp_xor2(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x4E, 0xF, 0xF, true));
}
__device__ __forceinline__ float dpp_next4(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x104, 0xF, 0xF, true));
}

// finish_distance with the partner values taken by DPP instead of LDS permutes. The sums are
// the same additions in the same association, so the result is bit-identical; only lanes
// m = 0 and 1 of each group hold meaningful h, and lane m = 0 the distance.
__device__ __forceinline__ float finish_distance_dpp(bool l2, float4 acc, const float (&qt)[8],
                                                     const float (&xt)[8], int tail) {
    // lanes m and m ^ 2: (acc0 + acc1) on m = 0, 1 and (acc2 + acc3) on m = 4, 5
    const float4 t = make_float4(acc.x + dpp_xor2(acc.x), acc.y + dpp_xor2(acc.y),
                                 acc.z + dpp_xor2(acc.z), acc.w + dpp_xor2(acc.w));
    // lanes m and m + 4: v[l] = (acc0 + acc1) + (acc2 + acc3) on m = 0 (l < 4) and 1 (l >= 4)
    const float4 v = make_float4(t.x + dpp_next4(t.x), t.y + dpp_next4(t.y), t.z + dpp_next4(t.z),
                                 t.w + dpp_next4(t.w));
    // lanes m and m ^ 1: h[i] = v[i] + v[i + 4] on m = 0
    const float4 h = make_float4(v.x + dpp_xor1(v.x), v.y + dpp_xor1(v.y), v.z + dpp_xor1(v.z),
                                 v.w + dpp_xor1(v.w));
    float sum = (h.x + h.z) + (h.y + h.w);
#pragma unroll
    for (int e = 0; e < 7; ++e) {
        if (e < tail) {
            if (l2) {
                const float d = qt[e] - xt[e];
                const float sq = d * d;
                sum = sum + sq;
            } else {
                const float p = qt[e] * xt[e];
                sum = sum + p;
            }
        }
    }
    return l2 ? sum : 1.0f - sum;
}

// distances() for the register-resident beam: the candidates' ids are read from LDS first,
// then every row's loads are issued under one predicate per row, then the arithmetic runs with
// DPP reductions. FIXED: the runtime shape equals NB/N8/TAIL (dimensions 128 and 100), so
// every block test folds away. MET: 1 squared L2, 0 the inner-product form, -1 a.l2 at run time.
template <int NB, int N8, int TAIL, int R, bool FIXED, int MET = -1>
__device__ __forceinline__ void distances_rows(const RowSource& a, const std::uint32_t* ids,
                                               std::uint32_t cnt, const float4 (&qb)[NB > 0 ? NB : 1],
                                               const float4 (&q8)[N8 > 0 ? N8 : 1],
                                               const float (&qt)[8], int nb, int n8, int tail,
                                               int first_tail, float* out) {
    const int lane = lane_id();
    const int g = lane >> 3;
    const int m = lane & 7;
    if (FIXED) {
        nb = NB;
        n8 = N8;
        tail = TAIL;
    }
    std::uint32_t rid[R];
    bool ok[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const std::uint32_t c = static_cast<std::uint32_t>(r * 8 + g);
        ok[r] = c < cnt;
        rid[r] = ok[r] ? ids[c] : 0u;
    }
    float4 x[R][NB > 0 ? NB : 1];
    float4 x8[R][N8 > 0 ? N8 : 1];
    float xt[R][8];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const float* row = a.vectors + static_cast<std::size_t>(rid[r]) * a.ld;
#pragma unroll
        for (int i = 0; i < NB; ++i) {
            x[r][i] = make_float4(0.f, 0.f, 0.f, 0.f);
        }
#pragma unroll
        for (int t = 0; t < N8; ++t) {
            x8[r][t] = make_float4(0.f, 0.f, 0.f, 0.f);
        }
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            xt[r][e] = 0.f;
        }
        if (ok[r]) {
#pragma unroll
            for (int i = 0; i < NB; ++i) {
                if (FIXED || i < nb) {
                    x[r][i] = *reinterpret_cast<const float4*>(row + 32 * i + 4 * m);
                }
            }
#pragma unroll
            for (int t = 0; t < N8; ++t) {
                if ((FIXED || t < n8) && m < 2) {
                    x8[r][t] = *reinterpret_cast<const float4*>(row + 32 * nb + 8 * t + 4 * m);
                }
            }
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                if (e < TAIL && (FIXED || e < tail) && m == 0) {
                    xt[r][e] = row[first_tail + e];
                }
            }
        }
    }
    const bool l2 = MET >= 0 ? MET == 1 : a.l2 != 0;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        if (static_cast<std::uint32_t>(r * 8) < cnt) {
            float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
            for (int i = 0; i < NB; ++i) {
                if (FIXED || i < nb) {
                    if (l2) {
                        const float4 d = f4_sub(qb[i], x[r][i]);
                        acc = f4_fma(d, d, acc);
                    } else {
                        acc = f4_fma(qb[i], x[r][i], acc);
                    }
                }
            }
#pragma unroll
            for (int t = 0; t < N8; ++t) {
                if (FIXED || t < n8) {
                    if (l2) {
                        const float4 d = f4_sub(q8[t], x8[r][t]);
                        acc = f4_fma(d, d, acc);
                    } else {
                        acc = f4_fma(q8[t], x8[r][t], acc);
                    }
                }
            }
            const float d = finish_distance_dpp(l2, acc, qt, xt[r], TAIL > 0 ? tail : 0);
            if (m == 0 && ok[r]) {
                out[r * 8 + g] = d;
            }
        }
    }
}


// distances_rows for squared L2 at dimension 128 with rows held as bytes: lane m of a row's
// 8-lane group loads its 16 elements (32i + 4m + j, i and j in 0..3) as one 16-byte load from
// byte 16m of the row (pack_rows_u8's layout) and converts them to the floats the FP32 rows
// hold; the arithmetic is then distances_rows<4, 0, 0, R, true>'s for L2, so every distance is
// bit-identical. All R rounds' loads are issued unconditionally (a slot past cnt reloads the
// first candidate's row, a cache hit) so that no branch sits between them and the compiler's
// waits can release round r while later rounds are still in flight.
template <int R>
__device__ __forceinline__ void distances_rows_u8_l2(const unsigned char* rows, const std::uint32_t* ids,
                                                     std::uint32_t cnt, const float4 (&qb)[4], float* out) {
    const int lane = lane_id();
    const int g = lane >> 3;
    const int m = lane & 7;
    std::uint32_t rid[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const std::uint32_t c = static_cast<std::uint32_t>(r * 8 + g);
        rid[r] = ids[c < cnt ? c : 0u];
    }
    uint4 w[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        w[r] = *reinterpret_cast<const uint4*>(rows + static_cast<std::size_t>(rid[r]) * 128 + 16 * m);
    }
    // every row load is in flight before the first conversion (the scheduler would otherwise
    // start round 0's arithmetic, and its wait, ahead of the later rounds' loads)
    __builtin_amdgcn_sched_barrier(0);
    float xt[8];
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        xt[e] = 0.f;
    }
#pragma unroll
    for (int r = 0; r < R; ++r) {
        if (static_cast<std::uint32_t>(r * 8) < cnt) {
            const std::uint32_t wi[4] = {w[r].x, w[r].y, w[r].z, w[r].w};
            float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                // byte j of the word converts exactly (v_cvt_f32_ubyte<j>)
                const float4 x = make_float4(static_cast<float>(wi[i] & 0xFFu), static_cast<float>((wi[i] >> 8) & 0xFFu),
                                             static_cast<float>((wi[i] >> 16) & 0xFFu), static_cast<float>(wi[i] >> 24));
                const float4 d = f4_sub(qb[i], x);
                acc = f4_fma(d, d, acc);
            }
            const float dist = finish_distance_dpp(true, acc, xt, xt, 0);
            const std::uint32_t c = static_cast<std::uint32_t>(r * 8 + g);
            if (m == 0 && c < cnt) {
                out[c] = dist;
            }
        }
    }
}

// Integer DPP partner moves within an 8-lane group (as dpp_xor1/dpp_xor2/dpp_next4).
__device__ __forceinline__ int dpp_xor1_i(int v) { return __builtin_amdgcn_update_dpp(0, v, 0xB1, 0xF, 0xF, true); }
__device__ __forceinline__ int dpp_xor2_i(int v) { return __builtin_amdgcn_update_dpp(0, v, 0x4E, 0xF, 0xF, true); }
__device__ __forceinline__ int dpp_next4_i(int v) { return __builtin_amdgcn_update_dpp(0, v, 0x104, 0xF, 0xF, true); }

// The sum of squares of a query held as byte words by the 8 lanes of a group (every lane).
__device__ __forceinline__ std::uint32_t bytes_sum_sq(const std::uint32_t (&qw)[4]) {
    std::uint32_t qq = 0;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        qq = __builtin_amdgcn_udot4(qw[i], qw[i], qq, false);
    }
    qq += static_cast<std::uint32_t>(__shfl_xor(static_cast<int>(qq), 1, 8));
    qq += static_cast<std::uint32_t>(__shfl_xor(static_cast<int>(qq), 2, 8));
    qq += static_cast<std::uint32_t>(__shfl_xor(static_cast<int>(qq), 4, 8));
    return qq;
}

// The query as byte words in the row copy's lane order (byte j of word i is element
// 32i + 4m + j of lane m's slice), when every component of it is an integer in [0, 255]:
// returns false (wave-uniform) otherwise. sum_sq: the query's sum of squares (every lane).
__device__ __forceinline__ bool query_as_bytes(const float4 (&qb)[4], std::uint32_t (&qw)[4], std::uint32_t& sum_sq) {
    bool ok = true;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const float v[4] = {qb[i].x, qb[i].y, qb[i].z, qb[i].w};
        std::uint32_t w = 0;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            ok = ok && v[j] >= 0.0f && v[j] <= 255.0f && v[j] == __builtin_floorf(v[j]);
            w |= static_cast<std::uint32_t>(__builtin_fminf(__builtin_fmaxf(v[j], 0.0f), 255.0f)) << (8 * j);
        }
        qw[i] = w;
    }
    sum_sq = bytes_sum_sq(qw);
    return __ballot(!ok) == 0;
}

// distances_rows_u8_l2 for a query whose components are all integers in [0, 255]
// (query_as_bytes): with both operands bytes, every partial sum of the reference's float32
// order is an integer of at most 128 * 255^2 = 8,323,200 < 2^24, so that order computes the
// exact integer sum of (q - x)^2 — which equals |q|^2 + |x|^2 - 2 q.x, computed here exactly in
// integers (v_dot4_u32_u8: four byte products per instruction) and converted once. Every
// distance is therefore bit-identical to the float path's, with an eighth of its arithmetic.
template <int R>
__device__ __forceinline__ void distances_rows_u8_l2_int(const unsigned char* rows, const std::uint32_t* ids,
                                                         std::uint32_t cnt, const std::uint32_t (&qw)[4],
                                                         std::uint32_t qq, float* out) {
    const int lane = lane_id();
    const int g = lane >> 3;
    const int m = lane & 7;
    std::uint32_t rid[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const std::uint32_t c = static_cast<std::uint32_t>(r * 8 + g);
        rid[r] = ids[c < cnt ? c : 0u];
    }
    uint4 w[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        w[r] = *reinterpret_cast<const uint4*>(rows + static_cast<std::size_t>(rid[r]) * 128 + 16 * m);
    }
    __builtin_amdgcn_sched_barrier(0);  // every row load in flight before the arithmetic
#pragma unroll
    for (int r = 0; r < R; ++r) {
        if (static_cast<std::uint32_t>(r * 8) < cnt) {
            const std::uint32_t wi[4] = {w[r].x, w[r].y, w[r].z, w[r].w};
            std::uint32_t qx = 0, xx = 0;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                qx = __builtin_amdgcn_udot4(qw[i], wi[i], qx, false);
                xx = __builtin_amdgcn_udot4(wi[i], wi[i], xx, false);
            }
            int p = static_cast<int>(xx) - 2 * static_cast<int>(qx);  // |p| < 2^25: no overflow
            p += dpp_xor1_i(p);
            p += dpp_xor2_i(p);
            p += dpp_next4_i(p);  // lane m = 0: the group's 8 lanes
            const std::uint32_t c = static_cast<std::uint32_t>(r * 8 + g);
            if (m == 0 && c < cnt) {
                out[c] = static_cast<float>(static_cast<int>(qq) + p);
            }
        }
    }
}

}  // namespace vdx::dev
