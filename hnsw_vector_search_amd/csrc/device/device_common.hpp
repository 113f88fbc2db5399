#pragma once

// Device helpers shared by the kernels: wave64 primitives, sortable result keys, and the
// exact-order distance of one candidate computed by a group of 8 lanes.

#include <hip/hip_runtime.h>

#include <cstdint>

namespace vdx::dev {

constexpr int kWave = 64;
constexpr std::uint32_t kExpanded = 0x80000000u;  // flag bit in a key's id half
constexpr std::uint64_t kKeyMask = ~static_cast<std::uint64_t>(kExpanded);

__device__ __forceinline__ int lane_id() { return __lane_id(); }

// Orders the lanes' LDS accesses: a write by one lane is visible to a later read by another.
// LDS requests of one wave complete in order; this keeps the compiler from moving them.
__device__ __forceinline__ void wave_lds_fence() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

__device__ __forceinline__ std::uint64_t ballot(bool p) { return __ballot(p); }

__device__ __forceinline__ std::uint32_t lanes_below(std::uint64_t mask) {
    return __builtin_amdgcn_mbcnt_hi(static_cast<std::uint32_t>(mask >> 32),
                                     __builtin_amdgcn_mbcnt_lo(static_cast<std::uint32_t>(mask), 0u));
}

__device__ __forceinline__ std::uint32_t broadcast(std::uint32_t v, int src) {
    return static_cast<std::uint32_t>(__shfl(static_cast<int>(v), src, kWave));
}

__device__ __forceinline__ float broadcast_f(float v, int src) { return __shfl(v, src, kWave); }

// float -> u32 preserving order (finite values and infinities; -0 is canonicalised first).
__device__ __forceinline__ std::uint32_t sortable(float d) {
    const std::uint32_t b = __float_as_uint(d + 0.0f);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}

__device__ __forceinline__ float unsortable(std::uint32_t s) {
    return __uint_as_float((s & 0x80000000u) ? (s & 0x7FFFFFFFu) : ~s);
}

// Key of a result: (distance, id) ordered as vecdex's `closer`, as one u64.
__device__ __forceinline__ std::uint64_t make_key(float d, std::uint32_t id) {
    return (static_cast<std::uint64_t>(sortable(d)) << 32) | id;
}

__device__ __forceinline__ float key_distance(std::uint64_t k) {
    return unsortable(static_cast<std::uint32_t>(k >> 32));
}

__device__ __forceinline__ std::uint32_t key_id(std::uint64_t k) {
    return static_cast<std::uint32_t>(k) & ~kExpanded;
}

__device__ __forceinline__ float4 f4_sub(float4 a, float4 b) {
    return make_float4(a.x - b.x, a.y - b.y, a.z - b.z, a.w - b.w);
}

__device__ __forceinline__ float4 f4_fma(float4 a, float4 b, float4 c) {
    return make_float4(__builtin_fmaf(a.x, b.x, c.x), __builtin_fmaf(a.y, b.y, c.y),
                       __builtin_fmaf(a.z, b.z, c.z), __builtin_fmaf(a.w, b.w, c.w));
}

__device__ __forceinline__ float4 f4_add(float4 a, float4 b) {
    return make_float4(a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w);
}

__device__ __forceinline__ float4 f4_shfl_xor(float4 v, int mask) {
    return make_float4(__shfl_xor(v.x, mask, 8), __shfl_xor(v.y, mask, 8),
                       __shfl_xor(v.z, mask, 8), __shfl_xor(v.w, mask, 8));
}

// The query as one 8-lane group holds it: lane m of the group keeps elements 32i + 4m .. +3
// of every full 32-block i (slots 4m..4m+3 of the 32 accumulator slots), the elements of the
// 8-float blocks that fall on slots 4m..4m+3 (lanes 0 and 1 only), and the tail.
template <int NB>
struct QueryRegs {
    float4 blk[NB];
    float4 b8[3];
    float tail[8];
};

struct DimShape {
    int n8;    // 8-float blocks after the full 32-blocks (0..3)
    int tail;  // elements after those (0..7)
    int first_tail;
};

template <int NB>
__device__ __forceinline__ void load_query(QueryRegs<NB>& q, const float* query, DimShape s,
                                           int m) {
#pragma unroll
    for (int i = 0; i < NB; ++i) {
        const float* p = query + 32 * i + 4 * m;
        q.blk[i] = make_float4(p[0], p[1], p[2], p[3]);
    }
#pragma unroll
    for (int t = 0; t < 3; ++t) {
        q.b8[t] = make_float4(0.f, 0.f, 0.f, 0.f);
        if (t < s.n8 && m < 2) {
            const float* p = query + 32 * NB + 8 * t + 4 * m;
            q.b8[t] = make_float4(p[0], p[1], p[2], p[3]);
        }
    }
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        q.tail[e] = e < s.tail ? query[s.first_tail + e] : 0.f;
    }
}

// Finishes one candidate's distance from the per-lane partial accumulators `acc` (slots
// 4m..4m+3) of its 8-lane group, in the reduction order of the reference's AVX2 kernels, then
// adds the tail elements (held by lane 0 in `xt`, the query's in `qt`) left to right.
// Lane m == 0 of the group holds the result.
__device__ __forceinline__ float finish_distance(bool l2, float4 acc, const float (&qt)[8],
                                                 const float (&xt)[8], int tail) {
    const float4 t = f4_add(acc, f4_shfl_xor(acc, 2));  // (acc0 + acc1), (acc2 + acc3)
    const float4 v = f4_add(t, f4_shfl_xor(t, 4));      // v[l] on lanes 0 (l<4) and 1 (l>=4)
    const float4 h = f4_add(v, f4_shfl_xor(v, 1));      // h[i] = v[i] + v[i+4]
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

}  // namespace vdx::dev
