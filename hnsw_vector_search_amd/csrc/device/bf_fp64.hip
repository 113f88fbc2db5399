// The float64 pass of the brute force: float64 distances of a window of candidates per query,
// the re-rank of candidates whose float32 distances tie within one ulp, and the largest error
// of the returned float32 distances against float64.
//
// The float64 distance is the reference's independent yardstick (vecdex
// tests/brute_force_test.cpp:189-213): sum (x - y)^2 for L2, 1 - sum x y for the inner
// product, 1 - dot / sqrt(|x|^2 |y|^2) for cosine, over the stored float32 components widened
// to double (the order of a double sum of 128 products moves it by ~1e-16 relative, far below
// any float32 gap).
//
// Ties: a window holds a query's first m results in `closer` order (float32 distance, then id).
// Walking it, a candidate joins the previous one's group when its float32 distance is at most
// one ulp above the previous candidate's (so a group is a chain of one-ulp steps); each group
// is re-ordered by (float64 distance, id) and the first `keep` candidates are returned with
// their own float32 distances. A group that reaches the end of a window shorter than the base
// may continue past it: the query is flagged so the host widens the window.

#include <hip/hip_runtime.h>

#include <cstdint>

#include "bf_fp64.hpp"
#include "device_common.hpp"

namespace vdx::dev {

namespace {

constexpr int kFp64Waves = 4;

__device__ __forceinline__ double dshfl_xor(double v, int mask) { return __shfl_xor(v, mask, 8); }

// The next float above a finite x (one ulp up; from either zero, the smallest subnormal).
__device__ __forceinline__ float next_up(float x) {
    if (x == 0.0f) return __uint_as_float(1u);
    const std::uint32_t u = __float_as_uint(x);
    return __uint_as_float((u >> 31) ? u - 1u : u + 1u);
}

}  // namespace

// One wave per query; 8 lanes per candidate, lane j summing components j, j + 8, ...
__global__ __launch_bounds__(256) void bf_fp64_distances_kernel(BfFp64Args a) {
    const int lane = lane_id();
    const std::uint32_t q = blockIdx.x * kFp64Waves + (threadIdx.x >> 6);
    if (q >= a.nq) return;  // the whole wave (no barrier below)
    const std::uint32_t cnt = min(a.cand_count[q], a.m);
    const float* qp = a.queries + static_cast<std::size_t>(q) * a.qld;
    const Neighbor* cand = a.cand + static_cast<std::size_t>(q) * a.m;
    double* out = a.d64 + static_cast<std::size_t>(q) * a.m;
    const int g = lane >> 3, j = lane & 7;
    for (std::uint32_t c0 = 0; c0 < cnt; c0 += 8) {
        const std::uint32_t c = c0 + static_cast<std::uint32_t>(g);
        double l2 = 0.0, dot = 0.0, xx = 0.0, yy = 0.0;
        if (c < cnt) {
            const float* row = a.vectors + static_cast<std::size_t>(cand[c].id) * a.ld;
            for (std::uint32_t t = static_cast<std::uint32_t>(j); t < a.dim; t += 8) {
                const double x = qp[t], y = row[t];
                l2 += (x - y) * (x - y);
                dot += x * y;
                xx += x * x;
                yy += y * y;
            }
        }
#pragma unroll
        for (int off = 1; off < 8; off <<= 1) {
            l2 += dshfl_xor(l2, off);
            dot += dshfl_xor(dot, off);
            xx += dshfl_xor(xx, off);
            yy += dshfl_xor(yy, off);
        }
        if (c < cnt && j == 0) {
            out[c] = a.metric == 0 ? l2 : (a.metric == 1 ? 1.0 - dot : 1.0 - dot / sqrt(xx * yy));
        }
    }
}

// One wave per query, one wave per workgroup: the window is staged in LDS, lane 0 walks it and
// re-orders each one-ulp group by insertion (groups are a few candidates), then all lanes write
// the first `keep` and measure their float32 error against float64.
__global__ __launch_bounds__(64) void bf_fp64_ties_kernel(BfFp64Args a) {
    extern __shared__ __align__(16) unsigned char lds[];
    Neighbor* wn = reinterpret_cast<Neighbor*>(lds);
    double* wd = reinterpret_cast<double*>(lds + 8u * a.m);
    const int lane = lane_id();
    const std::uint32_t q = blockIdx.x;
    if (q >= a.nq) return;
    const std::uint32_t cnt = min(a.cand_count[q], a.m);
    const Neighbor* cand = a.cand + static_cast<std::size_t>(q) * a.m;
    const double* d64 = a.d64 + static_cast<std::size_t>(q) * a.m;
    for (std::uint32_t i = static_cast<std::uint32_t>(lane); i < cnt; i += kWave) {
        wn[i] = cand[i];
        wd[i] = d64[i];
    }
    wave_lds_fence();
    const std::uint32_t keep = min(a.keep, cnt);
    if (a.ties != 0 && lane == 0) {
        unsigned long long tied = 0, moved = 0, overflow = 0;
        std::uint32_t i = 0;
        while (i < keep) {
            std::uint32_t e = i + 1;
            while (e < cnt && wn[e].distance <= next_up(wn[e - 1].distance)) {
                ++e;
            }
            if (e - i >= 2) {
                tied += e - i;
                if (e == cnt && cnt < a.n) {
                    overflow = 1;  // the group may go on past the window
                }
                for (std::uint32_t s = i + 1; s < e; ++s) {  // insertion by (float64, id)
                    const Neighbor vn = wn[s];
                    const double vd = wd[s];
                    std::uint32_t t = s;
                    while (t > i && (wd[t - 1] > vd || (wd[t - 1] == vd && wn[t - 1].id > vn.id))) {
                        wn[t] = wn[t - 1];
                        wd[t] = wd[t - 1];
                        --t;
                    }
                    if (t != s && t < keep) moved = 1;  // the first keep changed
                    wn[t] = vn;
                    wd[t] = vd;
                }
            }
            i = e;
        }
        if (a.stats != nullptr) {
            if (overflow) atomicAdd(a.stats + 0, 1ull);
            if (tied) atomicAdd(a.stats + 1, tied);
            if (moved) atomicAdd(a.stats + 2, 1ull);
        }
    }
    wave_lds_fence();
    double err = 0.0, rel = 0.0;
    Neighbor* out = a.out + static_cast<std::size_t>(q) * a.out_stride;
    for (std::uint32_t i = static_cast<std::uint32_t>(lane); i < keep; i += kWave) {
        if (a.out != nullptr) out[i] = wn[i];
        const double e = fabs(static_cast<double>(wn[i].distance) - wd[i]);
        err = fmax(err, e);
        rel = fmax(rel, e / (1.0 + fabs(wd[i])));  // the reference's tolerance form, 1e-5 (1 + |d|)
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        err = fmax(err, __shfl_xor(err, off, kWave));
        rel = fmax(rel, __shfl_xor(rel, off, kWave));
    }
    if (lane == 0) {
        if (a.out_count != nullptr) a.out_count[q] = keep;
        if (a.stats != nullptr) {
            // non-negative doubles order as their bit patterns
            atomicMax(a.stats + 3, static_cast<unsigned long long>(__double_as_longlong(err)));
            atomicMax(a.stats + 4, static_cast<unsigned long long>(__double_as_longlong(rel)));
        }
    }
}

hipError_t launch_bf_fp64(const BfFp64Args& a, hipStream_t stream) {
    if (a.nq == 0) return hipSuccess;
    hipLaunchKernelGGL(bf_fp64_distances_kernel, dim3((a.nq + kFp64Waves - 1) / kFp64Waves), dim3(64 * kFp64Waves),
                       0, stream, a);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(bf_fp64_ties_kernel, dim3(a.nq), dim3(64), bf_fp64_lds_bytes(a.m), stream, a);
    return hipGetLastError();
}

}  // namespace vdx::dev
