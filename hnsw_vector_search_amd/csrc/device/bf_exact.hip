// Exact brute force on the GPU: every distance in the library's exact summation order, top-k by
// (distance, id). It is BruteForceIndex::search (vecdex src/brute_force.cpp:20-37, 68-84) batched
// over queries, bit-equal to it. Two uses: the VDX_PREC_EXACT mode (and ground truth), and the
// exact re-rank of the candidates the MFMA shortlist keeps (rerank kernel below).
//
// scan: a workgroup takes 8 queries per wave (one per 8-lane group, query in registers) and one
// slab of the base; rows are staged through LDS 32 at a time and shared by all its queries.
// Each query keeps its best k keys sorted in LDS; a partial list per (slab, query) goes to HBM
// and merge_slabs picks the final k.

#include <hip/hip_runtime.h>

#include "bf_args.hpp"
#include "device_common.hpp"
#include "gather_distances.hpp"

namespace vdx::dev {

namespace {

constexpr int kRowsPerChunk = 32;

// Inserts `key` into the sorted list L[0, count) capped at cap; returns the new count.
__device__ __forceinline__ std::uint32_t insert_sorted(std::uint64_t* L, std::uint32_t count,
                                                       std::uint32_t cap, std::uint64_t key) {
    std::uint32_t i = count < cap ? count : cap - 1;
    while (i > 0 && L[i - 1] > key) {
        L[i] = L[i - 1];
        --i;
    }
    L[i] = key;
    return count < cap ? count + 1 : cap;
}

}  // namespace

// L2: squared L2 (a.l2 != 0), else 1 - <q, x>, fixed at compile time so the per-row arithmetic
// carries no metric branch; the group's partial sums are reduced with DPP moves (the same
// additions in the same association as the reference's AVX2 horizontal sum).
template <int NB, int N8, int TAIL, bool L2>
__global__ __launch_bounds__(256) void bf_exact_scan_kernel(BfScanArgs a) {
    extern __shared__ __align__(16) unsigned char lds[];
    const int lane = lane_id();
    const int wave = threadIdx.x >> 6;
    const int g = lane >> 3;
    const int m = lane & 7;
    const int waves = blockDim.x >> 6;
    const int qlocal = wave * 8 + g;
    const std::uint32_t q = blockIdx.x * (waves * 8) + qlocal;
    const bool active = q < a.nq;

    float* rows = reinterpret_cast<float*>(lds);
    std::uint64_t* top =
        reinterpret_cast<std::uint64_t*>(lds + sizeof(float) * kRowsPerChunk * a.ld) +
        static_cast<std::size_t>(qlocal) * a.kcap;

    const int nb = static_cast<int>(a.nb);
    const int n8 = static_cast<int>(a.n8);
    const int tail = static_cast<int>(a.tail);
    const int first_tail = static_cast<int>(a.first_tail);
    constexpr bool l2 = L2;

    float4 qb[NB > 0 ? NB : 1];
    float4 q8[N8 > 0 ? N8 : 1];
    float qt[8];
    const float* qp = a.queries + static_cast<std::size_t>(active ? q : 0) * a.qld;
#pragma unroll
    for (int i = 0; i < NB; ++i) {
        qb[i] = make_float4(0.f, 0.f, 0.f, 0.f);
        if (i < nb) {
            const float* p = qp + 32 * i + 4 * m;
            qb[i] = make_float4(p[0], p[1], p[2], p[3]);
        }
    }
#pragma unroll
    for (int t = 0; t < N8; ++t) {
        q8[t] = make_float4(0.f, 0.f, 0.f, 0.f);
        if (t < n8 && m < 2) {
            const float* p = qp + 32 * nb + 8 * t + 4 * m;
            q8[t] = make_float4(p[0], p[1], p[2], p[3]);
        }
    }
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        qt[e] = (e < TAIL && e < tail) ? qp[first_tail + e] : 0.f;
    }

    const std::uint64_t slab_rows = (static_cast<std::uint64_t>(a.n) + gridDim.y - 1) / gridDim.y;
    const std::uint32_t r0 = static_cast<std::uint32_t>(min<std::uint64_t>(a.n, slab_rows * blockIdx.y));
    const std::uint32_t r1 = static_cast<std::uint32_t>(min<std::uint64_t>(a.n, slab_rows * (blockIdx.y + 1)));
    const std::uint32_t cap = min<std::uint32_t>(a.k, a.n);
    std::uint32_t count = 0;
    std::uint64_t worst = ~0ull;  // lane m == 0 of each group tracks its list's worst key

    const int tid = threadIdx.x;
    const int nthreads = blockDim.x;
    const std::uint32_t row_f4 = a.ld / 4;
    for (std::uint32_t base = r0; base < r1; base += kRowsPerChunk) {
        const std::uint32_t nrows = min<std::uint32_t>(kRowsPerChunk, r1 - base);
        __syncthreads();
        const std::uint32_t total_f4 = nrows * row_f4;
        for (std::uint32_t i = tid; i < total_f4; i += nthreads) {
            reinterpret_cast<float4*>(rows)[i] =
                reinterpret_cast<const float4*>(a.vectors + static_cast<std::size_t>(base) * a.ld)[i];
        }
        __syncthreads();
        for (std::uint32_t r = 0; r < nrows; ++r) {
            const float* row = rows + r * a.ld;
            float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
            for (int i = 0; i < NB; ++i) {
                if (i < nb) {
                    const float4 x = *reinterpret_cast<const float4*>(row + 32 * i + 4 * m);
                    if (l2) {
                        const float4 d = f4_sub(qb[i], x);
                        acc = f4_fma(d, d, acc);
                    } else {
                        acc = f4_fma(qb[i], x, acc);
                    }
                }
            }
#pragma unroll
            for (int t = 0; t < N8; ++t) {
                if (t < n8) {
                    float4 x = make_float4(0.f, 0.f, 0.f, 0.f);
                    if (m < 2) {
                        x = *reinterpret_cast<const float4*>(row + 32 * nb + 8 * t + 4 * m);
                    }
                    if (l2) {
                        const float4 d = f4_sub(q8[t], x);
                        acc = f4_fma(d, d, acc);
                    } else {
                        acc = f4_fma(q8[t], x, acc);
                    }
                }
            }
            float xt[8];
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                xt[e] = (e < TAIL && e < tail) ? row[first_tail + e] : 0.f;
            }
            const float d = finish_distance_dpp(l2, acc, qt, xt, TAIL > 0 ? tail : 0);
            if (m == 0 && active) {
                const std::uint64_t key = make_key(d, base + r);
                if (count < cap || key < worst) {
                    count = insert_sorted(top, count, cap, key);
                    worst = top[count - 1];
                }
            }
        }
    }
    if (m == 0 && active) {
        std::uint64_t* dst = a.partial + (static_cast<std::size_t>(blockIdx.y) * a.nq + q) * a.kcap;
        for (std::uint32_t i = 0; i < count; ++i) {
            dst[i] = top[i];
        }
        a.partial_count[static_cast<std::size_t>(blockIdx.y) * a.nq + q] = count;
    }
}

// One thread per query: a P-way merge of the slabs' sorted lists.
__global__ void bf_merge_slabs_kernel(BfScanArgs a, std::uint32_t slabs) {
    const std::uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= a.nq) {
        return;
    }
    std::uint32_t pos[64];
    for (std::uint32_t p = 0; p < slabs; ++p) {
        pos[p] = 0;
    }
    const std::uint32_t want = min<std::uint32_t>(a.k, a.n);
    Neighbor* out = a.out + static_cast<std::size_t>(q) * a.out_stride;
    std::uint32_t written = 0;
    for (; written < want; ++written) {
        std::uint64_t best = ~0ull;
        std::uint32_t from = slabs;
        for (std::uint32_t p = 0; p < slabs; ++p) {
            const std::size_t row = static_cast<std::size_t>(p) * a.nq + q;
            if (pos[p] < a.partial_count[row]) {
                const std::uint64_t key = a.partial[row * a.kcap + pos[p]];
                if (key < best) {
                    best = key;
                    from = p;
                }
            }
        }
        if (from == slabs) {
            break;
        }
        ++pos[from];
        out[written] = Neighbor{key_distance(best), key_id(best)};
    }
    a.out_count[q] = written;
}

hipError_t launch_bf_exact_scan(const BfScanArgs& args, unsigned qgroups, unsigned slabs,
                                unsigned waves_per_group, std::size_t lds_bytes,
                                hipStream_t stream, int shape) {
    const dim3 grid(qgroups, slabs);
    const dim3 block(waves_per_group * kWave);
#define VDX_SCAN_LAUNCH(L2)                                                                       \
    switch (shape) {                                                                              \
        case 0:                                                                                   \
            hipLaunchKernelGGL((bf_exact_scan_kernel<4, 0, 0, L2>), grid, block, lds_bytes, stream, args); \
            break;                                                                                \
        case 1:                                                                                   \
            hipLaunchKernelGGL((bf_exact_scan_kernel<3, 0, 4, L2>), grid, block, lds_bytes, stream, args); \
            break;                                                                                \
        default:                                                                                  \
            hipLaunchKernelGGL((bf_exact_scan_kernel<8, 3, 7, L2>), grid, block, lds_bytes, stream, args); \
            break;                                                                                \
    }
    if (args.l2 != 0) {
        VDX_SCAN_LAUNCH(true)
    } else {
        VDX_SCAN_LAUNCH(false)
    }
#undef VDX_SCAN_LAUNCH
    hipError_t err = hipGetLastError();
    if (err != hipSuccess) {
        return err;
    }
    hipLaunchKernelGGL(bf_merge_slabs_kernel, dim3((args.nq + 127) / 128), dim3(128), 0, stream,
                       args, slabs);
    return hipGetLastError();
}

}  // namespace vdx::dev
