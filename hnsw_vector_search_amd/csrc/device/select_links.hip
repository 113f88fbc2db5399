// The neighbour choice of HnswIndex::insert on the GPU (the heuristic of select_neighbors,
// Malkov & Yashunin's Algorithm 4): for each inserted node, walk its layer-0 search results from
// closest to farthest and keep one only if it is closer to the node than to every kept one,
// until M are kept; with at most M results all are kept. One wave per node; a candidate's
// distances to the kept ones are computed together, 8 lanes each, in the exact summation order
// of the host's distance, so every comparison — and the choice — equals the host's.

#include "select_links.hpp"

#include <algorithm>

#include "gather_distances.hpp"

namespace vdx::dev {

namespace {

template <int NB, int N8, int TAIL, bool FIXED>
__global__ __launch_bounds__(256) void select_links_kernel(SelectArgs a) {
    __shared__ std::uint32_t kid_all[4][64];
    __shared__ float dk_all[4][64];
    const int lane = lane_id();
    const int wave = threadIdx.x >> 6;
    const int m8 = lane & 7;
    std::uint32_t* kid = kid_all[wave];
    float* dk = dk_all[wave];
    const int nb = FIXED ? NB : static_cast<int>(a.nb);
    const int n8 = FIXED ? N8 : static_cast<int>(a.n8);
    const int tail = FIXED ? TAIL : static_cast<int>(a.tail);
    const int first_tail = static_cast<int>(a.first_tail);
    const RowSource src{a.vectors, a.ld, a.l2};
    const std::uint64_t waves = static_cast<std::uint64_t>(gridDim.x) * (blockDim.x >> 6);
    for (std::uint64_t i = static_cast<std::uint64_t>(blockIdx.x) * (blockDim.x >> 6) + wave; i < a.count;
         i += waves) {
        const Neighbor* row = a.found + i * a.found_stride;
        Neighbor* out = a.chosen + i * a.m;
        const std::uint32_t cnt = __builtin_amdgcn_readfirstlane(a.found_count[i]);
        if (cnt <= a.m) {
            if (static_cast<std::uint32_t>(lane) < cnt) {
                out[lane] = row[lane];
            }
            if (lane == 0) {
                a.chosen_count[i] = cnt;
            }
            continue;
        }
        std::uint32_t nk = 0;
        for (std::uint32_t j = 0; j < cnt && nk < a.m; ++j) {
            const std::uint32_t cid = __builtin_amdgcn_readfirstlane(row[j].id);
            const float cdist = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(row[j].distance)));
            bool diverse = true;
            if (nk > 0) {
                // the candidate's vector as the "query", spread over the 8 lanes of each group
                const float* qp = a.vectors + static_cast<std::size_t>(cid) * a.ld;
                float4 qb[NB > 0 ? NB : 1];
                float4 q8[N8 > 0 ? N8 : 1];
                float qt[8];
#pragma unroll
                for (int b = 0; b < NB; ++b) {
                    qb[b] = make_float4(0.f, 0.f, 0.f, 0.f);
                    if (FIXED || b < nb) {
                        const float* p = qp + 32 * b + 4 * m8;
                        qb[b] = make_float4(p[0], p[1], p[2], p[3]);
                    }
                }
#pragma unroll
                for (int s = 0; s < N8; ++s) {
                    q8[s] = make_float4(0.f, 0.f, 0.f, 0.f);
                    if ((FIXED || s < n8) && m8 < 2) {
                        const float* p = qp + 32 * nb + 8 * s + 4 * m8;
                        q8[s] = make_float4(p[0], p[1], p[2], p[3]);
                    }
                }
#pragma unroll
                for (int e = 0; e < 8; ++e) {
                    qt[e] = (e < TAIL && (FIXED || e < tail)) ? qp[first_tail + e] : 0.f;
                }
                wave_lds_fence();
                for (std::uint32_t k0 = 0; k0 < nk; k0 += 32) {
                    distances_rows<NB, N8, TAIL, 4, FIXED>(src, kid + k0, min(32u, nk - k0), qb, q8, qt, nb, n8, tail,
                                                            first_tail, dk + k0);
                }
                wave_lds_fence();
                // closer to some kept neighbour than to the node: not diverse (strict '<')
                const bool near = static_cast<std::uint32_t>(lane) < nk && dk[lane] < cdist;
                diverse = ballot(near) == 0;
            }
            if (diverse) {
                if (lane == 0) {
                    kid[nk] = cid;
                    out[nk] = Neighbor{cdist, cid};
                }
                ++nk;
            }
            wave_lds_fence();
        }
        if (lane == 0) {
            a.chosen_count[i] = nk;
        }
    }
}

}  // namespace

hipError_t launch_select_links(const SelectArgs& a, BeamShape shape, hipStream_t stream) {
    if (a.count == 0) return hipSuccess;
    const std::uint64_t blocks = std::min<std::uint64_t>((a.count + 3) / 4, 65536);
    const dim3 grid(static_cast<unsigned>(blocks)), block(256);
    switch (shape) {
        case BeamShape::D128:
            hipLaunchKernelGGL((select_links_kernel<4, 0, 0, true>), grid, block, 0, stream, a);
            break;
        case BeamShape::D100:
            hipLaunchKernelGGL((select_links_kernel<3, 0, 4, true>), grid, block, 0, stream, a);
            break;
        default:
            hipLaunchKernelGGL((select_links_kernel<8, 3, 7, false>), grid, block, 0, stream, a);
            break;
    }
    return hipGetLastError();
}

}  // namespace vdx::dev
