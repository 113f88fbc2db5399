// K1: the HNSW query path on the GPU, one wavefront per query, persistent over a work queue.
//
// Replaces HnswIndex::search_unchecked (vecdex src/hnsw.cpp:504-522): greedy descent through
// the upper layers (greedy_closest, :416-431), then the layer-0 beam (search_layer, :438-482),
// then the closest min(k, |results|) in `closer` order. Results are bit-identical to vecdex's
// CPU search on the same graph:
//
//  * Distances follow vecdex's AVX2 summation order exactly (device_common.hpp), so every
//    distance, and hence every comparison, is the one the CPU makes.
//  * The two heaps of search_layer become one sorted list R of the best ef' = max(ef, k) keys
//    (distance, id) with an "expanded" bit. A candidate evicted from `results` can never be
//    expanded (when it reaches the front of `candidates` the stop test fires), so candidates
//    restricted to R lose nothing, the stop test never fires on a member of R, and the search
//    is: expand the closest unexpanded member of R until none is left.
//  * The fresh neighbours of one expansion are merged into R as a batch: admitting them one at
//    a time against a shrinking worst gives the same top-ef' set under a strict total order.
//  * The visited set is a lossy hash cache in LDS: it may forget a node (false negative), never
//    invent one. A forgotten node that is met again gets the same distance; if it is in R the
//    merge drops it as an exact duplicate key, and if it is not, R has been full since it was
//    last rejected or evicted and its worst has only improved, so it is rejected again. Results
//    are therefore exactly those of the exact visited list; forgetting only costs a recompute.
//
// Memory per wave (LDS): R (8 B x ef_cap), fresh ids/distances/keys (64 each), the visited cache.

#include <hip/hip_runtime.h>

#include "beam_args.hpp"
#include "device_common.hpp"
#include "gather_distances.hpp"
#include "mlp_stop.hpp"

namespace vdx::dev {

namespace {

constexpr std::uint32_t kNoNode = 0xFFFFFFFFu;

constexpr std::uint32_t kEmpty = 0xFFFFFFFFu;

__device__ __forceinline__ float wave_min(float v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        v = fminf(v, __shfl_xor(v, off, kWave));
    }
    return v;
}

__device__ __forceinline__ std::uint32_t wave_min_u32(std::uint32_t v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        v = min(v, static_cast<std::uint32_t>(__shfl_xor(static_cast<int>(v), off, kWave)));
    }
    return v;
}

__device__ __forceinline__ std::uint32_t uniform(std::uint32_t v) {
    return __builtin_amdgcn_readfirstlane(v);
}

// Merges the keys of cnt <= 64 fresh candidates (lane j holds candidate j's key in `f`, valid
// for j < cnt) into the sorted list R[0, size) bounded to efp entries. Updates size and hint
// (every entry below hint is expanded). F is LDS scratch for 64 keys.
__device__ __forceinline__ void merge(std::uint64_t* R, std::uint32_t& size, std::uint32_t efp,
                                      std::uint32_t& hint, std::uint64_t f, std::uint32_t cnt,
                                      std::uint64_t* F, unsigned long long& dups,
                                      std::uint32_t& inserted, std::uint32_t& first_pos) {
    inserted = 0;
    first_pos = 0xFFFFFFFFu;
    const std::uint32_t lane = static_cast<std::uint32_t>(lane_id());
    bool keep = lane < cnt;
    if (keep && size == efp) {
        keep = f < (R[size - 1] & kKeyMask);
    }
    std::uint32_t lb = 0;
    bool dup = false;
    if (keep) {
        std::uint32_t lo = 0, hi = size;
        while (lo < hi) {
            const std::uint32_t mid = (lo + hi) >> 1;
            if ((R[mid] & kKeyMask) < f) {
                lo = mid + 1;
            } else {
                hi = mid;
            }
        }
        lb = lo;
        if (lb < size && (R[lb] & kKeyMask) == f) {
            keep = false;  // already in R: a node the lossy visited cache forgot
            dup = true;
        }
    }
    dups += static_cast<unsigned long long>(__popcll(ballot(dup)));
    const std::uint64_t kept = ballot(keep);
    if (kept == 0) {
        return;
    }
    std::uint32_t rank = 0;
    for (std::uint32_t t = 0; t < cnt; ++t) {
        const std::uint64_t ft = __shfl(f, t, kWave);
        rank += (((kept >> t) & 1u) != 0 && ft < f) ? 1u : 0u;
    }
    const std::uint32_t pos = lb + rank;
    if (keep && pos >= efp) {
        keep = false;  // falls off the end; only ever the largest of the batch
    }
    const std::uint64_t ins = ballot(keep);
    const std::uint32_t nins = static_cast<std::uint32_t>(__popcll(ins));
    if (nins == 0) {
        return;
    }
    if (keep) {
        F[rank] = f;  // F[0, nins) sorted
    }
    const std::uint32_t lo_move = wave_min_u32(keep ? lb : 0xFFFFFFFFu);
    const std::uint32_t min_pos = wave_min_u32(keep ? pos : 0xFFFFFFFFu);
    inserted = nins;
    first_pos = min_pos;
    wave_lds_fence();
    // Shift the members of R at or after lo_move right by the number of inserted keys below
    // them, top chunk first, so no member is overwritten before it has been read.
    if (size > lo_move) {
        std::int32_t b = static_cast<std::int32_t>(((size - 1) / kWave) * kWave);
        for (; b >= static_cast<std::int32_t>(lo_move & ~(kWave - 1u)); b -= kWave) {
            const std::uint32_t i = static_cast<std::uint32_t>(b) + lane;
            if (i >= lo_move && i < size) {
                const std::uint64_t key = R[i];
                const std::uint64_t km = key & kKeyMask;
                std::uint32_t lo = 0, hi = nins;
                while (lo < hi) {
                    const std::uint32_t mid = (lo + hi) >> 1;
                    if (F[mid] < km) {
                        lo = mid + 1;
                    } else {
                        hi = mid;
                    }
                }
                const std::uint32_t ni = i + lo;
                if (ni < efp) {
                    R[ni] = key;
                }
            }
            wave_lds_fence();
        }
    }
    if (keep) {
        R[pos] = f;
    }
    wave_lds_fence();
    size = min(efp, size + nins);
    hint = min(hint, min_pos);
}

}  // namespace

template <int NB, int N8, int TAIL, int R, int MODE>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void hnsw_beam_kernel(BeamArgs a) {
    extern __shared__ __align__(16) unsigned char lds[];
    const int lane = lane_id();
    const int wave = threadIdx.x >> 6;
    const int m = lane & 7;
    unsigned char* base = lds + static_cast<std::size_t>(wave) * a.lds_per_wave;
    std::uint64_t* Rl = reinterpret_cast<std::uint64_t*>(base);
    std::uint64_t* F = reinterpret_cast<std::uint64_t*>(base + 8u * a.ef_cap);
    std::uint32_t* fid = reinterpret_cast<std::uint32_t*>(F + 64);
    float* fd = reinterpret_cast<float*>(fid + 64);
    float* feat = fd + 64;
    std::uint32_t* vis = reinterpret_cast<std::uint32_t*>(feat + 16);
    MlpWeights* model = nullptr;
    if constexpr (MODE == kBeamLearned) {
        // The stop model, staged once per workgroup after the waves' regions.
        model = reinterpret_cast<MlpWeights*>(lds + static_cast<std::size_t>(blockDim.x >> 6) * a.lds_per_wave);
        const std::uint32_t* src_words = static_cast<const std::uint32_t*>(a.mlp);
        std::uint32_t* dst_words = reinterpret_cast<std::uint32_t*>(model);
        for (std::uint32_t i = threadIdx.x; i < sizeof(MlpWeights) / 4; i += blockDim.x) {
            dst_words[i] = src_words[i];
        }
        __syncthreads();
    }
    const std::uint32_t vslots = 1u << a.visited_log2;
    const std::uint32_t vshift = 32u - a.visited_log2;
    std::uint32_t* bits = a.vbits == nullptr
                              ? nullptr
                              : a.vbits + static_cast<std::size_t>(blockIdx.x * (blockDim.x >> 6) + wave) * a.vbit_words;
    const int nb = static_cast<int>(a.nb);
    const int n8 = static_cast<int>(a.n8);
    const int tail = static_cast<int>(a.tail);
    const int first_tail = static_cast<int>(a.first_tail);
    const RowSource src{a.vectors, a.ld, a.l2};

    unsigned long long n_exp = 0, n_dist = 0, n_up = 0, n_dup = 0, n_stop = 0, n_pref = 0;

    std::uint32_t begin = 0, end = a.nq;
    if (a.range != nullptr) {
        begin = uniform(a.range[0]);
        end = uniform(a.range[1]);
    }
    for (;;) {
        std::uint32_t t = 0;
        if (lane == 0) {
            t = atomicAdd(a.queue, 1u);
        }
        t = uniform(broadcast(t, 0));
        if (t >= end - begin) {
            break;
        }
        const std::uint32_t q = a.order != nullptr ? uniform(a.order[begin + t]) : t;
        // The host sized R and the output rows; clamping keeps a caller's larger per-query
        // value from reaching past them.
        const std::uint32_t k = min(uniform(a.ks ? a.ks[q] : a.k), a.k_cap);
        const std::uint32_t ef = uniform(a.efs ? a.efs[q] : a.ef);
        const std::uint32_t efp = min(max(ef, k), a.ef_cap);
        if (max(ef, k) > a.ef_cap && lane == 0 && a.stats != nullptr) {
            atomicAdd(a.stats + 6, 1ull);  // runs at the launch's R capacity
        }
        Neighbor* out = a.out + static_cast<std::size_t>(q) * a.out_stride;
        if (a.n == 0) {
            if (lane == 0) {
                a.out_count[q] = 0;
            }
            continue;
        }

        // The query, spread over the 8 lanes of each group (every group holds a full copy).
        const float* qp = a.queries + static_cast<std::size_t>(q) * a.qld;
        float4 qb[NB > 0 ? NB : 1];
        float4 q8[N8 > 0 ? N8 : 1];
        float qt[8];
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

        // Entry point and greedy descent (src/hnsw.cpp:513-516, 416-431), or a given layer-0
        // entry (src/hnsw.cpp:324-335).
        std::uint32_t cur = a.entry_point;
        std::uint32_t top = a.max_level;
        if (a.entries != nullptr) {
            const std::uint32_t e = uniform(a.entries[q]);
            if (e < a.n) {
                cur = e;
                top = 0;
            }
        }
        if (lane == 0) {
            fid[0] = cur;
        }
        wave_lds_fence();
        distances<NB, N8, TAIL, R>(src, fid, 1, qb, q8, qt, nb, n8, tail, first_tail, fd);
        wave_lds_fence();
        float cur_d = fd[0];
        ++n_dist;
        for (std::uint32_t layer = top; layer > a.layer; --layer) {
            for (;;) {
                const std::uint32_t* blk =
                    a.upper + (static_cast<std::size_t>(a.upper_offset[cur]) + (layer - 1)) * a.bu;
                const std::uint32_t cnt = uniform(blk[0]);
                std::uint32_t id = 0;
                if (static_cast<std::uint32_t>(lane) < cnt) {
                    id = blk[1 + lane];
                    fid[lane] = id;
                }
                ++n_up;
                n_dist += cnt;
                wave_lds_fence();
                for (std::uint32_t c0 = 0; c0 < cnt; c0 += 8 * R) {
                    distances<NB, N8, TAIL, R>(src, fid + c0, min(8u * R, cnt - c0), qb, q8, qt, nb,
                                               n8, tail, first_tail, fd + c0);
                }
                wave_lds_fence();
                const float d = static_cast<std::uint32_t>(lane) < cnt ? fd[lane] : __int_as_float(0x7f800000);
                const float best = wave_min(d);
                if (!(best < cur_d)) {
                    break;
                }
                // The first neighbour in link order at the minimum: the one the sequential
                // scan with a strict '<' settles on.
                const std::uint64_t at = ballot(static_cast<std::uint32_t>(lane) < cnt && d == best);
                const int first = __ffsll(static_cast<long long>(at)) - 1;
                cur = uniform(broadcast(id, first));
                cur_d = best;
            }
        }

        // Layer 0 (src/hnsw.cpp:438-482) with ef' = max(ef, k) (src/hnsw.cpp:517).
        if (bits != nullptr) {
            // the exact visited set (vecdex's VisitedList): this wave's bitmap row, cleared, the
            // stores complete before the test-and-set atomics of the search touch it
            for (std::uint32_t i = static_cast<std::uint32_t>(lane) * 4; i < a.vbit_words; i += kWave * 4) {
                *reinterpret_cast<uint4*>(bits + i) = make_uint4(0u, 0u, 0u, 0u);
            }
            __threadfence();
            if (lane == 0) {
                atomicOr(bits + (cur >> 5), 1u << (cur & 31u));
            }
        } else {
            for (std::uint32_t i = static_cast<std::uint32_t>(lane) * 4; i < vslots; i += kWave * 4) {
                *reinterpret_cast<uint4*>(vis + i) = make_uint4(kEmpty, kEmpty, kEmpty, kEmpty);
            }
        }
        wave_lds_fence();
        if (lane == 0) {
            if (bits == nullptr) {
                vis[(cur * 2654435761u) >> vshift] = cur;
            }
            Rl[0] = make_key(cur_d, cur);
        }
        wave_lds_fence();
        std::uint32_t size = 1;
        std::uint32_t hint = 0;
        std::uint32_t pref_node = kNoNode, pref_cnt = 0, pref_id = 0;
        const std::uint32_t lcap = a.layer == 0 ? a.M2 : (a.M2 >> 1);
        const auto block_of = [&](std::uint32_t nd) -> const std::uint32_t* {
            return a.layer == 0
                       ? a.layer0 + static_cast<std::size_t>(nd) * a.b0
                       : a.upper + (static_cast<std::size_t>(a.upper_offset[nd]) + (a.layer - 1)) * a.bu;
        };
        // search-state features (trace and learned modes)
        const float inv_scale = 1.0f / fmaxf(fabsf(cur_d), 1e-6f);
        const float inv_efp = 1.0f / static_cast<float>(efp);
        const float inv_m2 = 1.0f / static_cast<float>(a.M2);
        std::uint32_t steps = 0, improved_hist = 0, since_improve = 0, countdown = 1;
        std::uint32_t truth_id = 0xFFFFFFFFu;
        std::uint32_t qexp = 0;
        // the query's expansion budget (per query when given, 0 there meaning none)
        std::uint32_t qmax = a.max_exp;
        if (a.max_exps != nullptr) {
            qmax = uniform(a.max_exps[q]);
            qmax = qmax == 0 ? 0xFFFFFFFFu : qmax;
        }
        if constexpr (MODE == kBeamTrace) {
            if (lane < 10) {
                truth_id = a.truth[static_cast<std::size_t>(q) * 10 + lane];
            }
        }
        for (;;) {
            // The closest unexpanded member of R.
            std::int32_t p = -1;
            for (std::uint32_t b = hint; b < size; b += kWave) {
                const std::uint32_t i = b + lane;
                const bool open = i < size && (static_cast<std::uint32_t>(Rl[i]) & kExpanded) == 0;
                const std::uint64_t bal = ballot(open);
                if (bal != 0) {
                    p = static_cast<std::int32_t>(b) + __ffsll(static_cast<long long>(bal)) - 1;
                    break;
                }
            }
            if (p < 0) {
                break;
            }
            if (qexp >= qmax) {
                // the expansion budget ends the search with an open member left
                if (lane == 0 && a.stats != nullptr) {
                    atomicAdd(a.stats + 4, 1ull);
                }
                break;
            }
            ++qexp;
            const std::uint64_t kp = Rl[p];
            const std::uint32_t node = uniform(key_id(kp));
            wave_lds_fence();
            if (lane == 0) {
                Rl[p] = kp | kExpanded;
            }
            hint = static_cast<std::uint32_t>(p) + 1;
            ++n_exp;

            // the searched layer's block (layer 0 for queries; an insertion may search an upper
            // one): prefetched during the previous expansion when it named this node
            std::uint32_t cnt, id;
            if (node == pref_node) {
                cnt = uniform(pref_cnt);
                id = pref_id;
                ++n_pref;
            } else {
                const std::uint32_t* blk = block_of(node);
                cnt = uniform(blk[0]);
                id = static_cast<std::uint32_t>(lane) < cnt ? blk[1 + lane] : 0u;
            }
            const bool valid = static_cast<std::uint32_t>(lane) < cnt;
            bool fresh = false;
            if (valid) {
                if (bits != nullptr) {
                    // test-and-set: one atomic per neighbour, the ids of a block being distinct
                    const std::uint32_t bit = 1u << (id & 31u);
                    fresh = (atomicOr(bits + (id >> 5), bit) & bit) == 0;
                } else {
                    const std::uint32_t slot = (id * 2654435761u) >> vshift;
                    fresh = vis[slot] != id;
                    if (fresh) {
                        vis[slot] = id;
                    }
                }
            }
            const std::uint64_t fb = ballot(fresh);
            const std::uint32_t nf = static_cast<std::uint32_t>(__popcll(fb));
            if (fresh) {
                fid[lanes_below(fb)] = id;
            }
            wave_lds_fence();
            // Prefetch the block of the next open member of R, the next expansion unless a fresh
            // key lands ahead of it: its latency passes under this expansion's rows and merge.
            // (Node `node`'s own block again when none is open: the load is unconditional.)
            {
                std::int32_t p2 = -1;
                for (std::uint32_t b = hint; b < size; b += kWave) {
                    const std::uint32_t i = b + lane;
                    const bool open = i < size && (static_cast<std::uint32_t>(Rl[i]) & kExpanded) == 0;
                    const std::uint64_t bal = ballot(open);
                    if (bal != 0) {
                        p2 = static_cast<std::int32_t>(b) + __ffsll(static_cast<long long>(bal)) - 1;
                        break;
                    }
                }
                const std::uint32_t next = p2 >= 0 ? uniform(key_id(Rl[p2])) : node;
                pref_node = p2 >= 0 ? next : kNoNode;
                const std::uint32_t* nb2 = block_of(next);
                pref_cnt = nb2[0];
                pref_id = static_cast<std::uint32_t>(lane) < lcap ? nb2[1 + lane] : 0u;
            }
            n_dist += nf;
            for (std::uint32_t f0 = 0; f0 < nf; f0 += 8 * R) {
                const std::uint32_t fc = min(8u * R, nf - f0);
                distances<NB, N8, TAIL, R>(src, fid + f0, fc, qb, q8, qt, nb, n8, tail, first_tail,
                                           fd + f0);
            }
            wave_lds_fence();
            std::uint64_t f = 0;
            if (static_cast<std::uint32_t>(lane) < nf) {
                f = make_key(fd[lane], fid[lane]);
            }
            std::uint32_t inserted = 0, first_pos = 0;
            merge(Rl, size, efp, hint, f, nf, F, n_dup, inserted, first_pos);
            if constexpr (MODE != kBeamPlain) {
                ++steps;
                const std::uint32_t kk = min(k, size);
                const bool improved = inserted > 0 && first_pos < kk;
                improved_hist = ((improved_hist << 1) | (improved ? 1u : 0u)) & 0xFFu;
                since_improve = improved ? 0u : since_improve + 1u;
                // the learned stop consults the model every mlp_every expansions once allowed
                bool consult = true;
                if constexpr (MODE == kBeamLearned) {
                    consult = false;
                    if (steps >= a.min_steps && size >= k && --countdown == 0) {
                        consult = true;
                        countdown = a.mlp_every;
                    }
                }
                if (consult) {
                const float feature = step_feature(lane, key_distance(Rl[0]), key_distance(Rl[kk - 1]),
                                                   key_distance(Rl[size - 1]), key_distance(kp), inv_scale, steps,
                                                   size, improved_hist, static_cast<std::uint32_t>(p), inserted, nf,
                                                   since_improve, inv_efp, efp, inv_m2, k, a.l2 != 0);
                if constexpr (MODE == kBeamTrace) {
                    // top-10 hits of the current result against the ground truth
                    const std::uint32_t kt = min(10u, size);
                    const std::uint32_t rid = static_cast<std::uint32_t>(lane) < kt ? key_id(Rl[lane]) : 0xFFFFFFFEu;
                    bool found = false;
                    for (int j = 0; j < 10; ++j) {
                        found = found || (static_cast<std::uint32_t>(__shfl(static_cast<int>(truth_id), j, kWave)) == rid);
                    }
                    const std::uint32_t hits = static_cast<std::uint32_t>(__popcll(ballot(found && static_cast<std::uint32_t>(lane) < kt)));
                    if (steps <= a.trace_cap) {
                        const std::size_t row = static_cast<std::size_t>(q) * a.trace_cap + (steps - 1);
                        if (lane < kFeatures) {
                            a.trace_feat[row * kFeatures + lane] = feature;
                        }
                        if (lane == 0) {
                            a.trace_hits[row] = static_cast<std::uint8_t>(hits);
                        }
                    }
                    if (lane == 0) {
                        a.trace_final[q] = hits;  // the last step's value is the result's
                    }
                }
                if constexpr (MODE == kBeamLearned) {
                    const float z = mlp_logit_lanes(*model, feature, lane, reinterpret_cast<std::uint16_t*>(feat),
                                                    reinterpret_cast<std::uint16_t*>(fd));
                    if (z > a.stop_logit) {
                        ++n_stop;
                        break;  // stop: the model judges further expansions useless
                    }
                }
                }  // consult
            }
        }
        if constexpr (MODE == kBeamTrace) {
            if (lane == 0) {
                a.trace_steps[q] = steps;
            }
        }

        const std::uint32_t count = min(k, size);
        for (std::uint32_t i = lane; i < count; i += kWave) {
            const std::uint64_t key = Rl[i];
            out[i] = Neighbor{key_distance(key), key_id(key)};
        }
        if (lane == 0) {
            a.out_count[q] = count;
        }
        wave_lds_fence();
    }
    if (lane == 0 && a.stats != nullptr) {
        atomicAdd(a.stats + 0, n_exp);
        atomicAdd(a.stats + 1, n_dist);
        atomicAdd(a.stats + 2, n_up);
        atomicAdd(a.stats + 3, n_dup);
        atomicAdd(a.stats + 4, n_stop);
        atomicAdd(a.stats + 5, n_pref);
    }
}

std::size_t beam_lds_shared(int mode) { return mode == kBeamLearned ? sizeof(MlpWeights) : 0; }

namespace {

template <int MODE>
const void* kernel_for(BeamShape shape) {
    switch (shape) {
        case BeamShape::D128:
        case BeamShape::D128U8:
            return reinterpret_cast<const void*>(&hnsw_beam_kernel<4, 0, 0, 4, MODE>);
        case BeamShape::D100:
            return reinterpret_cast<const void*>(&hnsw_beam_kernel<3, 0, 4, 4, MODE>);
        default:
            return reinterpret_cast<const void*>(&hnsw_beam_kernel<8, 3, 7, 1, MODE>);
    }
}

}  // namespace

const void* hnsw_beam_kernel_ptr(BeamShape shape, int mode) {
    return mode == kBeamTrace ? kernel_for<kBeamTrace>(shape)
                              : (mode == kBeamLearned ? kernel_for<kBeamLearned>(shape)
                                                      : kernel_for<kBeamPlain>(shape));
}

hipError_t launch_hnsw_beam(const BeamArgs& args, unsigned grid, unsigned waves_per_group,
                            std::size_t lds_bytes, hipStream_t stream, BeamShape shape, int mode) {
    void* params[] = {const_cast<BeamArgs*>(&args)};
    return hipLaunchKernel(hnsw_beam_kernel_ptr(shape, mode), dim3(grid), dim3(waves_per_group * kWave),
                           params, lds_bytes, stream);
}

}  // namespace vdx::dev

namespace vdx::dev {

const void* hnsw_beam_reg_ptr_c1(BeamShape shape, int mode);
const void* hnsw_beam_reg_ptr_c2(BeamShape shape, int mode);
const void* hnsw_beam_reg_ptr_c4(BeamShape shape, int mode);
const void* hnsw_beam_reg_ptr_c8(BeamShape shape, int mode);

const void* hnsw_beam_reg_kernel_ptr(BeamShape shape, int mode, int chunks) {
    switch (chunks) {
        case 1: return hnsw_beam_reg_ptr_c1(shape, mode);
        case 2: return hnsw_beam_reg_ptr_c2(shape, mode);
        case 4: return hnsw_beam_reg_ptr_c4(shape, mode);
        default: return hnsw_beam_reg_ptr_c8(shape, mode);
    }
}

hipError_t launch_hnsw_beam_reg(const BeamArgs& args, unsigned grid, unsigned waves_per_group,
                                std::size_t lds_bytes, hipStream_t stream, BeamShape shape, int mode,
                                int chunks) {
    void* params[] = {const_cast<BeamArgs*>(&args)};
    return hipLaunchKernel(hnsw_beam_reg_kernel_ptr(shape, mode, chunks), dim3(grid),
                           dim3(waves_per_group * kWave), params, lds_bytes, stream);
}

}  // namespace vdx::dev
