#pragma once

// K1 for ef' = max(ef, k) <= 512: the HNSW query path with the result list R in registers.
//
// Same algorithm, same results as hnsw_beam.hip (vecdex src/hnsw.cpp:504-522: greedy descent,
// layer-0 beam, closest min(k, |results|) in `closer` order; bit-identical to the CPU), laid out
// for the wavefront instead of LDS:
//
//  * R, the sorted best ef' keys, lives in CH chunks of 64 keys, lane l of chunk c holding
//    R[64c + l] (entries past the end hold an all-ones sentinel whose "expanded" bit is set).
//    The closest unexpanded member is one ballot per chunk.
//  * Merging one expansion's fresh keys: keys that cannot enter (R full and not closer than its
//    worst) drop at once; for each remaining key j (a scalar loop over a ballot mask), ballots
//    over R count the members below it and find an exact duplicate (a node the lossy visited
//    cache forgot), and every lane counts how many inserted keys fall below its own R member and
//    its own fresh key. Those counts are the final positions: R and the new keys are scattered
//    to an LDS buffer by position and read back as chunks.
//  * The link block of the next expansion is prefetched: the block of the next unexpanded
//    member of R is loaded together with this expansion's rows; after the merge
//    ranks, if a fresh key lands ahead of that member, its block is loaded instead. The next
//    expansion uses the prefetched block only if it names the node it selects.
//  * Row distances reduce across the 8 lanes of a candidate with DPP moves (same additions,
//    same association as the reference's AVX2 horizontal sum).
//
// LDS per wave: the merge buffer (8 B x 64 CH), fresh ids and distances (64 each), features,
// the visited cache.

#include <hip/hip_runtime.h>

#include "beam_args.hpp"
#include "device_common.hpp"
#include "gather_distances.hpp"
#include "mlp_stop.hpp"

namespace vdx::dev {

namespace reg {

constexpr std::uint32_t kNone = 0xFFFFFFFFu;
constexpr std::uint64_t kSentinel = ~static_cast<std::uint64_t>(0);

__device__ __forceinline__ std::uint32_t uniform(std::uint32_t v) {
    return __builtin_amdgcn_readfirstlane(v);
}

// Lane l's 64-bit value, l uniform.
__device__ __forceinline__ std::uint64_t readlane64(std::uint64_t v, std::uint32_t l) {
    const std::uint32_t lo = __builtin_amdgcn_readlane(static_cast<int>(static_cast<std::uint32_t>(v)), l);
    const std::uint32_t hi = __builtin_amdgcn_readlane(static_cast<int>(static_cast<std::uint32_t>(v >> 32)), l);
    return (static_cast<std::uint64_t>(hi) << 32) | lo;
}

__device__ __forceinline__ bool is_open(std::uint64_t key) {
    return (static_cast<std::uint32_t>(key) & kExpanded) == 0;
}

__device__ __forceinline__ float wave_min(float v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        v = fminf(v, __shfl_xor(v, off, kWave));
    }
    return v;
}

// R[i] for a uniform index i.
template <int CH>
__device__ __forceinline__ std::uint64_t r_at(const std::uint64_t (&Rk)[CH], std::uint32_t i) {
    std::uint64_t v = 0;
#pragma unroll
    for (int c = 0; c < CH; ++c) {
        if (static_cast<std::uint32_t>(c) == (i >> 6)) {
            v = readlane64(Rk[c], i & 63u);
        }
    }
    return v;
}

// The first open member of R at or after chunk `from`: (chunk, lane), chunk -1 if none.
template <int CH>
__device__ __forceinline__ void first_open(const std::uint64_t (&Rk)[CH], int from, int& pc,
                                           std::uint32_t& pl) {
    pc = -1;
    pl = 0;
#pragma unroll
    for (int c = 0; c < CH; ++c) {
        if (pc < 0 && c >= from) {
            const std::uint64_t bal = ballot(is_open(Rk[c]));
            if (bal != 0) {
                pc = c;
                pl = static_cast<std::uint32_t>(__ffsll(static_cast<long long>(bal)) - 1);
            }
        }
    }
}

}  // namespace reg

// U8 (dimension 128, squared L2 only): rows come from the 8-bit copy a.rows_u8
// (distances_rows_u8_l2). MET: 1 squared L2, 0 the inner-product form, -1 read from a.l2.
// QI (with U8): every query of the launch has all its components integers in [0, 255]
// (a.qflag, a.qflag_epoch), so the rows' distances take exact integer arithmetic (distances_rows_u8_l2_int)
// with the query held as 4 byte words.
template <int NB, int N8, int TAIL, int RR, bool FIXED, int CH, int MODE, bool U8 = false, int MET = -1,
          bool QI = false>
__device__ __forceinline__ void beam_reg_body(const BeamArgs& a) {
    static_assert(!U8 || (NB == 4 && N8 == 0 && TAIL == 0 && FIXED), "8-bit rows are dimension 128");
    static_assert(!QI || (U8 && MODE != kBeamTrace), "integer queries: the search over the 8-bit rows");
    using namespace reg;
    extern __shared__ __align__(16) unsigned char lds[];
    const int lane = lane_id();
    const int wave = threadIdx.x >> 6;
    const int m = lane & 7;
    unsigned char* base = lds + static_cast<std::size_t>(wave) * a.lds_per_wave;
    std::uint64_t* B = reinterpret_cast<std::uint64_t*>(base);
    std::uint32_t* fid = reinterpret_cast<std::uint32_t*>(B + 64 * CH);
    float* fd = reinterpret_cast<float*>(fid + 64);
    float* feat = fd + 64;
    // the plain search keeps no features: the visited cache follows the fresh distances
    std::uint32_t* vis = reinterpret_cast<std::uint32_t*>(MODE == kBeamPlain ? feat : feat + 16);
    MlpWeights* model = nullptr;
    if constexpr (MODE == kBeamLearned) {
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
    const int nb = FIXED ? NB : static_cast<int>(a.nb);
    const int n8 = FIXED ? N8 : static_cast<int>(a.n8);
    const int tail = FIXED ? TAIL : static_cast<int>(a.tail);
    const int first_tail = static_cast<int>(a.first_tail);
    const RowSource src{a.vectors, a.ld, a.l2};
    const std::uint32_t b0 = a.b0;

    std::uint32_t begin = 0, end = a.nq;
    if (a.range != nullptr) {
        begin = uniform(a.range[0]);
        end = uniform(a.range[1]);
    }

    unsigned long long n_exp = 0, n_dist = 0, n_up = 0, n_dup = 0, n_stop = 0, n_pref = 0;

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
        const std::uint32_t k = min(uniform(a.ks ? a.ks[q] : a.k), a.k_cap);
        const std::uint32_t ef = uniform(a.efs ? a.efs[q] : a.ef);
        const std::uint32_t efp = min(max(ef, k), static_cast<std::uint32_t>(64 * CH));
        // a per-query ef above the launch's R capacity runs at the capacity: counted at once
        // (rare; no counter kept live through the search), so a caller whose `ef` bound was
        // too small can see it
        if (max(ef, k) > static_cast<std::uint32_t>(64 * CH) && lane == 0 && a.stats != nullptr) {
            atomicAdd(a.stats + 6, 1ull);
        }
        Neighbor* out = a.out + static_cast<std::size_t>(q) * a.out_stride;

        // The query, spread over the 8 lanes of each group (every group holds a full copy).
        const float* qp = a.queries + static_cast<std::size_t>(q) * a.qld;
        float4 qb[NB > 0 ? NB : 1];
        float4 q8[N8 > 0 ? N8 : 1];
        float qt[8];
#pragma unroll
        for (int i = 0; i < NB; ++i) {
            qb[i] = make_float4(0.f, 0.f, 0.f, 0.f);
            if ((FIXED || i < nb) && (!QI || a.qbytes == nullptr)) {
                const float* p = qp + 32 * i + 4 * m;
                qb[i] = make_float4(p[0], p[1], p[2], p[3]);
            }
        }
        // QI: the query as byte words (the floats are dead from here), or as the host staged
        // it, 128 bytes in element order (lane m's words: elements 32i + 4m .. 32i + 4m + 3)
        std::uint32_t qw[4] = {0u, 0u, 0u, 0u};
        std::uint32_t qq = 0;
        if constexpr (QI) {
            if (a.qbytes != nullptr) {
                const std::uint32_t* qwp = reinterpret_cast<const std::uint32_t*>(a.qbytes + static_cast<std::size_t>(q) * 128) + m;
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    qw[i] = qwp[8 * i];
                }
                qq = bytes_sum_sq(qw);
            } else {
                (void)query_as_bytes(qb, qw, qq);
            }
        }
#pragma unroll
        for (int s = 0; s < N8; ++s) {
            q8[s] = make_float4(0.f, 0.f, 0.f, 0.f);
            if ((FIXED || s < n8) && m < 2) {
                const float* p = qp + 32 * nb + 8 * s + 4 * m;
                q8[s] = make_float4(p[0], p[1], p[2], p[3]);
            }
        }
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            qt[e] = (e < TAIL && (FIXED || e < tail)) ? qp[first_tail + e] : 0.f;
        }
        // Distances of cnt candidates (ids in LDS) to the query, into LDS.
        const auto row_distances = [&](const std::uint32_t* ids, std::uint32_t cnt, float* dst) {
            if constexpr (QI) {
                distances_rows_u8_l2_int<RR>(a.rows_u8, ids, cnt, qw, qq, dst);
            } else if constexpr (U8) {
                distances_rows_u8_l2<RR>(a.rows_u8, ids, cnt, qb, dst);
            } else {
                distances_rows<NB, N8, TAIL, RR, FIXED, MET>(src, ids, cnt, qb, q8, qt, nb, n8, tail, first_tail, dst);
            }
        };

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
        // cur's first upper block (loaded with its row): with each hop's neighbours' first
        // blocks loaded together with their rows, a hop costs two dependent loads (its block,
        // then the neighbours' rows and first blocks), not three. Not in the float search over
        // the byte rows, where the extra register would cost a wave per SIMD.
        constexpr bool kOffsets = QI || !U8;
        std::uint32_t cur_off = top > a.layer ? uniform(a.upper_offset[cur]) : 0u;
        if (lane == 0) {
            fid[0] = cur;
        }
        wave_lds_fence();
        row_distances(fid, 1, fd);
        wave_lds_fence();
        float cur_d = fd[0];
        ++n_dist;
        for (std::uint32_t layer = top; layer > a.layer; --layer) {
            for (;;) {
                const std::uint32_t* blk = a.upper + (static_cast<std::size_t>(cur_off) + (layer - 1)) * a.bu;
                const std::uint32_t cnt = uniform(blk[0]);
                std::uint32_t id = 0;
                if (static_cast<std::uint32_t>(lane) < cnt) {
                    id = blk[1 + lane];
                    fid[lane] = id;
                }
                // a neighbour on this layer has upper blocks down to layer 1 (the upload's link
                // rules), so its first block index is what the next hop or layer needs
                std::uint32_t nb_off = 0;
                if constexpr (kOffsets) {
                    nb_off = static_cast<std::uint32_t>(lane) < cnt ? a.upper_offset[id] : 0u;
                }
                ++n_up;
                n_dist += cnt;
                wave_lds_fence();
                for (std::uint32_t c0 = 0; c0 < cnt; c0 += 8 * RR) {
                    row_distances(fid + c0, min(8u * RR, cnt - c0), fd + c0);
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
                if constexpr (kOffsets) {
                    cur_off = uniform(broadcast(nb_off, first));
                } else {
                    cur_off = uniform(a.upper_offset[cur]);
                }
                cur_d = best;
            }
        }

        // The searched layer's link block of a node (layer 0 for queries); an insertion may
        // search an upper layer, whose blocks hold M links.
        const std::uint32_t entry_node = cur;
        const std::uint32_t lcap = a.layer == 0 ? a.M2 : (a.M2 >> 1);
        const auto block_of = [&](std::uint32_t nd) -> const std::uint32_t* {
            return a.layer == 0
                       ? a.layer0 + static_cast<std::size_t>(nd) * b0
                       : a.upper + (static_cast<std::size_t>(a.upper_offset[nd]) + (a.layer - 1)) * a.bu;
        };

        // Layer 0 (src/hnsw.cpp:438-482) with ef' = max(ef, k) (src/hnsw.cpp:517).
        for (std::uint32_t i = static_cast<std::uint32_t>(lane) * 4; i < vslots; i += kWave * 4) {
            *reinterpret_cast<uint4*>(vis + i) = make_uint4(kNone, kNone, kNone, kNone);
        }
        wave_lds_fence();
        if (lane == 0) {
            vis[(cur * 2654435761u) >> vshift] = cur;
        }
        std::uint64_t Rk[CH];
#pragma unroll
        for (int c = 0; c < CH; ++c) {
            Rk[c] = kSentinel;
        }
        if (lane == 0) {
            Rk[0] = make_key(cur_d, cur);
        }
        std::uint32_t size = 1;
        int hint_c = 0;
        std::uint32_t pref_node = kNone;
        std::uint32_t pref_cnt = 0, pref_id = 0;
        // search-state features (trace and learned modes)
        const float inv_scale = 1.0f / fmaxf(fabsf(cur_d), 1e-6f);
        const float inv_efp = 1.0f / static_cast<float>(efp);
        const float inv_m2 = 1.0f / static_cast<float>(a.M2);
        std::uint32_t steps = 0, improved_hist = 0, since_improve = 0, countdown = 1;
        std::uint32_t truth_id = kNone;
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
            // The closest unexpanded member of R; mark it expanded.
            int pc;
            std::uint32_t pl;
            first_open<CH>(Rk, hint_c, pc, pl);
            if (pc < 0) {
                break;
            }
            if (qexp >= qmax) {
                // the expansion budget ends the search with an open member left (counted at once:
                // no counter kept live through the search)
                if (lane == 0 && a.stats != nullptr) {
                    atomicAdd(a.stats + 4, 1ull);
                }
                break;
            }
            ++qexp;
            std::uint64_t kp = 0;
#pragma unroll
            for (int c = 0; c < CH; ++c) {
                if (c == pc) {
                    kp = readlane64(Rk[c], pl);
                    if (static_cast<std::uint32_t>(lane) == pl) {
                        Rk[c] |= kExpanded;
                    }
                }
            }
            hint_c = pc;
            const std::uint32_t p = static_cast<std::uint32_t>(pc) * 64u + pl;
            const std::uint32_t node = key_id(kp);
            ++n_exp;

            // Its link block: prefetched, or read now.
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
            // Fresh neighbours: not in the visited cache.
            const bool valid = static_cast<std::uint32_t>(lane) < cnt;
            bool fresh = false;
            if (valid) {
                const std::uint32_t slot = (id * 2654435761u) >> vshift;
                fresh = vis[slot] != id;
                if (fresh) {
                    vis[slot] = id;
                }
            }
            const std::uint64_t fb = ballot(fresh);
            const std::uint32_t nf = static_cast<std::uint32_t>(__popcll(fb));
            if (fresh) {
                fid[lanes_below(fb)] = id;
            }
            wave_lds_fence();
            // Prefetch the block of the next open member of R (the next expansion unless a
            // fresh key lands ahead of it). Issued after this block's ids are consumed and just
            // before the row loads, its latency overlaps theirs.
            // The load is unconditional (node 0's block when R has no open member) so no
            // branch merges stale and fresh stage registers.
            std::uint64_t spec_key = kSentinel;
            {
                int sc;
                std::uint32_t sl;
                first_open<CH>(Rk, pc, sc, sl);
                std::uint64_t ks = 0;
#pragma unroll
                for (int c = 0; c < CH; ++c) {
                    if (c == sc) {
                        ks = readlane64(Rk[c], sl);
                    }
                }
                spec_key = sc >= 0 ? (ks & kKeyMask) : kSentinel;
                pref_node = sc >= 0 ? key_id(ks) : kNone;
                const std::uint32_t* nb2 = block_of(sc >= 0 ? key_id(ks) : entry_node);
                pref_cnt = nb2[0];
                pref_id = static_cast<std::uint32_t>(lane) < lcap ? nb2[1 + lane] : 0u;
            }

            n_dist += nf;
            // The first 8 RR rows outside the loop: in straight-line code the compiler's waits
            // leave the prefetched block in flight behind the row loads (at a loop header it
            // waits for every load already issued). More rows than that only for M > 4 RR.
            if (nf > 0) {
                row_distances(fid, min(8u * RR, nf), fd);
            }
            for (std::uint32_t f0 = 8 * RR; f0 < nf; f0 += 8 * RR) {
                row_distances(fid + f0, min(8u * RR, nf - f0), fd + f0);
            }
            wave_lds_fence();
            std::uint64_t f = kSentinel;
            if (static_cast<std::uint32_t>(lane) < nf) {
                f = make_key(fd[lane], fid[lane]);
            }

            // Merge: keys that cannot enter drop first (R full and not closer than its worst).
            const std::uint64_t worst = r_at<CH>(Rk, size - 1) & kKeyMask;
            const bool cand = static_cast<std::uint32_t>(lane) < nf && (size < efp || f < worst);
            // One pass over the candidates: for each, the members of R below it (its rank in R,
            // written to its own lane) and whether R holds it already; every lane counts the
            // kept candidates below its own key and below its R entries. Everything else
            // follows after the pass: an R entry moves up by the kept candidates not below it,
            // and the closest kept candidate is the one with none below it.
            std::uint64_t rem = ballot(cand);
            std::uint64_t Rm[CH];
            std::uint32_t r_below[CH];
#pragma unroll
            for (int c = 0; c < CH; ++c) {
                Rm[c] = Rk[c] & kKeyMask;
                r_below[c] = 0;
            }
            std::uint32_t lb = 0, below = 0, nins = 0;
            std::uint64_t dupmask = 0;
            while (rem != 0) {
                const std::uint32_t j = static_cast<std::uint32_t>(__builtin_ctzll(rem));
                rem &= ~(1ull << j);
                const std::uint64_t fj = readlane64(f, j);
                bool eq = false;
                std::uint32_t lt = 0;
                bool r_lt[CH];
#pragma unroll
                for (int c = 0; c < CH; ++c) {
                    eq |= ballot(Rm[c] == fj) != 0;
                    r_lt[c] = Rm[c] < fj;
                    lt += static_cast<std::uint32_t>(__popcll(ballot(r_lt[c])));
                }
                if (eq) {
                    dupmask |= 1ull << j;  // already in R: a node the visited cache forgot
                    continue;
                }
#pragma unroll
                for (int c = 0; c < CH; ++c) {
                    r_below[c] += r_lt[c] ? 1u : 0u;
                }
                below += fj < f ? 1u : 0u;
                lb = static_cast<std::uint32_t>(lane) == j ? lt : lb;
                ++nins;
            }
            std::uint32_t gt[CH];
#pragma unroll
            for (int c = 0; c < CH; ++c) {
                gt[c] = nins - r_below[c];  // kept candidates below this R entry (none equals it)
            }
            n_dup += static_cast<unsigned long long>(__popcll(dupmask));
            const bool kept = cand && ((dupmask >> lane) & 1u) == 0;
            std::uint32_t fmin_pos = kNone;
            std::uint64_t fmin = kSentinel;
            if (nins > 0) {
                const int jmin = __builtin_ctzll(ballot(kept && below == 0));
                fmin = readlane64(f, static_cast<std::uint32_t>(jmin));
                fmin_pos = static_cast<std::uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(lb), jmin));
            }
            const std::uint32_t pos = lb + below;
            const bool ins = kept && pos < efp;
            std::uint32_t inserted = 0;
            if (nins > 0) {
                // A fresh key ahead of the prefetched member is the next expansion: fetch its
                // block now, while the merge completes.
                if (fmin_pos < efp && fmin < spec_key) {
                    pref_node = key_id(fmin);
                    const std::uint32_t* nb2 = block_of(pref_node);
                    pref_cnt = nb2[0];
                    pref_id = static_cast<std::uint32_t>(lane) < lcap ? nb2[1 + lane] : 0u;
                }
                inserted = static_cast<std::uint32_t>(__popcll(ballot(ins)));
#pragma unroll
                for (int c = 0; c < CH; ++c) {
                    const std::uint32_t i = static_cast<std::uint32_t>(c * 64 + lane);
                    if (i < size && i + gt[c] < efp) {
                        B[i + gt[c]] = Rk[c];
                    }
                }
                if (ins) {
                    B[pos] = f;
                }
                wave_lds_fence();
                size = min(efp, size + nins);
#pragma unroll
                for (int c = 0; c < CH; ++c) {
                    const std::uint32_t i = static_cast<std::uint32_t>(c * 64 + lane);
                    Rk[c] = i < size ? B[i] : kSentinel;
                }
                wave_lds_fence();
                if (fmin_pos < efp) {
                    hint_c = min(hint_c, static_cast<int>(fmin_pos >> 6));
                }
            }
            if constexpr (MODE != kBeamPlain) {
                ++steps;
                const std::uint32_t kk = min(k, size);
                const bool improved = inserted > 0 && fmin_pos < kk;
                improved_hist = ((improved_hist << 1) | (improved ? 1u : 0u)) & 0xFFu;
                since_improve = improved ? 0u : since_improve + 1u;
                // the learned stop consults the model every mlp_every expansions once allowed
                // (min_steps taken, R holding k); the features are computed only then
                bool consult = true;
                if constexpr (MODE == kBeamLearned) {
                    consult = false;
                    if (steps >= a.min_steps && size >= k && --countdown == 0) {
                        consult = true;
                        countdown = a.mlp_every;
                    }
                }
                if (consult) {
                const float feature =
                    step_feature(lane, key_distance(r_at<CH>(Rk, 0)), key_distance(r_at<CH>(Rk, kk - 1)),
                                 key_distance(r_at<CH>(Rk, size - 1)), key_distance(kp), inv_scale, steps, size,
                                 improved_hist, p, inserted, nf, since_improve, inv_efp, efp, inv_m2, k, a.l2 != 0);
                if constexpr (MODE == kBeamTrace) {
                    // top-10 hits of the current result against the ground truth
                    const std::uint32_t kt = min(10u, size);
                    const std::uint32_t rid = static_cast<std::uint32_t>(lane) < kt ? key_id(Rk[0]) : 0xFFFFFFFEu;
                    bool found = false;
                    for (int j = 0; j < 10; ++j) {
                        found = found || (static_cast<std::uint32_t>(__shfl(static_cast<int>(truth_id), j, kWave)) == rid);
                    }
                    const std::uint32_t hits = static_cast<std::uint32_t>(
                        __popcll(ballot(found && static_cast<std::uint32_t>(lane) < kt)));
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
#pragma unroll
        for (int c = 0; c < CH; ++c) {
            const std::uint32_t i = static_cast<std::uint32_t>(c * 64 + lane);
            if (i < count) {
                out[i] = Neighbor{key_distance(Rk[c]), key_id(Rk[c])};
            }
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

// The kernels. The plain search is held to 128 VGPRs (4 waves per SIMD): the occupancy it gains
// outweighs the few spilled registers at R capacities above 64 (measured). The trace and learned
// modes over the FP32 rows keep their registers (the stop model's LDS bounds their occupancy
// near there anyway).
template <int NB, int N8, int TAIL, int RR, bool FIXED, int CH, int MODE>
__global__ __launch_bounds__(256) void hnsw_beam_reg_kernel(BeamArgs a) {
    beam_reg_body<NB, N8, TAIL, RR, FIXED, CH, MODE>(a);
}

template <int NB, int N8, int TAIL, int RR, bool FIXED, int CH>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void hnsw_beam_reg_plain_kernel(
    BeamArgs a) {
    // the metric fixed per copy of the body (one uniform branch per launch instead of one per
    // row block): 2-3% faster over the FP32 rows (measured, L2 and cosine)
    if (a.l2 != 0) {
        beam_reg_body<NB, N8, TAIL, RR, FIXED, CH, kBeamPlain, false, 1>(a);
    } else {
        beam_reg_body<NB, N8, TAIL, RR, FIXED, CH, kBeamPlain, false, 0>(a);
    }
}

// The plain search over the 8-bit rows for a launch with a query that is not byte-valued
// (*qflag == qflag_epoch after query_bytes_check; the float arithmetic over the converted bytes).
// Held to 72 VGPRs at R capacity 64: 7 waves per SIMD, the occupancy its LDS allows, for six
// spilled registers (at 80 VGPRs and 6 waves: fractional queries, ef 64 0.587 -> 0.574 ms and
// ef 36 0.421 -> 0.418 ms).
template <int CH>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(CH <= 1 ? 7 : 4))) void hnsw_beam_reg_u8_kernel(
    BeamArgs a) {
    // a launch whose queries are all byte-valued is served by hnsw_beam_reg_u8int_kernel,
    // enqueued with this one
    if (a.qflag != nullptr && *a.qflag != a.qflag_epoch) {
        return;
    }
    beam_reg_body<4, 0, 0, 4, true, CH, kBeamPlain, true, 1, false>(a);
}

// The plain search over the 8-bit rows when every query of the launch has all its components
// integers in [0, 255] (*qflag != qflag_epoch after query_bytes_check): exact integer
// arithmetic with the query held as 4 byte words, so the same distances in a kernel of its own
// (sharing one kernel with the float path cost both bodies registers and a wave per SIMD).
// Held to 72 VGPRs (7 waves per SIMD) up to R capacity 256: the occupancy its LDS allows (seven
// 4-wave workgroups per CU), with no spilled register at R capacity 64 (64 VGPRs spilled six:
// ef 36 0.386 -> 0.377 ms, the expansion-budget launch 0.354 -> 0.344 ms; 80 VGPRs and 6 waves
// had been slower still).
// Enqueued together with hnsw_beam_reg_plain_kernel<U8>: whichever the flag does not select
// returns at once.
template <int CH>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(CH <= 4 ? 7 : 4))) void hnsw_beam_reg_u8int_kernel(
    BeamArgs a) {
    if (a.qflag == nullptr || *a.qflag == a.qflag_epoch) {
        return;
    }
    beam_reg_body<4, 0, 0, 4, true, CH, kBeamPlain, true, 1, true>(a);
}

// The learned stop over the 8-bit rows: integer arithmetic when every query is byte-valued, the
// float arithmetic otherwise, enqueued together like the plain pair. The model staged in LDS per
// workgroup (11 KB) leaves room for 5 four-wave workgroups per CU, so the kernels are held to the
// registers of 5 waves per SIMD up to R capacity 256.
template <int CH>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(CH <= 4 ? 5 : 4))) void
hnsw_beam_reg_u8int_learned_kernel(BeamArgs a) {
    if (a.qflag == nullptr || *a.qflag == a.qflag_epoch) {
        return;
    }
    beam_reg_body<4, 0, 0, 4, true, CH, kBeamLearned, true, 1, true>(a);
}

template <int CH>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(CH <= 4 ? 5 : 4))) void
hnsw_beam_reg_u8_learned_kernel(BeamArgs a) {
    if (a.qflag != nullptr && *a.qflag != a.qflag_epoch) {
        return;
    }
    beam_reg_body<4, 0, 0, 4, true, CH, kBeamLearned, true, 1, false>(a);
}

// Kernel pointers of one chunk count, by shape and mode (kBeamU8Int, kBeamU8IntLearned: the
// integer kernels above).
template <int CH>
const void* hnsw_beam_reg_ptr(BeamShape shape, int mode) {
#define VDX_REG_KERNEL(MODE)                                                                     \
    switch (shape) {                                                                             \
        case BeamShape::D128:                                                                    \
        case BeamShape::D128U8:                                                                  \
            return reinterpret_cast<const void*>(&hnsw_beam_reg_kernel<4, 0, 0, 4, true, CH, MODE>); \
        case BeamShape::D100:                                                                    \
            return reinterpret_cast<const void*>(&hnsw_beam_reg_kernel<3, 0, 4, 4, true, CH, MODE>); \
        default:                                                                                 \
            return reinterpret_cast<const void*>(&hnsw_beam_reg_kernel<8, 3, 7, 1, false, CH, MODE>); \
    }
    if (mode == kBeamU8Int) {
        return reinterpret_cast<const void*>(&hnsw_beam_reg_u8int_kernel<CH>);
    }
    if (mode == kBeamU8IntLearned) {
        return reinterpret_cast<const void*>(&hnsw_beam_reg_u8int_learned_kernel<CH>);
    }
    if (mode == kBeamLearned && shape == BeamShape::D128U8) {
        return reinterpret_cast<const void*>(&hnsw_beam_reg_u8_learned_kernel<CH>);
    }
    if (mode == kBeamTrace) {
        VDX_REG_KERNEL(kBeamTrace)
    }
    if (mode == kBeamLearned) {
        VDX_REG_KERNEL(kBeamLearned)
    }
#undef VDX_REG_KERNEL
    switch (shape) {
        case BeamShape::D128:
            return reinterpret_cast<const void*>(&hnsw_beam_reg_plain_kernel<4, 0, 0, 4, true, CH>);
        case BeamShape::D128U8:
            return reinterpret_cast<const void*>(&hnsw_beam_reg_u8_kernel<CH>);
        case BeamShape::D100:
            return reinterpret_cast<const void*>(&hnsw_beam_reg_plain_kernel<3, 0, 4, 4, true, CH>);
        default:
            return reinterpret_cast<const void*>(&hnsw_beam_reg_plain_kernel<8, 3, 7, 1, false, CH>);
    }
}

}  // namespace vdx::dev
