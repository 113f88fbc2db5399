// Recall measurement (vecdex src/recall.cpp:11-51, include/vecdex/recall.hpp:12-33): the
// id-based recall@k and the distance-threshold recall@k that every evaluation of the search
// path reports, with the reference's argument checks and messages, plus means over a results
// matrix for batched evaluation.

#include <algorithm>
#include <cmath>

#include "capi_common.hpp"

namespace vdx {
namespace {

// Hits of the first min(k, found_count) results among the first k true ids. Each true id
// counts at most once, so a repeated found id is one hit (src/recall.cpp:21-31).
double recall_one(const Neighbor* found, std::uint64_t found_count, const std::int32_t* truth,
                  std::uint64_t truth_count, std::uint64_t k) {
    if (k == 0) {
        throw std::invalid_argument("recall_at_k: k must be positive");
    }
    if (truth_count < k) {
        throw std::invalid_argument(fmt("recall_at_k: truth has %llu ids, fewer than k = %llu",
                                        static_cast<unsigned long long>(truth_count),
                                        static_cast<unsigned long long>(k)));
    }
    const std::uint64_t top = std::min(k, found_count);
    std::uint64_t hits = 0;
    for (std::uint64_t t = 0; t < k; ++t) {
        // An id is compared by value: a negative true id never equals a uint32 result id.
        if (truth[t] < 0) {
            continue;
        }
        const auto want = static_cast<std::uint32_t>(truth[t]);
        for (std::uint64_t f = 0; f < top; ++f) {
            if (found[f].id == want) {
                ++hits;
                break;
            }
        }
    }
    return static_cast<double>(hits) / static_cast<double>(k);
}

double threshold_recall_one(const Neighbor* found, std::uint64_t found_count, float kth_true,
                            std::uint64_t k, float tolerance) {
    if (k == 0) {
        throw std::invalid_argument("threshold_recall_at_k: k must be positive");
    }
    if (!std::isfinite(kth_true) || !std::isfinite(tolerance) || tolerance < 0.0F) {
        throw std::invalid_argument(
            "threshold_recall_at_k: distance and tolerance must be finite, tolerance >= 0");
    }
    // The slack is relative above magnitude 1 and absolute below (src/recall.cpp:45-46).
    const float limit = kth_true + tolerance * std::max(std::fabs(kth_true), 1.0F);
    const std::uint64_t top = std::min(k, found_count);
    std::uint64_t hits = 0;
    for (std::uint64_t f = 0; f < top; ++f) {
        hits += found[f].distance <= limit ? 1 : 0;
    }
    return static_cast<double>(hits) / static_cast<double>(k);
}

const Neighbor* as_neighbors(const vdx_neighbor* p) { return reinterpret_cast<const Neighbor*>(p); }

}  // namespace
}  // namespace vdx

using vdx::guarded;
using vdx::require;

extern "C" {

int vdx_recall_at_k(const vdx_neighbor* found, uint64_t found_count, const int32_t* truth,
                    uint64_t truth_count, uint64_t k, double* out) {
    return guarded([&] {
        require(out != nullptr, "recall_at_k: out is null");
        require(found_count == 0 || found != nullptr, "recall_at_k: found is null");
        require(truth_count == 0 || truth != nullptr, "recall_at_k: truth is null");
        *out = vdx::recall_one(vdx::as_neighbors(found), found_count, truth, truth_count, k);
    });
}

int vdx_threshold_recall_at_k(const vdx_neighbor* found, uint64_t found_count, float kth_true,
                              uint64_t k, float tolerance, double* out) {
    return guarded([&] {
        require(out != nullptr, "threshold_recall_at_k: out is null");
        require(found_count == 0 || found != nullptr, "threshold_recall_at_k: found is null");
        *out = vdx::threshold_recall_one(vdx::as_neighbors(found), found_count, kth_true, k, tolerance);
    });
}

int vdx_mean_recall_at_k(const vdx_neighbor* found, uint64_t found_stride,
                         const uint32_t* found_counts, const int32_t* truth, uint64_t truth_stride,
                         uint64_t nq, uint64_t k, double* out) {
    return guarded([&] {
        require(out != nullptr, "recall_at_k: out is null");
        require(nq == 0 || (found != nullptr && truth != nullptr), "recall_at_k: null rows");
        double sum = 0.0;
        for (uint64_t q = 0; q < nq; ++q) {
            const uint64_t count = found_counts ? std::min<uint64_t>(found_counts[q], found_stride)
                                                : found_stride;
            sum += vdx::recall_one(vdx::as_neighbors(found + q * found_stride), count,
                                   truth + q * truth_stride, truth_stride, k);
        }
        *out = nq ? sum / static_cast<double>(nq) : 0.0;
    });
}

int vdx_mean_threshold_recall_at_k(const vdx_neighbor* found, uint64_t found_stride,
                                   const uint32_t* found_counts, const float* kth_true, uint64_t nq,
                                   uint64_t k, float tolerance, double* out) {
    return guarded([&] {
        require(out != nullptr, "threshold_recall_at_k: out is null");
        require(nq == 0 || (found != nullptr && kth_true != nullptr),
                "threshold_recall_at_k: null rows");
        double sum = 0.0;
        for (uint64_t q = 0; q < nq; ++q) {
            const uint64_t count = found_counts ? std::min<uint64_t>(found_counts[q], found_stride)
                                                : found_stride;
            sum += vdx::threshold_recall_one(vdx::as_neighbors(found + q * found_stride), count,
                                             kth_true[q], k, tolerance);
        }
        *out = nq ? sum / static_cast<double>(nq) : 0.0;
    });
}

}  // extern "C"
