#include "validate.hpp"

#include <cmath>
#include <cstring>
#include <limits>

#include "cpu_distance.hpp"

namespace vdx {
namespace {

float norm(const float* v, std::size_t dim) noexcept {
    return std::sqrt(cpu::inner_product(v, v, dim));
}

}  // namespace

bool all_finite(const float* values, std::size_t count) noexcept {
    for (std::size_t i = 0; i < count; ++i) {
        if (!std::isfinite(values[i])) {
            return false;
        }
    }
    return true;
}

bool can_normalize(const float* v, std::size_t dim) noexcept {
    const float n = norm(v, dim);
    return n > 0.0F && std::isfinite(n);
}

void normalize(float* v, std::size_t dim) noexcept {
    const float n = norm(v, dim);
    for (std::size_t i = 0; i < dim; ++i) {
        v[i] /= n;
    }
}

void check_new_vectors(const std::string& where, const IndexShape& index, std::size_t given_dim,
                       const float* values, std::size_t count) {
    if (given_dim != index.dim) {
        throw std::invalid_argument(fmt("%s: got dimension %zu, index has %zu", where.c_str(),
                                        given_dim, index.dim));
    }
    if (count > std::numeric_limits<Id>::max() - index.size) {
        throw std::length_error(fmt("%s: too many vectors for 32-bit ids", where.c_str()));
    }
    if (!all_finite(values, count * index.dim)) {
        throw std::invalid_argument(
            fmt("%s: vectors must not contain NaN or inf", where.c_str()));
    }
    if (index.metric == Metric::Cosine) {
        for (std::size_t r = 0; r < count; ++r) {
            if (!can_normalize(values + r * index.dim, index.dim)) {
                throw std::invalid_argument(
                    fmt("%s: vector %zu has zero length, which the cosine metric can't normalize",
                        where.c_str(), r));
            }
        }
    }
}

void check_query(const std::string& where, const IndexShape& index, const float* query,
                 std::size_t query_dim, std::size_t k) {
    if (query_dim != index.dim) {
        throw std::invalid_argument(fmt("%s: query has dimension %zu, index has %zu",
                                        where.c_str(), query_dim, index.dim));
    }
    if (k == 0) {
        throw std::invalid_argument(fmt("%s: k must be positive", where.c_str()));
    }
    if (!all_finite(query, query_dim)) {
        throw std::invalid_argument(fmt("%s: query must not contain NaN or inf", where.c_str()));
    }
    if (index.metric == Metric::Cosine && !can_normalize(query, query_dim)) {
        throw std::invalid_argument(
            fmt("%s: the cosine metric can't normalize a zero query", where.c_str()));
    }
}

void prepare_query(Metric metric, const float* query, std::size_t dim, float* out) noexcept {
    if (out != query) {
        std::memcpy(out, query, dim * sizeof(float));
    }
    if (metric == Metric::Cosine) {
        normalize(out, dim);
    }
}

}  // namespace vdx
