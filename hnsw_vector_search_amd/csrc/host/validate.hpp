#pragma once

// Boundary validation and cosine normalization, with vecdex's rules and messages
// (src/vector_ops.cpp:14-86), so the GPU path accepts and rejects exactly what vecdex does.
// Every check runs on the host before anything is copied to the device; a rejected call
// changes nothing.

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "common.hpp"

namespace vdx {

bool all_finite(const float* values, std::size_t count) noexcept;

// sqrt(<v, v>) > 0 and finite, with <v, v> in the library's summation order
// (src/vector_ops.cpp:14-27).
bool can_normalize(const float* v, std::size_t dim) noexcept;

// v /= sqrt(<v, v>) element by element (src/vector_ops.cpp:29-34).
void normalize(float* v, std::size_t dim) noexcept;

struct IndexShape {
    Metric metric;
    std::size_t dim;
    std::size_t size;
};

// src/vector_ops.cpp:36-58. `count` rows of `given_dim` floats each, back to back.
void check_new_vectors(const std::string& where, const IndexShape& index, std::size_t given_dim,
                       const float* values, std::size_t count);

// src/vector_ops.cpp:60-76.
void check_query(const std::string& where, const IndexShape& index, const float* query,
                 std::size_t query_dim, std::size_t k);

// The query as the index compares it: a normalized copy under cosine, the query otherwise
// (src/vector_ops.cpp:78-86). Writes dim floats to `out`.
void prepare_query(Metric metric, const float* query, std::size_t dim, float* out) noexcept;

}  // namespace vdx
