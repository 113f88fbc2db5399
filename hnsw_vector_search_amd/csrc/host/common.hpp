#pragma once

// Core types shared by the host code and the C ABI: ids, result records, the ranking order,
// metrics, and the error types the C ABI maps to status codes.

#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <utility>

namespace vdx {

// A stored vector's id: its position in insertion order (vecdex include/vecdex/neighbor.hpp:10-12).
using Id = std::uint32_t;

// One search result, layout-identical to vecdex::Neighbor {float distance; Id id}
// (include/vecdex/neighbor.hpp:13-18), so result arrays cross the C ABI unconverted.
struct Neighbor {
    float distance;
    Id id;
    friend bool operator==(const Neighbor&, const Neighbor&) = default;
};
static_assert(sizeof(Neighbor) == 8);

// The total order of every heap, sort and merge: smaller distance, ties by smaller id
// (include/vecdex/neighbor.hpp:22-24).
constexpr bool closer(const Neighbor& a, const Neighbor& b) noexcept {
    return a.distance < b.distance || (a.distance == b.distance && a.id < b.id);
}

// The metric codes are the FORMAT v1 header codes (docs/FORMAT.md:38): 0 L2, 1 inner product
// (1 - <a, b>), 2 cosine (1 - <a, b> on unit vectors).
enum class Metric : std::uint32_t { L2 = 0, InnerProduct = 1, Cosine = 2 };

inline bool valid_metric(std::uint32_t code) noexcept { return code <= 2; }

// printf-style formatting into a std::string; this toolchain's library has no <format>.
template <typename... Args>
std::string fmt(const char* format, Args... args) {
    const int size = std::snprintf(nullptr, 0, format, args...);
    if (size <= 0) {
        return std::string(format);
    }
    std::string out(static_cast<std::size_t>(size) + 1, '\0');
    std::snprintf(out.data(), out.size(), format, args...);
    out.resize(static_cast<std::size_t>(size));
    return out;
}

// A failure on the device or in the HIP runtime; the C ABI reports it as VDX_DEVICE_ERROR.
class device_error : public std::runtime_error {
public:
    using std::runtime_error::runtime_error;
};

}  // namespace vdx
