#pragma once

// Streaming XXH64 (the 64-bit xxHash), the checksum of FORMAT v1 files (docs/FORMAT.md:98-102).
// Written from the published xxHash specification (doc/xxhash_spec.md in the xxHash repository);
// pinned by xxHash's own sanity vectors in the tests.

#include <cstddef>
#include <cstdint>

namespace vdx {

class Xxh64 {
public:
    explicit Xxh64(std::uint64_t seed = 0) noexcept;
    void update(const void* data, std::size_t size) noexcept;
    std::uint64_t digest() const noexcept;

private:
    void round4(const unsigned char* stripe) noexcept;

    std::uint64_t seed_;
    std::uint64_t v_[4];
    unsigned char buf_[32];
    std::size_t buffered_ = 0;
    std::uint64_t total_ = 0;
};

std::uint64_t xxh64(const void* data, std::size_t size, std::uint64_t seed = 0) noexcept;

}  // namespace vdx
