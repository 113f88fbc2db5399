#include "xxh64.hpp"

#include <cstring>

namespace vdx {
namespace {

constexpr std::uint64_t P1 = 0x9E3779B185EBCA87ULL;
constexpr std::uint64_t P2 = 0xC2B2AE3D27D4EB4FULL;
constexpr std::uint64_t P3 = 0x165667B19E3779F9ULL;
constexpr std::uint64_t P4 = 0x85EBCA77C2B2AE63ULL;
constexpr std::uint64_t P5 = 0x27D4EB2F165667C5ULL;

inline std::uint64_t rotl(std::uint64_t x, int r) noexcept { return (x << r) | (x >> (64 - r)); }

inline std::uint64_t le64(const unsigned char* p) noexcept {
    std::uint64_t v;
    std::memcpy(&v, p, 8);  // the host is little-endian, as the format is
    return v;
}

inline std::uint32_t le32(const unsigned char* p) noexcept {
    std::uint32_t v;
    std::memcpy(&v, p, 4);
    return v;
}

// One lane step of the specification: acc + lane * P2, rotated by 31, times P1.
inline std::uint64_t lane_step(std::uint64_t acc, std::uint64_t lane) noexcept {
    return rotl(acc + lane * P2, 31) * P1;
}

}  // namespace

Xxh64::Xxh64(std::uint64_t seed) noexcept : seed_(seed) {
    v_[0] = seed + P1 + P2;
    v_[1] = seed + P2;
    v_[2] = seed;
    v_[3] = seed - P1;
}

void Xxh64::round4(const unsigned char* stripe) noexcept {
    for (int i = 0; i < 4; ++i) {
        v_[i] = lane_step(v_[i], le64(stripe + 8 * i));
    }
}

void Xxh64::update(const void* data, std::size_t size) noexcept {
    const auto* p = static_cast<const unsigned char*>(data);
    total_ += size;
    if (buffered_ != 0) {
        const std::size_t need = 32 - buffered_;
        const std::size_t take = size < need ? size : need;
        std::memcpy(buf_ + buffered_, p, take);
        buffered_ += take;
        p += take;
        size -= take;
        if (buffered_ < 32) {
            return;
        }
        round4(buf_);
        buffered_ = 0;
    }
    while (size >= 32) {
        round4(p);
        p += 32;
        size -= 32;
    }
    if (size != 0) {
        std::memcpy(buf_, p, size);
        buffered_ = size;
    }
}

std::uint64_t Xxh64::digest() const noexcept {
    std::uint64_t h;
    if (total_ >= 32) {
        h = rotl(v_[0], 1) + rotl(v_[1], 7) + rotl(v_[2], 12) + rotl(v_[3], 18);
        for (int i = 0; i < 4; ++i) {
            h ^= lane_step(0, v_[i]);
            h = h * P1 + P4;
        }
    } else {
        h = seed_ + P5;
    }
    h += total_;
    const unsigned char* p = buf_;
    std::size_t left = buffered_;
    while (left >= 8) {
        h ^= lane_step(0, le64(p));
        h = rotl(h, 27) * P1 + P4;
        p += 8;
        left -= 8;
    }
    if (left >= 4) {
        h ^= static_cast<std::uint64_t>(le32(p)) * P1;
        h = rotl(h, 23) * P2 + P3;
        p += 4;
        left -= 4;
    }
    while (left > 0) {
        h ^= static_cast<std::uint64_t>(*p) * P5;
        h = rotl(h, 11) * P1;
        ++p;
        --left;
    }
    h ^= h >> 33;
    h *= P2;
    h ^= h >> 29;
    h *= P3;
    h ^= h >> 32;
    return h;
}

std::uint64_t xxh64(const void* data, std::size_t size, std::uint64_t seed) noexcept {
    Xxh64 h(seed);
    h.update(data, size);
    return h.digest();
}

}  // namespace vdx
