#pragma once

// FORMAT v2: this library's aligned, mappable index file (a new version of docs/FORMAT.md's
// format under its versioning rule: same magic, version at offset 8). v1 packs the sections with
// no padding and is meant to be streamed into memory the index owns; v2 lays every section out
// as the device copy holds it, each at a 4 KiB-aligned offset, so a file can be mapped, and a
// loader copies the sections to HBM as they are, checking them chunk by chunk on the way.
//
//   offset 0     header (4096 bytes)
//     0  magic \x89VDXHNSW        8  u32 version = 2     12 u32 metric
//    16  u64 dim  24 u64 M  32 u64 ef_construction  40 u64 seed  48 u64 n
//    56  u64 max level  64 u64 entry point  72 u64 U (upper blocks)
//    80  u64 ld (vector row length in floats: dim rounded up to a multiple of 4)
//    88  u64 alignment = 4096
//    96  5 x {u64 offset, u64 bytes, u64 XXH64 of the section's bytes}, in section order
//   216  u64 XXH64 of header bytes [0, 216); bytes 224..4095 are zero
//   sections, in order, each at the first multiple of 4096 at or after the previous one's end:
//     levels        n x u8
//     upper offsets n x u32   (node i's first upper block: levels[0] + ... + levels[i-1])
//     layer 0       n x (1 + 2M) x u32
//     upper         U x (1 + M) x u32
//     vectors       n x ld x f32   (components dim..ld-1 are +0.0)
//   the file ends where the vectors section ends. Gaps between sections are zero.
//
// Every check of a v1 load applies (ranges, levels, finite vectors, link rules), plus: the
// layout is exactly the one above, both checksum kinds match, offsets equal the prefix sums of
// the levels, and padding components are +0.0.

#include <cstdint>
#include <string>
#include <vector>

#include "xxh64.hpp"

namespace vdx {

constexpr std::uint64_t kV2Align = 4096;
constexpr std::uint64_t kV2HeaderBytes = 4096;
constexpr int kV2Sections = 5;
enum V2Section : int { kV2Levels = 0, kV2Offsets = 1, kV2Layer0 = 2, kV2Upper = 3, kV2Vectors = 4 };

struct V2Header {
    std::uint32_t metric = 0;
    std::uint64_t dim = 0, M = 0, ef_construction = 0, seed = 0, n = 0, max_level = 0, entry_point = 0;
    std::uint64_t upper_blocks = 0, ld = 0;
    std::uint64_t offset[kV2Sections] = {}, bytes[kV2Sections] = {}, checksum[kV2Sections] = {};
};

// The layout a header with these fields must have (offsets and sizes filled in).
void v2_layout(V2Header& h, const std::string& path);

// The file open for reading, its header checked (magic, version, ranges, header checksum, the
// exact layout, the file size). Section bytes are copied out with read() into memory the caller
// owns (a file that shrinks meanwhile is a runtime error, not a fault); V2Checker checks those
// copies, so what a loader keeps is what was checked.
class V2File {
public:
    explicit V2File(const std::string& path);
    ~V2File();
    V2File(const V2File&) = delete;
    V2File& operator=(const V2File&) = delete;
    const V2Header& header() const noexcept { return h_; }
    // bytes [at, at + bytes) of section s into dst
    void read(int s, std::uint64_t at, void* dst, std::uint64_t bytes) const;
    const std::string& path() const noexcept { return path_; }

private:
    void read_bytes(std::uint64_t at, void* dst, std::uint64_t bytes) const;
    std::string path_;
    int fd_ = -1;
    std::uint64_t size_ = 0;
    V2Header h_;
};

// The content checks of a v2 load, in an order that lets a loader stream: levels and offsets
// first (whole), then each of layer 0, upper and vectors in consecutive chunks of whole blocks
// or rows, then finish(), which compares the streamed checksums. Every failure throws a
// std::runtime_error starting with the path.
class V2Checker {
public:
    explicit V2Checker(const V2File& file);
    void levels_and_offsets();
    void layer0(const std::uint32_t* blocks, std::uint64_t first_node, std::uint64_t nodes);
    void upper(const std::uint32_t* blocks, std::uint64_t first_block, std::uint64_t count);
    void vectors(const float* rows, std::uint64_t first_row, std::uint64_t count);
    void finish();
    const std::vector<std::uint8_t>& levels() const noexcept { return levels_; }
    const std::vector<std::uint32_t>& offsets() const noexcept { return offsets_; }

private:
    void links(const std::uint32_t* b, std::uint64_t node, std::uint64_t layer, std::uint64_t cap);
    void advance(int s, std::uint64_t expect_first, std::uint64_t first, const void* data, std::uint64_t bytes);

    const V2File& map_;
    const V2Header& h_;
    std::vector<std::uint8_t> levels_;
    std::vector<std::uint32_t> offsets_;
    std::vector<std::uint32_t> marks_;
    std::uint32_t epoch_ = 0;
    std::uint64_t next_[kV2Sections] = {};  // next unit (node, block or row) each section expects
    Xxh64 hash_[kV2Sections];
    std::uint64_t upper_node_ = 0;          // node owning the next upper block
};

}  // namespace vdx
