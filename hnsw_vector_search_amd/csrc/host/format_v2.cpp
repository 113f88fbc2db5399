// FORMAT v2 (format_v2.hpp): writer, reader with streaming checks, and the host graph's
// v2 load. The checks are v1's (docs/FORMAT.md:104-117, src/hnsw_io.cpp:409-504) with v1's
// messages where they coincide, plus the layout, checksum and padding rules of v2.

#include "format_v2.hpp"

#include <fcntl.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstring>
#include <limits>

#include "file_io.hpp"
#include "graph.hpp"
#include "validate.hpp"

namespace vdx {
namespace {

constexpr unsigned char kMagic[8] = {0x89, 'V', 'D', 'X', 'H', 'N', 'S', 'W'};
constexpr std::uint32_t kVersion2 = 2;
constexpr std::uint64_t kTableAt = 96, kHeaderSumAt = 216, kHeaderUsed = 224;
constexpr const char* kSectionName[kV2Sections] = {"levels", "upper offsets", "layer-0 links", "upper links",
                                                   "vectors"};
using ull = unsigned long long;

std::uint64_t round_up(std::uint64_t v, std::uint64_t a, const std::string& path) {
    return mul_checked(add_checked(v, a - 1, path) / a, a, path);
}

template <typename T>
T read_at(const unsigned char* p, std::uint64_t off) {
    T v;
    std::memcpy(&v, p + off, sizeof(T));
    return v;
}

template <typename T>
void put_at(unsigned char* p, std::uint64_t off, T v) {
    std::memcpy(p + off, &v, sizeof(T));
}

void encode_header(const V2Header& h, unsigned char* out) {
    std::memset(out, 0, kV2HeaderBytes);
    std::memcpy(out, kMagic, 8);
    put_at<std::uint32_t>(out, 8, kVersion2);
    put_at<std::uint32_t>(out, 12, h.metric);
    const std::uint64_t fields[] = {h.dim, h.M, h.ef_construction, h.seed, h.n, h.max_level, h.entry_point,
                                    h.upper_blocks, h.ld, kV2Align};
    for (int i = 0; i < 10; ++i) put_at<std::uint64_t>(out, 16 + 8 * i, fields[i]);
    for (int s = 0; s < kV2Sections; ++s) {
        put_at<std::uint64_t>(out, kTableAt + 24 * s, h.offset[s]);
        put_at<std::uint64_t>(out, kTableAt + 24 * s + 8, h.bytes[s]);
        put_at<std::uint64_t>(out, kTableAt + 24 * s + 16, h.checksum[s]);
    }
    put_at<std::uint64_t>(out, kHeaderSumAt, xxh64(out, kHeaderSumAt));
}

}  // namespace

void v2_layout(V2Header& h, const std::string& path) {
    const std::uint64_t b0 = 1 + 2 * h.M, bu = 1 + h.M;
    h.bytes[kV2Levels] = h.n;
    h.bytes[kV2Offsets] = mul_checked(h.n, 4, path);
    h.bytes[kV2Layer0] = mul_checked(mul_checked(h.n, b0, path), 4, path);
    h.bytes[kV2Upper] = mul_checked(mul_checked(h.upper_blocks, bu, path), 4, path);
    h.bytes[kV2Vectors] = mul_checked(mul_checked(h.n, h.ld, path), 4, path);
    std::uint64_t at = kV2HeaderBytes;
    for (int s = 0; s < kV2Sections; ++s) {
        h.offset[s] = at;
        at = round_up(add_checked(at, h.bytes[s], path), kV2Align, path);
    }
}

V2File::V2File(const std::string& path) : path_(path) {
    fd_ = ::open(path.c_str(), O_RDONLY);
    if (fd_ < 0) {
        fail_path(path, std::strerror(errno));
    }
    struct Close {
        V2File* f;
        bool armed = true;
        ~Close() {
            if (armed) {
                ::close(f->fd_);
                f->fd_ = -1;
            }
        }
    } close_on_error{this};
    struct stat st {};
    if (::fstat(fd_, &st) != 0 || !S_ISREG(st.st_mode)) {
        fail_path(path, "not a regular file");
    }
    size_ = static_cast<std::uint64_t>(st.st_size);
    if (size_ < kV2HeaderBytes) {
        fail_path(path, fmt("%llu bytes is too small for a version 2 vecdex index", static_cast<ull>(size_)));
    }
    (void)::posix_fadvise(fd_, 0, 0, POSIX_FADV_SEQUENTIAL);
    std::vector<unsigned char> header(kV2HeaderBytes);
    read_bytes(0, header.data(), kV2HeaderBytes);

    const unsigned char* b = header.data();
    if (std::memcmp(b, kMagic, 8) != 0) {
        fail_path(path, "not a vecdex HNSW index (wrong magic bytes)");
    }
    const std::uint32_t version = read_at<std::uint32_t>(b, 8);
    if (version != kVersion2) {
        fail_path(path, fmt("format version %u: this loader reads version 2 files (HnswIndex::load reads "
                            "versions 1 and 2)",
                            version));
    }
    if (read_at<std::uint64_t>(b, kHeaderSumAt) != xxh64(b, kHeaderSumAt)) {
        fail_path(path, "header checksum mismatch: the file is corrupted");
    }
    for (std::uint64_t i = kHeaderUsed; i < kV2HeaderBytes; ++i) {
        if (b[i] != 0) {
            fail_path(path, "the header's reserved bytes are not zero");
        }
    }
    h_.metric = read_at<std::uint32_t>(b, 12);
    h_.dim = read_at<std::uint64_t>(b, 16);
    h_.M = read_at<std::uint64_t>(b, 24);
    h_.ef_construction = read_at<std::uint64_t>(b, 32);
    h_.seed = read_at<std::uint64_t>(b, 40);
    h_.n = read_at<std::uint64_t>(b, 48);
    h_.max_level = read_at<std::uint64_t>(b, 56);
    h_.entry_point = read_at<std::uint64_t>(b, 64);
    h_.upper_blocks = read_at<std::uint64_t>(b, 72);
    h_.ld = read_at<std::uint64_t>(b, 80);
    const std::uint64_t align = read_at<std::uint64_t>(b, 88);
    if (!valid_metric(h_.metric)) {
        fail_path(path, fmt("unknown metric code %u", h_.metric));
    }
    if (h_.dim == 0) {
        fail_path(path, "dimension is 0");
    }
    if (h_.M < 2) {
        fail_path(path, fmt("M is %llu, below the minimum of 2", static_cast<ull>(h_.M)));
    }
    if (h_.M > GraphParams::kMaxM) {
        fail_path(path, fmt("M is %llu, above the maximum of %zu", static_cast<ull>(h_.M), GraphParams::kMaxM));
    }
    if (h_.ef_construction == 0) {
        fail_path(path, "ef_construction is 0");
    }
    if (h_.n > std::numeric_limits<Id>::max()) {
        fail_path(path, fmt("%llu nodes is more than 32-bit ids can address", static_cast<ull>(h_.n)));
    }
    if (h_.max_level > 255) {
        fail_path(path, fmt("max level %llu is above 255", static_cast<ull>(h_.max_level)));
    }
    if (h_.n == 0 && (h_.max_level != 0 || h_.entry_point != 0 || h_.upper_blocks != 0)) {
        fail_path(path, "an empty index must have max level, entry point and upper blocks 0");
    }
    if (h_.n > 0 && h_.entry_point >= h_.n) {
        fail_path(path, fmt("entry point %llu is not one of the %llu nodes", static_cast<ull>(h_.entry_point),
                            static_cast<ull>(h_.n)));
    }
    if (h_.upper_blocks > std::numeric_limits<std::uint32_t>::max()) {
        fail_path(path, "too many upper-layer blocks for this build's 32-bit offsets");
    }
    if (h_.ld != round_up(h_.dim, 4, path)) {
        fail_path(path, fmt("vector row length %llu is not the dimension %llu rounded up to a multiple of 4",
                            static_cast<ull>(h_.ld), static_cast<ull>(h_.dim)));
    }
    if (align != kV2Align) {
        fail_path(path, fmt("section alignment %llu is not %llu", static_cast<ull>(align),
                            static_cast<ull>(kV2Align)));
    }
    V2Header expect = h_;
    v2_layout(expect, path);
    for (int s = 0; s < kV2Sections; ++s) {
        h_.offset[s] = read_at<std::uint64_t>(b, kTableAt + 24 * s);
        h_.bytes[s] = read_at<std::uint64_t>(b, kTableAt + 24 * s + 8);
        h_.checksum[s] = read_at<std::uint64_t>(b, kTableAt + 24 * s + 16);
        if (h_.offset[s] != expect.offset[s] || h_.bytes[s] != expect.bytes[s]) {
            fail_path(path, fmt("the %s section is at %llu (%llu bytes), the header's sizes put it at %llu "
                                "(%llu bytes)",
                                kSectionName[s], static_cast<ull>(h_.offset[s]), static_cast<ull>(h_.bytes[s]),
                                static_cast<ull>(expect.offset[s]), static_cast<ull>(expect.bytes[s])));
        }
    }
    const std::uint64_t end = h_.offset[kV2Vectors] + h_.bytes[kV2Vectors];
    if (end != size_) {
        fail_path(path, fmt("the header describes %llu bytes but the file has %llu; it is truncated or corrupted",
                            static_cast<ull>(end), static_cast<ull>(size_)));
    }
    close_on_error.armed = false;
}

V2File::~V2File() {
    if (fd_ >= 0) {
        ::close(fd_);
    }
}

void V2File::read_bytes(std::uint64_t at, void* dst, std::uint64_t bytes) const {
    auto* out = static_cast<unsigned char*>(dst);
    while (bytes > 0) {
        const std::size_t want = static_cast<std::size_t>(std::min<std::uint64_t>(bytes, std::uint64_t{1} << 30));
        const ssize_t got = ::pread(fd_, out, want, static_cast<off_t>(at));
        if (got < 0 && errno == EINTR) {
            continue;
        }
        if (got < 0) {
            fail_path(path_, std::string("read failed: ") + std::strerror(errno));
        }
        if (got == 0) {
            fail_path(path_, "the file ended early; it is truncated or was changed while loading");
        }
        out += got;
        at += static_cast<std::uint64_t>(got);
        bytes -= static_cast<std::uint64_t>(got);
    }
}

void V2File::read(int s, std::uint64_t at, void* dst, std::uint64_t bytes) const {
    if (at > h_.bytes[s] || bytes > h_.bytes[s] - at) {
        throw std::logic_error("V2File::read past the end of a section");
    }
    read_bytes(h_.offset[s] + at, dst, bytes);
}

V2Checker::V2Checker(const V2File& file) : map_(file), h_(file.header()) {}

void V2Checker::levels_and_offsets() {
    const std::string& path = map_.path();
    const std::uint64_t n = h_.n;
    levels_.resize(n);
    map_.read(kV2Levels, 0, levels_.data(), n);
    hash_[kV2Levels].update(levels_.data(), n);
    offsets_.resize(n);
    map_.read(kV2Offsets, 0, offsets_.data(), n * 4);
    hash_[kV2Offsets].update(offsets_.data(), n * 4);
    std::uint64_t blocks = 0;
    for (std::uint64_t i = 0; i < n; ++i) {
        if (levels_[i] > h_.max_level) {
            fail_path(path, fmt("node %llu has level %u, above the max level %llu", static_cast<ull>(i),
                                static_cast<unsigned>(levels_[i]), static_cast<ull>(h_.max_level)));
        }
        if (offsets_[i] != blocks) {
            fail_path(path, fmt("node %llu's first upper block is recorded as %u, its levels put it at %llu",
                                static_cast<ull>(i), offsets_[i], static_cast<ull>(blocks)));
        }
        blocks += levels_[i];
    }
    if (blocks != h_.upper_blocks) {
        fail_path(path, fmt("the levels add up to %llu upper-layer blocks, but the header says %llu",
                            static_cast<ull>(blocks), static_cast<ull>(h_.upper_blocks)));
    }
    if (n > 0 && levels_[h_.entry_point] != h_.max_level) {
        fail_path(path, fmt("entry point %llu has level %u, not the max level %llu", static_cast<ull>(h_.entry_point),
                            static_cast<unsigned>(levels_[h_.entry_point]), static_cast<ull>(h_.max_level)));
    }
    next_[kV2Levels] = next_[kV2Offsets] = n;
    marks_.assign(n, 0u);
}

void V2Checker::advance(int s, std::uint64_t expect_first, std::uint64_t first, const void* data,
                        std::uint64_t bytes) {
    if (levels_.size() != h_.n || next_[kV2Levels] != h_.n) {
        throw std::logic_error("V2Checker: levels_and_offsets() must come first");
    }
    if (first != expect_first) {
        throw std::logic_error("V2Checker: chunks must be consecutive");
    }
    hash_[s].update(data, bytes);
}

// The link rules of check_links (src/hnsw_io.cpp:472-504), with its messages.
void V2Checker::links(const std::uint32_t* b, std::uint64_t node, std::uint64_t layer, std::uint64_t cap) {
    const std::string& path = map_.path();
    const std::uint64_t count = b[0];
    if (count > cap) {
        fail_path(path, fmt("node %llu has %llu links on layer %llu (max %llu)", static_cast<ull>(node),
                            static_cast<ull>(count), static_cast<ull>(layer), static_cast<ull>(cap)));
    }
    if (++epoch_ == 0) {
        std::fill(marks_.begin(), marks_.end(), 0u);
        epoch_ = 1;
    }
    for (std::uint64_t i = 0; i < count; ++i) {
        const std::uint32_t id = b[1 + i];
        if (id >= h_.n) {
            fail_path(path, fmt("node %llu links to node %u, but there are only %llu nodes", static_cast<ull>(node), id,
                                static_cast<ull>(h_.n)));
        }
        if (id == node) {
            fail_path(path, fmt("node %llu links to itself on layer %llu", static_cast<ull>(node), static_cast<ull>(layer)));
        }
        if (levels_[id] < layer) {
            fail_path(path, fmt("node %llu links to node %u on layer %llu, above its level %u", static_cast<ull>(node), id,
                                static_cast<ull>(layer), static_cast<unsigned>(levels_[id])));
        }
        if (marks_[id] == epoch_) {
            fail_path(path, fmt("node %llu links to node %u twice on layer %llu", static_cast<ull>(node), id,
                                static_cast<ull>(layer)));
        }
        marks_[id] = epoch_;
    }
}

void V2Checker::layer0(const std::uint32_t* blocks, std::uint64_t first_node, std::uint64_t nodes) {
    const std::uint64_t b0 = 1 + 2 * h_.M;
    if (first_node + nodes > h_.n) throw std::logic_error("V2Checker: layer-0 chunk past the end");
    advance(kV2Layer0, next_[kV2Layer0], first_node, blocks, nodes * b0 * 4);
    for (std::uint64_t i = 0; i < nodes; ++i) {
        links(blocks + i * b0, first_node + i, 0, 2 * h_.M);
    }
    next_[kV2Layer0] += nodes;
}

void V2Checker::upper(const std::uint32_t* blocks, std::uint64_t first_block, std::uint64_t count) {
    const std::uint64_t bu = 1 + h_.M;
    if (first_block + count > h_.upper_blocks) throw std::logic_error("V2Checker: upper chunk past the end");
    advance(kV2Upper, next_[kV2Upper], first_block, blocks, count * bu * 4);
    for (std::uint64_t i = 0; i < count; ++i) {
        const std::uint64_t blk = first_block + i;
        while (offsets_[upper_node_] + levels_[upper_node_] <= blk) {
            ++upper_node_;  // blocks are in node order; level-0 nodes own none
        }
        links(blocks + i * bu, upper_node_, blk - offsets_[upper_node_] + 1, h_.M);
    }
    next_[kV2Upper] += count;
}

void V2Checker::vectors(const float* rows, std::uint64_t first_row, std::uint64_t count) {
    const std::string& path = map_.path();
    if (first_row + count > h_.n) throw std::logic_error("V2Checker: vector chunk past the end");
    advance(kV2Vectors, next_[kV2Vectors], first_row, rows, count * h_.ld * 4);
    for (std::uint64_t r = 0; r < count; ++r) {
        const float* v = rows + r * h_.ld;
        if (!all_finite(v, h_.dim)) {
            fail_path(path, "a stored vector contains NaN or infinity");
        }
        for (std::uint64_t c = h_.dim; c < h_.ld; ++c) {
            std::uint32_t bits;
            std::memcpy(&bits, v + c, 4);
            if (bits != 0) {
                fail_path(path, fmt("vector %llu has a nonzero padding component", static_cast<ull>(first_row + r)));
            }
        }
    }
    next_[kV2Vectors] += count;
}

void V2Checker::finish() {
    const std::uint64_t units[kV2Sections] = {h_.n, h_.n, h_.n, h_.upper_blocks, h_.n};
    for (int s = 0; s < kV2Sections; ++s) {
        if (next_[s] != units[s]) {
            throw std::logic_error("V2Checker: a section was not checked in full");
        }
        if (hash_[s].digest() != h_.checksum[s]) {
            fail_path(map_.path(), fmt("%s section checksum mismatch: the file is corrupted", kSectionName[s]));
        }
    }
}

void HnswGraph::save_v2(const std::string& path) const {
    V2Header h;
    h.metric = static_cast<std::uint32_t>(metric_);
    h.dim = dim_;
    h.M = params_.M;
    h.ef_construction = params_.ef_construction;
    h.seed = params_.seed;
    h.n = size();
    h.max_level = size() == 0 ? 0 : max_level_;
    h.entry_point = size() == 0 ? 0 : entry_point_;
    h.upper_blocks = upper_blocks();
    h.ld = (dim_ + 3) / 4 * 4;
    v2_layout(h, path);
    save_atomically(path, [&](std::ofstream& out) {
        std::vector<unsigned char> header(kV2HeaderBytes, 0);
        out.write(reinterpret_cast<const char*>(header.data()), kV2HeaderBytes);  // rewritten at the end
        std::uint64_t at = kV2HeaderBytes;
        const std::vector<char> zeros(kV2Align, 0);
        const auto section = [&](int s, const auto& write_body) {
            out.write(zeros.data(), static_cast<std::streamsize>(h.offset[s] - at));
            Xxh64 hash;
            write_body([&](const void* p, std::size_t bytes) {
                hash.update(p, bytes);
                out.write(static_cast<const char*>(p), static_cast<std::streamsize>(bytes));
            });
            h.checksum[s] = hash.digest();
            at = h.offset[s] + h.bytes[s];
        };
        section(kV2Levels, [&](auto put) { put(levels_.data(), levels_.size()); });
        section(kV2Offsets, [&](auto put) { put(upper_offset_.data(), upper_offset_.size() * 4); });
        section(kV2Layer0, [&](auto put) { put(layer0_.data(), layer0_.size() * 4); });
        section(kV2Upper, [&](auto put) { put(upper_.data(), upper_.size() * 4); });
        section(kV2Vectors, [&](auto put) {
            std::vector<float> row(h.ld, 0.0f);
            for (std::size_t i = 0; i < size(); ++i) {
                std::memcpy(row.data(), vector_of(static_cast<Id>(i)), dim_ * 4);
                put(row.data(), h.ld * 4);
            }
        });
        encode_header(h, header.data());
        out.seekp(0);
        out.write(reinterpret_cast<const char*>(header.data()), kV2HeaderBytes);
    });
}

std::unique_ptr<HnswGraph> HnswGraph::load_v2(const std::string& path) {
    // Every section is read into memory this load owns and checked there, so what is checked is
    // exactly what the graph keeps, whatever happens to the file meanwhile.
    V2File file(path);
    const V2Header& h = file.header();
    V2Checker check(file);
    check.levels_and_offsets();
    auto g = std::make_unique<HnswGraph>(h.dim, static_cast<Metric>(h.metric),
                                         GraphParams{h.M, h.ef_construction, h.seed});
    g->layer0_.resize(h.bytes[kV2Layer0] / 4);
    file.read(kV2Layer0, 0, g->layer0_.data(), h.bytes[kV2Layer0]);
    check.layer0(g->layer0_.data(), 0, h.n);
    g->upper_.resize(h.bytes[kV2Upper] / 4);
    file.read(kV2Upper, 0, g->upper_.data(), h.bytes[kV2Upper]);
    check.upper(g->upper_.data(), 0, h.upper_blocks);
    g->vectors_.resize(h.n * h.dim);
    const std::uint64_t chunk = std::max<std::uint64_t>(1, (std::uint64_t{16} << 20) / (h.ld * 4));
    std::vector<float> rows(std::min(chunk, h.n) * h.ld);
    for (std::uint64_t r0 = 0; r0 < h.n; r0 += chunk) {
        const std::uint64_t m = std::min(chunk, h.n - r0);
        file.read(kV2Vectors, r0 * h.ld * 4, rows.data(), m * h.ld * 4);
        check.vectors(rows.data(), r0, m);
        for (std::uint64_t i = 0; i < m; ++i) {
            std::memcpy(g->vectors_.data() + (r0 + i) * h.dim, rows.data() + i * h.ld, h.dim * 4);
        }
    }
    check.finish();
    g->levels_ = check.levels();
    g->upper_offset_ = check.offsets();
    g->touched_.assign(h.n, 0u);
    g->entry_point_ = static_cast<Id>(h.entry_point);
    g->max_level_ = h.max_level;
    g->rng_.discard(h.n);  // each stored node drew its level once
    return g;
}

}  // namespace vdx
