// FORMAT v1 (docs/FORMAT.md): the on-disk index of vecdex, written and read here from the
// specification. load() treats the file as hostile and checks it in the specification's order
// (docs/FORMAT.md:104-117, src/hnsw_io.cpp:199-284, 409-504), with the same messages, each
// starting with the path.

#include <fcntl.h>
#include <unistd.h>

#include <array>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <limits>
#include <random>
#include <system_error>

#include "file_io.hpp"
#include "graph.hpp"
#include "validate.hpp"
#include "xxh64.hpp"

namespace vdx {
namespace {

namespace fs = std::filesystem;

constexpr unsigned char kMagic[8] = {0x89, 'V', 'D', 'X', 'H', 'N', 'S', 'W'};
constexpr std::uint32_t kVersion = 1;
constexpr std::uint64_t kHeaderBytes = 80;
constexpr std::uint64_t kChecksumBytes = 8;
constexpr std::size_t kChunk = std::size_t{1} << 20;

[[noreturn]] void fail(const std::string& path, const std::string& what) { fail_path(path, what); }

// Streams bytes to a file and hashes them on the way.
class HashingWriter {
public:
    explicit HashingWriter(std::ofstream& out) : out_(out) {}
    void write(const void* data, std::size_t size) {
        const auto* p = static_cast<const char*>(data);
        for (std::size_t done = 0; done < size; done += kChunk) {
            const std::size_t n = std::min(kChunk, size - done);
            hash_.update(p + done, n);
            out_.write(p + done, static_cast<std::streamsize>(n));
        }
    }
    void u32(std::uint32_t v) { write(&v, 4); }
    void u64(std::uint64_t v) { write(&v, 8); }
    std::uint64_t digest() const { return hash_.digest(); }

private:
    std::ofstream& out_;
    Xxh64 hash_;
};

class HashingReader {
public:
    HashingReader(std::ifstream& in, const std::string& path) : in_(in), path_(path) {}
    void read(void* data, std::size_t size) {
        auto* p = static_cast<char*>(data);
        for (std::size_t done = 0; done < size; done += kChunk) {
            const std::size_t n = std::min(kChunk, size - done);
            in_.read(p + done, static_cast<std::streamsize>(n));
            if (in_.gcount() != static_cast<std::streamsize>(n)) {
                fail(path_, "unexpected end of file");
            }
            hash_.update(p + done, n);
        }
    }
    std::uint32_t u32() {
        std::uint32_t v = 0;
        read(&v, 4);
        return v;
    }
    std::uint64_t u64() {
        std::uint64_t v = 0;
        read(&v, 8);
        return v;
    }
    std::uint64_t digest() const { return hash_.digest(); }

private:
    std::ifstream& in_;
    const std::string& path_;
    Xxh64 hash_;
};

}  // namespace

void HnswGraph::save(const std::string& path) const {
    const bool empty = size() == 0;
    save_atomically(path, [&](std::ofstream& out) {
        HashingWriter w(out);
        w.write(kMagic, 8);
        w.u32(kVersion);
        w.u32(static_cast<std::uint32_t>(metric_));
        w.u64(dim_);
        w.u64(params_.M);
        w.u64(params_.ef_construction);
        w.u64(params_.seed);
        w.u64(size());
        w.u64(empty ? 0 : max_level_);
        w.u64(empty ? 0 : entry_point_);
        w.u64(upper_blocks());
        w.write(levels_.data(), levels_.size());
        w.write(layer0_.data(), layer0_.size() * sizeof(Id));
        w.write(upper_.data(), upper_.size() * sizeof(Id));
        w.write(vectors_.data(), vectors_.size() * sizeof(float));
        w.u64(w.digest());
    });
}

std::unique_ptr<HnswGraph> HnswGraph::load(const std::string& path) {
    std::error_code ec;
    const std::uintmax_t file_size = fs::file_size(path, ec);
    if (ec) {
        fail(path, ec.message());
    }
    if (file_size < kHeaderBytes + kChecksumBytes) {
        fail(path, fmt("%llu bytes is too small for a vecdex index",
                       static_cast<unsigned long long>(file_size)));
    }
    std::vector<char> stream_buffer(kChunk);
    std::ifstream in;
    in.rdbuf()->pubsetbuf(stream_buffer.data(), static_cast<std::streamsize>(kChunk));
    in.open(path, std::ios::binary);
    if (!in) {
        fail(path, "cannot open file");
    }
    HashingReader r(in, path);
    unsigned char magic[8];
    r.read(magic, 8);
    const std::uint32_t version = r.u32();
    const std::uint32_t metric = r.u32();
    const std::uint64_t dim = r.u64();
    const std::uint64_t M = r.u64();
    const std::uint64_t efc = r.u64();
    const std::uint64_t seed = r.u64();
    const std::uint64_t nodes = r.u64();
    const std::uint64_t max_level = r.u64();
    const std::uint64_t entry_point = r.u64();
    const std::uint64_t upper_blocks = r.u64();
    using ull = unsigned long long;

    if (std::memcmp(magic, kMagic, 8) != 0) {
        fail(path, "not a vecdex HNSW index (wrong magic bytes)");
    }
    if (version == 2) {
        in.close();
        return load_v2(path);
    }
    if (version != kVersion) {
        fail(path, fmt("format version %u is not supported (this build reads versions 1 and 2)", version));
    }
    if (!valid_metric(metric)) {
        fail(path, fmt("unknown metric code %u", metric));
    }
    if (dim == 0) {
        fail(path, "dimension is 0");
    }
    if (M < 2) {
        fail(path, fmt("M is %llu, below the minimum of 2", static_cast<ull>(M)));
    }
    if (M > GraphParams::kMaxM) {
        fail(path, fmt("M is %llu, above the maximum of %zu", static_cast<ull>(M),
                       GraphParams::kMaxM));
    }
    if (efc == 0) {
        fail(path, "ef_construction is 0");
    }
    if (nodes > std::numeric_limits<Id>::max()) {
        fail(path, fmt("%llu nodes is more than 32-bit ids can address", static_cast<ull>(nodes)));
    }
    if (max_level > 255) {
        fail(path, fmt("max level %llu is above 255", static_cast<ull>(max_level)));
    }
    if (nodes == 0 && (max_level != 0 || entry_point != 0 || upper_blocks != 0)) {
        fail(path, "an empty index must have max level, entry point and upper blocks 0");
    }
    if (nodes > 0 && entry_point >= nodes) {
        fail(path, fmt("entry point %llu is not one of the %llu nodes", static_cast<ull>(entry_point),
                       static_cast<ull>(nodes)));
    }
    const std::uint64_t b0 = 1 + 2 * M;
    const std::uint64_t bu = 1 + M;
    std::uint64_t expected = kHeaderBytes + kChecksumBytes;
    expected = add_checked(expected, nodes, path);
    expected = add_checked(expected, mul_checked(mul_checked(nodes, b0, path), 4, path), path);
    expected = add_checked(expected, mul_checked(mul_checked(upper_blocks, bu, path), 4, path), path);
    expected = add_checked(expected, mul_checked(mul_checked(nodes, dim, path), 4, path), path);
    if (expected != file_size) {
        fail(path, fmt("the header describes %llu bytes but the file has %llu; it is truncated or "
                       "corrupted",
                       static_cast<ull>(expected), static_cast<ull>(file_size)));
    }

    auto g = std::make_unique<HnswGraph>(dim, static_cast<Metric>(metric),
                                         GraphParams{M, efc, seed});
    g->levels_.resize(nodes);
    g->layer0_.resize(nodes * b0);
    g->upper_.resize(upper_blocks * bu);
    g->vectors_.resize(nodes * dim);
    r.read(g->levels_.data(), g->levels_.size());
    r.read(g->layer0_.data(), g->layer0_.size() * sizeof(Id));
    r.read(g->upper_.data(), g->upper_.size() * sizeof(Id));
    r.read(g->vectors_.data(), g->vectors_.size() * sizeof(float));
    const std::uint64_t computed = r.digest();
    if (r.u64() != computed) {
        fail(path, "checksum mismatch: the file is corrupted");
    }

    std::uint64_t blocks = 0;
    g->upper_offset_.resize(nodes);
    g->touched_.assign(nodes, 0u);
    for (std::size_t node = 0; node < nodes; ++node) {
        if (g->levels_[node] > max_level) {
            fail(path, fmt("node %zu has level %u, above the max level %llu", node,
                           static_cast<unsigned>(g->levels_[node]), static_cast<ull>(max_level)));
        }
        g->upper_offset_[node] = static_cast<std::uint32_t>(blocks);
        blocks += g->levels_[node];
    }
    if (blocks != upper_blocks) {
        fail(path, fmt("the levels add up to %llu upper-layer blocks, but the header says %llu",
                       static_cast<ull>(blocks), static_cast<ull>(upper_blocks)));
    }
    if (nodes > 0 && g->levels_[entry_point] != max_level) {
        fail(path, fmt("entry point %llu has level %u, not the max level %llu",
                       static_cast<ull>(entry_point), static_cast<unsigned>(g->levels_[entry_point]),
                       static_cast<ull>(max_level)));
    }
    if (upper_blocks > std::numeric_limits<std::uint32_t>::max()) {
        fail(path, "too many upper-layer blocks for this build's 32-bit offsets");
    }
    if (!all_finite(g->vectors_.data(), g->vectors_.size())) {
        fail(path, "a stored vector contains NaN or infinity");
    }
    g->entry_point_ = static_cast<Id>(entry_point);
    g->max_level_ = max_level;
    g->check_links(path);
    g->rng_.discard(nodes);  // each stored node drew its level once
    return g;
}

// Every link block checked before any link is followed (src/hnsw_io.cpp:472-504): the count
// before the ids it covers, each id before it is used as an index.
void HnswGraph::check_links(const std::string& path) const {
    const std::size_t n = size();
    std::vector<std::uint32_t> marks(n, 0);
    std::uint32_t epoch = 0;
    for (Id node = 0; node < n; ++node) {
        const std::size_t level = levels_[node];
        for (std::size_t layer = 0; layer <= level; ++layer) {
            const Id* b = block(node, layer);
            const std::size_t count = b[0];
            if (count > max_links(layer)) {
                fail(path, fmt("node %u has %zu links on layer %zu (max %zu)", node, count, layer,
                               max_links(layer)));
            }
            if (++epoch == 0) {
                std::fill(marks.begin(), marks.end(), 0U);
                epoch = 1;
            }
            for (std::size_t i = 0; i < count; ++i) {
                const Id id = b[1 + i];
                if (id >= n) {
                    fail(path, fmt("node %u links to node %u, but there are only %zu nodes", node,
                                   id, n));
                }
                if (id == node) {
                    fail(path, fmt("node %u links to itself on layer %zu", node, layer));
                }
                if (levels_[id] < layer) {
                    fail(path, fmt("node %u links to node %u on layer %zu, above its level %u",
                                   node, id, layer, static_cast<unsigned>(levels_[id])));
                }
                if (marks[id] == epoch) {
                    fail(path, fmt("node %u links to node %u twice on layer %zu", node, id, layer));
                }
                marks[id] = epoch;
            }
        }
    }
}

}  // namespace vdx
