#pragma once

// The host-side HNSW graph: storage in exactly the FORMAT v1 section layout (docs/FORMAT.md),
// the multithreaded CPU build (a restatement of vecdex's HnswIndex::add, src/hnsw.cpp:223-409),
// the structural checker (src/hnsw.cpp:527-594) and FORMAT v1 save/load (src/hnsw_io.cpp).
//
// Storage is flat and in file order, so the GPU upload reads the sections in place:
//   levels   u8[n]
//   layer0   u32[n * (1 + 2M)]        blocks [count, id_1 .. id_2M]
//   upper    u32[U * (1 + M)]         per node in id order, one block per layer 1..level
//   vectors  f32[n * dim]             normalized under cosine
// plus upper_offset[n], node i's first upper block (the prefix sum of levels).
//
// Search is not here: queries go to the GPU. The build's own layer search (used while
// inserting) is a private member.

#include <atomic>
#include <cstddef>
#include <cstdint>
#include <memory>
#include <mutex>
#include <random>
#include <string>
#include <vector>

#include "common.hpp"

namespace vdx {

struct GraphParams {
    std::size_t M = 16;
    std::size_t ef_construction = 200;
    std::uint64_t seed = 42;
    static constexpr std::size_t kMaxM = 10'000;
};

struct GraphReport {
    std::vector<std::string> violations;
    std::size_t reachable = 0;
    std::size_t max_level = 0;
    double mean_degree0 = 0.0;
};

class HnswGraph {
public:
    HnswGraph(std::size_t dim, Metric metric, GraphParams params);
    ~HnswGraph();
    HnswGraph(const HnswGraph&) = delete;
    HnswGraph& operator=(const HnswGraph&) = delete;

    std::size_t dim() const noexcept { return dim_; }
    std::size_t size() const noexcept { return levels_.size(); }
    Metric metric() const noexcept { return metric_; }
    const GraphParams& params() const noexcept { return params_; }
    std::size_t max_level() const noexcept { return max_level_; }
    Id entry_point() const noexcept { return entry_point_; }
    std::size_t upper_blocks() const noexcept { return upper_.size() / (1 + params_.M); }

    const std::vector<std::uint8_t>& levels() const noexcept { return levels_; }
    const std::vector<Id>& layer0() const noexcept { return layer0_; }
    const std::vector<Id>& upper() const noexcept { return upper_; }
    const std::vector<std::uint32_t>& upper_offset() const noexcept { return upper_offset_; }
    const std::vector<float>& vectors() const noexcept { return vectors_; }

    // Appends `rows` vectors and links them (src/hnsw.cpp:247-291). num_threads 0 means one
    // per hardware thread; 1 builds exactly what adding the rows one by one builds.
    void add(const float* data, std::size_t rows, std::size_t num_threads);

    GraphReport inspect() const;
    std::size_t memory_bytes() const noexcept;

    void save(const std::string& path) const;
    static std::unique_ptr<HnswGraph> load(const std::string& path);

    // Bumped by every add(); a device copy remembers the version it was made from.
    std::uint64_t version() const noexcept { return version_; }
    // Nodes whose link blocks (any layer) were written by an add() after `version`: the nodes
    // it stored and the existing nodes it linked back to, in id order.
    std::vector<Id> changed_since(std::uint64_t version) const;

    struct Scratch;
    class ScratchPool;
    struct BuildLocks;

private:
    enum class Locking { None, PerNode };

    std::size_t max_links(std::size_t layer) const noexcept {
        return layer == 0 ? 2 * params_.M : params_.M;
    }
    const Id* block(Id node, std::size_t layer) const noexcept;
    Id* block(Id node, std::size_t layer) noexcept;
    const float* vector_of(Id node) const noexcept { return vectors_.data() + node * dim_; }
    float distance(const float* a, const float* b) const noexcept;

    std::size_t random_level();
    void store(const float* vector);
    template <Locking L>
    void insert(Id node);
    template <Locking L>
    void link_back(Id node, std::size_t layer, const std::vector<Neighbor>& chosen);
    std::vector<Neighbor> select_neighbors(std::vector<Neighbor> candidates, std::size_t m) const;
    template <Locking L>
    const Id* read_links(Id node, std::size_t layer, std::vector<Id>& copy,
                         std::size_t& count) const;
    template <Locking L>
    Neighbor greedy_closest(const float* query, Neighbor entry, std::size_t layer,
                            Scratch& scratch) const;
    template <Locking L>
    void search_layer(const float* query, Neighbor entry, std::size_t ef, std::size_t layer,
                      Scratch& scratch) const;
    void check_links(const std::string& path) const;

    std::size_t dim_;
    Metric metric_;
    GraphParams params_;
    double level_multiplier_;
    std::mt19937_64 rng_;

    std::vector<std::uint8_t> levels_;
    std::vector<Id> layer0_;
    std::vector<Id> upper_;
    std::vector<std::uint32_t> upper_offset_;
    std::vector<float> vectors_;
    std::vector<std::uint32_t> touched_;  // per node: the version that last wrote its links
    Id entry_point_ = 0;
    std::size_t max_level_ = 0;
    std::uint64_t version_ = 0;

    std::unique_ptr<ScratchPool> scratch_;
    std::unique_ptr<BuildLocks> build_locks_;
};

// Calls fn(i) for every i in [0, count) on up to `threads` threads, the caller among them,
// claiming items from a shared counter; the first exception is rethrown after all threads stop
// (the scheduling of vecdex's parallel_for, src/hnsw.cpp:84-127).
template <typename Fn>
void parallel_for(std::size_t count, std::size_t threads, const Fn& fn);

std::size_t resolve_threads(std::size_t num_threads) noexcept;

}  // namespace vdx

#include "parallel_for.hpp"
