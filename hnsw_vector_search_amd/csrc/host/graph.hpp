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

class HnswGraph;

// Runs the layer-0 insertion searches of a batch of nodes on a copy of the graph held
// elsewhere (the GPU): for each node, the greedy descent from the entry point through the upper
// layers and search_layer on layer 0 with breadth ef (src/hnsw.cpp:315-335), from the node's
// own stored vector, over the graph as of the last sync.
class BatchSearcher {
public:
    virtual ~BatchSearcher() = default;
    // Makes the copy current: `changed` lists, ascending, the nodes stored or whose link
    // blocks were written since the previous sync (all of them for the first).
    virtual void sync(const HnswGraph& g, const std::vector<Id>& changed) = 0;
    // out[i * ef ..] receives the count[i] results of nodes[i], in any order. entries[i] is a
    // layer-0 entry that replaces the descent (insert's layer-0 search starts from its layer-1
    // result, src/hnsw.cpp:324-335), or kDescend.
    // With m > 0 (only when selects()), out[i * m ..] receives select_neighbors(results, m)
    // instead: the neighbour choice made where the results are. `layer` > 0 (only when
    // upper_layers()) searches that layer, entries being nodes on it; kDescend then descends from
    // the top to layer + 1.
    static constexpr Id kDescend = 0xFFFFFFFFu;
    virtual void search(const std::vector<Id>& nodes, const std::vector<Id>& entries, std::size_t layer,
                        std::size_t ef, std::size_t m, std::vector<Neighbor>& out,
                        std::vector<std::uint32_t>& count) = 0;
    virtual bool selects() const { return false; }
    virtual bool upper_layers() const { return false; }
};

// How add_batched splits the rows: the first `host_prefix` nodes of the graph are inserted on
// the host; after that, batches of batch_fraction x (nodes linked so far), clamped to
// [batch_min, batch_max].
struct BatchPlan {
    std::size_t host_prefix = 16384;
    std::size_t batch_min = 1024;
    std::size_t batch_max = 131072;
    double batch_fraction = 0.05;
};

struct BuildStats {
    std::size_t batches = 0, searched = 0, host_inserted = 0;
    double host_insert_s = 0, sync_s = 0, search_s = 0, link_s = 0, total_s = 0;
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
    // FORMAT v1 or v2 (dispatching on the version field).
    static std::unique_ptr<HnswGraph> load(const std::string& path);
    // FORMAT v2: sections aligned and laid out as the device copy holds them.
    void save_v2(const std::string& path) const;
    static std::unique_ptr<HnswGraph> load_v2(const std::string& path);

    // add() with the layer-0 insertion searches done in batches by `searcher`: per batch, the
    // upper-layer part of each insert runs on the host (descent, and for nodes with upper
    // layers their layer searches and neighbour choices), the searcher's copy is synced, one
    // batched search returns every node's layer-0 candidates (from its descent, or from its
    // layer-1 result), and each node is linked as insert links it (neighbour selection,
    // link-back with pruning, entry point). Nodes of one batch do not see each other as
    // candidates. `synced_version` is the version the searcher's copy was last synced at; the
    // copy is current when this returns. A batch of one node with one thread builds exactly what
    // add() with one thread builds.
    void add_batched(const float* data, std::size_t rows, std::size_t num_threads, BatchSearcher& searcher,
                     const BatchPlan& plan, std::uint64_t synced_version, BuildStats* stats = nullptr);

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

    void reserve_sections(std::size_t total);
    std::size_t random_level();
    void store(const float* vector);
    template <Locking L>
    void insert(Id node);
    template <Locking L>
    void link_back(Id node, std::size_t layer, const std::vector<Neighbor>& chosen);
    struct UpperPlan {
        std::vector<std::vector<Neighbor>> chosen;  // per layer 1..top (index = layer)
        Id entry0 = 0;                              // where the layer-0 search starts
    };
    void plan_upper(Id node, UpperPlan& plan);
    void upper_by_searcher(BatchSearcher& searcher, const std::vector<Id>& high, std::vector<UpperPlan>& plans,
                           std::size_t ef, std::size_t m, std::vector<Neighbor>& found,
                           std::vector<std::uint32_t>& count);
    void link_searched(Id node, const Neighbor* found, std::size_t count, bool chosen, UpperPlan* plan);
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
