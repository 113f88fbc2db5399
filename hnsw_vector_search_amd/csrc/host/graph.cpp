// The CPU graph build, restated from vecdex's HnswIndex (src/hnsw.cpp:131-409, 416-482 for the
// build's own layer search) so that the same vectors, parameters and seed give the same graph.
// The single-threaded build is pinned by vecdex's golden fingerprints
// (tests/hnsw_parallel_test.cpp:146-169), checked in tests/test_graph_build.py.

#include "graph.hpp"

#include <algorithm>
#include <bit>
#include <chrono>
#include <cassert>
#include <cmath>

#include "cpu_distance.hpp"
#include "validate.hpp"

namespace vdx {

// Epoch-stamped visited marks: O(1) reset per search (src/search_scratch.hpp:21-47).
class VisitedMarks {
public:
    void reset(std::size_t n) {
        if (marks_.size() < n) {
            marks_.resize(n, 0);
        }
        if (++epoch_ == 0) {
            std::fill(marks_.begin(), marks_.end(), 0U);
            epoch_ = 1;
        }
    }
    bool insert(Id id) noexcept {
        if (marks_[id] == epoch_) {
            return false;
        }
        marks_[id] = epoch_;
        return true;
    }

private:
    std::vector<std::uint32_t> marks_;
    std::uint32_t epoch_ = 0;
};

struct HnswGraph::Scratch {
    VisitedMarks visited;
    std::vector<Neighbor> candidates;  // min-heap under `farther`
    std::vector<Neighbor> results;     // max-heap under `closer`
    std::vector<Id> fresh;
    std::vector<Id> link_copy;
};

// Hands each concurrent insert its own scratch and keeps it for the next (src/search_scratch.hpp:61-105).
class HnswGraph::ScratchPool {
public:
    std::unique_ptr<Scratch> take() {
        const std::lock_guard<std::mutex> lock(mutex_);
        if (free_.empty()) {
            return std::make_unique<Scratch>();
        }
        std::unique_ptr<Scratch> s = std::move(free_.back());
        free_.pop_back();
        return s;
    }
    void give(std::unique_ptr<Scratch> s) {
        const std::lock_guard<std::mutex> lock(mutex_);
        free_.push_back(std::move(s));
    }

private:
    std::mutex mutex_;
    std::vector<std::unique_ptr<Scratch>> free_;
};

// Locks of the threaded build: a node's link blocks are guarded by one of a fixed set of
// striped mutexes (no code path holds two of them at once, so sharing a stripe only
// serializes), and the entry point by its own.
struct HnswGraph::BuildLocks {
    static constexpr std::size_t kStripes = 1u << 16;
    explicit BuildLocks(std::size_t nodes) : stripes(std::min(kStripes, std::bit_ceil(std::max<std::size_t>(nodes, 1)))),
                                             locks(new std::mutex[stripes]) {}
    std::mutex& link(Id node) { return locks[node & (stripes - 1)]; }
    std::size_t stripes;
    std::unique_ptr<std::mutex[]> locks;
    std::mutex entry;
};

namespace {

constexpr bool farther(const Neighbor& a, const Neighbor& b) noexcept { return closer(b, a); }

GraphParams validated(const GraphParams& p) {
    if (p.M < 2) {
        throw std::invalid_argument("HnswIndex: M must be at least 2");
    }
    if (p.M > GraphParams::kMaxM) {
        throw std::invalid_argument(fmt("HnswIndex: M must be at most %zu", GraphParams::kMaxM));
    }
    if (p.ef_construction == 0) {
        throw std::invalid_argument("HnswIndex: ef_construction must be positive");
    }
    return p;
}

// RAII loan of a scratch from the pool.
class Lease {
public:
    explicit Lease(HnswGraph::ScratchPool& pool) : pool_(pool), scratch_(pool.take()) {}
    ~Lease() { pool_.give(std::move(scratch_)); }
    Lease(const Lease&) = delete;
    Lease& operator=(const Lease&) = delete;
    HnswGraph::Scratch& operator*() const noexcept { return *scratch_; }

private:
    HnswGraph::ScratchPool& pool_;
    std::unique_ptr<HnswGraph::Scratch> scratch_;
};

}  // namespace

HnswGraph::HnswGraph(std::size_t dim, Metric metric, GraphParams params)
    : dim_(dim),
      metric_(metric),
      params_(validated(params)),
      level_multiplier_(1.0 / std::log(static_cast<double>(params_.M))),
      rng_(params_.seed),
      scratch_(std::make_unique<ScratchPool>()) {
    if (dim == 0) {
        throw std::invalid_argument("Matrix: dim must be positive");
    }
}

HnswGraph::~HnswGraph() = default;

const Id* HnswGraph::block(Id node, std::size_t layer) const noexcept {
    if (layer == 0) {
        return layer0_.data() + static_cast<std::size_t>(node) * (1 + 2 * params_.M);
    }
    return upper_.data() +
           (static_cast<std::size_t>(upper_offset_[node]) + (layer - 1)) * (1 + params_.M);
}

Id* HnswGraph::block(Id node, std::size_t layer) noexcept {
    return const_cast<Id*>(static_cast<const HnswGraph&>(*this).block(node, layer));
}

float HnswGraph::distance(const float* a, const float* b) const noexcept {
    return cpu::metric_distance(static_cast<unsigned>(metric_), a, b, dim_);
}

std::size_t HnswGraph::memory_bytes() const noexcept {
    return vectors_.size() * sizeof(float) + levels_.size() + layer0_.size() * sizeof(Id) +
           upper_.size() * sizeof(Id) + upper_offset_.size() * sizeof(std::uint32_t);
}

// level = floor(-ln(U) * mL), U in (0, 1] from the top 53 bits of mt19937_64
// (src/hnsw.cpp:219-228): platform-independent levels for a given seed.
std::size_t HnswGraph::random_level() {
    const double u = 1.0 - static_cast<double>(rng_() >> 11U) * 0x1.0p-53;
    return static_cast<std::size_t>(-std::log(u) * level_multiplier_);
}

void HnswGraph::store(const float* vector) {
    vectors_.insert(vectors_.end(), vector, vector + dim_);
    if (metric_ == Metric::Cosine) {
        normalize(vectors_.data() + (size() * dim_), dim_);
    }
    const std::size_t level = random_level();
    const std::size_t first_block = upper_.size() / (1 + params_.M);
    if (first_block > UINT32_MAX - level) {
        throw std::length_error("HnswIndex::add: too many upper-layer blocks for 32-bit offsets");
    }
    levels_.push_back(static_cast<std::uint8_t>(level));
    upper_offset_.push_back(static_cast<std::uint32_t>(first_block));
    touched_.push_back(static_cast<std::uint32_t>(version_));
    layer0_.resize(layer0_.size() + 1 + 2 * params_.M, Id{0});
    upper_.resize(upper_.size() + level * (1 + params_.M), Id{0});
}

// Room for `total` nodes, growing by at least half of what is there: a graph built by many
// adds then copies each section a bounded number of times, not once per add.
void HnswGraph::reserve_sections(std::size_t total) {
    const auto grow = [](auto& v, std::size_t need) {
        if (need > v.capacity()) {
            v.reserve(std::max(need, v.capacity() + v.capacity() / 2));
        }
    };
    grow(vectors_, total * dim_);
    grow(levels_, total);
    grow(upper_offset_, total);
    grow(touched_, total);
    grow(layer0_, total * (1 + 2 * params_.M));
}

std::vector<Id> HnswGraph::changed_since(std::uint64_t version) const {
    std::vector<Id> out;
    for (std::size_t i = 0; i < touched_.size(); ++i) {
        if (touched_[i] > version) {
            out.push_back(static_cast<Id>(i));
        }
    }
    return out;
}

void HnswGraph::add(const float* data, std::size_t rows, std::size_t num_threads) {
    check_new_vectors("HnswIndex::add", IndexShape{metric_, dim_, size()}, dim_, data, rows);
    ++version_;
    const std::size_t first = size();
    const std::size_t total = first + rows;
    reserve_sections(total);
    const std::size_t threads = std::min(resolve_threads(num_threads), rows);
    if (threads <= 1) {
        for (std::size_t r = 0; r < rows; ++r) {
            store(data + r * dim_);
            insert<Locking::None>(static_cast<Id>(size() - 1));
        }
        return;
    }
    // Every row is stored before any is linked, so no container grows under a thread's feet;
    // the levels are drawn in serial order.
    build_locks_ = std::make_unique<BuildLocks>(total);
    for (std::size_t r = 0; r < rows; ++r) {
        store(data + r * dim_);
    }
    std::size_t next = first;
    if (next == 0) {
        insert<Locking::None>(0);
        next = 1;
    }
    parallel_for(total - next, threads,
                 [this, next](std::size_t i) { insert<Locking::PerNode>(static_cast<Id>(next + i)); });
    build_locks_.reset();
}

void HnswGraph::add_batched(const float* data, std::size_t rows, std::size_t num_threads,
                            BatchSearcher& searcher, const BatchPlan& plan, std::uint64_t synced_version,
                            BuildStats* stats) {
    using clock = std::chrono::steady_clock;
    BuildStats local;
    BuildStats& st = stats != nullptr ? *stats : local;
    st = BuildStats{};
    const auto start = clock::now();
    auto lap = clock::now();
    const auto since = [&lap] {
        const auto now = clock::now();
        const double s = std::chrono::duration<double>(now - lap).count();
        lap = now;
        return s;
    };
    check_new_vectors("HnswIndex::add", IndexShape{metric_, dim_, size()}, dim_, data, rows);
    if (plan.batch_min == 0 || plan.batch_max < plan.batch_min || !(plan.batch_fraction >= 0.0)) {
        throw std::invalid_argument("HnswIndex::add: batch sizes must satisfy 1 <= batch_min <= batch_max "
                                    "and batch_fraction >= 0");
    }
    ++version_;
    const std::size_t first = size();
    const std::size_t total = first + rows;
    reserve_sections(total);
    // As in add(): every row is stored (levels drawn in serial order) before any is linked.
    for (std::size_t r = 0; r < rows; ++r) {
        store(data + r * dim_);
    }
    if (rows == 0) {
        return;
    }
    const std::size_t threads = std::max<std::size_t>(1, std::min(resolve_threads(num_threads), rows));
    build_locks_ = std::make_unique<BuildLocks>(total);
    struct Release {
        HnswGraph* g;
        ~Release() { g->build_locks_.reset(); }
    } release{this};
    std::size_t next = first;
    if (next == 0) {
        insert<Locking::None>(0);
        next = 1;
    }
    const std::size_t prefix_end = std::min(total, std::max(next, plan.host_prefix));
    since();
    if (prefix_end > next) {
        parallel_for(prefix_end - next, threads,
                     [this, next](std::size_t i) { insert<Locking::PerNode>(static_cast<Id>(next + i)); });
        st.host_inserted += prefix_end - next;
        next = prefix_end;
    }
    st.host_insert_s += since();
    // Stamps written after a sync carry a later version than the one synced.
    std::uint64_t synced = synced_version;
    const auto sync = [&] {
        searcher.sync(*this, changed_since(synced));
        synced = version_;
        ++version_;
        st.sync_s += since();
    };
    const std::size_t ef = std::max(params_.ef_construction, params_.M);
    const std::size_t m = searcher.selects() ? params_.M : 0;  // the searcher makes the choice
    const std::size_t stride = m > 0 ? m : ef;
    std::vector<Id> nodes, entries, high;
    std::vector<std::size_t> high_at;
    std::vector<UpperPlan> plans;
    std::vector<Neighbor> found;
    std::vector<std::uint32_t> count;
    while (next < total) {
        const auto scaled = static_cast<std::size_t>(plan.batch_fraction * static_cast<double>(next));
        const std::size_t batch = std::clamp(scaled, plan.batch_min, plan.batch_max);
        const std::size_t end = std::min(total, next + batch);
        nodes.clear();
        entries.clear();
        high.clear();
        high_at.clear();
        for (std::size_t id = next; id < end; ++id) {
            nodes.push_back(static_cast<Id>(id));
            entries.push_back(BatchSearcher::kDescend);
            if (levels_[id] > 0) {
                high.push_back(static_cast<Id>(id));
                high_at.push_back(id - next);
            }
        }
        st.host_insert_s += since();
        sync();
        // The upper-layer part of the inserts (nothing is written, so no locks): by the searcher
        // layer by layer, top down, or on the host threads.
        plans.resize(high.size());
        if (!high.empty() && searcher.upper_layers()) {
            upper_by_searcher(searcher, high, plans, ef, m, found, count);
            st.search_s += since();
        } else if (!high.empty()) {
            parallel_for(high.size(), std::min(threads, high.size()),
                         [this, &high, &plans](std::size_t i) { plan_upper(high[i], plans[i]); });
            st.host_inserted += high.size();
            st.host_insert_s += since();
        }
        for (std::size_t i = 0; i < high.size(); ++i) {
            entries[high_at[i]] = plans[i].entry0;
        }
        searcher.search(nodes, entries, 0, ef, m, found, count);
        st.search_s += since();
        st.searched += nodes.size();
        std::vector<UpperPlan*> plan_of(nodes.size(), nullptr);
        for (std::size_t i = 0; i < high.size(); ++i) {
            plan_of[high_at[i]] = &plans[i];
        }
        parallel_for(nodes.size(), std::min(threads, nodes.size()),
                     [this, &nodes, &found, &count, &plan_of, stride, m](std::size_t i) {
                         link_searched(nodes[i], found.data() + i * stride, count[i], m > 0, plan_of[i]);
                     });
        st.link_s += since();
        ++st.batches;
        next = end;
    }
    sync();
    st.total_s = std::chrono::duration<double>(clock::now() - start).count();
}

// The first half of insert (src/hnsw.cpp:315-335) for a node with upper layers, against the
// graph as it stands: the descent to its level, then for each of its layers down to 1 the layer
// search and the neighbour choice; the layer-0 search is left to the batch's searcher, from the
// layer-1 result.
void HnswGraph::plan_upper(Id node, UpperPlan& plan) {
    const std::size_t level = levels_[node];
    Lease lease(*scratch_);
    Scratch& scratch = *lease;
    const float* query = vector_of(node);
    Neighbor entry{distance(query, vector_of(entry_point_)), entry_point_};
    for (std::size_t layer = max_level_; layer > level; --layer) {
        entry = greedy_closest<Locking::None>(query, entry, layer, scratch);
    }
    const std::size_t ef = std::max(params_.ef_construction, params_.M);
    const std::size_t top = std::min(level, max_level_);
    plan.chosen.assign(top + 1, {});
    for (std::size_t layer = top; layer >= 1; --layer) {
        search_layer<Locking::None>(query, entry, ef, layer, scratch);
        entry = *std::min_element(scratch.results.begin(), scratch.results.end(), closer);
        plan.chosen[layer] = select_neighbors(scratch.results, params_.M);
    }
    plan.entry0 = entry.id;
}

// plan_upper for a batch's nodes with upper layers, the layer searches done by the searcher: for
// each layer from the batch's top down, one search call over the nodes that reach it, each from
// the best result of the layer above (or, on its own top layer, from the descent), and the
// neighbour choice among the results (by the searcher when it selects).
void HnswGraph::upper_by_searcher(BatchSearcher& searcher, const std::vector<Id>& high, std::vector<UpperPlan>& plans,
                                  std::size_t ef, std::size_t m, std::vector<Neighbor>& found,
                                  std::vector<std::uint32_t>& count) {
    std::size_t top_all = 0;
    std::vector<std::size_t> top(high.size());
    for (std::size_t i = 0; i < high.size(); ++i) {
        top[i] = std::min<std::size_t>(levels_[high[i]], max_level_);
        top_all = std::max(top_all, top[i]);
        plans[i].chosen.assign(top[i] + 1, {});
        plans[i].entry0 = BatchSearcher::kDescend;  // the layer-0 search descends if no upper layer was searched
    }
    const std::size_t stride = m > 0 ? m : ef;
    std::vector<Id> nodes, entries;
    std::vector<std::size_t> which;
    for (std::size_t layer = top_all; layer >= 1; --layer) {
        nodes.clear();
        entries.clear();
        which.clear();
        for (std::size_t i = 0; i < high.size(); ++i) {
            if (top[i] >= layer) {
                nodes.push_back(high[i]);
                entries.push_back(plans[i].entry0);
                which.push_back(i);
            }
        }
        searcher.search(nodes, entries, layer, ef, m, found, count);
        for (std::size_t j = 0; j < nodes.size(); ++j) {
            const Neighbor* row = found.data() + j * stride;
            std::vector<Neighbor> chosen(row, row + count[j]);
            if (m == 0) {
                chosen = select_neighbors(std::move(chosen), params_.M);
            }
            UpperPlan& p = plans[which[j]];
            // the closest result: select_neighbors keeps it first
            p.entry0 = chosen.empty() ? p.entry0 : chosen.front().id;
            p.chosen[layer] = std::move(chosen);
        }
    }
}

// The linking half of insert for a node whose layer-0 search results are `found` (or, with
// `chosen`, already its layer-0 choice) and, for a node with upper layers, whose upper choices
// are in `plan`.
void HnswGraph::link_searched(Id node, const Neighbor* found, std::size_t count, bool chosen, UpperPlan* plan) {
    std::vector<Neighbor> chosen0 = chosen ? std::vector<Neighbor>(found, found + count)
                                           : select_neighbors(std::vector<Neighbor>(found, found + count), params_.M);
    const std::size_t top = plan != nullptr ? plan->chosen.size() - 1 : 0;
    {
        const std::lock_guard<std::mutex> lock(build_locks_->link(node));
        touched_[node] = static_cast<std::uint32_t>(version_);
        for (std::size_t layer = 0; layer <= top; ++layer) {
            const std::vector<Neighbor>& chosen = layer == 0 ? chosen0 : plan->chosen[layer];
            Id* b = block(node, layer);
            b[0] = static_cast<Id>(chosen.size());
            for (std::size_t i = 0; i < chosen.size(); ++i) {
                b[1 + i] = chosen[i].id;
            }
        }
    }
    for (std::size_t layer = 0; layer <= top; ++layer) {
        link_back<Locking::PerNode>(node, layer, layer == 0 ? chosen0 : plan->chosen[layer]);
    }
    const std::size_t level = levels_[node];
    if (level > 0) {
        const std::lock_guard<std::mutex> lock(build_locks_->entry);
        if (level > max_level_) {
            max_level_ = level;
            entry_point_ = node;
        }
    }
}

template <HnswGraph::Locking L>
const Id* HnswGraph::read_links(Id node, std::size_t layer, std::vector<Id>& copy,
                                std::size_t& count) const {
    if constexpr (L == Locking::PerNode) {
        const std::lock_guard<std::mutex> lock(build_locks_->link(node));
        const Id* b = block(node, layer);
        copy.assign(b + 1, b + 1 + b[0]);
        count = copy.size();
        return copy.data();
    } else {
        const Id* b = block(node, layer);
        count = b[0];
        return b + 1;
    }
}

template <HnswGraph::Locking L>
Neighbor HnswGraph::greedy_closest(const float* query, Neighbor entry, std::size_t layer,
                                   Scratch& scratch) const {
    for (bool moved = true; moved;) {
        moved = false;
        std::size_t count = 0;
        const Id* ids = read_links<L>(entry.id, layer, scratch.link_copy, count);
        for (std::size_t i = 0; i < count; ++i) {
            const float d = distance(query, vector_of(ids[i]));
            if (d < entry.distance) {
                entry = {d, ids[i]};
                moved = true;
            }
        }
    }
    return entry;
}

template <HnswGraph::Locking L>
void HnswGraph::search_layer(const float* query, Neighbor entry, std::size_t ef, std::size_t layer,
                             Scratch& scratch) const {
    VisitedMarks& visited = scratch.visited;
    std::vector<Neighbor>& candidates = scratch.candidates;
    std::vector<Neighbor>& results = scratch.results;
    std::vector<Id>& fresh = scratch.fresh;
    visited.reset(size());
    visited.insert(entry.id);
    candidates.assign(1, entry);
    results.assign(1, entry);
    while (!candidates.empty()) {
        std::pop_heap(candidates.begin(), candidates.end(), farther);
        const Neighbor current = candidates.back();
        candidates.pop_back();
        if (results.size() >= ef && closer(results.front(), current)) {
            break;
        }
        fresh.clear();
        std::size_t count = 0;
        const Id* ids = read_links<L>(current.id, layer, scratch.link_copy, count);
        for (std::size_t i = 0; i < count; ++i) {
            if (visited.insert(ids[i])) {
                fresh.push_back(ids[i]);
                __builtin_prefetch(vector_of(ids[i]));
            }
        }
        for (const Id id : fresh) {
            const Neighbor next{distance(query, vector_of(id)), id};
            if (results.size() < ef || closer(next, results.front())) {
                candidates.push_back(next);
                std::push_heap(candidates.begin(), candidates.end(), farther);
                results.push_back(next);
                std::push_heap(results.begin(), results.end(), closer);
                if (results.size() > ef) {
                    std::pop_heap(results.begin(), results.end(), closer);
                    results.pop_back();
                }
            }
        }
    }
}

template <HnswGraph::Locking L>
void HnswGraph::insert(Id node) {
    const std::size_t level = levels_[node];
    if (node == 0) {
        entry_point_ = node;
        max_level_ = level;
        return;
    }
    // A node that raises the top layer holds the entry lock until it is linked.
    std::unique_lock<std::mutex> entry_lock;
    if constexpr (L == Locking::PerNode) {
        entry_lock = std::unique_lock<std::mutex>(build_locks_->entry);
    }
    const Id entry_point = entry_point_;
    const std::size_t max_level = max_level_;
    if (entry_lock.owns_lock() && level <= max_level) {
        entry_lock.unlock();
    }
    Lease lease(*scratch_);
    Scratch& scratch = *lease;
    const float* query = vector_of(node);

    Neighbor entry{distance(query, vector_of(entry_point)), entry_point};
    for (std::size_t layer = max_level; layer > level; --layer) {
        entry = greedy_closest<L>(query, entry, layer, scratch);
    }
    const std::size_t ef = std::max(params_.ef_construction, params_.M);
    const std::size_t top = std::min(level, max_level);
    std::vector<std::vector<Neighbor>> chosen(top + 1);
    for (std::size_t i = 0; i <= top; ++i) {
        const std::size_t layer = top - i;
        search_layer<L>(query, entry, ef, layer, scratch);
        entry = *std::min_element(scratch.results.begin(), scratch.results.end(), closer);
        chosen[layer] = select_neighbors(scratch.results, params_.M);
    }
    {
        std::unique_lock<std::mutex> lock;
        if constexpr (L == Locking::PerNode) {
            lock = std::unique_lock<std::mutex>(build_locks_->link(node));
        }
        touched_[node] = static_cast<std::uint32_t>(version_);
        for (std::size_t layer = 0; layer <= top; ++layer) {
            Id* b = block(node, layer);
            b[0] = static_cast<Id>(chosen[layer].size());
            for (std::size_t i = 0; i < chosen[layer].size(); ++i) {
                b[1 + i] = chosen[layer][i].id;
            }
        }
    }
    for (std::size_t layer = 0; layer <= top; ++layer) {
        link_back<L>(node, layer, chosen[layer]);
    }
    if (level > max_level) {
        max_level_ = level;
        entry_point_ = node;
    }
}

template <HnswGraph::Locking L>
void HnswGraph::link_back(Id node, std::size_t layer, const std::vector<Neighbor>& chosen) {
    for (const Neighbor& neighbor : chosen) {
        std::unique_lock<std::mutex> lock;
        if constexpr (L == Locking::PerNode) {
            lock = std::unique_lock<std::mutex>(build_locks_->link(neighbor.id));
        }
        Id* b = block(neighbor.id, layer);
        touched_[neighbor.id] = static_cast<std::uint32_t>(version_);
        const std::size_t count = b[0];
        if (count < max_links(layer)) {
            b[1 + count] = node;
            b[0] = static_cast<Id>(count + 1);
            continue;
        }
        std::vector<Neighbor> candidates;
        candidates.reserve(count + 1);
        const float* base = vector_of(neighbor.id);
        for (std::size_t i = 0; i < count; ++i) {
            candidates.push_back({distance(base, vector_of(b[1 + i])), b[1 + i]});
        }
        candidates.push_back({neighbor.distance, node});
        const std::vector<Neighbor> kept = select_neighbors(std::move(candidates), max_links(layer));
        b[0] = static_cast<Id>(kept.size());
        for (std::size_t i = 0; i < kept.size(); ++i) {
            b[1 + i] = kept[i].id;
        }
    }
}

// The neighbor-selection heuristic (Malkov & Yashunin, Algorithm 4): walking the candidates
// from closest to farthest, keep one only if it is closer to the base than to every kept one.
std::vector<Neighbor> HnswGraph::select_neighbors(std::vector<Neighbor> candidates,
                                                  std::size_t m) const {
    std::sort(candidates.begin(), candidates.end(), closer);
    if (candidates.size() <= m) {
        return candidates;
    }
    std::vector<Neighbor> kept;
    kept.reserve(m);
    for (const Neighbor& c : candidates) {
        if (kept.size() == m) {
            break;
        }
        const float* v = vector_of(c.id);
        bool diverse = true;
        for (const Neighbor& other : kept) {
            if (distance(v, vector_of(other.id)) < c.distance) {
                diverse = false;
                break;
            }
        }
        if (diverse) {
            kept.push_back(c);
        }
    }
    return kept;
}

GraphReport HnswGraph::inspect() const {
    GraphReport report;
    report.max_level = max_level_;
    const std::size_t n = size();
    if (n == 0) {
        return report;
    }
    std::size_t highest = 0;
    std::size_t layer0_links = 0;
    std::vector<Id> sorted;
    for (Id node = 0; node < n; ++node) {
        const std::size_t level = levels_[node];
        highest = std::max(highest, level);
        for (std::size_t layer = 0; layer <= level; ++layer) {
            const Id* b = block(node, layer);
            const std::size_t count = b[0];
            if (layer == 0) {
                layer0_links += count;
            }
            if (count > max_links(layer)) {
                report.violations.push_back(fmt("node %u has %zu links on layer %zu (max %zu)",
                                                node, count, layer, max_links(layer)));
                continue;
            }
            sorted.assign(b + 1, b + 1 + count);
            std::sort(sorted.begin(), sorted.end());
            if (std::adjacent_find(sorted.begin(), sorted.end()) != sorted.end()) {
                report.violations.push_back(
                    fmt("node %u has a duplicate link on layer %zu", node, layer));
            }
            for (std::size_t i = 0; i < count; ++i) {
                const Id id = b[1 + i];
                if (id == node) {
                    report.violations.push_back(
                        fmt("node %u links to itself on layer %zu", node, layer));
                } else if (id >= n) {
                    report.violations.push_back(fmt("node %u links to missing node %u", node, id));
                } else if (levels_[id] < layer) {
                    report.violations.push_back(fmt(
                        "node %u links to node %u on layer %zu, above its level", node, id, layer));
                }
            }
        }
    }
    if (highest != max_level_) {
        report.violations.push_back(
            fmt("max level is %zu but the tallest node reaches %zu", max_level_, highest));
    }
    if (levels_[entry_point_] != max_level_) {
        report.violations.push_back(fmt("entry point %u is not on the top layer", entry_point_));
    }
    std::vector<bool> seen(n, false);
    std::vector<Id> frontier{entry_point_};
    seen[entry_point_] = true;
    for (std::size_t i = 0; i < frontier.size(); ++i) {
        const Id* b = block(frontier[i], 0);
        const std::size_t count = std::min<std::size_t>(b[0], max_links(0));
        for (std::size_t j = 0; j < count; ++j) {
            const Id id = b[1 + j];
            if (id < n && !seen[id]) {
                seen[id] = true;
                frontier.push_back(id);
            }
        }
    }
    report.reachable = frontier.size();
    report.mean_degree0 = static_cast<double>(layer0_links) / static_cast<double>(n);
    return report;
}

template void HnswGraph::insert<HnswGraph::Locking::None>(Id);
template void HnswGraph::insert<HnswGraph::Locking::PerNode>(Id);

}  // namespace vdx
