// The host half of the C ABI: the restated graph build, FORMAT v1 and validation.

#include <algorithm>
#include <cstring>
#include <memory>

#include "batched_build.hpp"
#include "capi_common.hpp"
#include "cpu_distance.hpp"
#include "graph.hpp"
#include "validate.hpp"

namespace vdx {
namespace {
thread_local std::string g_last_error;
}

void set_last_error(const std::string& message) { g_last_error = message; }

}  // namespace vdx

struct vdx_host_graph {
    std::unique_ptr<vdx::HnswGraph> graph;
};

using vdx::guarded;
using vdx::require;

namespace vdx {
namespace {

// The graph's section view (FORMAT order), as vdx_host_graph_view fills it.
void fill_view(const HnswGraph& h, vdx_graph_view* out) {
    std::memset(out, 0, sizeof(*out));
    out->metric = static_cast<uint32_t>(h.metric());
    out->dim = h.dim();
    out->n = h.size();
    out->M = h.params().M;
    out->max_level = h.size() == 0 ? 0 : h.max_level();
    out->entry_point = h.size() == 0 ? 0 : h.entry_point();
    out->upper_blocks = h.upper_blocks();
    out->levels = h.levels().data();
    out->layer0 = h.layer0().data();
    out->upper = h.upper().data();
    out->vectors = h.vectors().data();
}

// A BatchSearcher over C callbacks.
class HookSearcher final : public BatchSearcher {
public:
    HookSearcher(const vdx_batch_searcher& hooks, bool errors_set) : hooks_(hooks), errors_set_(errors_set) {}
    void sync(const HnswGraph& g, const std::vector<Id>& changed) override {
        vdx_graph_view view;
        fill_view(g, &view);
        check(hooks_.sync(hooks_.ctx, &view, changed.data(), changed.size()), "sync");
    }
    void search(const std::vector<Id>& nodes, const std::vector<Id>& entries, std::size_t layer, std::size_t ef,
                std::size_t m, std::vector<Neighbor>& out, std::vector<std::uint32_t>& count) override {
        const std::size_t stride = m > 0 ? m : ef;
        out.resize(nodes.size() * stride);
        count.assign(nodes.size(), 0u);
        static_assert(sizeof(Neighbor) == sizeof(vdx_neighbor));
        check(hooks_.search(hooks_.ctx, nodes.data(), entries.data(), nodes.size(), static_cast<std::uint32_t>(layer),
                            static_cast<std::uint32_t>(ef), static_cast<std::uint32_t>(m),
                            reinterpret_cast<vdx_neighbor*>(out.data()), count.data()),
              "search");
        for (const std::uint32_t c : count) {
            if (c > stride) {
                throw std::runtime_error("HnswIndex::add: the batch searcher returned more results than asked for");
            }
        }
    }
    bool selects() const override { return (hooks_.caps & VDX_SEARCHER_SELECTS) != 0; }
    bool upper_layers() const override { return (hooks_.caps & VDX_SEARCHER_UPPER_LAYERS) != 0; }

private:
    void check(int rc, const char* which) const {
        if (rc == VDX_OK) return;
        if (errors_set_) throw_status(rc);
        throw std::runtime_error(fmt("HnswIndex::add: the batch searcher's %s callback returned status %d", which, rc));
    }

    vdx_batch_searcher hooks_;
    bool errors_set_;
};

}  // namespace

std::uint64_t host_graph_add_batched(vdx_host_graph* g, const float* data, std::uint64_t rows, std::uint64_t dim,
                                     std::uint64_t num_threads, const vdx_batch_searcher& searcher,
                                     bool errors_set, const vdx_build_plan* plan, std::uint64_t synced_version,
                                     vdx_build_stats* stats) {
    require(g != nullptr, "HnswIndex::add: graph is null");
    require(rows == 0 || data != nullptr, "HnswIndex::add: data is null");
    require(searcher.sync != nullptr && searcher.search != nullptr, "HnswIndex::add: searcher callback is null");
    if (dim != g->graph->dim()) {
        throw std::invalid_argument(fmt("HnswIndex::add: got dimension %llu, index has %zu",
                                        static_cast<unsigned long long>(dim), g->graph->dim()));
    }
    HookSearcher hooks(searcher, errors_set);
    BatchPlan p;
    if (plan != nullptr) {
        p.host_prefix = plan->host_prefix;
        p.batch_min = plan->batch_min;
        p.batch_max = plan->batch_max;
        p.batch_fraction = plan->batch_fraction;
    }
    BuildStats st;
    g->graph->add_batched(data, rows, num_threads, hooks, p, synced_version, &st);
    if (stats != nullptr) {
        *stats = vdx_build_stats{st.batches, st.searched, st.host_inserted, st.host_insert_s,
                                 st.sync_s,  st.search_s, st.link_s,        st.total_s};
    }
    return g->graph->version();
}

}  // namespace vdx

extern "C" {

const char* vdx_last_error(void) { return vdx::g_last_error.c_str(); }

const char* vdx_version(void) { return "0.1.0"; }

int vdx_host_graph_create(uint64_t dim, uint32_t metric, uint64_t M, uint64_t ef_construction,
                          uint64_t seed, vdx_host_graph** out) {
    return guarded([&] {
        require(out != nullptr, "vdx_host_graph_create: out is null");
        require(vdx::valid_metric(metric), "vdx_host_graph_create: unknown metric code");
        auto g = std::make_unique<vdx_host_graph>();
        g->graph = std::make_unique<vdx::HnswGraph>(
            dim, static_cast<vdx::Metric>(metric), vdx::GraphParams{M, ef_construction, seed});
        *out = g.release();
    });
}

int vdx_host_graph_add(vdx_host_graph* g, const float* data, uint64_t rows, uint64_t given_dim,
                       uint64_t num_threads) {
    return guarded([&] {
        require(g != nullptr, "vdx_host_graph_add: graph is null");
        require(rows == 0 || data != nullptr, "vdx_host_graph_add: data is null");
        if (given_dim != g->graph->dim()) {
            throw std::invalid_argument(vdx::fmt("HnswIndex::add: got dimension %llu, index has %zu",
                                                 static_cast<unsigned long long>(given_dim),
                                                 g->graph->dim()));
        }
        g->graph->add(data, rows, num_threads);
    });
}

int vdx_host_graph_save(const vdx_host_graph* g, const char* path) {
    return guarded([&] {
        require(g != nullptr && path != nullptr, "vdx_host_graph_save: null argument");
        g->graph->save(path);
    });
}

int vdx_host_graph_save_v2(const vdx_host_graph* g, const char* path) {
    return guarded([&] {
        require(g != nullptr && path != nullptr, "vdx_host_graph_save_v2: null argument");
        g->graph->save_v2(path);
    });
}

int vdx_host_graph_load(const char* path, vdx_host_graph** out) {
    return guarded([&] {
        require(out != nullptr && path != nullptr, "vdx_host_graph_load: null argument");
        auto g = std::make_unique<vdx_host_graph>();
        g->graph = vdx::HnswGraph::load(path);
        *out = g.release();
    });
}

int vdx_host_graph_add_batched(vdx_host_graph* g, const float* data, uint64_t rows, uint64_t dim,
                               uint64_t num_threads, const vdx_batch_searcher* searcher, const vdx_build_plan* plan,
                               uint64_t synced_version, vdx_build_stats* stats) {
    return guarded([&] {
        require(searcher != nullptr, "vdx_host_graph_add_batched: searcher is null");
        vdx::host_graph_add_batched(g, data, rows, dim, num_threads, *searcher, false, plan, synced_version, stats);
    });
}

int vdx_host_graph_view(const vdx_host_graph* g, vdx_graph_view* out) {
    return guarded([&] {
        require(g != nullptr && out != nullptr, "vdx_host_graph_view: null argument");
        vdx::fill_view(*g->graph, out);
    });
}

int vdx_host_graph_params(const vdx_host_graph* g, uint64_t* ef_construction, uint64_t* seed,
                          uint64_t* version) {
    return guarded([&] {
        require(g != nullptr, "vdx_host_graph_params: graph is null");
        if (ef_construction) *ef_construction = g->graph->params().ef_construction;
        if (seed) *seed = g->graph->params().seed;
        if (version) *version = g->graph->version();
    });
}

int vdx_host_graph_changes(const vdx_host_graph* g, uint64_t since_version, uint32_t* nodes,
                           uint64_t capacity, uint64_t* count) {
    return guarded([&] {
        require(g != nullptr && count != nullptr, "vdx_host_graph_changes: null argument");
        const std::vector<vdx::Id> changed = g->graph->changed_since(since_version);
        *count = changed.size();
        if (nodes != nullptr) {
            require(capacity >= changed.size(), "vdx_host_graph_changes: capacity is below the count");
            std::copy(changed.begin(), changed.end(), nodes);  // (no memcpy from an empty vector's null data)
        }
    });
}

uint64_t vdx_host_graph_memory_bytes(const vdx_host_graph* g) {
    return g == nullptr ? 0 : g->graph->memory_bytes();
}

int vdx_host_graph_inspect(const vdx_host_graph* g, vdx_graph_report* out, char* first_violation,
                           uint64_t buflen) {
    return guarded([&] {
        require(g != nullptr && out != nullptr, "vdx_host_graph_inspect: null argument");
        const vdx::GraphReport r = g->graph->inspect();
        out->reachable = r.reachable;
        out->max_level = r.max_level;
        out->violation_count = r.violations.size();
        out->mean_degree0 = r.mean_degree0;
        if (first_violation != nullptr && buflen > 0) {
            const std::string& msg = r.violations.empty() ? std::string() : r.violations.front();
            const std::size_t n = std::min<std::size_t>(msg.size(), buflen - 1);
            std::memcpy(first_violation, msg.data(), n);
            first_violation[n] = '\0';
        }
    });
}

void vdx_host_graph_free(vdx_host_graph* g) { delete g; }

int vdx_host_check_new_vectors(const char* where, uint32_t metric, uint64_t dim,
                               uint64_t index_size, const float* data, uint64_t rows,
                               uint64_t given_dim) {
    return guarded([&] {
        require(vdx::valid_metric(metric), "vdx_host_check_new_vectors: unknown metric code");
        require(rows == 0 || data != nullptr, "vdx_host_check_new_vectors: data is null");
        vdx::check_new_vectors(where ? where : "add", {static_cast<vdx::Metric>(metric), dim, index_size},
                               given_dim, data, rows);
    });
}

int vdx_host_normalize_rows(float* data, uint64_t rows, uint64_t dim) {
    return guarded([&] {
        require(rows == 0 || data != nullptr, "vdx_host_normalize_rows: data is null");
        for (uint64_t r = 0; r < rows; ++r) {
            vdx::normalize(data + r * dim, dim);
        }
    });
}

float vdx_host_distance(uint32_t metric, const float* a, const float* b, uint64_t dim) {
    return vdx::cpu::metric_distance(metric, a, b, dim);
}

}  // extern "C"
