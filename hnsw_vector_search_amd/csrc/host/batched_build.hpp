#pragma once

// The host graph's batched add (HnswGraph::add_batched) behind the C handle. The layer-0
// searcher is a vdx_batch_searcher (plain C callbacks), so the GPU half of the library, built
// by another compiler, shares no C++ class layout with the host half.

#include <cstdint>

#include "../../../include/vdxgpu.h"

namespace vdx {

// Throws as HnswIndex::add does. `errors_set`: the callbacks record their failure message (as
// the library's own entry points do), so it is kept; otherwise a failed callback is reported by
// its status. Returns the graph's version afterwards (the searcher's copy is current).
std::uint64_t host_graph_add_batched(vdx_host_graph* g, const float* data, std::uint64_t rows, std::uint64_t dim,
                                     std::uint64_t num_threads, const vdx_batch_searcher& searcher,
                                     bool errors_set, const vdx_build_plan* plan, std::uint64_t synced_version,
                                     vdx_build_stats* stats);

}  // namespace vdx
