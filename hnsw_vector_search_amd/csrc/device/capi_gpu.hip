// The GPU half of the C ABI (include/vdxgpu.h): device contexts, the HBM copies of a graph and
// of a brute-force base, and the search entry points that replace HnswIndex::search/search_batch
// and BruteForceIndex::search.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <thread>
#include <vector>

#include "../host/batched_build.hpp"
#include "../host/capi_common.hpp"
#include "../host/format_v2.hpp"
#include "../host/parallel_for.hpp"
#include "../host/validate.hpp"
#include "beam_args.hpp"
#include "beam_host.hpp"
#include "bf_args.hpp"
#include "bf_fp64.hpp"
#include "bf_host.hpp"
#include "fill_copy.hpp"
#include "gpu_index.hpp"
#include "gpu_state.hpp"
#include "select_links.hpp"

using vdx::guarded;
using vdx::require;
using vdx::gpu::check;
using vdx::gpu::DeviceBuffer;
using vdx::gpu::DeviceGuard;
using vdx::gpu::WorkspaceLease;

namespace {

constexpr std::uint32_t kMaxGpuDim = 287;    // 8 full 32-blocks, 3 8-blocks, 7 tail elements
constexpr std::uint32_t kMaxBeamEf = 8192;   // R capacity per wave in LDS
constexpr std::uint32_t kMaxGpuM = 32;       // a layer-0 block fits one wave
constexpr std::uint32_t kMaxBfK = 2048;
constexpr std::uint32_t kRingSlots = vdx_gpu_hnsw::kRingSlots;

int bf_shape(std::uint64_t dim) { return dim == 128 ? 0 : (dim == 100 ? 1 : 2); }

// Copies `rows` host rows of `dim` floats to device rows of `ld` floats (zero padded). Rows
// that need padding are packed in pinned chunks: a strided copy from pageable memory would go
// row by row.
void copy_rows(float* dst, const float* src, std::uint64_t rows, std::uint64_t dim, std::uint64_t ld,
               hipStream_t stream) {
    if (rows == 0) {
        return;
    }
    if (ld == dim) {
        check(hipMemcpyAsync(dst, src, rows * dim * sizeof(float), hipMemcpyHostToDevice, stream),
              "hipMemcpyAsync(vectors)");
        return;
    }
    const std::uint64_t chunk = std::max<std::uint64_t>(1, (std::uint64_t{64} << 20) / (ld * sizeof(float)));
    vdx::gpu::PinnedBuffer staging;
    staging.reserve(std::min(rows, chunk) * ld * sizeof(float));
    float* st = staging.as<float>();
    for (std::uint64_t r0 = 0; r0 < rows; r0 += chunk) {
        const std::uint64_t m = std::min(chunk, rows - r0);
        for (std::uint64_t r = 0; r < m; ++r) {
            std::memcpy(st + r * ld, src + (r0 + r) * dim, dim * sizeof(float));
            std::memset(st + r * ld + dim, 0, (ld - dim) * sizeof(float));
        }
        check(hipMemcpyAsync(dst + r0 * ld, st, m * ld * sizeof(float), hipMemcpyHostToDevice, stream),
              "hipMemcpyAsync(vectors)");
        check(hipStreamSynchronize(stream), "hipStreamSynchronize(vectors)");  // before the chunk is reused
    }
}

// Copies `rows` rows of `dim` floats into a device buffer of rows of `ld` floats (zero padded).
void upload_rows(DeviceBuffer& dst, const float* src, std::uint64_t rows, std::uint64_t dim,
                 std::uint64_t ld, hipStream_t stream) {
    dst.resize(std::max<std::uint64_t>(1, rows) * ld * sizeof(float));
    copy_rows(dst.as<float>(), src, rows, dim, ld, stream);
}

// Queues one beam launch over prepared device queries.
// Keeps h's 8-bit row copy current after rows [first, h.n) changed: squared L2 at dimension 128
// only (the byte-row search's shape), and only while every stored value is an integer in
// [0, 255] (checked on the device as the rows are packed); otherwise the copy is dropped and the
// search reads the FP32 rows. The buffer grows with a quarter of headroom, keeping the rows
// before `first`. Synchronizes the stream.
void refresh_rows_u8(vdx_gpu_hnsw& h, std::uint64_t first, hipStream_t s) {
    if (h.dim != 128 || h.metric != 0 || (first > 0 && !h.u8) || h.n == 0) {
        h.u8 = false;
        h.rows_u8.release();
        return;
    }
    if (first == h.n) return;  // no new rows
    const std::uint64_t need = h.n * 128;
    if (need > h.rows_u8.bytes()) {
        DeviceBuffer grown(need + need / 4);
        if (first > 0) {
            check(hipMemcpyAsync(grown.as<void>(), h.rows_u8.as<void>(), first * 128, hipMemcpyDeviceToDevice, s),
                  "hipMemcpyAsync(grow rows_u8)");
        }
        check(hipStreamSynchronize(s), "hipStreamSynchronize(grow rows_u8)");
        h.rows_u8.swap(grown);
    }
    h.u8_flag.reserve(4);
    check(vdx::dev::launch_fill_u32(h.u8_flag.as<std::uint32_t>(), 0, 1, s), "fill(u8 flag)");
    check(vdx::dev::launch_pack_rows_u8(h.rows_u8.as<unsigned char>(), h.vectors.as<float>(), h.split.ld, first,
                                        h.n - first, h.u8_flag.as<std::uint32_t>(), s),
          "pack_rows_u8 launch");
    std::uint32_t bad = 1;
    check(hipMemcpyAsync(&bad, h.u8_flag.as<void>(), 4, hipMemcpyDeviceToHost, s), "hipMemcpyAsync(u8 flag)");
    check(hipStreamSynchronize(s), "hipStreamSynchronize(pack rows_u8)");
    h.u8 = bad == 0;
    if (!h.u8) h.rows_u8.release();
}

void enqueue_beam(vdx_gpu_hnsw& h, const float* d_queries, std::uint64_t nq, const std::uint32_t* d_ks,
                  std::uint32_t k, const std::uint32_t* d_efs, std::uint32_t ef,
                  std::uint32_t max_efp, vdx::dev::Neighbor* d_out, std::uint64_t out_stride,
                  std::uint32_t* d_count, std::uint32_t* d_queue, unsigned long long* d_stats,
                  const vdx_search_opts* opts, hipStream_t stream,
                  const vdx::dev::BeamArgs* trace = nullptr, const std::uint32_t* d_entries = nullptr,
                  std::uint32_t layer = 0, DeviceBuffer* bitmap = nullptr, int host_bytes = 0,
                  const std::uint8_t* d_qbytes = nullptr) {
    const int mode = vdx::gpu::beam_mode(opts, trace != nullptr);
    const vdx::gpu::BeamPlan plan = vdx::gpu::plan_beam(h, max_efp, nq, vdx::gpu::config_from(opts), mode);
    vdx::dev::BeamArgs a = vdx::gpu::beam_args(h, plan, mode, opts, d_queries, nq, d_ks, k, d_efs, ef, d_out,
                                               out_stride, d_count, d_queue, d_stats);
    if (plan.bitmap_bytes > 0) {
        require(bitmap != nullptr, "visited_form 2 (HBM bitmap) is not available on this path");
        bitmap->reserve(plan.bitmap_bytes);
        a.vbits = bitmap->as<std::uint32_t>();
        a.vbit_words = plan.bitmap_words;
    }
    if (trace != nullptr) {
        a.truth = trace->truth;
        a.trace_feat = trace->trace_feat;
        a.trace_hits = trace->trace_hits;
        a.trace_steps = trace->trace_steps;
        a.trace_final = trace->trace_final;
        a.trace_cap = trace->trace_cap;
    }
    a.entries = d_entries;
    a.layer = layer;
    bool int_only = false;
    if (plan.shape == vdx::dev::BeamShape::D128U8 && host_bytes != 0) {
        // the host checked the queries while staging them (host_bytes 1: every component an
        // integer in [0, 255]): the counter and the flag word after it are cleared, and the
        // integer kernel alone goes with epoch 1 (the flag 0 differs), or the float kernel alone
        check(vdx::dev::launch_fill_u32(d_queue, 0u, 2, stream), "fill(queue)");
        if (host_bytes == 1) {
            a.qflag = d_queue + 1;
            a.qflag_epoch = 1;
            // the integer kernel alone reads them (the LDS-resident beam reads the floats)
            require(d_qbytes == nullptr || plan.chunks > 0, "byte-staged queries need the register-resident beam");
            a.qbytes = d_qbytes;
            int_only = true;
        }
    } else if (plan.shape == vdx::dev::BeamShape::D128U8) {
        // the byte-query check also clears the work counter; its flag is the word after the
        // counter (every caller's counter word has a free word after it), stamped with this
        // launch's epoch, so it needs no clearing
        std::uint32_t epoch = h.query_epoch.fetch_add(1) + 1;
        if (epoch == 0) epoch = h.query_epoch.fetch_add(1) + 1;
        check(vdx::dev::launch_query_bytes_check(d_queries, nq * h.dim, d_queue + 1, epoch, d_queue, stream),
              "query_bytes_check launch");
        a.qflag = d_queue + 1;
        a.qflag_epoch = epoch;
    } else {
        check(hipMemsetAsync(d_queue, 0, sizeof(std::uint32_t), stream), "hipMemsetAsync(queue)");
    }
    vdx::gpu::launch_beam(h, plan, mode, a, stream, int_only);
}

// The link rules of HnswIndex::check_links (src/hnsw_io.cpp:472-504) for the blocks of one node:
// counts within capacity, ids naming nodes that exist on the block's layer, no self-links, no
// id twice in a block. The kernels rely on them: an upper link to a node below the layer would
// index past that node's blocks, and a repeated id would enter one expansion's fresh batch twice.
class LinkChecker {
public:
    // `marks` (one per node, grown here) and `epoch` persist between checkers of one index.
    LinkChecker(const vdx_graph_view& g, const char* where, std::vector<std::uint32_t>& marks,
                std::uint32_t& epoch)
        : g_(g), where_(where), marks_(marks), epoch_(epoch) {
        if (marks_.size() < g.n) {
            marks_.resize(g.n, 0u);
        }
    }

    // Node i's layer-0 block and its upper blocks, which start at upper block `off`.
    void node(std::uint64_t i, std::uint64_t off) {
        block(g_.layer0 + i * (1 + 2 * g_.M), i, 0, 2 * g_.M);
        for (std::uint64_t layer = 1; layer <= g_.levels[i]; ++layer) {
            block(g_.upper + (off + layer - 1) * (1 + g_.M), i, layer, g_.M);
        }
    }

private:
    void block(const std::uint32_t* b, std::uint64_t node, std::uint64_t layer, std::uint64_t cap) {
        using ull = unsigned long long;
        if (b[0] > cap) {
            throw std::invalid_argument(vdx::fmt("%s: node %llu has %u links on layer %llu (max %llu)", where_,
                                                 static_cast<ull>(node), b[0], static_cast<ull>(layer),
                                                 static_cast<ull>(cap)));
        }
        if (++epoch_ == 0) {
            std::fill(marks_.begin(), marks_.end(), 0u);
            epoch_ = 1;
        }
        for (std::uint32_t j = 0; j < b[0]; ++j) {
            const std::uint32_t id = b[1 + j];
            const char* why = id >= g_.n                ? "names a missing node"
                              : id == node              ? "is a self-link"
                              : g_.levels[id] < layer   ? "names a node below the layer"
                              : marks_[id] == epoch_    ? "repeats a node"
                                                        : nullptr;
            if (why != nullptr) {
                throw std::invalid_argument(vdx::fmt("%s: link %u of node %llu on layer %llu %s", where_, id,
                                                     static_cast<ull>(node), static_cast<ull>(layer), why));
            }
            marks_[id] = epoch_;
        }
    }

    const vdx_graph_view& g_;
    const char* where_;
    std::vector<std::uint32_t>& marks_;
    std::uint32_t& epoch_;
};

// Brings the device copy up to a graph that extends it (vdx_gpu_hnsw_update): the old nodes keep
// their levels (so their upper blocks keep their positions); new nodes are appended and the
// listed old nodes' blocks rewritten, all under the upload's link rules. Host work is
// proportional to what changed (plus one comparison of the old levels).
// `trusted`: the graph comes from this library's own build (whose link invariants hold by
// construction and are tested), so the per-link checks are skipped.
void update_index(vdx_gpu_hnsw* h, const vdx_graph_view* g, const std::uint32_t* nodes, std::uint64_t count,
                  bool trusted = false) {
    require(h != nullptr && g != nullptr, "vdx_gpu_hnsw_update: null argument");
    require(count == 0 || nodes != nullptr, "vdx_gpu_hnsw_update: null node list");
    require(g->metric == h->metric && g->dim == h->dim && g->M == h->M,
            "vdx_gpu_hnsw_update: the graph's metric, dimension or M differs from the device copy's");
    require(g->n >= h->n, "vdx_gpu_hnsw_update: the graph has fewer nodes than the device copy");
    require(g->n < 0x80000000ull, "vdx_gpu_hnsw_update: more than 2^31 - 1 nodes");
    require(g->n == 0 || (g->levels && g->layer0 && g->vectors), "vdx_gpu_hnsw_update: null section");
    require(g->upper_blocks == 0 || g->upper != nullptr, "vdx_gpu_hnsw_update: null upper section");
    require(g->n == 0 || g->entry_point < g->n, "vdx_gpu_hnsw_update: entry point out of range");
    if (g->n == 0) return;
    const std::uint64_t old_n = h->n;
    require(std::memcmp(g->levels, h->levels.data(), old_n) == 0,
            "vdx_gpu_hnsw_update: a level of a node in the device copy changed");
    // New nodes' upper blocks follow the old ones in id order. A new node's level may exceed
    // max_level while a build has stored it but not linked it yet: the kernels never read
    // levels, and the link rules below keep every followed link on its node's layers.
    std::vector<std::uint32_t> new_offsets(g->n - old_n);
    std::uint64_t acc = h->upper_blocks;
    for (std::uint64_t i = old_n; i < g->n; ++i) {
        new_offsets[i - old_n] = static_cast<std::uint32_t>(acc);
        acc += g->levels[i];
    }
    require(acc == g->upper_blocks, "vdx_gpu_hnsw_update: levels do not add up to upper_blocks");
    require(acc <= 0xFFFFFFFFull, "vdx_gpu_hnsw_update: too many upper blocks");
    require(g->levels[g->entry_point] == g->max_level, "vdx_gpu_hnsw_update: the entry point is not on the top layer");
    // The changed old nodes (ascending, below old_n); every new node counts as changed.
    std::vector<std::uint32_t> changed;
    for (std::uint64_t i = 0; i < count; ++i) {
        require(nodes[i] < g->n, "vdx_gpu_hnsw_update: a listed node does not exist");
        require(i == 0 || nodes[i] > nodes[i - 1], "vdx_gpu_hnsw_update: the node list is not ascending");
        if (nodes[i] < old_n) changed.push_back(nodes[i]);
    }
    if (!trusted) {
        LinkChecker check_links(*g, "vdx_gpu_hnsw_update", h->link_marks, h->link_epoch);
        for (const std::uint32_t i : changed) check_links.node(i, h->offsets[i]);
        for (std::uint64_t i = old_n; i < g->n; ++i) check_links.node(i, new_offsets[i - old_n]);
    }

    DeviceGuard guard(h->ctx->device);
    hipStream_t s = h->ctx->stream;
    ++h->generation;  // from here on buffers may move, even if a step below fails
    const std::uint64_t b0 = 1 + 2 * g->M, bu = 1 + g->M, ld = h->split.ld;
    // Grows a buffer to hold `need` bytes with a quarter of headroom, keeping its contents.
    const auto ensure = [&](DeviceBuffer& buf, std::uint64_t need) {
        if (need <= buf.bytes()) return;
        DeviceBuffer grown(need + need / 4);
        if (buf.bytes() > 0) {
            check(hipMemcpyAsync(grown.as<void>(), buf.as<void>(), buf.bytes(), hipMemcpyDeviceToDevice, s),
                  "hipMemcpyAsync(grow)");
        }
        check(hipStreamSynchronize(s), "hipStreamSynchronize(grow)");
        buf.swap(grown);
    };
    ensure(h->vectors, g->n * ld * 4);
    ensure(h->layer0, g->n * b0 * 4);
    ensure(h->upper, std::max<std::uint64_t>(1, g->upper_blocks) * bu * 4);
    ensure(h->upper_offset, g->n * 4);
    const std::uint64_t add = g->n - old_n;
    if (add > 0) {
        copy_rows(h->vectors.as<float>() + old_n * ld, g->vectors + old_n * g->dim, add, g->dim, ld, s);
        check(hipMemcpyAsync(h->layer0.as<std::uint32_t>() + old_n * b0, g->layer0 + old_n * b0, add * b0 * 4,
                             hipMemcpyHostToDevice, s),
              "hipMemcpyAsync(layer0)");
        check(hipMemcpyAsync(h->upper_offset.as<std::uint32_t>() + old_n, new_offsets.data(), add * 4,
                             hipMemcpyHostToDevice, s),
              "hipMemcpyAsync(offsets)");
        const std::uint64_t new_blocks = g->upper_blocks - h->upper_blocks;
        if (new_blocks > 0) {
            check(hipMemcpyAsync(h->upper.as<std::uint32_t>() + h->upper_blocks * bu, g->upper + h->upper_blocks * bu,
                                 new_blocks * bu * 4, hipMemcpyHostToDevice, s),
                  "hipMemcpyAsync(upper)");
        }
    }
    std::vector<std::uint32_t> ublocks;
    for (const std::uint32_t i : changed) {
        for (std::uint32_t l = 0; l < g->levels[i]; ++l) ublocks.push_back(h->offsets[i] + l);
    }
    const std::uint64_t words = changed.size() * (1 + b0) + ublocks.size() * (1 + bu);
    if (words > 0) {
        // The changed blocks with their ids, gathered on the host into pinned staging (headroom:
        // the updates of a batched build grow; the gather runs on several threads, since a
        // batch of a large build rewrites millions of blocks), one copy over, scattered on the
        // device. Copying only the changed blocks from pinned memory beats a whole-section copy
        // from pageable memory even when most blocks changed.
        h->staging.reserve(words * 6);
        std::uint32_t* st = h->staging.as<std::uint32_t>();
        std::memcpy(st, changed.data(), changed.size() * 4);
        std::uint32_t* l0 = st + changed.size();
        constexpr std::size_t kRowsPerTask = 4096;
        const std::size_t tasks = (changed.size() + kRowsPerTask - 1) / kRowsPerTask;
        vdx::parallel_for(tasks, std::min<std::size_t>(16, vdx::resolve_threads(0)), [&](std::size_t t) {
            const std::size_t end = std::min(changed.size(), (t + 1) * kRowsPerTask);
            for (std::size_t r = t * kRowsPerTask; r < end; ++r) {
                std::memcpy(l0 + r * b0, g->layer0 + static_cast<std::uint64_t>(changed[r]) * b0, b0 * 4);
            }
        });
        std::uint32_t* uids = l0 + changed.size() * b0;
        std::memcpy(uids, ublocks.data(), ublocks.size() * 4);
        std::uint32_t* ub = uids + ublocks.size();
        for (std::size_t r = 0; r < ublocks.size(); ++r) {
            std::memcpy(ub + r * bu, g->upper + static_cast<std::uint64_t>(ublocks[r]) * bu, bu * 4);
        }
        h->changes.reserve(words * 6);
        check(hipMemcpyAsync(h->changes.as<void>(), st, words * 4, hipMemcpyHostToDevice, s),
              "hipMemcpyAsync(changes)");
        const std::uint32_t* d = h->changes.as<std::uint32_t>();
        check(vdx::dev::launch_scatter_rows(h->layer0.as<std::uint32_t>(), d + changed.size(), d, changed.size(),
                                            static_cast<std::uint32_t>(b0), s),
              "scatter(layer0)");
        check(vdx::dev::launch_scatter_rows(h->upper.as<std::uint32_t>(), d + changed.size() * (1 + b0) + ublocks.size(),
                                            d + changed.size() * (1 + b0), ublocks.size(),
                                            static_cast<std::uint32_t>(bu), s),
              "scatter(upper)");
    }
    check(hipStreamSynchronize(s), "hipStreamSynchronize(update)");
    h->levels.insert(h->levels.end(), g->levels + old_n, g->levels + g->n);
    h->offsets.insert(h->offsets.end(), new_offsets.begin(), new_offsets.end());
    h->n = g->n;
    h->upper_blocks = g->upper_blocks;
    h->max_level = g->max_level;
    h->entry_point = g->entry_point;
    refresh_rows_u8(*h, old_n, s);
}

// The GPU as the searcher of a batched build: the device copy kept current by update_index, each
// batch's insertion searches one beam launch over the nodes' own vectors gathered from HBM.
struct BuildSearch {
    vdx_gpu_hnsw* h = nullptr;
    DeviceBuffer ids, queries, out, count, queue;  // ids: the nodes, then their layer-0 entries
    DeviceBuffer chosen, chosen_count;             // the neighbour choice (m per node)
    vdx::gpu::PinnedBuffer h_ids, h_out, h_count;
    ~BuildSearch() {
        // a call that failed part-way may have left copies from or to these buffers queued
        if (h != nullptr) {
            int previous = 0;
            const bool got = hipGetDevice(&previous) == hipSuccess;
            (void)hipSetDevice(h->ctx->device);
            (void)hipStreamSynchronize(h->ctx->stream);
            if (got) (void)hipSetDevice(previous);
        }
    }
};

int build_sync(void* ctx, const vdx_graph_view* view, const std::uint32_t* changed, std::uint64_t count) {
    return guarded([&] { update_index(static_cast<BuildSearch*>(ctx)->h, view, changed, count, true); });
}

int build_search(void* ctx, const std::uint32_t* nodes, const std::uint32_t* entries, std::uint64_t count,
                 std::uint32_t layer, std::uint32_t ef, std::uint32_t m, vdx_neighbor* out, std::uint32_t* counts) {
    return guarded([&] {
        BuildSearch& b = *static_cast<BuildSearch*>(ctx);
        vdx_gpu_hnsw& h = *b.h;
        if (ef > kMaxBeamEf) {
            throw std::invalid_argument(vdx::fmt("HnswIndex::add: ef_construction %u is above the GPU path's limit of %u",
                                                 ef, kMaxBeamEf));
        }
        require(count < 0x80000000ull, "HnswIndex::add: batch too large");
        for (std::uint64_t i = 0; i < count; ++i) {
            require(nodes[i] < h.n, "HnswIndex::add: a batch node is not in the device copy");
            require(entries == nullptr || entries[i] == 0xFFFFFFFFu ||
                        (entries[i] < h.n && h.levels[entries[i]] >= layer),
                    "HnswIndex::add: an entry is not a node of the searched layer");
            require(layer == 0 || h.levels[nodes[i]] >= layer, "HnswIndex::add: a node does not reach the searched layer");
        }
        require(layer <= h.max_level, "HnswIndex::add: the searched layer is above the graph's top");
        if (count == 0) return;
        DeviceGuard guard(h.ctx->device);
        hipStream_t s = h.ctx->stream;
        // sizes with headroom: batches grow
        const auto grow = [](std::uint64_t bytes) { return bytes + bytes / 2; };
        b.h_ids.reserve(grow(count * 8));
        require(m <= kMaxGpuM, "HnswIndex::add: M above the GPU path's limit");
        const std::uint64_t stride = m > 0 ? m : ef;
        b.h_out.reserve(grow(count * stride * sizeof(vdx_neighbor)));
        b.h_count.reserve(grow(count * 4));
        b.ids.reserve(grow(count * 8));
        b.queries.reserve(grow(count * h.dim * 4));
        b.out.reserve(grow(count * ef * sizeof(vdx_neighbor)));
        b.count.reserve(grow(count * 4));
        b.queue.reserve(64);
        std::uint32_t* staged = b.h_ids.as<std::uint32_t>();
        std::memcpy(staged, nodes, count * 4);
        if (entries != nullptr) {
            std::memcpy(staged + count, entries, count * 4);
        }
        check(hipMemcpyAsync(b.ids.as<void>(), staged, (entries != nullptr ? 8 : 4) * count, hipMemcpyHostToDevice, s),
              "hipMemcpyAsync(batch ids)");
        check(vdx::dev::launch_gather_rows(b.queries.as<float>(), static_cast<std::uint32_t>(h.dim), h.vectors.as<float>(),
                                           h.split.ld, b.ids.as<std::uint32_t>(), count,
                                           static_cast<std::uint32_t>(h.dim), s),
              "gather(batch vectors)");
        enqueue_beam(h, b.queries.as<float>(), count, nullptr, ef, nullptr, ef, ef, b.out.as<vdx::dev::Neighbor>(), ef,
                     b.count.as<std::uint32_t>(), b.queue.as<std::uint32_t>(), nullptr, nullptr, s, nullptr,
                     entries != nullptr ? b.ids.as<std::uint32_t>() + count : nullptr, layer);
        const void* d_res = b.out.as<void>();
        const void* d_cnt = b.count.as<void>();
        if (m > 0) {
            b.chosen.reserve(grow(count * m * sizeof(vdx_neighbor)));
            b.chosen_count.reserve(grow(count * 4));
            vdx::dev::SelectArgs sa{};
            sa.vectors = h.vectors.as<float>();
            sa.ld = h.split.ld;
            sa.l2 = h.metric == 0 ? 1u : 0u;
            sa.nb = h.split.nb;
            sa.n8 = h.split.n8;
            sa.tail = h.split.tail;
            sa.first_tail = h.split.first_tail;
            sa.found = b.out.as<vdx::dev::Neighbor>();
            sa.found_stride = ef;
            sa.found_count = b.count.as<std::uint32_t>();
            sa.m = m;
            sa.count = count;
            sa.chosen = b.chosen.as<vdx::dev::Neighbor>();
            sa.chosen_count = b.chosen_count.as<std::uint32_t>();
            check(vdx::dev::launch_select_links(sa, vdx::gpu::beam_shape(h.dim), s), "select_links launch");
            d_res = b.chosen.as<void>();
            d_cnt = b.chosen_count.as<void>();
        }
        check(hipMemcpyAsync(b.h_out.as<void>(), d_res, count * stride * sizeof(vdx_neighbor), hipMemcpyDeviceToHost, s),
              "hipMemcpyAsync(batch results)");
        check(hipMemcpyAsync(b.h_count.as<void>(), d_cnt, count * 4, hipMemcpyDeviceToHost, s),
              "hipMemcpyAsync(batch counts)");
        check(hipStreamSynchronize(s), "batch search");
        std::memcpy(out, b.h_out.as<void>(), count * stride * sizeof(vdx_neighbor));
        std::memcpy(counts, b.h_count.as<void>(), count * 4);
    });
}

// Queries per item of the host pool's validation and staging, and per item of its unpacking.
constexpr std::uint64_t kHostChunk = 128;
constexpr std::uint64_t kUnpackChunk = 1024;

// Checks every query (vecdex's check_query, its messages; the first bad row's error, as a
// sequential loop reports) and writes the prepared rows (normalized under cosine) to `staging`,
// in chunks on the context's host threads. `nonbyte`, when given, is set to 1 unless every
// component staged is an integer in [0, 255] (the 8-bit rows' integer kernel). With `bytes`
// as well (dimension 128), a batch whose every component is such an integer is staged there
// instead, one byte per component in order, and `staging` is left unwritten; any other batch
// is staged as floats in full.
void prepare_queries(vdx_gpu_ctx& ctx, const char* where, vdx::Metric metric, std::uint64_t dim,
                     std::uint64_t size, const float* queries, std::uint64_t nq, std::uint64_t query_dim,
                     const std::uint32_t* ks, std::uint32_t k, float* staging,
                     std::atomic<std::uint32_t>* nonbyte = nullptr, std::uint8_t* bytes = nullptr) {
    const vdx::IndexShape shape{metric, static_cast<std::size_t>(dim), static_cast<std::size_t>(size)};
    const std::string w(where);
    const std::size_t chunks = static_cast<std::size_t>((nq + kHostChunk - 1) / kHostChunk);
    // chunks staged as bytes (to be restaged as floats if another chunk is not byte-valued)
    std::unique_ptr<std::uint8_t[]> as_bytes(bytes != nullptr ? new std::uint8_t[chunks]() : nullptr);
    const auto stage = [&](std::uint64_t begin, std::uint64_t end) {
        if (bytes != nullptr && nonbyte != nullptr && metric != vdx::Metric::Cosine && query_dim == dim &&
            nonbyte->load(std::memory_order_relaxed) == 0) {
            // one pass that checks, then the bytes if every value is a byte (else the float
            // path below, which checks again)
            std::uint32_t bad = 0, nb = 0;
            for (std::uint64_t q = begin; q < end; ++q) {
                const std::uint32_t* src = reinterpret_cast<const std::uint32_t*>(queries + q * dim);
                const float* f = queries + q * dim;
                bad |= (ks ? ks[q] : k) == 0 ? 1u : 0u;
                for (std::uint64_t i = 0; i < dim; ++i) {
                    bad |= (src[i] & 0x7F800000u) == 0x7F800000u ? 1u : 0u;
                    const float c = (f[i] >= 0.0f && f[i] <= 255.0f) ? f[i] : 0.5f;
                    nb |= static_cast<float>(static_cast<int>(c)) != c ? 1u : 0u;
                }
            }
            if (bad == 0 && nb == 0) {
                for (std::uint64_t q = begin; q < end; ++q) {
                    const float* f = queries + q * dim;
                    std::uint8_t* dst = bytes + q * dim;
                    for (std::uint64_t i = 0; i < dim; ++i) {
                        dst[i] = static_cast<std::uint8_t>(static_cast<int>(f[i]));
                    }
                }
                as_bytes[begin / kHostChunk] = 1;
                return;
            }
        }
        if (metric != vdx::Metric::Cosine && query_dim == dim) {
            // Squared L2 and inner product stage the rows as they are: one pass copies them and
            // looks for a non-finite value (exponent all ones) and a zero k; a chunk with either
            // is redone row by row below, which raises check_query's error for its first bad row.
            bool clean = true;
            std::uint32_t nb = 0;
            for (std::uint64_t q = begin; q < end; ++q) {
                const std::uint32_t* src = reinterpret_cast<const std::uint32_t*>(queries + q * dim);
                std::uint32_t* dst = reinterpret_cast<std::uint32_t*>(staging + q * dim);
                std::uint32_t bad = (ks ? ks[q] : k) == 0 ? 1u : 0u;
                if (nonbyte != nullptr) {
                    const float* f = queries + q * dim;
                    for (std::uint64_t i = 0; i < dim; ++i) {
                        const std::uint32_t b = src[i];
                        dst[i] = b;
                        bad |= (b & 0x7F800000u) == 0x7F800000u ? 1u : 0u;
                        // a value outside [0, 255] (or NaN) becomes 0.5, never an integer
                        const float c = (f[i] >= 0.0f && f[i] <= 255.0f) ? f[i] : 0.5f;
                        nb |= static_cast<float>(static_cast<int>(c)) != c ? 1u : 0u;
                    }
                } else {
                    for (std::uint64_t i = 0; i < dim; ++i) {
                        const std::uint32_t b = src[i];
                        dst[i] = b;
                        bad |= (b & 0x7F800000u) == 0x7F800000u ? 1u : 0u;
                    }
                }
                clean = clean && bad == 0;
            }
            if (clean) {
                if (nb != 0) {
                    nonbyte->store(1u, std::memory_order_relaxed);
                }
                return;
            }
        }
        if (nonbyte != nullptr) {
            nonbyte->store(1u, std::memory_order_relaxed);
        }
        for (std::uint64_t q = begin; q < end; ++q) {
            const float* row = queries + q * query_dim;
            vdx::check_query(w, shape, row, query_dim, ks ? ks[q] : k);
            vdx::prepare_query(metric, row, dim, staging + q * dim);
        }
    };
    const auto chunk = [&](std::size_t c) {
        const std::uint64_t begin = c * kHostChunk;
        stage(begin, std::min<std::uint64_t>(nq, begin + kHostChunk));
    };
    if (ctx.host_pool) {
        ctx.host_pool->run(chunks, chunk);
    } else {
        for (std::size_t c = 0; c < chunks; ++c) chunk(c);
    }
    if (bytes != nullptr && nonbyte != nullptr && nonbyte->load() != 0) {
        // a mixed batch runs on floats: the chunks staged as bytes (checked) are copied too
        for (std::size_t c = 0; c < chunks; ++c) {
            if (as_bytes[c]) {
                const std::uint64_t b = c * kHostChunk, e = std::min<std::uint64_t>(nq, b + kHostChunk);
                std::memcpy(staging + b * dim, queries + b * dim, (e - b) * dim * sizeof(float));
            }
        }
    }
}

}  // namespace

extern "C" {

int vdx_gpu_open(int device, vdx_gpu_ctx** out) {
    return guarded([&] {
        require(out != nullptr, "vdx_gpu_open: out is null");
        int count = 0;
        check(hipGetDeviceCount(&count), "hipGetDeviceCount");
        if (device < 0 || device >= count) {
            throw std::invalid_argument(vdx::fmt("vdx_gpu_open: device %d does not exist (%d visible)",
                                                 device, count));
        }
        auto ctx = std::make_unique<vdx_gpu_ctx>();
        ctx->device = device;
        DeviceGuard guard(device);
        hipDeviceProp_t prop{};
        check(hipGetDeviceProperties(&prop, device), "hipGetDeviceProperties");
        ctx->cu_count = prop.multiProcessorCount;
        ctx->lds_per_block = prop.sharedMemPerBlock;
        check(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking), "hipStreamCreate");
        ctx->host_pool = std::make_unique<vdx::HostPool>(
            std::min<std::size_t>(16, std::max(1u, std::thread::hardware_concurrency())));
        vdx::gpu::noalloc_scratch_attach(device);  // with the device current
        *out = ctx.release();
    });
}

void vdx_gpu_close(vdx_gpu_ctx* ctx) {
    if (ctx != nullptr) {
        DeviceGuard guard(ctx->device);
        delete ctx;
    }
}

void* vdx_gpu_stream(vdx_gpu_ctx* ctx) { return ctx ? static_cast<void*>(ctx->stream) : nullptr; }

int vdx_gpu_alloc(vdx_gpu_ctx* ctx, uint64_t bytes, void** out) {
    return guarded([&] {
        require(ctx != nullptr && out != nullptr, "vdx_gpu_alloc: null argument");
        DeviceGuard guard(ctx->device);
        *out = nullptr;
        check(hipMalloc(out, std::max<uint64_t>(bytes, 1)), "hipMalloc");
    });
}

void vdx_gpu_free(vdx_gpu_ctx* ctx, void* ptr) {
    if (ctx != nullptr && ptr != nullptr) {
        DeviceGuard guard(ctx->device);
        (void)hipStreamSynchronize(ctx->stream);
        (void)hipFree(ptr);
    }
}

int vdx_gpu_copy_to_device(vdx_gpu_ctx* ctx, void* dst, const void* src, uint64_t bytes) {
    return guarded([&] {
        require(ctx != nullptr, "vdx_gpu_copy_to_device: null context");
        DeviceGuard guard(ctx->device);
        check(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream), "hipMemcpyAsync");
        check(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    });
}

int vdx_gpu_copy_device(vdx_gpu_ctx* ctx, void* dst, const void* src, uint64_t bytes) {
    return guarded([&] {
        require(ctx != nullptr, "vdx_gpu_copy_device: null context");
        require(bytes == 0 || (dst != nullptr && src != nullptr), "vdx_gpu_copy_device: null pointer");
        require(reinterpret_cast<std::uintptr_t>(dst) % 16 == 0 && reinterpret_cast<std::uintptr_t>(src) % 16 == 0,
                "vdx_gpu_copy_device: pointers must be 16-byte aligned");
        DeviceGuard guard(ctx->device);
        check(vdx::dev::launch_copy_bytes(dst, src, bytes, ctx->stream), "copy_bytes launch");
        check(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    });
}

int vdx_gpu_gather_probe(vdx_gpu_ctx* ctx, const void* table, uint64_t table_bytes, uint32_t row_bytes,
                         uint32_t rows_per_wave, uint32_t waves, void* sink, uint64_t* rows_read) {
    return guarded([&] {
        require(ctx != nullptr && table != nullptr && sink != nullptr, "vdx_gpu_gather_probe: null argument");
        require(row_bytes >= 16 && row_bytes <= 1024 && row_bytes % 16 == 0 && 64 % (row_bytes / 16) == 0,
                "vdx_gpu_gather_probe: row_bytes must be 16, 32, ..., 1024 (16 bytes a lane, lanes dividing 64)");
        require(waves >= 4 && waves % 4 == 0, "vdx_gpu_gather_probe: waves must be a positive multiple of 4");
        require(reinterpret_cast<std::uintptr_t>(table) % 16 == 0, "vdx_gpu_gather_probe: table must be 16-byte aligned");
        const std::uint64_t rows = table_bytes / row_bytes;
        require(rows >= 1 && rows <= 0xFFFFFFFFull, "vdx_gpu_gather_probe: the table holds no row or too many");
        const std::uint64_t round = (64 / (row_bytes / 16)) * 4;
        const std::uint64_t per_wave = (rows_per_wave + round - 1) / round * round;
        require(per_wave <= 0xFFFFFFFFull - round, "vdx_gpu_gather_probe: rows_per_wave too large");
        DeviceGuard guard(ctx->device);
        check(vdx::dev::launch_gather_probe(table, static_cast<std::uint32_t>(rows), row_bytes,
                                            static_cast<std::uint32_t>(per_wave), waves,
                                            static_cast<std::uint32_t*>(sink), ctx->stream),
              "gather_probe launch");
        check(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
        if (rows_read != nullptr) {
            *rows_read = per_wave * waves;
        }
    });
}

int vdx_gpu_copy_to_host(vdx_gpu_ctx* ctx, void* dst, const void* src, uint64_t bytes) {
    return guarded([&] {
        require(ctx != nullptr, "vdx_gpu_copy_to_host: null context");
        DeviceGuard guard(ctx->device);
        check(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream), "hipMemcpyAsync");
        check(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    });
}

int vdx_gpu_synchronize(vdx_gpu_ctx* ctx) {
    return guarded([&] {
        require(ctx != nullptr, "vdx_gpu_synchronize: null context");
        DeviceGuard guard(ctx->device);
        check(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    });
}

int vdx_gpu_event_record(vdx_gpu_ctx* ctx, uint32_t slot) {
    return guarded([&] {
        require(ctx != nullptr && slot < 16, "vdx_gpu_event_record: bad argument");
        DeviceGuard guard(ctx->device);
        if (ctx->events[slot] == nullptr) {
            check(hipEventCreate(&ctx->events[slot]), "hipEventCreate");
        }
        check(hipEventRecord(ctx->events[slot], ctx->stream), "hipEventRecord");
    });
}

int vdx_gpu_event_elapsed(vdx_gpu_ctx* ctx, uint32_t from, uint32_t to, float* ms) {
    return guarded([&] {
        require(ctx != nullptr && from < 16 && to < 16 && ms != nullptr,
                "vdx_gpu_event_elapsed: bad argument");
        require(ctx->events[from] != nullptr && ctx->events[to] != nullptr,
                "vdx_gpu_event_elapsed: event not recorded");
        DeviceGuard guard(ctx->device);
        check(hipEventSynchronize(ctx->events[to]), "hipEventSynchronize");
        check(hipEventElapsedTime(ms, ctx->events[from], ctx->events[to]), "hipEventElapsedTime");
    });
}

int vdx_gpu_hnsw_upload(vdx_gpu_ctx* ctx, const vdx_graph_view* g, vdx_gpu_hnsw** out) {
    return guarded([&] {
        require(ctx != nullptr && g != nullptr && out != nullptr, "vdx_gpu_hnsw_upload: null argument");
        require(vdx::valid_metric(g->metric), "vdx_gpu_hnsw_upload: unknown metric code");
        require(g->dim >= 1, "vdx_gpu_hnsw_upload: dimension is 0");
        if (g->dim > kMaxGpuDim) {
            throw std::invalid_argument(vdx::fmt(
                "vdx_gpu_hnsw_upload: dimension %llu is above the GPU path's limit of %u",
                static_cast<unsigned long long>(g->dim), kMaxGpuDim));
        }
        if (g->M < 2 || g->M > kMaxGpuM) {
            throw std::invalid_argument(vdx::fmt(
                "vdx_gpu_hnsw_upload: M = %llu is outside the GPU path's range [2, %u]",
                static_cast<unsigned long long>(g->M), kMaxGpuM));
        }
        require(g->n < 0x80000000ull, "vdx_gpu_hnsw_upload: more than 2^31 - 1 nodes");
        require(g->n == 0 || (g->levels && g->layer0 && g->vectors), "vdx_gpu_hnsw_upload: null section");
        require(g->upper_blocks == 0 || g->upper != nullptr, "vdx_gpu_hnsw_upload: null upper section");
        require(g->n == 0 || g->entry_point < g->n, "vdx_gpu_hnsw_upload: entry point out of range");
        // Offsets of each node's upper blocks; the levels must add up to U.
        std::vector<std::uint32_t> offsets(std::max<std::uint64_t>(1, g->n));
        std::uint64_t acc = 0;
        for (std::uint64_t i = 0; i < g->n; ++i) {
            offsets[i] = static_cast<std::uint32_t>(acc);
            require(g->levels[i] <= g->max_level, "vdx_gpu_hnsw_upload: a level is above max_level");
            acc += g->levels[i];
        }
        require(acc == g->upper_blocks, "vdx_gpu_hnsw_upload: levels do not add up to upper_blocks");
        require(acc <= 0xFFFFFFFFull, "vdx_gpu_hnsw_upload: too many upper blocks");
        require(g->n == 0 || g->levels[g->entry_point] == g->max_level,
                "vdx_gpu_hnsw_upload: the entry point is not on the top layer");
        // Every block the kernel may follow, with the checks of HnswIndex::check_links
        // (src/hnsw_io.cpp:472-504).
        auto h = std::make_unique<vdx_gpu_hnsw>();
        {
            LinkChecker check_links(*g, "vdx_gpu_hnsw_upload", h->link_marks, h->link_epoch);
            std::uint64_t off = 0;
            for (std::uint64_t i = 0; i < g->n; ++i) {
                check_links.node(i, off);
                off += g->levels[i];
            }
        }

        DeviceGuard guard(ctx->device);
        h->ctx = ctx;
        h->metric = g->metric;
        h->dim = g->dim;
        h->n = g->n;
        h->M = g->M;
        h->max_level = g->n ? g->max_level : 0;
        h->entry_point = g->n ? g->entry_point : 0;
        h->upper_blocks = g->upper_blocks;
        h->levels.assign(g->levels, g->levels + g->n);
        h->offsets.assign(offsets.begin(), offsets.begin() + g->n);
        h->split = vdx::gpu::split_dim(g->dim);
        hipStream_t s = ctx->stream;
        upload_rows(h->vectors, g->vectors, g->n, g->dim, h->split.ld, s);
        const std::uint64_t l0 = g->n * (1 + 2 * g->M);
        h->layer0.resize(std::max<std::uint64_t>(1, l0) * 4);
        if (l0) {
            check(hipMemcpyAsync(h->layer0.as<void>(), g->layer0, l0 * 4, hipMemcpyHostToDevice, s),
                  "hipMemcpyAsync(layer0)");
        }
        const std::uint64_t up = g->upper_blocks * (1 + g->M);
        h->upper.resize(std::max<std::uint64_t>(1, up) * 4);
        if (up) {
            check(hipMemcpyAsync(h->upper.as<void>(), g->upper, up * 4, hipMemcpyHostToDevice, s),
                  "hipMemcpyAsync(upper)");
        }
        h->upper_offset.resize(offsets.size() * 4);
        check(hipMemcpyAsync(h->upper_offset.as<void>(), offsets.data(), offsets.size() * 4,
                             hipMemcpyHostToDevice, s),
              "hipMemcpyAsync(offsets)");
        h->ring.resize(kRingSlots * 256);
        check(hipStreamSynchronize(s), "hipStreamSynchronize(upload)");
        refresh_rows_u8(*h, 0, s);
        *out = h.release();
    });
}

int vdx_gpu_hnsw_update(vdx_gpu_hnsw* h, const vdx_graph_view* g, const uint32_t* nodes, uint64_t count) {
    return guarded([&] { update_index(h, g, nodes, count); });
}

int vdx_gpu_hnsw_build_add(vdx_gpu_hnsw* index, vdx_host_graph* graph, const float* data, uint64_t rows,
                           uint64_t dim, uint64_t num_threads, const vdx_build_plan* plan, uint64_t synced_version,
                           vdx_build_stats* stats) {
    return guarded([&] {
        require(index != nullptr, "vdx_gpu_hnsw_build_add: index is null");
        BuildSearch b;
        b.h = index;
        const vdx_batch_searcher searcher{&b, &build_sync, &build_search,
                                          VDX_SEARCHER_SELECTS | VDX_SEARCHER_UPPER_LAYERS};
        vdx::host_graph_add_batched(graph, data, rows, dim, num_threads, searcher, true, plan, synced_version, stats);
    });
}

int vdx_gpu_hnsw_load(vdx_gpu_ctx* ctx, const char* path, vdx_gpu_hnsw** out) {
    return guarded([&] {
        require(ctx != nullptr && path != nullptr && out != nullptr, "vdx_gpu_hnsw_load: null argument");
        vdx::V2File map(path);
        const vdx::V2Header& hd = map.header();
        if (hd.dim > kMaxGpuDim) {
            throw std::invalid_argument(vdx::fmt("vdx_gpu_hnsw_load: dimension %llu is above the GPU path's limit of %u",
                                                 static_cast<unsigned long long>(hd.dim), kMaxGpuDim));
        }
        if (hd.M > kMaxGpuM) {
            throw std::invalid_argument(vdx::fmt("vdx_gpu_hnsw_load: M = %llu is outside the GPU path's range [2, %u]",
                                                 static_cast<unsigned long long>(hd.M), kMaxGpuM));
        }
        vdx::V2Checker verify(map);
        verify.levels_and_offsets();

        DeviceGuard guard(ctx->device);
        auto h = std::make_unique<vdx_gpu_hnsw>();
        h->ctx = ctx;
        h->metric = hd.metric;
        h->dim = hd.dim;
        h->n = hd.n;
        h->M = hd.M;
        h->max_level = hd.max_level;
        h->entry_point = hd.entry_point;
        h->upper_blocks = hd.upper_blocks;
        h->split = vdx::gpu::split_dim(hd.dim);
        require(h->split.ld == hd.ld, "vdx_gpu_hnsw_load: row length differs from the device layout");
        const std::uint64_t b0 = 1 + 2 * hd.M, bu = 1 + hd.M;
        h->vectors.resize(std::max<std::uint64_t>(1, hd.n) * hd.ld * 4);
        h->layer0.resize(std::max<std::uint64_t>(1, hd.n * b0) * 4);
        h->upper.resize(std::max<std::uint64_t>(1, hd.upper_blocks * bu) * 4);
        h->upper_offset.resize(std::max<std::uint64_t>(1, hd.n) * 4);
        h->ring.resize(kRingSlots * 256);
        hipStream_t s = ctx->stream;
        if (hd.n > 0) {
            check(hipMemcpyAsync(h->upper_offset.as<void>(), verify.offsets().data(), hd.n * 4, hipMemcpyHostToDevice, s),
                  "hipMemcpyAsync(offsets)");
        }
        // Two pinned chunks in turn: the file is read into one (and checked there, cache-hot)
        // while the other's copy to HBM is in flight.
        constexpr std::uint64_t kChunkBytes = std::uint64_t{64} << 20;
        vdx::gpu::PinnedBuffer pinned[2];
        hipEvent_t done[2] = {nullptr, nullptr};
        struct Drain {  // no copy may still read a pinned chunk when it is freed
            hipStream_t s;
            hipEvent_t* ev;
            ~Drain() {
                (void)hipStreamSynchronize(s);
                for (int i = 0; i < 2; ++i) {
                    if (ev[i] != nullptr) (void)hipEventDestroy(ev[i]);
                }
            }
        } drain{s, done};
        for (int i = 0; i < 2; ++i) {
            check(hipEventCreateWithFlags(&done[i], hipEventDisableTiming), "hipEventCreate");
        }
        std::uint64_t chunk_no = 0;
        bool in_flight[2] = {false, false};
        const auto stream_section = [&](int sec, std::uint64_t unit_bytes, std::uint64_t units, char* dst,
                                        const auto& check_chunk) {
            const std::uint64_t per = std::max<std::uint64_t>(1, kChunkBytes / unit_bytes);
            for (std::uint64_t u0 = 0; u0 < units; u0 += per, ++chunk_no) {
                const std::uint64_t m = std::min(per, units - u0);
                const std::uint64_t bytes = m * unit_bytes;
                const int slot = static_cast<int>(chunk_no & 1);
                if (in_flight[slot]) {
                    check(hipEventSynchronize(done[slot]), "hipEventSynchronize");
                }
                pinned[slot].reserve(std::min(units, per) * unit_bytes);
                map.read(sec, u0 * unit_bytes, pinned[slot].as<void>(), bytes);
                check_chunk(pinned[slot].as<void>(), u0, m);
                check(hipMemcpyAsync(dst + u0 * unit_bytes, pinned[slot].as<void>(), bytes, hipMemcpyHostToDevice, s),
                      "hipMemcpyAsync(section)");
                check(hipEventRecord(done[slot], s), "hipEventRecord");
                in_flight[slot] = true;
            }
        };
        stream_section(vdx::kV2Layer0, b0 * 4, hd.n, h->layer0.as<char>(), [&](void* p, std::uint64_t first, std::uint64_t m) {
            verify.layer0(static_cast<const std::uint32_t*>(p), first, m);
        });
        stream_section(vdx::kV2Upper, bu * 4, hd.upper_blocks, h->upper.as<char>(),
                       [&](void* p, std::uint64_t first, std::uint64_t m) {
                           verify.upper(static_cast<const std::uint32_t*>(p), first, m);
                       });
        stream_section(vdx::kV2Vectors, hd.ld * 4, hd.n, h->vectors.as<char>(),
                       [&](void* p, std::uint64_t first, std::uint64_t m) {
                           verify.vectors(static_cast<const float*>(p), first, m);
                       });
        verify.finish();
        check(hipStreamSynchronize(s), "hipStreamSynchronize(load)");
        h->levels = verify.levels();
        h->offsets = verify.offsets();
        refresh_rows_u8(*h, 0, s);
        *out = h.release();
    });
}

int vdx_gpu_hnsw_info(const vdx_gpu_hnsw* h, uint32_t* metric, uint64_t* dim, uint64_t* n, uint64_t* M) {
    return guarded([&] {
        require(h != nullptr, "vdx_gpu_hnsw_info: index is null");
        if (metric) *metric = h->metric;
        if (dim) *dim = h->dim;
        if (n) *n = h->n;
        if (M) *M = h->M;
    });
}

int vdx_gpu_hnsw_u8_rows(const vdx_gpu_hnsw* h, uint32_t* kept) {
    return guarded([&] {
        require(h != nullptr, "vdx_gpu_hnsw_u8_rows: index is null");
        require(kept != nullptr, "vdx_gpu_hnsw_u8_rows: kept is null");
        *kept = h->u8 ? 1u : 0u;
    });
}

void vdx_gpu_hnsw_free(vdx_gpu_hnsw* index) {
    if (index != nullptr) {
        DeviceGuard guard(index->ctx->device);
        delete index;
    }
}

int vdx_gpu_hnsw_search(vdx_gpu_hnsw* h, const float* queries, uint64_t nq, uint64_t query_dim,
                        const uint32_t* ks, uint32_t k, const uint32_t* efs, uint32_t ef,
                        vdx_neighbor* out, uint64_t out_stride, uint32_t* out_count,
                        const char* where, const vdx_search_opts* opts) {
    return guarded([&] {
        require(h != nullptr, "vdx_gpu_hnsw_search: index is null");
        const char* w = where ? where : "HnswIndex::search_batch";
        if (nq == 0) {
            return;
        }
        require(queries != nullptr && out != nullptr && out_count != nullptr,
                "vdx_gpu_hnsw_search: null buffer");
        std::uint32_t kmax = 0, efmax = 0;
        for (std::uint64_t q = 0; q < nq; ++q) {
            const std::uint32_t kq = ks ? ks[q] : k;
            const std::uint32_t eq = efs ? efs[q] : ef;
            kmax = std::max(kmax, kq);
            efmax = std::max(efmax, std::max(eq, kq));
        }
        require(out_stride >= kmax, "vdx_gpu_hnsw_search: out_stride is below k");
        DeviceGuard guard(h->ctx->device);
        WorkspaceLease ws(*h->ctx);
        ws->h_queries.reserve(nq * h->dim * sizeof(float));
        // Validation needs the dimension check first (query_dim may differ from dim).
        if (query_dim != h->dim) {
            vdx::check_query(w, {static_cast<vdx::Metric>(h->metric), h->dim, h->n}, queries,
                             query_dim, ks ? ks[0] : k);
        }
        const vdx::Metric metric = static_cast<vdx::Metric>(h->metric);
        // The host threads check and stage the queries in mapped pinned memory, which the search
        // kernel reads directly over PCIe (each query once, when a wave takes it up) and into
        // which it writes the results: no copy phase before or after the launch. Every query is
        // checked before anything reaches the caller's buffers, and the first bad row's error is
        // the one reported, as with one pass. Over the 8-bit rows the same pass tells whether
        // every query is byte-valued, which picks the integer kernel.
        hipStream_t s = ws->stream;
        const bool want_bytes = h->u8 && h->metric == 0 && h->dim == 128 && query_dim == 128;
        std::atomic<std::uint32_t> nonbyte{0};
        // Byte-valued queries cross PCIe as bytes (a quarter of the floats) when the integer
        // kernel is the one that will read them: the register-resident beam over the 8-bit rows.
        // An option error here is left for the launch to report, after the query checks.
        bool stage_bytes = false;
        if (want_bytes && efmax <= kMaxBeamEf) {
            try {
                const vdx::gpu::BeamPlan plan = vdx::gpu::plan_beam(
                    *h, efmax, nq, vdx::gpu::config_from(opts), vdx::gpu::beam_mode(opts, false));
                stage_bytes = plan.shape == vdx::dev::BeamShape::D128U8 && plan.chunks > 0;
            } catch (const std::exception&) {
                stage_bytes = false;
            }
        }
        if (stage_bytes) {
            ws->h_qbytes.reserve(nq * 128);
        }
        try {
            prepare_queries(*h->ctx, w, metric, h->dim, h->n, queries, nq, query_dim, ks, k,
                            ws->h_queries.as<float>(), want_bytes ? &nonbyte : nullptr,
                            stage_bytes ? ws->h_qbytes.as<std::uint8_t>() : nullptr);
        } catch (...) {
            // nothing in flight may outlive the call (the workspace goes back to the pool)
            (void)hipStreamSynchronize(s);
            throw;
        }
        if (h->n == 0) {
            std::fill(out_count, out_count + nq, 0u);
            std::memset(static_cast<void*>(out), 0, nq * out_stride * sizeof(vdx_neighbor));
            return;
        }
        if (efmax > kMaxBeamEf) {
            throw std::invalid_argument(vdx::fmt("%s: ef %u is above the GPU path's limit of %u", w,
                                                 efmax, kMaxBeamEf));
        }
        ws->counters.reserve(128);  // the queue counter, then the stats at byte 16
        ws->h_out.reserve(nq * kmax * sizeof(vdx_neighbor));
        ws->h_count.reserve((nq * 4 + 7) / 8 * 8 + sizeof(unsigned long long) * vdx::dev::kBeamStats);
        const std::uint32_t* d_ks = nullptr;
        const std::uint32_t* d_efs = nullptr;
        if (ks) {
            ws->ks.reserve(nq * 4);
            check(hipMemcpyAsync(ws->ks.as<void>(), ks, nq * 4, hipMemcpyHostToDevice, s), "hipMemcpyAsync(ks)");
            d_ks = ws->ks.as<std::uint32_t>();
        }
        if (efs) {
            ws->efs.reserve(nq * 4);
            check(hipMemcpyAsync(ws->efs.as<void>(), efs, nq * 4, hipMemcpyHostToDevice, s),
                  "hipMemcpyAsync(efs)");
            d_efs = ws->efs.as<std::uint32_t>();
        }
        unsigned long long* d_stats = nullptr;
        if (opts && opts->stats) {
            d_stats = reinterpret_cast<unsigned long long*>(ws->counters.as<char>() + 16);
            check(hipMemsetAsync(d_stats, 0, vdx::dev::kBeamStats * sizeof(unsigned long long), s),
                  "hipMemsetAsync(stats)");
        }
        // per-query budgets arrive in host memory here: the kernels read a device copy
        vdx_search_opts dopts{};
        const vdx_search_opts* kopts = opts;
        if (opts && opts->max_expansions_per_query) {
            ws->aux3.reserve(nq * 4);
            check(hipMemcpyAsync(ws->aux3.as<void>(), opts->max_expansions_per_query, nq * 4,
                                 hipMemcpyHostToDevice, s),
                  "hipMemcpyAsync(budgets)");
            dopts = *opts;
            dopts.max_expansions_per_query = ws->aux3.as<std::uint32_t>();
            kopts = &dopts;
        }
        enqueue_beam(*h, ws->h_queries.dev<float>(), nq, d_ks, k, d_efs, ef, efmax,
                     ws->h_out.dev<vdx::dev::Neighbor>(), kmax, ws->h_count.dev<std::uint32_t>(),
                     ws->counters.as<std::uint32_t>(), d_stats, kopts, s, nullptr, nullptr, 0, &ws->bitmap,
                     want_bytes ? (nonbyte.load() != 0 ? 2 : 1) : 0,
                     stage_bytes && nonbyte.load() == 0 ? ws->h_qbytes.dev<std::uint8_t>() : nullptr);
        // stats after the counts, 8-byte aligned
        unsigned long long* host_stats =
            reinterpret_cast<unsigned long long*>(ws->h_count.as<char>() + (nq * 4 + 7) / 8 * 8);
        if (d_stats) {
            check(hipMemcpyAsync(host_stats, d_stats, sizeof(unsigned long long) * vdx::dev::kBeamStats,
                                 hipMemcpyDeviceToHost, s),
                  "hipMemcpyAsync(stats)");
        }
        check(hipStreamSynchronize(s), "hnsw search");
        const vdx_neighbor* staged = ws->h_out.as<vdx_neighbor>();
        const std::uint32_t* staged_count = ws->h_count.as<std::uint32_t>();
        // ranges of kUnpackChunk queries: one copy where every row is full and the rows are
        // adjacent in both layouts, else row by row with the rest of each row zeroed
        const auto unpack = [&](std::size_t c) {
            const std::uint64_t begin = c * kUnpackChunk;
            const std::uint64_t end = std::min<std::uint64_t>(nq, begin + kUnpackChunk);
            std::memcpy(out_count + begin, staged_count + begin, (end - begin) * sizeof(std::uint32_t));
            bool full = out_stride == kmax;
            for (std::uint64_t q = begin; q < end && full; ++q) {
                full = staged_count[q] == kmax;
            }
            if (full) {
                std::memcpy(out + begin * kmax, staged + begin * kmax, (end - begin) * kmax * sizeof(vdx_neighbor));
                return;
            }
            for (std::uint64_t q = begin; q < end; ++q) {
                const std::uint32_t c_q = staged_count[q];
                std::memcpy(out + q * out_stride, staged + q * kmax, c_q * sizeof(vdx_neighbor));
                std::memset(static_cast<void*>(out + q * out_stride + c_q), 0, (out_stride - c_q) * sizeof(vdx_neighbor));
            }
        };
        h->ctx->host_pool->run(static_cast<std::size_t>((nq + kUnpackChunk - 1) / kUnpackChunk), unpack);
        if (d_stats) {
            for (int i = 0; i < vdx::dev::kBeamStats; ++i) {
                opts->stats[i] += host_stats[i];
            }
        }
    });
}

int vdx_gpu_hnsw_search_device(vdx_gpu_hnsw* h, const float* d_queries, uint64_t nq,
                               const uint32_t* d_ks, uint32_t k, const uint32_t* d_efs,
                               uint32_t ef, uint32_t max_k, vdx_neighbor* d_out,
                               uint64_t out_stride, uint32_t* d_out_count,
                               const vdx_search_opts* opts, void* stream) {
    return guarded([&] {
        require(h != nullptr, "vdx_gpu_hnsw_search_device: index is null");
        require(max_k >= 1 && (d_ks != nullptr || k <= max_k), "vdx_gpu_hnsw_search_device: k above max_k");
        require(out_stride >= max_k, "vdx_gpu_hnsw_search_device: out_stride is below max_k");
        if (nq == 0) {
            return;
        }
        // max(ef, k) over the batch sizes R; with per-query arrays the caller's bound is max_k
        // and `ef` must bound the ef array (values above run at the launch's R capacity and are
        // counted in VDX_STAT_CLAMPED).
        require(d_efs == nullptr || ef >= 1,
                "vdx_gpu_hnsw_search_device: with per-query efs, ef must bound them (it is 0)");
        const std::uint32_t efmax = std::max(ef, max_k);
        require(efmax <= kMaxBeamEf, "vdx_gpu_hnsw_search_device: ef above the GPU path's limit");
        DeviceGuard guard(h->ctx->device);
        hipStream_t s = stream ? static_cast<hipStream_t>(stream) : h->ctx->stream;
        const std::uint32_t slot = h->next_slot.fetch_add(1) % kRingSlots;
        // the slot is reused only once the launch that last took it has completed
        const std::lock_guard<std::mutex> slot_lock(h->ring_mutex[slot]);
        hipEvent_t& slot_done = h->ring_event[slot];
        if (slot_done == nullptr) {
            check(hipEventCreateWithFlags(&slot_done, hipEventDisableTiming), "hipEventCreate(ring slot)");
        } else {
            check(hipEventSynchronize(slot_done), "hipEventSynchronize(ring slot)");
        }
        char* base = h->ring.as<char>() + static_cast<std::size_t>(slot) * 256;
        std::uint32_t* counter = reinterpret_cast<std::uint32_t*>(base);
        unsigned long long* d_stats = reinterpret_cast<unsigned long long*>(base + 64);
        if (h->n == 0) {
            check(hipMemsetAsync(d_out_count, 0, nq * 4, s), "hipMemsetAsync(count)");
            return;
        }
        const bool want_stats = opts && opts->stats;
        if (want_stats) {
            check(hipMemsetAsync(d_stats, 0, vdx::dev::kBeamStats * sizeof(unsigned long long), s), "hipMemsetAsync(stats)");
        }
        if (opts != nullptr && opts->visited_form == 2) {
            const std::lock_guard<std::mutex> bitmap_lock(h->bitmap_mutex);
            if (h->bitmap_done == nullptr) {
                check(hipEventCreateWithFlags(&h->bitmap_done, hipEventDisableTiming), "hipEventCreate(bitmap)");
            } else {
                check(hipEventSynchronize(h->bitmap_done), "hipEventSynchronize(bitmap)");
            }
            enqueue_beam(*h, d_queries, nq, d_ks, k, d_efs, ef, efmax,
                         reinterpret_cast<vdx::dev::Neighbor*>(d_out), out_stride, d_out_count, counter,
                         want_stats ? d_stats : nullptr, opts, s, nullptr, nullptr, 0, &h->bitmap);
            check(hipEventRecord(h->bitmap_done, s), "hipEventRecord(bitmap)");
        } else {
            enqueue_beam(*h, d_queries, nq, d_ks, k, d_efs, ef, efmax,
                         reinterpret_cast<vdx::dev::Neighbor*>(d_out), out_stride, d_out_count, counter,
                         want_stats ? d_stats : nullptr, opts, s);
        }
        check(hipEventRecord(slot_done, s), "hipEventRecord(ring slot)");
        if (want_stats) {
            unsigned long long host_stats[vdx::dev::kBeamStats];
            check(hipMemcpyAsync(host_stats, d_stats, sizeof(host_stats), hipMemcpyDeviceToHost, s),
                  "hipMemcpyAsync(stats)");
            check(hipStreamSynchronize(s), "hnsw search stats");
            for (int i = 0; i < vdx::dev::kBeamStats; ++i) {
                opts->stats[i] += host_stats[i];
            }
        }
    });
}

int vdx_gpu_hnsw_trace_device(vdx_gpu_hnsw* h, const float* d_queries, uint64_t nq, uint32_t k,
                              uint32_t ef, const uint32_t* d_truth10, uint32_t max_steps,
                              float* d_features, uint8_t* d_hits, uint32_t* d_steps,
                              uint32_t* d_final) {
    return guarded([&] {
        require(h != nullptr, "vdx_gpu_hnsw_trace_device: index is null");
        require(k >= 1 && max_steps >= 1, "vdx_gpu_hnsw_trace_device: k and max_steps must be positive");
        require(d_truth10 && d_features && d_hits && d_steps && d_final,
                "vdx_gpu_hnsw_trace_device: null buffer");
        require(h->n > 0, "vdx_gpu_hnsw_trace_device: empty index");
        if (nq == 0) return;
        const std::uint32_t efmax = std::max(ef, k);
        require(efmax <= kMaxBeamEf, "vdx_gpu_hnsw_trace_device: ef above the GPU path's limit");
        DeviceGuard guard(h->ctx->device);
        WorkspaceLease ws(*h->ctx);
        hipStream_t s = ws->stream;
        ws->out.reserve(nq * k * sizeof(vdx_neighbor));
        ws->count.reserve(nq * 4);
        ws->counters.reserve(64);
        vdx::dev::BeamArgs t{};
        t.truth = d_truth10;
        t.trace_feat = d_features;
        t.trace_hits = d_hits;
        t.trace_steps = d_steps;
        t.trace_final = d_final;
        t.trace_cap = max_steps;
        enqueue_beam(*h, d_queries, nq, nullptr, k, nullptr, ef, efmax, ws->out.as<vdx::dev::Neighbor>(), k,
                     ws->count.as<std::uint32_t>(), ws->counters.as<std::uint32_t>(), nullptr, nullptr, s, &t);
        check(hipStreamSynchronize(s), "hnsw trace");
    });
}

// ------------------------------------------------------------------------------- brute force

int vdx_gpu_bf_upload(vdx_gpu_ctx* ctx, uint32_t metric, uint64_t dim, uint64_t n,
                      const float* vectors, vdx_gpu_bf** out) {
    return guarded([&] {
        require(ctx != nullptr && out != nullptr, "vdx_gpu_bf_upload: null argument");
        require(vdx::valid_metric(metric), "vdx_gpu_bf_upload: unknown metric code");
        require(dim >= 1, "vdx_gpu_bf_upload: dimension is 0");
        if (dim > kMaxGpuDim) {
            throw std::invalid_argument(vdx::fmt("vdx_gpu_bf_upload: dimension %llu is above the GPU path's limit of %u",
                                                 static_cast<unsigned long long>(dim), kMaxGpuDim));
        }
        require(n < 0x80000000ull, "vdx_gpu_bf_upload: more than 2^31 - 1 vectors");
        require(n == 0 || vectors != nullptr, "vdx_gpu_bf_upload: vectors is null");
        DeviceGuard guard(ctx->device);
        auto b = std::make_unique<vdx_gpu_bf>();
        b->ctx = ctx;
        b->metric = metric;
        b->dim = dim;
        b->n = n;
        b->split = vdx::gpu::split_dim(dim);
        upload_rows(b->vectors, vectors, n, dim, b->split.ld, ctx->stream);
        check(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize(upload)");
        *out = b.release();
    });
}

int vdx_gpu_bf_append(vdx_gpu_bf* b, const float* vectors, uint64_t count, uint64_t dim) {
    return guarded([&] {
        require(b != nullptr, "vdx_gpu_bf_append: index is null");
        require(dim == b->dim, "vdx_gpu_bf_append: rows of another dimension");
        require(count == 0 || vectors != nullptr, "vdx_gpu_bf_append: vectors is null");
        require(b->n + count < 0x80000000ull, "vdx_gpu_bf_append: more than 2^31 - 1 vectors");
        if (count == 0) return;
        DeviceGuard guard(b->ctx->device);
        hipStream_t s = b->ctx->stream;
        const std::uint64_t ld = b->split.ld;
        const std::uint64_t need = (b->n + count) * ld * 4;
        if (need > b->vectors.bytes()) {  // grow with a quarter of headroom, keeping the rows
            DeviceBuffer grown(need + need / 4);
            if (b->n > 0) {
                check(hipMemcpyAsync(grown.as<void>(), b->vectors.as<void>(), b->n * ld * 4, hipMemcpyDeviceToDevice, s),
                      "hipMemcpyAsync(grow)");
            }
            check(hipStreamSynchronize(s), "hipStreamSynchronize(grow)");
            b->vectors.swap(grown);
        }
        copy_rows(b->vectors.as<float>() + b->n * ld, vectors, count, dim, ld, s);
        check(hipStreamSynchronize(s), "hipStreamSynchronize(append)");
        {
            // the MFMA operands were made for the old rows: made again on next use
            const std::lock_guard<std::mutex> lock(b->prep_mutex);
            for (auto& p : b->prep) {
                p.ready = false;
            }
        }
        b->n += count;
    });
}

void vdx_gpu_bf_free(vdx_gpu_bf* index) {
    if (index != nullptr) {
        DeviceGuard guard(index->ctx->device);
        delete index;
    }
}

}  // extern "C"

namespace vdx::gpu {

// Exact scan of every row for nq prepared device queries, results to d_out (rows of out_stride).
void bf_exact(vdx_gpu_bf& b, const float* d_queries, std::uint64_t nq, std::uint32_t k,
              vdx::dev::Neighbor* d_out, std::uint64_t out_stride, std::uint32_t* d_count,
              Workspace& ws, hipStream_t s) {
    const std::uint32_t kcap = std::min<std::uint64_t>(k, std::max<std::uint64_t>(1, b.n));
    const std::size_t limit = b.ctx->lds_per_block;
    const auto lds_for = [&](std::uint32_t w) {
        return sizeof(float) * 32 * b.split.ld + static_cast<std::size_t>(w) * 8 * kcap * 8;
    };
    std::uint32_t waves = 4;
    while (waves > 1 && lds_for(waves) > limit) {
        --waves;
    }
    const std::size_t lds = lds_for(waves);
    require(lds <= limit, "BruteForceIndex::search: k too large for the GPU path");
    const std::uint64_t qgroups = (nq + waves * 8 - 1) / (waves * 8);
    std::uint64_t slabs = std::max<std::uint64_t>(1, (2 * static_cast<std::uint64_t>(b.ctx->cu_count) * 2 + qgroups - 1) / qgroups);
    slabs = std::min<std::uint64_t>(slabs, 64);
    slabs = std::min<std::uint64_t>(slabs, std::max<std::uint64_t>(1, b.n / 256));
    ws.aux0.reserve(slabs * nq * kcap * 8);
    ws.aux1.reserve(slabs * nq * 4);
    vdx::dev::BfScanArgs a{};
    a.vectors = b.vectors.as<float>();
    a.ld = b.split.ld;
    a.dim = static_cast<std::uint32_t>(b.dim);
    a.n = static_cast<std::uint32_t>(b.n);
    a.l2 = b.metric == 0 ? 1u : 0u;
    a.nb = b.split.nb;
    a.n8 = b.split.n8;
    a.tail = b.split.tail;
    a.first_tail = b.split.first_tail;
    a.queries = d_queries;
    a.qld = static_cast<std::uint32_t>(b.dim);
    a.nq = static_cast<std::uint32_t>(nq);
    a.k = k;
    a.kcap = kcap;
    a.partial = ws.aux0.as<std::uint64_t>();
    a.partial_count = ws.aux1.as<std::uint32_t>();
    a.out = d_out;
    a.out_stride = out_stride;
    a.out_count = d_count;
    check(vdx::dev::launch_bf_exact_scan(a, static_cast<unsigned>(qgroups), static_cast<unsigned>(slabs),
                                         waves, lds, s, bf_shape(b.dim)),
          "bf_exact_scan launch");
}

}  // namespace vdx::gpu

extern "C" {

int vdx_gpu_bf_gemm_shape(vdx_gpu_bf* b, uint32_t shape) {
    return guarded([&] {
        require(b != nullptr, "vdx_gpu_bf_gemm_shape: index is null");
        require(shape <= 1, "vdx_gpu_bf_gemm_shape: shape must be 0 or 1");
        b->gemm_shape.store(static_cast<int>(shape));
    });
}

int vdx_gpu_bf_tune(vdx_gpu_bf* b, const float* d_queries, uint64_t nq, uint32_t k, uint32_t reps,
                    float* ms, uint32_t* shape) {
    return guarded([&] {
        require(b != nullptr, "vdx_gpu_bf_tune: index is null");
        require(k >= 1 && nq >= 1, "vdx_gpu_bf_tune: empty sample");
        require(k <= kMaxBfK || b->n <= kMaxBfK, "vdx_gpu_bf_tune: k above the GPU path's limit");
        DeviceGuard guard(b->ctx->device);
        WorkspaceLease ws(*b->ctx);
        hipStream_t s = ws->stream;
        const std::uint64_t keep = std::min<std::uint64_t>(k, b->n);
        DeviceBuffer out(nq * std::max<std::uint64_t>(keep, 1) * sizeof(vdx_neighbor)), count(nq * 4);
        hipEvent_t t0 = nullptr, t1 = nullptr;
        check(hipEventCreate(&t0), "hipEventCreate");
        check(hipEventCreate(&t1), "hipEventCreate");
        float best[2] = {0.f, 0.f};
        const int before = b->gemm_shape.load();
        try {
            for (int sh = 0; sh < 2; ++sh) {
                b->gemm_shape.store(sh);
                for (std::uint32_t r = 0; r <= std::max(1u, reps); ++r) {  // the first run prepares operands
                    check(hipEventRecord(t0, s), "hipEventRecord");
                    vdx::gpu::bf_mfma(*b, VDX_PREC_BF16, d_queries, nq, k, out.as<vdx::dev::Neighbor>(),
                                      std::max<std::uint64_t>(keep, 1), count.as<std::uint32_t>(), *ws, s, nullptr);
                    check(hipEventRecord(t1, s), "hipEventRecord");
                    check(hipEventSynchronize(t1), "hipEventSynchronize");
                    float t = 0.f;
                    check(hipEventElapsedTime(&t, t0, t1), "hipEventElapsedTime");
                    if (r == 1 || (r > 1 && t < best[sh])) best[sh] = t;
                }
            }
        } catch (...) {
            b->gemm_shape.store(before);
            (void)hipEventDestroy(t0);
            (void)hipEventDestroy(t1);
            throw;
        }
        (void)hipEventDestroy(t0);
        (void)hipEventDestroy(t1);
        const int pick = best[1] < best[0] ? 1 : 0;
        b->gemm_shape.store(pick);
        if (ms) {
            ms[0] = best[0];
            ms[1] = best[1];
        }
        if (shape) *shape = static_cast<std::uint32_t>(pick);
    });
}

int vdx_gpu_bf_search_device(vdx_gpu_bf* b, const float* d_queries, uint64_t nq, uint32_t k,
                             uint32_t precision, vdx_neighbor* d_out, uint64_t out_stride,
                             uint32_t* d_out_count, vdx_bf_stats* stats) {
    return guarded([&] {
        require(b != nullptr, "vdx_gpu_bf_search_device: index is null");
        require(k >= 1, "BruteForceIndex::search: k must be positive");
        require(out_stride >= std::min<std::uint64_t>(k, b->n), "vdx_gpu_bf_search_device: out_stride is below k");
        if (k > kMaxBfK && b->n > kMaxBfK) {
            throw std::invalid_argument(vdx::fmt("BruteForceIndex::search: k %u is above the GPU path's limit of %u", k, kMaxBfK));
        }
        if (stats) {
            std::memset(stats, 0, sizeof(*stats));
        }
        if (nq == 0) {
            return;
        }
        DeviceGuard guard(b->ctx->device);
        WorkspaceLease ws(*b->ctx);
        hipStream_t s = ws->stream;
        if (b->n == 0) {
            check(hipMemsetAsync(d_out_count, 0, nq * 4, s), "hipMemsetAsync(count)");
            check(hipStreamSynchronize(s), "bf search");
            return;
        }
        const std::uint32_t mode = precision & 0xFFu;
        const bool fp64_ties = (precision & VDX_PREC_FP64_TIES) != 0;
        const bool fp64_error = (precision & VDX_PREC_FP64_ERROR) != 0;
        if ((precision & ~(0xFFu | VDX_PREC_FP64_TIES | VDX_PREC_FP64_ERROR)) != 0 || mode > VDX_PREC_EXACT) {
            throw std::invalid_argument("BruteForceIndex::search: unknown precision mode");
        }
        const auto run_mode = [&](std::uint32_t kk, vdx::dev::Neighbor* out, std::uint64_t stride, std::uint32_t* count) {
            if (mode == VDX_PREC_EXACT) {
                vdx::gpu::bf_exact(*b, d_queries, nq, kk, out, stride, count, *ws, s);
            } else {
                vdx::gpu::bf_mfma(*b, static_cast<int>(mode), d_queries, nq, kk, out, stride, count, *ws, s, stats);
            }
        };
        if (!fp64_ties && !fp64_error) {
            run_mode(k, reinterpret_cast<vdx::dev::Neighbor*>(d_out), out_stride, d_out_count);
            check(hipStreamSynchronize(s), "bf search");
            return;
        }
        // The float64 pass over a window of each query's first m results: the k wanted, plus (for
        // the tie re-rank) as many again (at least 8) so a tie group at the k-th result is whole;
        // a query whose group reaches the window's end doubles the window, up to 1024.
        const std::uint64_t keep = std::min<std::uint64_t>(k, b->n);
        if (fp64_ties && keep > vdx::dev::kFp64MaxWindow / 2) {
            throw std::invalid_argument(vdx::fmt("BruteForceIndex::search: the float64 tie re-rank takes k up to %u",
                                                 vdx::dev::kFp64MaxWindow / 2));
        }
        std::uint64_t m = fp64_ties ? std::min<std::uint64_t>(b->n, keep + std::max<std::uint64_t>(keep, 8)) : keep;
        unsigned long long counters[5] = {};
        for (;;) {
            ws->fp64[0].reserve(nq * m * sizeof(vdx::dev::Neighbor));
            ws->fp64[1].reserve(nq * 4);
            ws->fp64[2].reserve(nq * m * sizeof(double));
            ws->fp64[3].reserve(sizeof(counters));
            check(hipMemsetAsync(ws->fp64[3].as<void>(), 0, sizeof(counters), s), "hipMemsetAsync(fp64 counters)");
            auto* window = ws->fp64[0].as<vdx::dev::Neighbor>();
            run_mode(static_cast<std::uint32_t>(m), window, m, ws->fp64[1].as<std::uint32_t>());
            vdx::dev::BfFp64Args a{};
            a.vectors = b->vectors.as<float>();
            a.ld = b->split.ld;
            a.dim = static_cast<std::uint32_t>(b->dim);
            a.metric = b->metric;
            a.n = static_cast<std::uint32_t>(b->n);
            a.queries = d_queries;
            a.qld = static_cast<std::uint32_t>(b->dim);
            a.nq = static_cast<std::uint32_t>(nq);
            a.cand = window;
            a.cand_count = ws->fp64[1].as<std::uint32_t>();
            a.m = static_cast<std::uint32_t>(m);
            a.d64 = ws->fp64[2].as<double>();
            a.out = reinterpret_cast<vdx::dev::Neighbor*>(d_out);
            a.out_stride = out_stride;
            a.out_count = d_out_count;
            a.keep = static_cast<std::uint32_t>(keep);
            a.ties = fp64_ties ? 1u : 0u;
            a.stats = ws->fp64[3].as<unsigned long long>();
            check(vdx::dev::launch_bf_fp64(a, s), "bf_fp64 launch");
            check(hipMemcpyAsync(counters, ws->fp64[3].as<void>(), sizeof(counters), hipMemcpyDeviceToHost, s),
                  "hipMemcpyAsync(fp64 counters)");
            check(hipStreamSynchronize(s), "bf search (float64 pass)");
            if (!fp64_ties || counters[0] == 0 || m >= b->n || m >= vdx::dev::kFp64MaxWindow) break;
            m = std::min<std::uint64_t>({b->n, 2 * m, vdx::dev::kFp64MaxWindow});
        }
        if (stats) {
            double err = 0.0, rel = 0.0;
            std::memcpy(&err, &counters[3], sizeof(err));
            std::memcpy(&rel, &counters[4], sizeof(rel));
            stats->fp64_unresolved = fp64_ties ? counters[0] : 0;
            stats->fp64_ties = counters[1];
            stats->fp64_reordered = counters[2];
            stats->fp64_window = m;
            stats->max_err_fp64 = err;
            stats->max_rel_err_fp64 = rel;
        }
        check(hipStreamSynchronize(s), "bf search");
    });
}

int vdx_gpu_bf_search(vdx_gpu_bf* b, const float* queries, uint64_t nq, uint64_t query_dim,
                      uint32_t k, uint32_t precision, vdx_neighbor* out, uint64_t out_stride,
                      uint32_t* out_count, const char* where, vdx_bf_stats* stats) {
    return guarded([&] {
        require(b != nullptr, "vdx_gpu_bf_search: index is null");
        const char* w = where ? where : "BruteForceIndex::search";
        if (nq == 0) {
            return;
        }
        require(queries != nullptr && out != nullptr && out_count != nullptr, "vdx_gpu_bf_search: null buffer");
        DeviceGuard guard(b->ctx->device);
        std::vector<float> prepared(nq * b->dim);
        if (query_dim != b->dim) {
            vdx::check_query(w, {static_cast<vdx::Metric>(b->metric), b->dim, b->n}, queries, query_dim, k);
        }
        prepare_queries(*b->ctx, w, static_cast<vdx::Metric>(b->metric), b->dim, b->n, queries, nq, query_dim,
                        nullptr, k, prepared.data());
        const std::uint64_t keep = std::min<std::uint64_t>(k, b->n);
        require(out_stride >= keep, "vdx_gpu_bf_search: out_stride is below k");
        if (b->n == 0) {
            std::fill(out_count, out_count + nq, 0u);
            return;
        }
        DeviceBuffer d_q(nq * b->dim * sizeof(float));
        DeviceBuffer d_out(nq * keep * sizeof(vdx_neighbor));
        DeviceBuffer d_count(nq * 4);
        check(hipMemcpy(d_q.as<void>(), prepared.data(), nq * b->dim * sizeof(float), hipMemcpyHostToDevice),
              "hipMemcpy(queries)");
        const int rc = vdx_gpu_bf_search_device(b, d_q.as<float>(), nq, k, precision, d_out.as<vdx_neighbor>(),
                                                keep, d_count.as<std::uint32_t>(), stats);
        if (rc != VDX_OK) {
            const std::string msg = vdx_last_error();
            if (rc == VDX_INVALID_ARGUMENT) throw std::invalid_argument(msg);
            if (rc == VDX_OUT_OF_MEMORY) throw std::bad_alloc();
            if (rc == VDX_DEVICE_ERROR) throw vdx::device_error(msg);
            throw std::runtime_error(msg);
        }
        std::vector<vdx_neighbor> staged(nq * keep);
        check(hipMemcpy(staged.data(), d_out.as<void>(), staged.size() * sizeof(vdx_neighbor), hipMemcpyDeviceToHost),
              "hipMemcpy(out)");
        check(hipMemcpy(out_count, d_count.as<void>(), nq * 4, hipMemcpyDeviceToHost), "hipMemcpy(count)");
        for (std::uint64_t q = 0; q < nq; ++q) {
            std::memcpy(out + q * out_stride, staged.data() + q * keep, out_count[q] * sizeof(vdx_neighbor));
        }
    });
}

}  // extern "C"
