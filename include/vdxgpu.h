/* libvdxgpu: the C ABI of the MI355X-native vecdex query path.
 *
 * Plain pointers and sizes only; no C++ or torch types cross it and no exception escapes it.
 * Every function returns a vdx_status (0 = success); on failure vdx_last_error() holds the
 * message of the C++ exception vecdex would have thrown, and the status names its type:
 *   VDX_INVALID_ARGUMENT  std::invalid_argument  (Python ValueError)
 *   VDX_RUNTIME_ERROR     std::runtime_error     (Python RuntimeError)
 *   VDX_LENGTH_ERROR      std::length_error
 *   VDX_OUT_OF_MEMORY     std::bad_alloc, or hipErrorOutOfMemory on the device
 *   VDX_DEVICE_ERROR      any other HIP failure
 * A rejected call changes nothing (vecdex include/vecdex/hnsw.hpp:60-62).
 *
 * Each entry point names the vecdex interface it replaces or serves (file:line in vecdex).
 */
#ifndef VDXGPU_H
#define VDXGPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#if defined(__GNUC__)
#define VDX_API __attribute__((visibility("default")))
#else
#define VDX_API
#endif

enum vdx_status {
    VDX_OK = 0,
    VDX_INVALID_ARGUMENT = 1,
    VDX_RUNTIME_ERROR = 2,
    VDX_LENGTH_ERROR = 3,
    VDX_OUT_OF_MEMORY = 4,
    VDX_DEVICE_ERROR = 5
};

/* Layout-identical to vecdex::Neighbor {float distance; Id id} (include/vecdex/neighbor.hpp:13-18). */
typedef struct vdx_neighbor {
    float distance;
    uint32_t id;
} vdx_neighbor;

/* A graph in FORMAT v1 section layout (docs/FORMAT.md:13-96): what HnswIndex holds in
 * levels_, layer0_, upper_ (concatenated in id order) and vectors_
 * (include/vecdex/hnsw.hpp:170-178). The caller keeps ownership; upload copies it to HBM. */
typedef struct vdx_graph_view {
    uint32_t metric; /* FORMAT code: 0 L2, 1 inner product, 2 cosine */
    uint32_t reserved;
    uint64_t dim;
    uint64_t n;
    uint64_t M;
    uint64_t max_level;
    uint64_t entry_point;
    uint64_t upper_blocks;   /* U: the sum of all levels */
    const uint8_t* levels;   /* n */
    const uint32_t* layer0;  /* n * (1 + 2M) */
    const uint32_t* upper;   /* U * (1 + M) */
    const float* vectors;    /* n * dim; unit length under cosine */
} vdx_graph_view;

/* The message of the calling thread's last failure. */
VDX_API const char* vdx_last_error(void);
VDX_API const char* vdx_version(void);

/* ================================================================================= host side
 * A restatement of vecdex's graph build and file format, so an index can be built or loaded
 * on the host without vecdex itself. Not accelerated (graph build stays on the CPU). */

typedef struct vdx_host_graph vdx_host_graph;

/* HnswIndex::HnswIndex(dim, metric, params) (include/vecdex/hnsw.hpp:67). */
VDX_API int vdx_host_graph_create(uint64_t dim, uint32_t metric, uint64_t M, uint64_t ef_construction,
                          uint64_t seed, vdx_host_graph** out);
/* HnswIndex::add(MatrixView, num_threads) (include/vecdex/hnsw.hpp:87; src/hnsw.cpp:259-291). */
VDX_API int vdx_host_graph_add(vdx_host_graph* g, const float* data, uint64_t rows, uint64_t given_dim,
                       uint64_t num_threads);
/* HnswIndex::save / load (include/vecdex/hnsw.hpp:115,123; src/hnsw_io.cpp:341-467). load reads
 * FORMAT v1 and this library's FORMAT v2 (dispatching on the version field). */
VDX_API int vdx_host_graph_save(const vdx_host_graph* g, const char* path);
VDX_API int vdx_host_graph_load(const char* path, vdx_host_graph** out);
/* FORMAT v2 (no counterpart in vecdex, whose v1 is not mappable, README.md:329): every section
 * 4 KiB-aligned and laid out as the device copy holds it (vectors padded to rows of dim rounded
 * up to 4), per-section XXH64, written atomically and durably as save does. */
VDX_API int vdx_host_graph_save_v2(const vdx_host_graph* g, const char* path);
/* Zero-copy view of the graph's sections, valid until the next add or free. */
VDX_API int vdx_host_graph_view(const vdx_host_graph* g, vdx_graph_view* out);
VDX_API int vdx_host_graph_params(const vdx_host_graph* g, uint64_t* ef_construction, uint64_t* seed,
                          uint64_t* version);
VDX_API uint64_t vdx_host_graph_memory_bytes(const vdx_host_graph* g);
/* HnswIndex::add (src/hnsw.cpp:247-291) with the layer-0 insertion searches run in batches by a
 * searcher over its own copy of the graph (no counterpart in vecdex, which searches node by
 * node). Per batch: sync() brings the searcher's copy up to the graph (`changed`: ascending ids
 * of the nodes stored or rewritten since the last sync); the layers above 0 of the batch's nodes
 * that have them are searched top down (by the searcher with VDX_SEARCHER_UPPER_LAYERS, else on
 * the host), each from the layer above's best result; search() then returns each node's layer-0
 * search_layer results at breadth ef = max(ef_construction, M), any order, starting from
 * entries[i] (its layer-1 result, src/hnsw.cpp:324-335) or, where entries[i] is UINT32_MAX,
 * from the greedy descent (src/hnsw.cpp:315-323); each node is linked as insert links it. Nodes
 * of one batch are not candidates of each other. Callbacks run on the calling thread and return
 * a vdx status. */
typedef struct vdx_batch_searcher {
    void* ctx;
    int (*sync)(void* ctx, const vdx_graph_view* graph, const uint32_t* changed, uint64_t count);
    /* The beam on `layer` (0 unless caps has VDX_SEARCHER_UPPER_LAYERS): from entries[i], a node
     * on that layer, or where it is UINT32_MAX from the greedy descent down to layer + 1.
     * m == 0: out[i * ef ..] receives node i's results. m > 0 (only with VDX_SEARCHER_SELECTS):
     * out[i * m ..] receives insert's neighbour choice among them (select_neighbors(results, m),
     * src/hnsw.cpp:350-409 heuristic, closest first). */
    int (*search)(void* ctx, const uint32_t* nodes, const uint32_t* entries, uint64_t count, uint32_t layer,
                  uint32_t ef, uint32_t m, vdx_neighbor* out, uint32_t* counts);
    uint32_t caps; /* VDX_SEARCHER_* bits */
} vdx_batch_searcher;
enum {
    VDX_SEARCHER_SELECTS = 1,      /* search() also makes the neighbour choice */
    VDX_SEARCHER_UPPER_LAYERS = 2  /* search() also searches layers above 0 (else the host does) */
};
/* Batch sizes: the first host_prefix nodes of the graph are inserted on the host; then batches
 * of batch_fraction x (nodes so far), clamped to [batch_min, batch_max]. NULL: 16384, 1024,
 * 131072, 0.05. batch_min = batch_max = 1 with host_prefix 0 and one thread builds exactly
 * what vdx_host_graph_add with one thread builds (given exact searches). */
typedef struct vdx_build_plan {
    uint64_t host_prefix, batch_min, batch_max;
    double batch_fraction;
} vdx_build_plan;
/* Where a batched add spent its time (seconds, wall clock) and how it was split. */
typedef struct vdx_build_stats {
    uint64_t batches, searched, host_inserted;  /* batches; nodes whose layer-0 search the searcher ran;
                                                   nodes inserted (the prefix) or with upper layers
                                                   searched on the host */
    double host_insert_s, sync_s, search_s, link_s, total_s;
} vdx_build_stats;
/* `synced_version`: the graph version the searcher's copy reflects (vdx_host_graph_params);
 * the first sync lists everything changed since. `stats` may be NULL. */
VDX_API int vdx_host_graph_add_batched(vdx_host_graph* g, const float* data, uint64_t rows, uint64_t dim,
                                       uint64_t num_threads, const vdx_batch_searcher* searcher,
                                       const vdx_build_plan* plan, uint64_t synced_version,
                                       vdx_build_stats* stats);

/* Nodes whose link blocks an add() after `since_version` wrote (the nodes it stored and those it
 * linked back to), ascending; with nodes NULL only the count. No counterpart in vecdex: it
 * drives the incremental device update (vdx_gpu_hnsw_update). */
VDX_API int vdx_host_graph_changes(const vdx_host_graph* g, uint64_t since_version, uint32_t* nodes,
                                   uint64_t capacity, uint64_t* count);

/* HnswIndex::inspect (include/vecdex/hnsw.hpp:102; src/hnsw.cpp:527-594). */
typedef struct vdx_graph_report {
    uint64_t reachable;
    uint64_t max_level;
    uint64_t violation_count;
    double mean_degree0;
} vdx_graph_report;
VDX_API int vdx_host_graph_inspect(const vdx_host_graph* g, vdx_graph_report* out, char* first_violation,
                           uint64_t buflen);
VDX_API void vdx_host_graph_free(vdx_host_graph* g);

/* detail::check_new_vectors (src/vector_ops.cpp:36-58), with `where` as the message prefix. */
VDX_API int vdx_host_check_new_vectors(const char* where, uint32_t metric, uint64_t dim,
                               uint64_t index_size, const float* data, uint64_t rows,
                               uint64_t given_dim);
/* detail::normalize on each row (src/vector_ops.cpp:29-34), as BruteForceIndex::store does
 * under cosine (src/brute_force.cpp:58-66). */
VDX_API int vdx_host_normalize_rows(float* data, uint64_t rows, uint64_t dim);
/* The library's distance (the arithmetic every kernel reproduces). */
VDX_API float vdx_host_distance(uint32_t metric, const float* a, const float* b, uint64_t dim);

/* recall_at_k (src/recall.cpp:11-34): |first k of found ∩ first k of truth| / k; a duplicate
 * found id counts once. Rejects k == 0 and truth_count < k (the reference's messages). */
VDX_API int vdx_recall_at_k(const vdx_neighbor* found, uint64_t found_count, const int32_t* truth,
                            uint64_t truth_count, uint64_t k, double* out);
/* threshold_recall_at_k (src/recall.cpp:36-51): the fraction of the first k found whose distance
 * is <= kth_true + tolerance * max(|kth_true|, 1). Rejects k == 0, a non-finite kth_true or
 * tolerance, and a negative tolerance. */
VDX_API int vdx_threshold_recall_at_k(const vdx_neighbor* found, uint64_t found_count,
                                      float kth_true, uint64_t k, float tolerance, double* out);
/* TEXMEX .fvecs / .ivecs (src/io.cpp:30-87): the shape from the first record and the file size,
 * then every record read with its dimension checked, into dst (rows * dim 4-byte values; float
 * for .fvecs, int32 for .ivecs). Errors are runtime errors with the path first. */
VDX_API int vdx_vecs_info(const char* path, uint64_t* rows, uint64_t* dim);
VDX_API int vdx_vecs_read(const char* path, void* dst, uint64_t rows, uint64_t dim);
/* Means of the two over nq result rows (no counterpart in the reference, which evaluates one
 * query at a time in bench/hnsw_eval.cpp:513-520): row q holds found_counts[q] results (all
 * found_stride when found_counts is NULL) at found + q * found_stride; truth rows are
 * truth_stride ids (recall) or one k-th true distance per query (threshold recall). */
VDX_API int vdx_mean_recall_at_k(const vdx_neighbor* found, uint64_t found_stride,
                                 const uint32_t* found_counts, const int32_t* truth,
                                 uint64_t truth_stride, uint64_t nq, uint64_t k, double* out);
VDX_API int vdx_mean_threshold_recall_at_k(const vdx_neighbor* found, uint64_t found_stride,
                                           const uint32_t* found_counts, const float* kth_true,
                                           uint64_t nq, uint64_t k, float tolerance, double* out);

/* ================================================================================= GPU side */

typedef struct vdx_gpu_ctx vdx_gpu_ctx;

/* Opens HIP device `device`; each context owns one stream. */
VDX_API int vdx_gpu_open(int device, vdx_gpu_ctx** out);
VDX_API void vdx_gpu_close(vdx_gpu_ctx* ctx);
/* The context's stream (a hipStream_t), for callers that enqueue around the device entry points. */
VDX_API void* vdx_gpu_stream(vdx_gpu_ctx* ctx);

/* Device memory and timing plumbing on the context's stream, so callers that keep inputs
 * resident in HBM need no second HIP runtime in their process. Copies are synchronous. */
VDX_API int vdx_gpu_alloc(vdx_gpu_ctx* ctx, uint64_t bytes, void** out);
VDX_API void vdx_gpu_free(vdx_gpu_ctx* ctx, void* ptr);
VDX_API int vdx_gpu_copy_to_device(vdx_gpu_ctx* ctx, void* dst, const void* src, uint64_t bytes);
VDX_API int vdx_gpu_copy_to_host(vdx_gpu_ctx* ctx, void* dst, const void* src, uint64_t bytes);
/* Device-to-device copy by a plain kernel (16-byte aligned pointers): also the reference
 * transfer of known size for calibrating HBM traffic counters. No counterpart in vecdex. */
VDX_API int vdx_gpu_copy_device(vdx_gpu_ctx* ctx, void* dst, const void* src, uint64_t bytes);
/* Random row gathers of known size (no counterpart in vecdex): `waves` waves (a multiple of 4)
 * each read rows_per_wave rows (rounded up to whole rounds; *rows_read = the total) of row_bytes
 * (16..1024, 16 bytes a lane, lanes dividing 64) at hashed ids among the table_bytes / row_bytes
 * rows of `table` (device memory, 16-byte aligned), four rows per lane group in flight; one word
 * per wave goes to `sink` (4 * waves bytes of device memory). The rate random gathers reach with
 * no dependent loads between them: the ceiling the search's row and link-block reads face.
 * Synchronous on the context's stream (time it with vdx_gpu_event_record around it). */
VDX_API int vdx_gpu_gather_probe(vdx_gpu_ctx* ctx, const void* table, uint64_t table_bytes, uint32_t row_bytes,
                                 uint32_t rows_per_wave, uint32_t waves, void* sink, uint64_t* rows_read);
VDX_API int vdx_gpu_synchronize(vdx_gpu_ctx* ctx);
/* Records event `slot` (0..15) on the context's stream; elapsed milliseconds between two
 * recorded events once both completed. */
VDX_API int vdx_gpu_event_record(vdx_gpu_ctx* ctx, uint32_t slot);
VDX_API int vdx_gpu_event_elapsed(vdx_gpu_ctx* ctx, uint32_t from, uint32_t to, float* ms);

/* ---- HNSW: replaces HnswIndex::search / search_batch (include/vecdex/hnsw.hpp:93-99;
 *      src/hnsw.cpp:484-522) on a graph uploaded once. */

typedef struct vdx_gpu_hnsw vdx_gpu_hnsw;

VDX_API int vdx_gpu_hnsw_upload(vdx_gpu_ctx* ctx, const vdx_graph_view* graph, vdx_gpu_hnsw** out);
/* Brings the device copy up to a graph that extends the uploaded one (HnswIndex::add appends
 * nodes and rewrites the link blocks of the nodes it links back to; no counterpart in vecdex,
 * whose index has no device copy): new nodes' rows, blocks and offsets are appended and the
 * blocks of the listed changed nodes (vdx_host_graph_changes) rewritten, all checked with the
 * upload's link rules. Device buffers grow with headroom. Not concurrent with searches on it. */
VDX_API int vdx_gpu_hnsw_update(vdx_gpu_hnsw* index, const vdx_graph_view* graph, const uint32_t* nodes,
                                uint64_t count);
/* vdx_host_graph_add_batched with the GPU as the searcher: `index` is the device copy of `graph`
 * as of graph version `synced_version`; it is kept current (vdx_gpu_hnsw_update per batch) and
 * each batch's searches are one beam launch over the nodes' vectors already in HBM. */
VDX_API int vdx_gpu_hnsw_build_add(vdx_gpu_hnsw* index, vdx_host_graph* graph, const float* data, uint64_t rows,
                                   uint64_t dim, uint64_t num_threads, const vdx_build_plan* plan,
                                   uint64_t synced_version, vdx_build_stats* stats);
/* A FORMAT v2 file mapped and streamed to HBM with no host graph: each section goes over in
 * chunks through pinned buffers, every chunk checked (v1's load checks, v2's layout, checksums
 * and padding) before its copy is queued; on any failure nothing is kept. */
VDX_API int vdx_gpu_hnsw_load(vdx_gpu_ctx* ctx, const char* path, vdx_gpu_hnsw** out);
/* The device copy's shape. */
VDX_API int vdx_gpu_hnsw_info(const vdx_gpu_hnsw* index, uint32_t* metric, uint64_t* dim, uint64_t* n,
                              uint64_t* M);
/* No vecdex counterpart: *kept = 1 if the device copy also holds its rows as bytes (squared
 * L2, dimension 128, every stored value an integer in [0, 255], as SIFT-like bases are; kept
 * current by vdx_gpu_hnsw_update and dropped once a row that is not so arrives). The plain search
 * reads that copy unless vdx_search_opts.row_format = 1: the same floats, so the same results, at
 * a quarter of the row traffic. */
VDX_API int vdx_gpu_hnsw_u8_rows(const vdx_gpu_hnsw* index, uint32_t* kept);
VDX_API void vdx_gpu_hnsw_free(vdx_gpu_hnsw* index);

/* Knobs of one search launch. Zero-initialise for defaults (autotuned). */
typedef struct vdx_search_opts {
    uint32_t visited_log2;     /* LDS visited-cache slots = 2^visited_log2; 0: autotune */
    uint32_t waves_per_group;  /* waves per workgroup; 0: autotune */
    uint32_t termination;      /* 0: fixed ef (exact vecdex semantics); 1: learned stop */
    uint32_t min_steps;        /* learned stop: expansions before the model is consulted */
    const void* mlp;           /* learned stop model (a vdx_gpu_mlp*), termination == 1 */
    float mlp_threshold;       /* stop when the model's probability exceeds this, in (0, 1) */
    float reserved_f;
    uint64_t* stats;           /* optional host array of VDX_STAT_COUNT counters, summed */
    uint32_t row_format;       /* 0: the plain and learned L2 searches read the 8-bit row copy where
                                  the index keeps one (dimension 128, every stored value an integer
                                  in [0, 255]: the same floats, so the same distances and results);
                                  1: always the FP32 rows */
    uint32_t mlp_every;        /* learned stop: consult the model every this many expansions once
                                  allowed (0 or 1: after every expansion) */
    uint32_t max_expansions;   /* expansion budget: a query stops after this many layer-0 link
                                  blocks (0: no budget, vecdex's search). A stopped query returns
                                  the closest k found so far; under either termination. No vecdex
                                  counterpart: a serving bound on the slowest queries of a batch */
    const uint32_t* max_expansions_per_query; /* optional per-query budgets (0: none for that
                                  query) replacing max_expansions, indexed like the queries: a host
                                  array for vdx_gpu_hnsw_search, a device array for the device and
                                  batch entry points. No vecdex counterpart */
    uint32_t visited_form;     /* the visited set of searches with max(ef, k) above 512: 0 or 1 the
                                  LDS cache (lossy, exact duplicate removal in the merge), 2 an
                                  exact bitmap in HBM per resident wave (n bits, cleared per query;
                                  frees the LDS for more waves). Results are the same either way;
                                  vdx_gpu_hnsw_search_uneven_noalloc takes 0 or 1 only */
} vdx_search_opts;

enum vdx_stat {
    VDX_STAT_EXPANSIONS = 0,   /* layer-0 link blocks read */
    VDX_STAT_DISTANCES = 1,    /* distances computed (layer 0 and above) */
    VDX_STAT_UPPER_BLOCKS = 2, /* upper-layer link blocks read */
    VDX_STAT_DUPLICATES = 3,   /* recomputed members of the result list the lossy visited cache let through */
    VDX_STAT_EARLY_STOPS = 4,  /* queries stopped before the fixed-ef stop, by the learned model or
                                  by the expansion budget */
    VDX_STAT_LINK_PREFETCHED = 5, /* layer-0 link blocks that were already prefetched when needed */
    VDX_STAT_CLAMPED = 6,      /* queries whose max(ef, k) exceeded the launch's result-list capacity
                                  (the device entry point's `ef` bound) and ran at that capacity */
    VDX_STAT_COUNT = 8
};

/* search_batch over `nq` host queries of `query_dim` floats each (row-major, contiguous).
 * k and ef: per-query arrays `ks`/`efs` (may be NULL), else the uniform `k`/`ef`; ef below k acts
 * as k. Each query is validated like detail::check_query with `where` as the message prefix
 * ("HnswIndex::search_batch" if NULL) and normalized under cosine, on the context's host threads,
 * into pinned staging that the search kernel reads over PCIe (and into which it writes the
 * results): the call makes no copy to or from the device.
 * Results: out[q * out_stride + i] for i < out_count[q] <= k, closest first, ties by smaller id
 * — element q equals HnswIndex::search(row q, k, ef); the rest of each row is zeroed. */
VDX_API int vdx_gpu_hnsw_search(vdx_gpu_hnsw* index, const float* queries, uint64_t nq,
                        uint64_t query_dim, const uint32_t* ks, uint32_t k, const uint32_t* efs,
                        uint32_t ef, vdx_neighbor* out, uint64_t out_stride, uint32_t* out_count,
                        const char* where, const vdx_search_opts* opts);

/* The same on device memory, asynchronous on `stream` (the context's stream if NULL):
 * d_queries are already prepared (validated; unit length under cosine), nq x dim floats.
 * d_ks / d_efs may be NULL. max_k bounds every k (it sizes nothing; it is checked). The launch's
 * result-list capacity is sized from max(ef, max_k): with d_efs, `ef` must be at least every
 * d_efs value (ef = 0 is rejected); a query whose max(ef, k) is above the capacity runs at the
 * capacity and is counted in VDX_STAT_CLAMPED (opts.stats). Each call takes one of 64 queue
 * counter slots in turn and first waits for the launch that last used that slot, so any number
 * of calls may be in flight on different streams. */
VDX_API int vdx_gpu_hnsw_search_device(vdx_gpu_hnsw* index, const float* d_queries, uint64_t nq,
                               const uint32_t* d_ks, uint32_t k, const uint32_t* d_efs,
                               uint32_t ef, uint32_t max_k, vdx_neighbor* d_out,
                               uint64_t out_stride, uint32_t* d_out_count,
                               const vdx_search_opts* opts, void* stream);

/* ---- Learned adaptive termination (no vecdex counterpart; an addition of this path).
 * A 16 -> 64 -> 64 -> 1 MLP (ReLU, sigmoid), BF16 with FP32 accumulation, reads 16 search-state
 * features after every layer-0 expansion and stops the beam when its probability exceeds a
 * threshold. It can only stop earlier than the fixed-ef search it runs inside. */

#define VDX_MLP_PARAMS 5345
#define VDX_TRACE_FEATURES 16

typedef struct vdx_gpu_mlp vdx_gpu_mlp;

/* params: VDX_MLP_PARAMS floats: mean[16], inv_std[16] (feature standardisation), W1[16][64],
 * b1[64], W2[64][64], b2[64], w3[64], b3. */
VDX_API int vdx_gpu_mlp_create(vdx_gpu_ctx* ctx, const float* params, uint64_t count, vdx_gpu_mlp** out);
VDX_API int vdx_gpu_mlp_params(vdx_gpu_mlp* mlp, float* params, uint64_t count);
VDX_API void vdx_gpu_mlp_free(vdx_gpu_mlp* mlp);
/* Mini-batch Adam (beta 0.9, 0.999) on binary cross-entropy over n device rows of 16 features
 * and labels (0/1). epoch_loss (optional, host, `epochs` floats) gets each epoch's mean loss. */
VDX_API int vdx_gpu_mlp_train(vdx_gpu_mlp* mlp, const float* d_features, const float* d_labels,
                              uint64_t n, uint32_t epochs, uint32_t batch, float lr, uint64_t seed,
                              float* epoch_loss);
/* Logits for n device feature rows into d_logits. */
VDX_API int vdx_gpu_mlp_infer(vdx_gpu_mlp* mlp, const float* d_features, uint64_t n, float* d_logits);

/* Fixed-ef search that records, after each layer-0 expansion of query q (up to max_steps), the
 * 16 features into d_features[q][step][16] and the current top-10 hits against d_truth10[q][10]
 * into d_hits[q][step]; d_steps[q] = expansions taken, d_final[q] = hits of the result.
 * Synchronous. */
VDX_API int vdx_gpu_hnsw_trace_device(vdx_gpu_hnsw* index, const float* d_queries, uint64_t nq,
                                      uint32_t k, uint32_t ef, const uint32_t* d_truth10,
                                      uint32_t max_steps, float* d_features, uint8_t* d_hits,
                                      uint32_t* d_steps, uint32_t* d_final);

/* ---- Uneven batches (no counterpart in vecdex, whose search_batch takes one k and ef,
 *      src/hnsw.cpp:490-501): queries with their own (ef, k) in one call. A classify kernel
 *      groups the queries by R capacity class (64, 128, 256, 512, and one LDS-kernel class
 *      above 512); each class runs as its own persistent launch on its own stream, forked
 *      from and joined to the caller's stream. Each query's result equals its standalone
 *      search at its (ef, k). Per-query k above max_k is clamped to max_k. */
typedef struct vdx_gpu_batch vdx_gpu_batch;
typedef struct vdx_batch_class {
    uint32_t ef_cap;          /* R capacity of the class */
    uint32_t queries;         /* queries of the tuning sample in the class */
    uint32_t visited_log2;    /* launch shape in use */
    uint32_t waves_per_group;
    uint32_t blocks_per_cu;   /* resident workgroups per CU the grid fills (0: as many as fit) */
    float ms;                 /* the tuned launch time on the sample (0 before tuning) */
    uint32_t visited_form;    /* 1: the LDS visited cache; 2: the HBM bitmap (R capacity above 512) */
} vdx_batch_class;
/* Device buffers for batches of up to max_nq queries with k <= max_k and ef <= max_ef; opts
 * selects termination (a learned model must outlive the batch) and, when its visited_log2 or
 * waves_per_group are set, fixes those shapes. Free the batch before its index. */
VDX_API int vdx_gpu_hnsw_batch_create(vdx_gpu_hnsw* index, uint64_t max_nq, uint32_t max_k, uint32_t max_ef,
                                      const vdx_search_opts* opts, vdx_gpu_batch** out);
VDX_API void vdx_gpu_hnsw_batch_free(vdx_gpu_batch* batch);
/* Autotuner: times each class's launch on a sample batch over visited-set forms (LDS cache
 * sizes; for R capacities above 512 also the HBM bitmap), waves per workgroup and resident
 * workgroups per CU, keeping the fastest shape per class. Synchronous. */
VDX_API int vdx_gpu_hnsw_batch_tune(vdx_gpu_batch* batch, const float* d_queries, uint64_t nq,
                                    const uint32_t* d_ks, uint32_t k, const uint32_t* d_efs, uint32_t ef,
                                    vdx_neighbor* d_out, uint64_t out_stride, uint32_t* d_out_count,
                                    uint32_t reps);
/* Enqueues the batch on `stream` (the index context's stream if NULL) and returns. With
 * use_graph, the launch sequence is captured into a HIP graph on the first call with these
 * buffers and sizes and replayed by later identical calls. Calls on one batch must not overlap. */
VDX_API int vdx_gpu_hnsw_batch_run(vdx_gpu_batch* batch, const float* d_queries, uint64_t nq,
                                   const uint32_t* d_ks, uint32_t k, const uint32_t* d_efs, uint32_t ef,
                                   vdx_neighbor* d_out, uint64_t out_stride, uint32_t* d_out_count,
                                   int use_graph, void* stream);
/* The classes and their launch shapes (count = number of classes; up to `capacity` written). */
VDX_API int vdx_gpu_hnsw_batch_classes(const vdx_gpu_batch* batch, vdx_batch_class* out, uint32_t capacity,
                                       uint32_t* count);

/* No counterpart in vecdex: for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc). vdx_gpu_hnsw_batch_run's uneven
 * batch without a batch object: the same classes as a batch created with (nq, max_k, max_ef,
 * opts), the same per-query results, enqueued on `stream` (the index context's stream if NULL)
 * as one classify launch and one persistent launch per class, in turn on that stream, in default
 * launch shapes. One call's scratch (work counters, class ranges, the query order) lives in a
 * fixed device array of the library, 4 MiB per GPU, located once per device by vdx_gpu_open:
 * the call makes no hipMalloc/hipFree and no other allocation on the device, and clears nothing
 * with memset (the classify kernel zeroes the counters). nq above VDX_NOALLOC_MAX_QUERIES is
 * rejected, as is opts.stats (not collected). Every call on a GPU shares that GPU's array: calls
 * that overlap on one GPU (on different streams, or from different threads) race and must be
 * serialized by the caller; calls queued on one stream are ordered by it. */
#define VDX_NOALLOC_MAX_QUERIES 1048576u
VDX_API int vdx_gpu_hnsw_search_uneven_noalloc(vdx_gpu_hnsw* index, const float* d_queries, uint64_t nq,
                                               const uint32_t* d_ks, uint32_t k, const uint32_t* d_efs,
                                               uint32_t ef, uint32_t max_k, uint32_t max_ef,
                                               vdx_neighbor* d_out, uint64_t out_stride,
                                               uint32_t* d_out_count, const vdx_search_opts* opts,
                                               void* stream);

/* ---- Brute force: replaces BruteForceIndex::search (include/vecdex/brute_force.hpp:36;
 *      src/brute_force.cpp:68-84), batched over queries. */

typedef struct vdx_gpu_bf vdx_gpu_bf;

enum vdx_precision {
    VDX_PREC_FP32 = 0, /* FP32 MFMA shortlist, exact re-rank */
    VDX_PREC_BF16 = 1, /* BF16 MFMA shortlist (FP32 accumulate), exact re-rank */
    VDX_PREC_FP8 = 2,  /* FP8-e4m3 MFMA with per-vector scales, exact re-rank */
    VDX_PREC_EXACT = 3, /* no GEMM: every distance in the exact order (slow; for checks) */
    /* Flags OR'ed into a mode above (no vecdex counterpart; vecdex's float64 yardstick is
     * tests/brute_force_test.cpp:189-213):
     * FP64_TIES: candidates whose float32 distances chain within one ulp of each other are
     *   re-ranked on the GPU by their float64 distance (ties by smaller id) and the first k of
     *   the re-ranked list returned, each with its own float32 distance (so results can differ
     *   from BruteForceIndex::search exactly where float32 cannot order them); k <= 512.
     * FP64_ERROR: the float64 distances of the returned neighbours are computed on the GPU and
     *   the largest |float32 - float64| reported in vdx_bf_stats (results unchanged). */
    VDX_PREC_FP64_TIES = 0x100,
    VDX_PREC_FP64_ERROR = 0x200
};

/* Vectors as BruteForceIndex stores them (unit length under cosine). */
VDX_API int vdx_gpu_bf_upload(vdx_gpu_ctx* ctx, uint32_t metric, uint64_t dim, uint64_t n,
                      const float* vectors, vdx_gpu_bf** out);
/* BruteForceIndex::add after upload (no counterpart in vecdex, whose index has no device copy):
 * appends `count` stored rows (normalized under cosine, as add stores them) to the device copy,
 * growing it with headroom; the MFMA operands are made again on next use. Not concurrent with
 * searches on the same handle. */
VDX_API int vdx_gpu_bf_append(vdx_gpu_bf* index, const float* vectors, uint64_t count, uint64_t dim);
VDX_API void vdx_gpu_bf_free(vdx_gpu_bf* index);

typedef struct vdx_bf_stats {
    uint64_t shortlist;        /* candidates kept per query in the first pass */
    uint64_t widened;          /* queries whose shortlist failed the bound and were redone */
    uint64_t exact_fallback;   /* queries finished by an exact scan */
    uint64_t fp64_ties;        /* VDX_PREC_FP64_TIES: candidates in one-ulp groups re-ranked in float64 */
    double max_abs_error;      /* max |GEMM distance - exact distance| over re-ranked candidates */
    uint64_t fp64_reordered;   /* VDX_PREC_FP64_TIES: queries whose order the re-rank changed */
    uint64_t fp64_unresolved;  /* VDX_PREC_FP64_TIES: queries whose tie group at the k-th result
                                  still reached the end of the largest window (1024 candidates) */
    uint64_t fp64_window;      /* candidates per query the float64 pass examined */
    double max_err_fp64;       /* FP64_TIES or FP64_ERROR: max |returned distance - float64 distance| */
    double max_rel_err_fp64;   /* the same divided by 1 + |float64 distance| */
} vdx_bf_stats;

/* The min(k, n) closest stored vectors of each query, ascending by (distance, id); the
 * distances are the exact ones (bit-equal to BruteForceIndex::search). Queries validated and
 * prepared as in vdx_gpu_hnsw_search ("BruteForceIndex::search" if where is NULL). */
VDX_API int vdx_gpu_bf_search(vdx_gpu_bf* index, const float* queries, uint64_t nq, uint64_t query_dim,
                      uint32_t k, uint32_t precision, vdx_neighbor* out, uint64_t out_stride,
                      uint32_t* out_count, const char* where, vdx_bf_stats* stats);

/* The BF16 filter GEMM's shape for 97-128-dimensional bases (no vecdex counterpart; results are
 * the same either way): 0, eight-wave workgroups of one 32-query block per wave (the default);
 * 1, four-wave workgroups of two 32-query blocks per wave (each base fragment read from LDS feeds
 * two MFMAs). vdx_gpu_bf_tune times both on a sample of prepared device queries (the whole BF16
 * search at k, best of `reps` after one warm-up each), keeps the faster, and reports the times
 * (ms[2], may be NULL) and the choice. */
VDX_API int vdx_gpu_bf_gemm_shape(vdx_gpu_bf* index, uint32_t shape);
VDX_API int vdx_gpu_bf_tune(vdx_gpu_bf* index, const float* d_queries, uint64_t nq, uint32_t k, uint32_t reps,
                            float* ms, uint32_t* shape);

/* Device-memory variant: prepared d_queries, results to d_out (rows of out_stride), on the
 * context's stream; returns after the results are complete. */
VDX_API int vdx_gpu_bf_search_device(vdx_gpu_bf* index, const float* d_queries, uint64_t nq, uint32_t k,
                             uint32_t precision, vdx_neighbor* d_out, uint64_t out_stride,
                             uint32_t* d_out_count, vdx_bf_stats* stats);

#ifdef __cplusplus
}
#endif

#endif /* VDXGPU_H */
