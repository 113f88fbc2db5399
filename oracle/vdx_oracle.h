/* The oracle: a plain-C restatement of vecdex's query path, used ONLY as the checker.
 *
 * Test infrastructure. Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * may load it; the product (hnsw_vector_search_amd) never calls it and has no CPU search path.
 *
 * What each function restates, and what pins it, is listed in vdx_oracle.c.
 */
#ifndef VDX_ORACLE_H
#define VDX_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct vdxo_neighbor {
    float distance;
    uint32_t id;
} vdxo_neighbor;

/* A loaded graph in FORMAT v1 section layout (docs/FORMAT.md). */
typedef struct vdxo_graph {
    uint32_t metric; /* 0 L2, 1 inner product, 2 cosine (FORMAT codes) */
    uint64_t dim;
    uint64_t n;
    uint64_t M;
    uint64_t max_level;
    uint64_t entry_point;
    const uint8_t* levels;
    const uint32_t* layer0;
    const uint32_t* upper;
    const float* vectors;
} vdxo_graph;

/* Distances. *_avx2order: the summation order of src/kernels_avx2.cpp:17-80, written as
 * scalar C with fmaf (bit-exact with the intrinsics). *_scalar: src/kernels_scalar.cpp:9-24. */
float vdxo_l2_avx2order(const float* a, const float* b, size_t dim);
float vdxo_ip_avx2order(const float* a, const float* b, size_t dim);
float vdxo_l2_scalar(const float* a, const float* b, size_t dim);
float vdxo_ip_scalar(const float* a, const float* b, size_t dim);
/* The same order with AVX2 intrinsics when the CPU has them (used by the timed CPU baseline);
 * returns 0 from vdxo_has_avx2 when it falls back to the scalar restatement. */
int vdxo_has_avx2(void);
float vdxo_l2_fast(const float* a, const float* b, size_t dim);
float vdxo_ip_fast(const float* a, const float* b, size_t dim);

/* v /= sqrt(<v, v>) (src/vector_ops.cpp:29-34), dot in the AVX2 order. */
void vdxo_normalize(float* v, size_t dim);

/* Exact k-NN in double precision over float inputs, ties by smaller id.
 * metric 0: sum (x-y)^2; 1: 1 - <x,y>; 2: 1 - <x,y>/sqrt(|x|^2 |y|^2).
 * Writes min(k, n) results per query into rows of k (ids, distances); returns that count. */
uint64_t vdxo_bruteforce_f64(const float* base, uint64_t n, uint64_t dim, uint32_t metric,
                             const float* queries, uint64_t nq, uint64_t k, uint32_t* out_ids,
                             double* out_dist, int threads);

/* BruteForceIndex::search restated (src/brute_force.cpp:20-37, 68-84): float distances in the
 * AVX2 order (metric 0: l2; 1, 2: 1 - dot over the given, already normalized vectors),
 * bounded max-heap, ascending by (distance, id). Rows of k; returns min(k, n). */
uint64_t vdxo_bruteforce_f32(const float* base, uint64_t n, uint64_t dim, uint32_t metric,
                             const float* queries, uint64_t nq, uint64_t k, vdxo_neighbor* out,
                             int threads);

/* HnswIndex::search_unchecked restated (src/hnsw.cpp:416-482, 504-522) over prepared queries
 * (normalized by the caller under cosine). Rows of `out_stride` results; out_count[q] gets the
 * number found. threads <= 0 means one per online CPU. Returns 0. */
int vdxo_hnsw_search_batch(const vdxo_graph* g, const float* queries, uint64_t nq, uint64_t k,
                           uint64_t ef, vdxo_neighbor* out, uint64_t out_stride,
                           uint32_t* out_count, int threads, int use_fast_kernel);
/* The same on any layer, with optional per-query entries on it (entries[q] = UINT32_MAX: descend
 * from the top to the layer above as above): the layer searches of HnswIndex::insert, each of
 * which starts from the result of the layer above (src/hnsw.cpp:315-335). */
int vdxo_hnsw_search_batch_from(const vdxo_graph* g, const float* queries, const uint32_t* entries,
                                uint64_t layer, uint64_t nq, uint64_t k, uint64_t ef, vdxo_neighbor* out,
                                uint64_t out_stride, uint32_t* out_count, int threads, int use_fast_kernel);

/* vdxo_hnsw_search_batch with an expansion budget (no vecdex counterpart): a query stops after
 * `max_expansions` layer-0 expansions (0: no budget) and returns the closest k found so far;
 * expansions[q] (if not NULL) gets the expansions it took. With a budget at or above a query's
 * own count the search is vecdex's. */
int vdxo_hnsw_search_batch_bounded(const vdxo_graph* g, const float* queries, uint64_t nq, uint64_t k,
                                   uint64_t ef, uint64_t max_expansions, vdxo_neighbor* out,
                                   uint64_t out_stride, uint32_t* out_count, uint32_t* expansions,
                                   int threads, int use_fast_kernel);

/* XXH64 from the xxHash specification (doc/xxhash_spec.md), one shot. */
uint64_t vdxo_xxh64(const void* data, size_t size, uint64_t seed);

#ifdef __cplusplus
}
#endif

#endif
