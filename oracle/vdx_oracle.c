/* The oracle (test infrastructure only; see vdx_oracle.h).
 *
 * Restates, in plain C, vecdex's query path:
 *   - distance kernels: src/kernels_avx2.cpp:17-80 (order), src/kernels_scalar.cpp:9-24;
 *     pinned by the known answers of tests/distance_test.cpp:16-58 and by exactness on
 *     integer data (tests/kernels_test.cpp:144-161), and cross-checked against an
 *     independent AVX2-intrinsics path and against float64 (tests/test_oracle.py);
 *   - normalize: src/vector_ops.cpp:14-34;
 *   - exact brute force: src/brute_force.cpp:20-37, 68-84 (float) and the float64 reference of
 *     tests/brute_force_test.cpp:178-200; pinned by the hand cases of
 *     tests/brute_force_test.cpp:46-73, 158-175 and by numpy float64 brute force;
 *   - HNSW search: src/hnsw.cpp:416-482 (greedy_closest, search_layer), 504-522
 *     (search_unchecked), src/search_scratch.hpp:21-47 (visited epochs); pinned by the golden
 *     fingerprints 6133164406513967666 / 15666754926265556722 of
 *     tests/hnsw_parallel_test.cpp:146-169 and by exactness when ef covers every node
 *     (tests/hnsw_test.cpp:101-111);
 *   - XXH64: the xxHash specification; pinned by xxHash's sanity vectors
 *     (tests/xxhash64_test.cpp:34-56).
 *
 * Built with -ffp-contract=off so that every multiply and add rounds as written.
 */
#include "vdx_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdatomic.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#if defined(__x86_64__)
#include <immintrin.h>
#endif

/* ---------------------------------------------------------------- distances */

static float reduce32(const float* acc) {
    float v[8], h[4];
    for (int l = 0; l < 8; ++l) v[l] = (acc[l] + acc[8 + l]) + (acc[16 + l] + acc[24 + l]);
    for (int l = 0; l < 4; ++l) h[l] = v[l] + v[l + 4];
    return (h[0] + h[2]) + (h[1] + h[3]);
}

float vdxo_l2_avx2order(const float* a, const float* b, size_t dim) {
    float acc[32] = {0};
    size_t i = 0;
    for (; dim - i >= 32; i += 32)
        for (int s = 0; s < 32; ++s) {
            float d = a[i + s] - b[i + s];
            acc[s] = fmaf(d, d, acc[s]);
        }
    for (; dim - i >= 8; i += 8)
        for (int l = 0; l < 8; ++l) {
            float d = a[i + l] - b[i + l];
            acc[l] = fmaf(d, d, acc[l]);
        }
    float sum = reduce32(acc);
    for (; i < dim; ++i) {
        float d = a[i] - b[i];
        float sq = d * d;
        sum = sum + sq;
    }
    return sum;
}

float vdxo_ip_avx2order(const float* a, const float* b, size_t dim) {
    float acc[32] = {0};
    size_t i = 0;
    for (; dim - i >= 32; i += 32)
        for (int s = 0; s < 32; ++s) acc[s] = fmaf(a[i + s], b[i + s], acc[s]);
    for (; dim - i >= 8; i += 8)
        for (int l = 0; l < 8; ++l) acc[l] = fmaf(a[i + l], b[i + l], acc[l]);
    float sum = reduce32(acc);
    for (; i < dim; ++i) {
        float p = a[i] * b[i];
        sum = sum + p;
    }
    return sum;
}

float vdxo_l2_scalar(const float* a, const float* b, size_t dim) {
    float sum = 0.0f;
    for (size_t i = 0; i < dim; ++i) {
        float d = a[i] - b[i];
        float sq = d * d;
        sum = sum + sq;
    }
    return sum;
}

float vdxo_ip_scalar(const float* a, const float* b, size_t dim) {
    float sum = 0.0f;
    for (size_t i = 0; i < dim; ++i) {
        float p = a[i] * b[i];
        sum = sum + p;
    }
    return sum;
}

#if defined(__x86_64__)
__attribute__((target("avx2,fma"))) static float hsum_avx2(__m256 a0, __m256 a1, __m256 a2,
                                                          __m256 a3) {
    __m256 v = _mm256_add_ps(_mm256_add_ps(a0, a1), _mm256_add_ps(a2, a3));
    __m128 h = _mm_add_ps(_mm256_castps256_ps128(v), _mm256_extractf128_ps(v, 1));
    __m128 p = _mm_add_ps(h, _mm_movehl_ps(h, h));
    return _mm_cvtss_f32(_mm_add_ss(p, _mm_shuffle_ps(p, p, 0x55)));
}

__attribute__((target("avx2,fma"))) static float l2_avx2(const float* a, const float* b,
                                                        size_t dim) {
    __m256 c0 = _mm256_setzero_ps(), c1 = c0, c2 = c0, c3 = c0;
    size_t i = 0;
    for (; dim - i >= 32; i += 32) {
        __m256 d0 = _mm256_sub_ps(_mm256_loadu_ps(a + i), _mm256_loadu_ps(b + i));
        __m256 d1 = _mm256_sub_ps(_mm256_loadu_ps(a + i + 8), _mm256_loadu_ps(b + i + 8));
        __m256 d2 = _mm256_sub_ps(_mm256_loadu_ps(a + i + 16), _mm256_loadu_ps(b + i + 16));
        __m256 d3 = _mm256_sub_ps(_mm256_loadu_ps(a + i + 24), _mm256_loadu_ps(b + i + 24));
        c0 = _mm256_fmadd_ps(d0, d0, c0);
        c1 = _mm256_fmadd_ps(d1, d1, c1);
        c2 = _mm256_fmadd_ps(d2, d2, c2);
        c3 = _mm256_fmadd_ps(d3, d3, c3);
    }
    for (; dim - i >= 8; i += 8) {
        __m256 d = _mm256_sub_ps(_mm256_loadu_ps(a + i), _mm256_loadu_ps(b + i));
        c0 = _mm256_fmadd_ps(d, d, c0);
    }
    float sum = hsum_avx2(c0, c1, c2, c3);
    for (; i < dim; ++i) {
        float d = a[i] - b[i];
        float sq = d * d;
        sum = sum + sq;
    }
    return sum;
}

__attribute__((target("avx2,fma"))) static float ip_avx2(const float* a, const float* b,
                                                        size_t dim) {
    __m256 c0 = _mm256_setzero_ps(), c1 = c0, c2 = c0, c3 = c0;
    size_t i = 0;
    for (; dim - i >= 32; i += 32) {
        c0 = _mm256_fmadd_ps(_mm256_loadu_ps(a + i), _mm256_loadu_ps(b + i), c0);
        c1 = _mm256_fmadd_ps(_mm256_loadu_ps(a + i + 8), _mm256_loadu_ps(b + i + 8), c1);
        c2 = _mm256_fmadd_ps(_mm256_loadu_ps(a + i + 16), _mm256_loadu_ps(b + i + 16), c2);
        c3 = _mm256_fmadd_ps(_mm256_loadu_ps(a + i + 24), _mm256_loadu_ps(b + i + 24), c3);
    }
    for (; dim - i >= 8; i += 8)
        c0 = _mm256_fmadd_ps(_mm256_loadu_ps(a + i), _mm256_loadu_ps(b + i), c0);
    float sum = hsum_avx2(c0, c1, c2, c3);
    for (; i < dim; ++i) {
        float p = a[i] * b[i];
        sum = sum + p;
    }
    return sum;
}
#endif

static int g_avx2 = -1;

int vdxo_has_avx2(void) {
    if (g_avx2 < 0) {
#if defined(__x86_64__)
        __builtin_cpu_init();
        g_avx2 = __builtin_cpu_supports("avx2") && __builtin_cpu_supports("fma");
#else
        g_avx2 = 0;
#endif
    }
    return g_avx2;
}

float vdxo_l2_fast(const float* a, const float* b, size_t dim) {
#if defined(__x86_64__)
    if (vdxo_has_avx2()) return l2_avx2(a, b, dim);
#endif
    return vdxo_l2_avx2order(a, b, dim);
}

float vdxo_ip_fast(const float* a, const float* b, size_t dim) {
#if defined(__x86_64__)
    if (vdxo_has_avx2()) return ip_avx2(a, b, dim);
#endif
    return vdxo_ip_avx2order(a, b, dim);
}

void vdxo_normalize(float* v, size_t dim) {
    float n = sqrtf(vdxo_ip_avx2order(v, v, dim));
    for (size_t i = 0; i < dim; ++i) v[i] /= n;
}

typedef float (*dist_fn)(const float*, const float*, size_t);

static float metric_dist(uint32_t metric, dist_fn l2, dist_fn ip, const float* a, const float* b,
                         size_t dim) {
    return metric == 0 ? l2(a, b, dim) : 1.0f - ip(a, b, dim);
}

/* ---------------------------------------------------------------- ordering and heaps */

static int closer(vdxo_neighbor a, vdxo_neighbor b) {
    return a.distance < b.distance || (a.distance == b.distance && a.id < b.id);
}

/* Binary heap whose front is the element for which `before(front, x)` holds for all x:
 * before = closer gives a min-heap (closest in front), before = farther a max-heap. */
typedef int (*order_fn)(vdxo_neighbor, vdxo_neighbor);

static int farther(vdxo_neighbor a, vdxo_neighbor b) { return closer(b, a); }

static void heap_push(vdxo_neighbor* h, size_t* size, vdxo_neighbor x, order_fn before) {
    size_t i = (*size)++;
    while (i > 0) {
        size_t parent = (i - 1) / 2;
        if (!before(x, h[parent])) break;
        h[i] = h[parent];
        i = parent;
    }
    h[i] = x;
}

static vdxo_neighbor heap_pop(vdxo_neighbor* h, size_t* size, order_fn before) {
    vdxo_neighbor top = h[0];
    vdxo_neighbor last = h[--(*size)];
    size_t n = *size, i = 0;
    for (;;) {
        size_t c = 2 * i + 1;
        if (c >= n) break;
        if (c + 1 < n && before(h[c + 1], h[c])) ++c;
        if (!before(h[c], last)) break;
        h[i] = h[c];
        i = c;
    }
    if (n > 0) h[i] = last;
    return top;
}

static int cmp_closer(const void* a, const void* b) {
    vdxo_neighbor x = *(const vdxo_neighbor*)a, y = *(const vdxo_neighbor*)b;
    return closer(x, y) ? -1 : (closer(y, x) ? 1 : 0);
}

/* ---------------------------------------------------------------- threads */

typedef struct {
    uint64_t count;
    atomic_ulong next;
    void (*fn)(void* ctx, uint64_t i, void* local);
    void* ctx;
    size_t local_bytes;
    void (*cleanup)(void* local);
} job_t;

static void* worker(void* arg) {
    job_t* job = (job_t*)arg;
    void* local = job->local_bytes ? calloc(1, job->local_bytes) : NULL;
    for (uint64_t i = atomic_fetch_add(&job->next, 1); i < job->count;
         i = atomic_fetch_add(&job->next, 1))
        job->fn(job->ctx, i, local);
    if (local && job->cleanup) job->cleanup(local);
    free(local);
    return NULL;
}

static void run_parallel(uint64_t count, int threads, void (*fn)(void*, uint64_t, void*), void* ctx,
                         size_t local_bytes, void (*cleanup)(void*)) {
    if (threads <= 0) threads = (int)sysconf(_SC_NPROCESSORS_ONLN);
    if (threads < 1) threads = 1;
    if ((uint64_t)threads > count) threads = count ? (int)count : 1;
    job_t job;
    job.count = count;
    atomic_init(&job.next, 0);
    job.fn = fn;
    job.ctx = ctx;
    job.local_bytes = local_bytes;
    job.cleanup = cleanup;
    pthread_t* tids = (pthread_t*)malloc(sizeof(pthread_t) * (size_t)threads);
    int started = 0;
    for (int t = 1; t < threads; ++t)
        if (pthread_create(&tids[started], NULL, worker, &job) == 0) ++started;
    worker(&job);
    for (int t = 0; t < started; ++t) pthread_join(tids[t], NULL);
    free(tids);
}

/* ---------------------------------------------------------------- float64 brute force */

typedef struct {
    double d;
    uint32_t id;
} dn_t;

static int dn_before(dn_t a, dn_t b) { return a.d < b.d || (a.d == b.d && a.id < b.id); }

typedef struct {
    const float* base;
    uint64_t n, dim, k, keep;
    uint32_t metric;
    const float* queries;
    uint32_t* out_ids;
    double* out_dist;
} bf64_ctx;

static void bf64_one(void* vctx, uint64_t q, void* local) {
    bf64_ctx* c = (bf64_ctx*)vctx;
    dn_t* heap = (dn_t*)local; /* max-heap of the best `keep`, worst in front */
    size_t size = 0;
    const float* x = c->queries + q * c->dim;
    double xx = 0.0;
    for (uint64_t t = 0; t < c->dim; ++t) xx += (double)x[t] * (double)x[t];
    for (uint64_t i = 0; i < c->n; ++i) {
        const float* y = c->base + i * c->dim;
        double l2 = 0.0, dot = 0.0, yy = 0.0;
        for (uint64_t t = 0; t < c->dim; ++t) {
            double a = x[t], b = y[t];
            l2 += (a - b) * (a - b);
            dot += a * b;
            yy += b * b;
        }
        double d = c->metric == 0 ? l2 : (c->metric == 1 ? 1.0 - dot : 1.0 - dot / sqrt(xx * yy));
        dn_t cand = {d, (uint32_t)i};
        if (size < c->keep) {
            size_t j = size++;
            while (j > 0) {
                size_t p = (j - 1) / 2;
                if (!dn_before(heap[p], cand)) break;
                heap[j] = heap[p];
                j = p;
            }
            heap[j] = cand;
        } else if (dn_before(cand, heap[0])) {
            size_t j = 0;
            for (;;) {
                size_t ch = 2 * j + 1;
                if (ch >= size) break;
                if (ch + 1 < size && dn_before(heap[ch], heap[ch + 1])) ++ch;
                if (!dn_before(cand, heap[ch])) break;
                heap[j] = heap[ch];
                j = ch;
            }
            heap[j] = cand;
        }
    }
    /* heap-sort ascending */
    for (size_t end = size; end > 1; --end) {
        dn_t top = heap[0], last = heap[end - 1];
        size_t j = 0, n = end - 1;
        for (;;) {
            size_t ch = 2 * j + 1;
            if (ch >= n) break;
            if (ch + 1 < n && dn_before(heap[ch], heap[ch + 1])) ++ch;
            if (!dn_before(last, heap[ch])) break;
            heap[j] = heap[ch];
            j = ch;
        }
        heap[j] = last;
        heap[end - 1] = top;
    }
    for (size_t j = 0; j < size; ++j) {
        c->out_ids[q * c->k + j] = heap[j].id;
        c->out_dist[q * c->k + j] = heap[j].d;
    }
}

uint64_t vdxo_bruteforce_f64(const float* base, uint64_t n, uint64_t dim, uint32_t metric,
                             const float* queries, uint64_t nq, uint64_t k, uint32_t* out_ids,
                             double* out_dist, int threads) {
    bf64_ctx c = {base, n, dim, k, k < n ? k : n, metric, queries, out_ids, out_dist};
    run_parallel(nq, threads, bf64_one, &c, sizeof(dn_t) * (c.keep + 1), NULL);
    return c.keep;
}

/* ---------------------------------------------------------------- float32 brute force */

typedef struct {
    const float* base;
    uint64_t n, dim, k, keep;
    uint32_t metric;
    const float* queries;
    vdxo_neighbor* out;
} bf32_ctx;

static void bf32_one(void* vctx, uint64_t q, void* local) {
    bf32_ctx* c = (bf32_ctx*)vctx;
    vdxo_neighbor* heap = (vdxo_neighbor*)local;
    size_t size = 0;
    const float* x = c->queries + q * c->dim;
    for (uint64_t i = 0; i < c->n; ++i) {
        vdxo_neighbor cand = {metric_dist(c->metric, vdxo_l2_fast, vdxo_ip_fast, x,
                                          c->base + i * c->dim, c->dim),
                              (uint32_t)i};
        if (size < c->keep) {
            heap_push(heap, &size, cand, farther);
        } else if (closer(cand, heap[0])) {
            heap_pop(heap, &size, farther);
            heap_push(heap, &size, cand, farther);
        }
    }
    qsort(heap, size, sizeof(vdxo_neighbor), cmp_closer);
    memcpy(c->out + q * c->k, heap, size * sizeof(vdxo_neighbor));
}

uint64_t vdxo_bruteforce_f32(const float* base, uint64_t n, uint64_t dim, uint32_t metric,
                             const float* queries, uint64_t nq, uint64_t k, vdxo_neighbor* out,
                             int threads) {
    bf32_ctx c = {base, n, dim, k, k < n ? k : n, metric, queries, out};
    run_parallel(nq, threads, bf32_one, &c, sizeof(vdxo_neighbor) * (c.keep + 1), NULL);
    return c.keep;
}

/* ---------------------------------------------------------------- HNSW search */

typedef struct {
    uint32_t* marks;
    uint32_t epoch;
    vdxo_neighbor* cand; /* min-heap (closest in front) */
    vdxo_neighbor* res;  /* max-heap (worst in front) */
    uint32_t* fresh;
    size_t cap;
} scratch_t;

typedef struct {
    const vdxo_graph* g;
    uint64_t* upper_offset;
    const float* queries;
    const uint32_t* entries; /* per query: an entry on `layer` (no descent), or UINT32_MAX; may be NULL */
    uint64_t layer;          /* the layer the beam searches: 0 for queries, any for an insertion */
    uint64_t k, ef;
    vdxo_neighbor* out;
    uint64_t out_stride;
    uint32_t* out_count;
    uint64_t max_expansions; /* expansions of the searched layer per query at most (UINT64_MAX: none) */
    uint32_t* expansions;    /* per query: expansions taken; may be NULL */
    dist_fn l2, ip;
} hnsw_ctx;

static const uint32_t* links_of(const hnsw_ctx* c, uint32_t node, uint64_t layer) {
    const vdxo_graph* g = c->g;
    if (layer == 0) return g->layer0 + (uint64_t)node * (1 + 2 * g->M);
    return g->upper + (c->upper_offset[node] + (layer - 1)) * (1 + g->M);
}

static void hnsw_one(void* vctx, uint64_t q, void* local) {
    hnsw_ctx* c = (hnsw_ctx*)vctx;
    const vdxo_graph* g = c->g;
    scratch_t* s = (scratch_t*)local;
    const uint64_t ef = c->ef > c->k ? c->ef : c->k;
    const float* query = c->queries + q * g->dim;
    if (g->n == 0) {
        c->out_count[q] = 0;
        return;
    }
    if (s->marks == NULL) {
        s->marks = (uint32_t*)calloc(g->n, sizeof(uint32_t));
        s->cap = 1024;
        s->cand = (vdxo_neighbor*)malloc(s->cap * sizeof(vdxo_neighbor));
        s->res = (vdxo_neighbor*)malloc(s->cap * sizeof(vdxo_neighbor));
        s->fresh = (uint32_t*)malloc((2 * g->M + 1) * sizeof(uint32_t));
    }
    /* room: candidates gain at most 2M per expansion; results stay <= ef + 1 */
    const uint64_t dim = g->dim;
    vdxo_neighbor entry = {metric_dist(g->metric, c->l2, c->ip, query,
                                       g->vectors + g->entry_point * dim, dim),
                           (uint32_t)g->entry_point};
    /* A given entry replaces the descent: insert's search of each layer starts from the best
     * result of the layer above (src/hnsw.cpp:324-335). */
    uint64_t top = g->max_level;
    if (c->entries != NULL && c->entries[q] != UINT32_MAX) {
        entry.id = c->entries[q];
        entry.distance = metric_dist(g->metric, c->l2, c->ip, query, g->vectors + (uint64_t)entry.id * dim, dim);
        top = 0;
    }
    /* greedy descent: move to any strictly closer neighbor, in link order (src/hnsw.cpp:416-431) */
    for (uint64_t layer = top; layer > c->layer; --layer) {
        for (int moved = 1; moved;) {
            moved = 0;
            const uint32_t* b = links_of(c, entry.id, layer);
            for (uint32_t i = 0; i < b[0]; ++i) {
                float d = metric_dist(g->metric, c->l2, c->ip, query, g->vectors + (uint64_t)b[1 + i] * dim,
                                      dim);
                if (d < entry.distance) {
                    entry.distance = d;
                    entry.id = b[1 + i];
                    moved = 1;
                }
            }
        }
    }
    /* the beam on the searched layer (src/hnsw.cpp:438-482) */
    if (++s->epoch == 0) {
        memset(s->marks, 0, g->n * sizeof(uint32_t));
        s->epoch = 1;
    }
    size_t ncand = 0, nres = 0;
    uint64_t taken = 0;
    s->marks[entry.id] = s->epoch;
    heap_push(s->cand, &ncand, entry, closer);
    heap_push(s->res, &nres, entry, farther);
    while (ncand > 0) {
        vdxo_neighbor cur = heap_pop(s->cand, &ncand, closer);
        if (nres >= ef && closer(s->res[0], cur)) break;
        /* the expansion budget (no vecdex counterpart): stop with the results found so far */
        if (taken >= c->max_expansions) break;
        ++taken;
        const uint32_t* b = links_of(c, cur.id, c->layer);
        uint32_t nf = 0;
        for (uint32_t i = 0; i < b[0]; ++i) {
            uint32_t id = b[1 + i];
            if (s->marks[id] != s->epoch) {
                s->marks[id] = s->epoch;
                s->fresh[nf++] = id;
                const char* p = (const char*)(g->vectors + (uint64_t)id * dim);
                for (uint64_t off = 0; off < dim * sizeof(float); off += 64) __builtin_prefetch(p + off);
            }
        }
        for (uint32_t i = 0; i < nf; ++i) {
            vdxo_neighbor next = {metric_dist(g->metric, c->l2, c->ip, query,
                                              g->vectors + (uint64_t)s->fresh[i] * dim, dim),
                                  s->fresh[i]};
            if (nres < ef || closer(next, s->res[0])) {
                if (ncand + 1 > s->cap || nres + 2 > s->cap) {
                    s->cap *= 2;
                    s->cand = (vdxo_neighbor*)realloc(s->cand, s->cap * sizeof(vdxo_neighbor));
                    s->res = (vdxo_neighbor*)realloc(s->res, s->cap * sizeof(vdxo_neighbor));
                }
                heap_push(s->cand, &ncand, next, closer);
                heap_push(s->res, &nres, next, farther);
                if (nres > ef) heap_pop(s->res, &nres, farther);
            }
        }
    }
    qsort(s->res, nres, sizeof(vdxo_neighbor), cmp_closer);
    uint64_t count = nres < c->k ? nres : c->k;
    memcpy(c->out + q * c->out_stride, s->res, count * sizeof(vdxo_neighbor));
    c->out_count[q] = (uint32_t)count;
    if (c->expansions != NULL) c->expansions[q] = (uint32_t)taken;
}

typedef struct {
    scratch_t s;
} hnsw_local;

static void hnsw_worker_entry(void* ctx, uint64_t q, void* local) {
    hnsw_one(ctx, q, local);
}

static void hnsw_cleanup(void* local) {
    scratch_t* s = (scratch_t*)local;
    free(s->marks);
    free(s->cand);
    free(s->res);
    free(s->fresh);
}

static int search_batch(const vdxo_graph* g, const float* queries, const uint32_t* entries, uint64_t layer,
                        uint64_t nq, uint64_t k, uint64_t ef, uint64_t max_expansions, vdxo_neighbor* out,
                        uint64_t out_stride, uint32_t* out_count, uint32_t* expansions, int threads,
                        int use_fast_kernel) {
    hnsw_ctx c;
    c.max_expansions = max_expansions == 0 ? UINT64_MAX : max_expansions;
    c.expansions = expansions;
    c.g = g;
    c.queries = queries;
    c.entries = entries;
    c.layer = layer;
    c.k = k;
    c.ef = ef;
    c.out = out;
    c.out_stride = out_stride;
    c.out_count = out_count;
    c.l2 = use_fast_kernel ? vdxo_l2_fast : vdxo_l2_avx2order;
    c.ip = use_fast_kernel ? vdxo_ip_fast : vdxo_ip_avx2order;
    c.upper_offset = (uint64_t*)malloc((g->n + 1) * sizeof(uint64_t));
    uint64_t acc = 0;
    for (uint64_t i = 0; i < g->n; ++i) {
        c.upper_offset[i] = acc;
        acc += g->levels[i];
    }
    /* Each worker's scratch lives in its zeroed `local` block; what it allocates lazily is freed
     * by hnsw_cleanup when the worker ends. */
    run_parallel(nq, threads, hnsw_worker_entry, &c, sizeof(hnsw_local), hnsw_cleanup);
    free(c.upper_offset);
    return 0;
}

int vdxo_hnsw_search_batch_from(const vdxo_graph* g, const float* queries, const uint32_t* entries,
                                uint64_t layer, uint64_t nq, uint64_t k, uint64_t ef, vdxo_neighbor* out,
                                uint64_t out_stride, uint32_t* out_count, int threads, int use_fast_kernel) {
    return search_batch(g, queries, entries, layer, nq, k, ef, 0, out, out_stride, out_count, NULL, threads,
                        use_fast_kernel);
}

int vdxo_hnsw_search_batch(const vdxo_graph* g, const float* queries, uint64_t nq, uint64_t k,
                           uint64_t ef, vdxo_neighbor* out, uint64_t out_stride,
                           uint32_t* out_count, int threads, int use_fast_kernel) {
    return search_batch(g, queries, NULL, 0, nq, k, ef, 0, out, out_stride, out_count, NULL, threads,
                        use_fast_kernel);
}

int vdxo_hnsw_search_batch_bounded(const vdxo_graph* g, const float* queries, uint64_t nq, uint64_t k,
                                   uint64_t ef, uint64_t max_expansions, vdxo_neighbor* out,
                                   uint64_t out_stride, uint32_t* out_count, uint32_t* expansions,
                                   int threads, int use_fast_kernel) {
    return search_batch(g, queries, NULL, 0, nq, k, ef, max_expansions, out, out_stride, out_count, expansions,
                        threads, use_fast_kernel);
}

/* ---------------------------------------------------------------- XXH64 */

static const uint64_t XP1 = 0x9E3779B185EBCA87ULL, XP2 = 0xC2B2AE3D27D4EB4FULL,
                      XP3 = 0x165667B19E3779F9ULL, XP4 = 0x85EBCA77C2B2AE63ULL,
                      XP5 = 0x27D4EB2F165667C5ULL;

static uint64_t xrotl(uint64_t x, int r) { return (x << r) | (x >> (64 - r)); }
static uint64_t xrd64(const unsigned char* p) {
    uint64_t v;
    memcpy(&v, p, 8);
    return v;
}
static uint32_t xrd32(const unsigned char* p) {
    uint32_t v;
    memcpy(&v, p, 4);
    return v;
}
static uint64_t xround(uint64_t acc, uint64_t lane) { return xrotl(acc + lane * XP2, 31) * XP1; }

uint64_t vdxo_xxh64(const void* data, size_t size, uint64_t seed) {
    const unsigned char* p = (const unsigned char*)data;
    const unsigned char* end = p + size;
    uint64_t h;
    if (size >= 32) {
        uint64_t v1 = seed + XP1 + XP2, v2 = seed + XP2, v3 = seed, v4 = seed - XP1;
        while (end - p >= 32) {
            v1 = xround(v1, xrd64(p));
            v2 = xround(v2, xrd64(p + 8));
            v3 = xround(v3, xrd64(p + 16));
            v4 = xround(v4, xrd64(p + 24));
            p += 32;
        }
        h = xrotl(v1, 1) + xrotl(v2, 7) + xrotl(v3, 12) + xrotl(v4, 18);
        uint64_t vs[4] = {v1, v2, v3, v4};
        for (int i = 0; i < 4; ++i) {
            h ^= xround(0, vs[i]);
            h = h * XP1 + XP4;
        }
    } else {
        h = seed + XP5;
    }
    h += (uint64_t)size;
    while (end - p >= 8) {
        h ^= xround(0, xrd64(p));
        h = xrotl(h, 27) * XP1 + XP4;
        p += 8;
    }
    if (end - p >= 4) {
        h ^= (uint64_t)xrd32(p) * XP1;
        h = xrotl(h, 23) * XP2 + XP3;
        p += 4;
    }
    while (p < end) {
        h ^= (uint64_t)(*p) * XP5;
        h = xrotl(h, 11) * XP1;
        ++p;
    }
    h ^= h >> 33;
    h *= XP2;
    h ^= h >> 29;
    h *= XP3;
    h ^= h >> 32;
    return h;
}
