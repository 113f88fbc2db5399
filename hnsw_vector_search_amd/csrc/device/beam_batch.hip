// Uneven batches: queries with their own (ef, k) served together, each by the beam kernel
// sized for its R capacity class, with per-class launch shapes chosen by measurement and the
// whole launch sequence replayable as one HIP graph.
//
// A batch of mixed ef sized as one launch would give every wave the LDS and registers of the
// largest ef' = max(ef, k) in it. Instead a classify kernel sorts the queries into classes by
// R capacity (64, 128, 256, 512: the register-resident kernel; above: the LDS kernel), writing
// each class's query list, its range and a zeroed work counter in device memory. One
// persistent launch per class then serves that class's list; the launches run on their own
// streams, forked from and joined back to the caller's stream, so a class that drains early
// frees its CUs for the others. Nothing on the path reads a count back to the host, and no
// memset node is recorded, so the sequence can be captured once and replayed.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <vector>

#include "../host/capi_common.hpp"
#include "beam_args.hpp"
#include "beam_host.hpp"
#include "device_common.hpp"
#include "fill_copy.hpp"
#include "gpu_index.hpp"
#include "gpu_state.hpp"

namespace vdx::dev {

constexpr int kMaxClasses = 5;

struct ClassifyArgs {
    std::uint32_t nq;
    const std::uint32_t* ks;
    std::uint32_t k;
    const std::uint32_t* efs;
    std::uint32_t ef;
    std::uint32_t k_cap;
    std::uint32_t ncls;
    std::uint32_t caps[kMaxClasses];  // ascending R capacities; the last covers every query
    std::uint32_t* order;             // nq: query indices grouped by class
    std::uint32_t* ranges;            // 2 per class: [begin, end) in order
    std::uint32_t* queues;            // 1 per class: zeroed
    std::uint32_t* qflag;             // the byte-query flag of the launches (zeroed), or null
};

__device__ __forceinline__ std::uint32_t class_of(const ClassifyArgs& a, std::uint32_t q) {
    const std::uint32_t k = min(a.ks ? a.ks[q] : a.k, a.k_cap);
    const std::uint32_t ef = a.efs ? a.efs[q] : a.ef;
    const std::uint32_t efp = max(ef, k);
    std::uint32_t c = 0;
    while (c + 1 < a.ncls && a.caps[c] < efp) {
        ++c;
    }
    return c;
}

// One workgroup of 1024 threads: a counting pass, the class offsets, a placing pass. Each wave
// aggregates its lanes per class with a ballot, so there is one LDS atomic per wave and class.
__global__ __launch_bounds__(1024) void beam_classify_kernel(ClassifyArgs a) {
    __shared__ std::uint32_t count[kMaxClasses], next[kMaxClasses];
    const std::uint32_t lane = static_cast<std::uint32_t>(lane_id());
    const std::uint32_t wave = threadIdx.x >> 6;
    const std::uint32_t waves = blockDim.x >> 6;
    if (threadIdx.x < kMaxClasses) {
        count[threadIdx.x] = 0;
    }
    __syncthreads();
    for (std::uint32_t q0 = wave * 64; q0 < a.nq; q0 += waves * 64) {
        const std::uint32_t q = q0 + lane;
        const std::uint32_t c = q < a.nq ? class_of(a, q) : kMaxClasses;
        for (std::uint32_t cc = 0; cc < a.ncls; ++cc) {
            const std::uint64_t m = ballot(c == cc);
            if (m != 0 && lane == static_cast<std::uint32_t>(__ffsll(static_cast<long long>(m)) - 1)) {
                atomicAdd(&count[cc], static_cast<std::uint32_t>(__popcll(m)));
            }
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        std::uint32_t run = 0;
        for (std::uint32_t c = 0; c < a.ncls; ++c) {
            next[c] = run;
            a.ranges[2 * c] = run;
            run += count[c];
            a.ranges[2 * c + 1] = run;
            a.queues[c] = 0;
        }
        if (a.qflag != nullptr) {
            *a.qflag = 0;
        }
    }
    __syncthreads();
    for (std::uint32_t q0 = wave * 64; q0 < a.nq; q0 += waves * 64) {
        const std::uint32_t q = q0 + lane;
        const std::uint32_t c = q < a.nq ? class_of(a, q) : kMaxClasses;
        for (std::uint32_t cc = 0; cc < a.ncls; ++cc) {
            const std::uint64_t m = ballot(c == cc);
            if (m == 0) {
                continue;
            }
            const std::uint32_t leader = static_cast<std::uint32_t>(__ffsll(static_cast<long long>(m)) - 1);
            std::uint32_t at = 0;
            if (lane == leader) {
                at = atomicAdd(&next[cc], static_cast<std::uint32_t>(__popcll(m)));
            }
            at = static_cast<std::uint32_t>(__shfl(static_cast<int>(at), static_cast<int>(leader), kWave));
            if (c == cc) {
                a.order[at + lanes_below(m)] = q;
            }
        }
    }
}

}  // namespace vdx::dev

// The per-call scratch of vdx_gpu_hnsw_search_uneven_noalloc: one call's classification (the
// work counters, the class ranges, the query order by class) for up to
// VDX_NOALLOC_MAX_QUERIES queries, 4 MiB on every GPU the library runs on, held for the life
// of the process. In host code `scratch` names a host-side stand-in: only the address that
// hipGetSymbolAddress returns for a device (noalloc_scratch_attach) may reach a kernel.
namespace {
constexpr std::uint32_t kScratchQueues = 0;   // kMaxClasses work counters
constexpr std::uint32_t kScratchQflag = 8;    // the byte-query flag
constexpr std::uint32_t kScratchRanges = 16;  // 2 x kMaxClasses words
constexpr std::uint32_t kScratchOrder = 64;   // the query order, VDX_NOALLOC_MAX_QUERIES words
constexpr int kScratchDevices = 64;
}  // namespace

static __device__ __align__(256) std::uint32_t scratch[kScratchOrder + VDX_NOALLOC_MAX_QUERIES];

namespace vdx::gpu {
namespace {
std::mutex g_scratch_mutex;
bool g_scratch_done[kScratchDevices] = {};
hipError_t g_scratch_status[kScratchDevices] = {};
std::uint32_t* g_scratch_ptr[kScratchDevices] = {};
}  // namespace

void noalloc_scratch_attach(int device) {
    if (device < 0 || device >= kScratchDevices) return;  // the entry point reports it
    const std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (g_scratch_done[device]) return;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));  // on the current device
    g_scratch_status[device] = e;
    g_scratch_ptr[device] = e == hipSuccess ? static_cast<std::uint32_t*>(p) : nullptr;
    g_scratch_done[device] = true;
}

namespace {
std::uint32_t* noalloc_scratch(int device) {
    if (device < 0 || device >= kScratchDevices) {
        throw std::runtime_error(vdx::fmt("vdx_gpu_hnsw_search_uneven_noalloc: device %d is beyond the %d this "
                                          "build keeps scratch addresses for", device, kScratchDevices));
    }
    const std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_done[device] || g_scratch_ptr[device] == nullptr) {
        throw std::runtime_error(vdx::fmt("vdx_gpu_hnsw_search_uneven_noalloc: the scratch array of device %d "
                                          "could not be located (%s)", device,
                                          g_scratch_done[device] ? hipGetErrorName(g_scratch_status[device])
                                                                 : "the device was never opened"));
    }
    return g_scratch_ptr[device];
}
}  // namespace
}  // namespace vdx::gpu

using vdx::guarded;
using vdx::require;
using vdx::gpu::check;
using vdx::gpu::DeviceBuffer;
using vdx::gpu::DeviceGuard;

struct vdx_gpu_batch {
    vdx_gpu_hnsw* index = nullptr;
    int device = -1;  // the destructor needs no other object (the index may be gone by then)
    std::uint64_t max_nq = 0;
    std::uint32_t max_k = 0, max_ef = 0;
    int mode = vdx::dev::kBeamPlain;
    vdx_search_opts opts{};
    bool has_opts = false;
    int ncls = 0;
    std::uint32_t caps[vdx::dev::kMaxClasses] = {};
    vdx::gpu::BeamConfig cfg[vdx::dev::kMaxClasses];
    vdx_batch_class info[vdx::dev::kMaxClasses] = {};
    DeviceBuffer order, ranges, queues;
    DeviceBuffer bits[vdx::dev::kMaxClasses];  // visited bitmaps of classes tuned to visited_form 2
    hipStream_t side[vdx::dev::kMaxClasses] = {};
    hipEvent_t fork = nullptr, join[vdx::dev::kMaxClasses] = {};
    // the captured launch sequence and the call it was captured for
    hipGraphExec_t exec = nullptr;
    std::uint64_t key[11] = {};

    ~vdx_gpu_batch() {
        if (device >= 0) {
            DeviceGuard guard(device);
            if (exec != nullptr) (void)hipGraphExecDestroy(exec);
            for (int c = 0; c < vdx::dev::kMaxClasses; ++c) {
                if (side[c] != nullptr) {
                    (void)hipStreamSynchronize(side[c]);
                    (void)hipStreamDestroy(side[c]);
                }
                if (join[c] != nullptr) (void)hipEventDestroy(join[c]);
            }
            if (fork != nullptr) (void)hipEventDestroy(fork);
        }
    }

    const vdx_search_opts* opts_ptr() const { return has_opts ? &opts : nullptr; }
};

namespace {

struct BatchCall {
    const float* d_queries;
    std::uint64_t nq;
    const std::uint32_t* d_ks;
    std::uint32_t k;
    const std::uint32_t* d_efs;
    std::uint32_t ef;
    vdx_neighbor* d_out;
    std::uint64_t out_stride;
    std::uint32_t* d_count;
};

void check_call(const vdx_gpu_batch& b, const BatchCall& c) {
    require(c.nq <= b.max_nq, "vdx_gpu_hnsw_batch: more queries than the batch was created for");
    require(c.d_queries != nullptr && c.d_out != nullptr && c.d_count != nullptr,
            "vdx_gpu_hnsw_batch: null buffer");
    require(c.d_ks != nullptr || (c.k >= 1 && c.k <= b.max_k), "vdx_gpu_hnsw_batch: k outside [1, max_k]");
    require(c.d_efs != nullptr || c.ef <= b.max_ef, "vdx_gpu_hnsw_batch: ef above max_ef");
    require(c.out_stride >= b.max_k, "vdx_gpu_hnsw_batch: out_stride is below max_k");
}

// Where one call's classification lives: the query indices grouped by class, two range words
// and one work counter per class (a batch's own buffers, or the no-allocation scratch).
struct ClassBuffers {
    std::uint32_t* order;
    std::uint32_t* ranges;
    std::uint32_t* queues;
    std::uint32_t* qflag;  // the byte-query flag word the integer kernel reads
    DeviceBuffer* bits = nullptr;  // per class, the visited bitmaps (reserved before enqueueing)
};

// The R-capacity classes of batches with k <= max_k and ef <= max_ef: 64, 128, 256, 512 (the
// register-resident kernel, up to the first that holds both), then one LDS-kernel class.
int class_caps(std::uint32_t max_k, std::uint32_t max_ef, std::uint32_t (&caps)[vdx::dev::kMaxClasses]) {
    const std::uint32_t top = std::max(max_ef, max_k);
    int n = 0;
    for (std::uint32_t cap = 64; cap <= vdx::dev::kRegMaxEf; cap *= 2) {
        caps[n++] = cap;
        if (cap >= top) break;
    }
    if (top > vdx::dev::kRegMaxEf) {
        caps[n++] = (top + 63u) / 64u * 64u;
    }
    return n;
}

// The classify launch, then (where the plain search reads the 8-bit rows) the check of every
// query for byte values, into the flag the classify kernel zeroed.
void enqueue_classify_into(const vdx_gpu_hnsw& h, int mode, const vdx::gpu::BeamConfig& cfg, const BatchCall& c,
                           std::uint32_t k_cap, int ncls, const std::uint32_t (&caps)[vdx::dev::kMaxClasses],
                           const ClassBuffers& buf, hipStream_t s) {
    vdx::dev::ClassifyArgs a{};
    a.nq = static_cast<std::uint32_t>(c.nq);
    a.ks = c.d_ks;
    a.k = c.k;
    a.efs = c.d_efs;
    a.ef = c.ef;
    a.k_cap = k_cap;
    a.ncls = static_cast<std::uint32_t>(ncls);
    for (int i = 0; i < ncls; ++i) a.caps[i] = caps[i];
    a.order = buf.order;
    a.ranges = buf.ranges;
    a.queues = buf.queues;
    const bool bytes = h.u8 && h.metric == 0 && h.dim == 128 &&
                       (mode == vdx::dev::kBeamPlain || mode == vdx::dev::kBeamLearned) && !cfg.f32_rows;
    a.qflag = bytes ? buf.qflag : nullptr;
    hipLaunchKernelGGL(vdx::dev::beam_classify_kernel, dim3(1), dim3(1024), 0, s, a);
    check(hipGetLastError(), "beam_classify launch");
    if (bytes) {
        // the classify kernel zeroed the flag (a replayed graph repeats the same epoch, 1)
        check(vdx::dev::launch_query_bytes_check(c.d_queries, c.nq * h.dim, buf.qflag, 1u, nullptr, s),
              "query_bytes_check launch");
    }
}

// The launch of class `cls` (R capacity `cap`) on stream s; its work counter must be zero.
void enqueue_class_into(vdx_gpu_hnsw& h, int mode, const vdx_search_opts* opts, std::uint32_t cap,
                        std::uint32_t k_cap, int cls, const vdx::gpu::BeamConfig& cfg, const BatchCall& c,
                        const ClassBuffers& buf, hipStream_t s) {
    const vdx::gpu::BeamPlan plan = vdx::gpu::plan_beam(h, cap, c.nq, cfg, mode);
    vdx::dev::BeamArgs a = vdx::gpu::beam_args(
        h, plan, mode, opts, c.d_queries, c.nq, c.d_ks, c.k, c.d_efs, c.ef,
        reinterpret_cast<vdx::dev::Neighbor*>(c.d_out), c.out_stride, c.d_count, buf.queues + cls, nullptr);
    a.k_cap = std::min<std::uint32_t>(a.k_cap, k_cap);
    a.order = buf.order;
    a.range = buf.ranges + 2 * cls;
    if (plan.shape == vdx::dev::BeamShape::D128U8) {
        a.qflag = buf.qflag;
        a.qflag_epoch = 1u;
    }
    if (plan.bitmap_bytes > 0) {
        require(buf.bits != nullptr, "visited_form 2 (HBM bitmap) is not available on this path");
        require(buf.bits[cls].bytes() >= plan.bitmap_bytes, "batch visited bitmap not reserved");
        a.vbits = buf.bits[cls].as<std::uint32_t>();
        a.vbit_words = plan.bitmap_words;
    }
    vdx::gpu::launch_beam(h, plan, mode, a, s);
}

// Grows the visited bitmaps the classes' plans need (before anything is enqueued or captured).
void reserve_bitmaps(vdx_gpu_batch& b, std::uint64_t nq, const vdx::gpu::BeamConfig* cfgs) {
    for (int cls = 0; cls < b.ncls; ++cls) {
        const vdx::gpu::BeamPlan plan = vdx::gpu::plan_beam(*b.index, b.caps[cls], nq, cfgs[cls], b.mode);
        if (plan.bitmap_bytes > 0) {
            b.bits[cls].reserve(plan.bitmap_bytes);
        }
    }
}

ClassBuffers buffers_of(vdx_gpu_batch& b) {
    return {b.order.as<std::uint32_t>(), b.ranges.as<std::uint32_t>(), b.queues.as<std::uint32_t>(),
            b.queues.as<std::uint32_t>() + vdx::dev::kMaxClasses, b.bits};
}

void enqueue_classify(vdx_gpu_batch& b, const BatchCall& c, hipStream_t s) {
    enqueue_classify_into(*b.index, b.mode, vdx::gpu::config_from(b.opts_ptr()), c, b.max_k, b.ncls, b.caps,
                          buffers_of(b), s);
}

void enqueue_class(vdx_gpu_batch& b, int cls, const vdx::gpu::BeamConfig& cfg, const BatchCall& c,
                   hipStream_t s) {
    enqueue_class_into(*b.index, b.mode, b.opts_ptr(), b.caps[cls], b.max_k, cls, cfg, c, buffers_of(b), s);
}

// classify, then every class on its own stream, joined back to s.
void enqueue_all(vdx_gpu_batch& b, const BatchCall& c, hipStream_t s) {
    enqueue_classify(b, c, s);
    check(hipEventRecord(b.fork, s), "hipEventRecord(fork)");
    for (int cls = b.ncls - 1; cls >= 0; --cls) {
        check(hipStreamWaitEvent(b.side[cls], b.fork, 0), "hipStreamWaitEvent(fork)");
        enqueue_class(b, cls, b.cfg[cls], c, b.side[cls]);
        check(hipEventRecord(b.join[cls], b.side[cls]), "hipEventRecord(join)");
        check(hipStreamWaitEvent(s, b.join[cls], 0), "hipStreamWaitEvent(join)");
    }
}

}  // namespace

extern "C" {

int vdx_gpu_hnsw_batch_create(vdx_gpu_hnsw* index, uint64_t max_nq, uint32_t max_k, uint32_t max_ef,
                              const vdx_search_opts* opts, vdx_gpu_batch** out) {
    return guarded([&] {
        require(index != nullptr && out != nullptr, "vdx_gpu_hnsw_batch_create: null argument");
        require(max_nq >= 1 && max_nq <= 0xFFFFFFFFull, "vdx_gpu_hnsw_batch_create: max_nq outside [1, 2^32 - 1]");
        require(max_k >= 1, "vdx_gpu_hnsw_batch_create: max_k must be positive");
        require(index->n > 0, "vdx_gpu_hnsw_batch_create: empty index");
        const std::uint32_t top = std::max(max_ef, max_k);
        require(top <= 8192, "vdx_gpu_hnsw_batch_create: ef above the GPU path's limit");
        auto b = std::make_unique<vdx_gpu_batch>();
        b->mode = vdx::gpu::beam_mode(opts, false);
        if (opts != nullptr) {
            b->opts = *opts;
            b->opts.stats = nullptr;
            b->has_opts = true;
        }
        b->max_nq = max_nq;
        b->max_k = max_k;
        b->max_ef = max_ef;
        b->ncls = class_caps(max_k, max_ef, b->caps);
        const vdx::gpu::BeamConfig base = vdx::gpu::config_from(opts);
        for (int c = 0; c < b->ncls; ++c) {
            b->cfg[c] = base;
            b->info[c].ef_cap = b->caps[c];
        }
        DeviceGuard guard(index->ctx->device);
        b->order.resize(max_nq * sizeof(std::uint32_t));
        b->ranges.resize(2 * vdx::dev::kMaxClasses * sizeof(std::uint32_t));
        b->queues.resize((vdx::dev::kMaxClasses + 1) * sizeof(std::uint32_t));  // the counters, the byte-query flag
        check(hipEventCreateWithFlags(&b->fork, hipEventDisableTiming), "hipEventCreate");
        for (int c = 0; c < b->ncls; ++c) {
            check(hipStreamCreateWithFlags(&b->side[c], hipStreamNonBlocking), "hipStreamCreate");
            check(hipEventCreateWithFlags(&b->join[c], hipEventDisableTiming), "hipEventCreate");
        }
        b->index = index;
        b->device = index->ctx->device;
        *out = b.release();
    });
}

void vdx_gpu_hnsw_batch_free(vdx_gpu_batch* b) { delete b; }

int vdx_gpu_hnsw_batch_run(vdx_gpu_batch* b, const float* d_queries, uint64_t nq, const uint32_t* d_ks,
                           uint32_t k, const uint32_t* d_efs, uint32_t ef, vdx_neighbor* d_out,
                           uint64_t out_stride, uint32_t* d_out_count, int use_graph, void* stream) {
    return guarded([&] {
        require(b != nullptr, "vdx_gpu_hnsw_batch_run: batch is null");
        const BatchCall c{d_queries, nq, d_ks, k, d_efs, ef, d_out, out_stride, d_out_count};
        if (nq == 0) return;
        check_call(*b, c);
        DeviceGuard guard(b->index->ctx->device);
        hipStream_t s = stream ? static_cast<hipStream_t>(stream) : b->index->ctx->stream;
        {
            // a bitmap that grows changes its address: a captured sequence holding the old one goes
            void* before[vdx::dev::kMaxClasses];
            for (int cls = 0; cls < b->ncls; ++cls) before[cls] = b->bits[cls].as<void>();
            check(hipStreamSynchronize(s), "hipStreamSynchronize");
            reserve_bitmaps(*b, nq, b->cfg);
            for (int cls = 0; cls < b->ncls; ++cls) {
                if (before[cls] != b->bits[cls].as<void>() && b->exec != nullptr) {
                    check(hipGraphExecDestroy(b->exec), "hipGraphExecDestroy");
                    b->exec = nullptr;
                }
            }
        }
        if (!use_graph) {
            enqueue_all(*b, c, s);
            return;
        }
        const std::uint64_t key[11] = {reinterpret_cast<std::uint64_t>(d_queries), nq,
                                       reinterpret_cast<std::uint64_t>(d_ks), k,
                                       reinterpret_cast<std::uint64_t>(d_efs), ef,
                                       reinterpret_cast<std::uint64_t>(d_out), out_stride,
                                       reinterpret_cast<std::uint64_t>(d_out_count),
                                       reinterpret_cast<std::uint64_t>(s), b->index->generation};
        if (b->exec == nullptr || std::memcmp(key, b->key, sizeof(key)) != 0) {
            if (b->exec != nullptr) {
                check(hipStreamSynchronize(s), "hipStreamSynchronize");
                check(hipGraphExecDestroy(b->exec), "hipGraphExecDestroy");
                b->exec = nullptr;
            }
            hipGraph_t graph = nullptr;
            check(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
            try {
                enqueue_all(*b, c, s);
            } catch (...) {
                (void)hipStreamEndCapture(s, &graph);
                if (graph != nullptr) (void)hipGraphDestroy(graph);
                throw;
            }
            check(hipStreamEndCapture(s, &graph), "hipStreamEndCapture");
            const hipError_t e = hipGraphInstantiate(&b->exec, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            check(e, "hipGraphInstantiate");
            std::memcpy(b->key, key, sizeof(key));
        }
        check(hipGraphLaunch(b->exec, s), "hipGraphLaunch");
    });
}

int vdx_gpu_hnsw_batch_tune(vdx_gpu_batch* b, const float* d_queries, uint64_t nq, const uint32_t* d_ks,
                            uint32_t k, const uint32_t* d_efs, uint32_t ef, vdx_neighbor* d_out,
                            uint64_t out_stride, uint32_t* d_out_count, uint32_t reps) {
    return guarded([&] {
        require(b != nullptr, "vdx_gpu_hnsw_batch_tune: batch is null");
        const BatchCall c{d_queries, nq, d_ks, k, d_efs, ef, d_out, out_stride, d_out_count};
        require(nq >= 1, "vdx_gpu_hnsw_batch_tune: no sample queries");
        check_call(*b, c);
        reps = std::max<uint32_t>(reps, 1);
        DeviceGuard guard(b->index->ctx->device);
        vdx_gpu_ctx& ctx = *b->index->ctx;
        hipStream_t s = ctx.stream;
        enqueue_classify(*b, c, s);
        std::uint32_t ranges[2 * vdx::dev::kMaxClasses] = {};
        check(hipMemcpyAsync(ranges, b->ranges.as<void>(), sizeof(ranges), hipMemcpyDeviceToHost, s),
              "hipMemcpyAsync(ranges)");
        check(hipStreamSynchronize(s), "hipStreamSynchronize");
        hipEvent_t t0 = nullptr, t1 = nullptr;
        check(hipEventCreate(&t0), "hipEventCreate");
        check(hipEventCreate(&t1), "hipEventCreate");
        const auto time_class = [&](int cls, const vdx::gpu::BeamConfig& cfg) {
            vdx::gpu::BeamConfig cfgs[vdx::dev::kMaxClasses];
            for (int i = 0; i < b->ncls; ++i) cfgs[i] = b->cfg[i];
            cfgs[cls] = cfg;
            reserve_bitmaps(*b, c.nq, cfgs);
            float best = 0.f;
            for (std::uint32_t r = 0; r < reps; ++r) {
                check(vdx::dev::launch_fill_u32(b->queues.as<std::uint32_t>() + cls, 0u, 1, s), "fill");
                check(hipEventRecord(t0, s), "hipEventRecord");
                enqueue_class(*b, cls, cfg, c, s);
                check(hipEventRecord(t1, s), "hipEventRecord");
                check(hipEventSynchronize(t1), "hipEventSynchronize");
                float ms = 0.f;
                check(hipEventElapsedTime(&ms, t0, t1), "hipEventElapsedTime");
                best = r == 0 ? ms : std::min(best, ms);
            }
            return best;
        };
        try {
            for (int cls = 0; cls < b->ncls; ++cls) {
                const std::uint32_t count = ranges[2 * cls + 1] - ranges[2 * cls];
                b->info[cls].queries = count;
                if (count == 0) continue;
                const vdx::gpu::BeamConfig fixed = vdx::gpu::config_from(b->opts_ptr());
                std::vector<std::uint32_t> vl2s, wavess, bpcs = {0, 2};
                if (fixed.visited_log2) {
                    vl2s = {fixed.visited_log2};
                } else if (b->caps[cls] <= 128) {
                    vl2s = {10, 11, 12};
                } else {
                    vl2s = {9, 10, 11};
                }
                wavess = fixed.waves ? std::vector<std::uint32_t>{fixed.waves} : std::vector<std::uint32_t>{2, 4};
                // the visited form: the LDS cache at each size, and (the LDS kernel's classes,
                // R capacity above 512) the exact bitmap in HBM, which frees that LDS for waves
                std::vector<std::pair<std::uint32_t, std::uint32_t>> forms;  // (form, visited_log2)
                if (fixed.visited_form != 2) {
                    for (const std::uint32_t v : vl2s) forms.push_back({1u, v});
                }
                if (b->caps[cls] > vdx::dev::kRegMaxEf && fixed.visited_form != 1) {
                    forms.push_back({2u, 10u});
                }
                float best_ms = -1.f;
                for (const auto& [form, v] : forms) {
                    for (const std::uint32_t w : wavess) {
                        for (const std::uint32_t bpc : bpcs) {
                            vdx::gpu::BeamConfig cfg;
                            cfg.visited_log2 = v;
                            cfg.visited_form = form;
                            cfg.waves = w;
                            cfg.blocks_per_cu = bpc;
                            const float ms = time_class(cls, cfg);
                            if (best_ms < 0.f || ms < best_ms) {
                                best_ms = ms;
                                b->cfg[cls] = cfg;
                            }
                        }
                    }
                }
                b->info[cls].ms = best_ms;
            }
        } catch (...) {
            (void)hipEventDestroy(t0);
            (void)hipEventDestroy(t1);
            throw;
        }
        (void)hipEventDestroy(t0);
        (void)hipEventDestroy(t1);
        for (int cls = 0; cls < b->ncls; ++cls) {
            const vdx::gpu::BeamPlan plan = vdx::gpu::plan_beam(*b->index, b->caps[cls], 1, b->cfg[cls], b->mode);
            b->info[cls].visited_log2 = plan.visited_log2;
            b->info[cls].waves_per_group = plan.waves;
            b->info[cls].blocks_per_cu = b->cfg[cls].blocks_per_cu;
            b->info[cls].visited_form = plan.bitmap_bytes > 0 ? 2u : 1u;
        }
        // a captured sequence holds the old shapes
        if (b->exec != nullptr) {
            check(hipGraphExecDestroy(b->exec), "hipGraphExecDestroy");
            b->exec = nullptr;
        }
    });
}

int vdx_gpu_hnsw_batch_classes(const vdx_gpu_batch* b, vdx_batch_class* out, uint32_t capacity,
                               uint32_t* count) {
    return guarded([&] {
        require(b != nullptr && count != nullptr, "vdx_gpu_hnsw_batch_classes: null argument");
        *count = static_cast<std::uint32_t>(b->ncls);
        for (int c = 0; c < b->ncls && static_cast<std::uint32_t>(c) < capacity; ++c) {
            vdx_batch_class info = b->info[c];
            if (info.visited_log2 == 0) {
                const vdx::gpu::BeamPlan plan = vdx::gpu::plan_beam(*b->index, b->caps[c], 1, b->cfg[c], b->mode);
                info.visited_log2 = plan.visited_log2;
                info.waves_per_group = plan.waves;
                info.blocks_per_cu = b->cfg[c].blocks_per_cu;
                info.visited_form = plan.bitmap_bytes > 0 ? 2u : 1u;
            }
            out[c] = info;
        }
    });
}

int vdx_gpu_hnsw_search_uneven_noalloc(vdx_gpu_hnsw* index, const float* d_queries, uint64_t nq,
                                       const uint32_t* d_ks, uint32_t k, const uint32_t* d_efs, uint32_t ef,
                                       uint32_t max_k, uint32_t max_ef, vdx_neighbor* d_out, uint64_t out_stride,
                                       uint32_t* d_out_count, const vdx_search_opts* opts, void* stream) {
    return guarded([&] {
        require(index != nullptr, "vdx_gpu_hnsw_search_uneven_noalloc: index is null");
        if (nq == 0) return;
        if (nq > VDX_NOALLOC_MAX_QUERIES) {
            throw std::invalid_argument(vdx::fmt("vdx_gpu_hnsw_search_uneven_noalloc: %llu queries is above the "
                                                 "scratch limit of %u per call",
                                                 static_cast<unsigned long long>(nq), VDX_NOALLOC_MAX_QUERIES));
        }
        require(max_k >= 1, "vdx_gpu_hnsw_search_uneven_noalloc: max_k must be positive");
        require(std::max(max_ef, max_k) <= 8192, "vdx_gpu_hnsw_search_uneven_noalloc: ef above the GPU path's limit");
        require(index->n > 0, "vdx_gpu_hnsw_search_uneven_noalloc: empty index");
        require(opts == nullptr || opts->stats == nullptr,
                "vdx_gpu_hnsw_search_uneven_noalloc: opts.stats is not collected by this entry point");
        require(d_queries != nullptr && d_out != nullptr && d_out_count != nullptr,
                "vdx_gpu_hnsw_search_uneven_noalloc: null buffer");
        require(d_ks != nullptr || (k >= 1 && k <= max_k), "vdx_gpu_hnsw_search_uneven_noalloc: k outside [1, max_k]");
        require(d_efs != nullptr || ef <= max_ef, "vdx_gpu_hnsw_search_uneven_noalloc: ef above max_ef");
        require(out_stride >= max_k, "vdx_gpu_hnsw_search_uneven_noalloc: out_stride is below max_k");
        const int mode = vdx::gpu::beam_mode(opts, false);
        const vdx::gpu::BeamConfig cfg = vdx::gpu::config_from(opts);
        require(cfg.visited_form != 2,
                "vdx_gpu_hnsw_search_uneven_noalloc: visited_form 2 (HBM bitmap) needs device memory this entry "
                "point does not allocate");
        std::uint32_t caps[vdx::dev::kMaxClasses] = {};
        const int ncls = class_caps(max_k, max_ef, caps);
        DeviceGuard guard(index->ctx->device);
        std::uint32_t* sc = vdx::gpu::noalloc_scratch(index->ctx->device);
        const ClassBuffers buf{sc + kScratchOrder, sc + kScratchRanges, sc + kScratchQueues, sc + kScratchQflag};
        hipStream_t s = stream ? static_cast<hipStream_t>(stream) : index->ctx->stream;
        const BatchCall c{d_queries, nq, d_ks, k, d_efs, ef, d_out, out_stride, d_out_count};
        // the classify kernel zeroes the work counters itself, so nothing is cleared here
        enqueue_classify_into(*index, mode, cfg, c, max_k, ncls, caps, buf, s);
        for (int cls = 0; cls < ncls; ++cls) {
            enqueue_class_into(*index, mode, opts, caps[cls], max_k, cls, cfg, c, buf, s);
        }
    });
}

}  // extern "C"
