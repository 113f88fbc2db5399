// Host-side planning and launch of the HNSW beam kernels (beam_host.hpp).

#include "beam_host.hpp"

#include <algorithm>
#include <cmath>
#include <stdexcept>

#include "../host/capi_common.hpp"
#include "gpu_state.hpp"

namespace vdx::gpu {

namespace {

// A larger visited cache saves recomputed distances but costs LDS, hence waves per CU. Measured
// (1M x 128 L2, 1.2M x 100 cosine; results identical throughout): 2^11 slots up to R capacity
// 256 over the FP32 rows; 2^10 over the 8-bit rows (a recomputed row costs a quarter of the
// bytes) and from capacity 512 up, where R itself takes LDS: cosine ef 512 / 640 / 1024 ran
// 10.8 / 40.9 / 91.1 ms per 10k queries with 2^11 / 2^13 / 2^13 slots and 10.0 / 18.0 / 30.5 ms
// with 2^10 (2^13 slots had left one wave per SIMD above capacity 512).
std::uint32_t default_visited_log2(std::uint32_t ef_cap, bool u8_rows) {
    return (u8_rows || ef_cap >= 512) ? 10 : 11;
}

}  // namespace

vdx::dev::BeamShape beam_shape(std::uint64_t dim) {
    if (dim == 128) return vdx::dev::BeamShape::D128;
    if (dim == 100) return vdx::dev::BeamShape::D100;
    return vdx::dev::BeamShape::Generic;
}

int beam_mode(const vdx_search_opts* opts, bool trace) {
    if (trace) {
        return vdx::dev::kBeamTrace;
    }
    if (opts != nullptr && opts->termination == 1) {
        vdx::require(opts->mlp != nullptr, "learned termination needs a model (opts.mlp)");
        vdx::require(opts->mlp_threshold > 0.f && opts->mlp_threshold < 1.f,
                     "learned termination needs a threshold in (0, 1)");
        return vdx::dev::kBeamLearned;
    }
    if (opts != nullptr && opts->termination > 1) {
        throw std::invalid_argument("unknown termination mode");
    }
    return vdx::dev::kBeamPlain;
}

BeamConfig config_from(const vdx_search_opts* opts) {
    BeamConfig c;
    if (opts != nullptr) {
        c.visited_log2 = opts->visited_log2;
        c.waves = opts->waves_per_group;
        vdx::require(opts->row_format <= 1, "row_format must be 0 (8-bit rows where kept) or 1 (FP32 rows)");
        c.f32_rows = opts->row_format == 1;
        vdx::require(opts->visited_form <= 2, "visited_form must be 0 (auto), 1 (LDS cache) or 2 (HBM bitmap)");
        c.visited_form = opts->visited_form;
    }
    return c;
}

BeamPlan plan_beam(const vdx_gpu_hnsw& h, std::uint32_t max_efp, std::uint64_t nq, const BeamConfig& cfg,
                   int mode) {
    BeamPlan p{};
    const std::uint32_t need = std::max<std::uint32_t>(1, (max_efp + 63u) / 64u);
    if (need * 64u <= vdx::dev::kRegMaxEf) {
        p.chunks = 1;
        while (static_cast<std::uint32_t>(p.chunks) < need) {
            p.chunks *= 2;
        }
        p.ef_cap = 64u * static_cast<std::uint32_t>(p.chunks);
    } else {
        p.chunks = 0;
        p.ef_cap = need * 64u;
    }
    p.shape = beam_shape(h.dim);
    if (p.shape == vdx::dev::BeamShape::D128 && h.u8 && h.metric == 0 && p.chunks > 0 &&
        (mode == vdx::dev::kBeamPlain || mode == vdx::dev::kBeamLearned) && !cfg.f32_rows) {
        p.shape = vdx::dev::BeamShape::D128U8;
    }
    p.visited_log2 = cfg.visited_log2 ? cfg.visited_log2
                                      : default_visited_log2(p.ef_cap, p.shape == vdx::dev::BeamShape::D128U8);
    vdx::require(p.visited_log2 >= 6 && p.visited_log2 <= 15, "visited_log2 must be within [6, 15]");
    p.waves = cfg.waves ? cfg.waves : 4;
    vdx::require(p.waves >= 1 && p.waves <= 4, "waves_per_group must be within [1, 4]");
    const bool bitmap = p.chunks == 0 && cfg.visited_form == 2;
    p.lds_per_wave = p.chunks > 0 ? vdx::dev::beam_reg_lds_per_wave(p.ef_cap, p.visited_log2, mode != vdx::dev::kBeamPlain)
                                  : vdx::dev::beam_lds_per_wave(p.ef_cap, p.visited_log2, bitmap);
    const std::size_t limit = h.ctx->lds_per_block;
    const std::size_t shared = vdx::dev::beam_lds_shared(mode);
    while (p.waves > 1 && p.lds_per_wave * p.waves + shared > limit) {
        --p.waves;
    }
    p.lds_bytes = p.lds_per_wave * p.waves + shared;
    vdx::require(p.lds_bytes <= limit, "ef too large for the GPU beam's LDS");
    const void* kernel = p.chunks > 0 ? vdx::dev::hnsw_beam_reg_kernel_ptr(p.shape, mode, p.chunks)
                                      : vdx::dev::hnsw_beam_kernel_ptr(p.shape, mode);
    int per_cu = 0;
    check(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, p.waves * 64, p.lds_bytes),
          "hipOccupancyMaxActiveBlocksPerMultiprocessor");
    per_cu = std::max(per_cu, 1);
    if (cfg.blocks_per_cu != 0) {
        per_cu = std::min<int>(per_cu, static_cast<int>(cfg.blocks_per_cu));
    }
    const std::uint64_t resident = static_cast<std::uint64_t>(per_cu) * h.ctx->cu_count;
    const std::uint64_t needed = (nq + p.waves - 1) / p.waves;
    p.grid = static_cast<std::uint32_t>(std::max<std::uint64_t>(1, std::min(resident, needed)));
    p.bitmap_words = 0;
    p.bitmap_bytes = 0;
    if (bitmap) {
        // rows of whole 16-byte stores (the clear), one per resident wave
        p.bitmap_words = static_cast<std::uint32_t>((h.n + 127) / 128 * 4);
        p.bitmap_bytes = static_cast<std::uint64_t>(p.grid) * p.waves * p.bitmap_words * 4;
    }
    p.grid_int = 0;
    if (p.shape == vdx::dev::BeamShape::D128U8) {
        int per_cu_int = 0;
        check(hipOccupancyMaxActiveBlocksPerMultiprocessor(
                  &per_cu_int,
                  vdx::dev::hnsw_beam_reg_kernel_ptr(
                      p.shape, mode == vdx::dev::kBeamLearned ? vdx::dev::kBeamU8IntLearned : vdx::dev::kBeamU8Int,
                      p.chunks),
                  p.waves * 64, p.lds_bytes),
              "hipOccupancyMaxActiveBlocksPerMultiprocessor");
        per_cu_int = std::max(per_cu_int, 1);
        if (cfg.blocks_per_cu != 0) {
            per_cu_int = std::min<int>(per_cu_int, static_cast<int>(cfg.blocks_per_cu));
        }
        const std::uint64_t resident_int = static_cast<std::uint64_t>(per_cu_int) * h.ctx->cu_count;
        p.grid_int = static_cast<std::uint32_t>(std::max<std::uint64_t>(1, std::min(resident_int, needed)));
    }
    return p;
}

vdx::dev::BeamArgs beam_args(const vdx_gpu_hnsw& h, const BeamPlan& plan, int mode,
                             const vdx_search_opts* opts, const float* d_queries, std::uint64_t nq,
                             const std::uint32_t* d_ks, std::uint32_t k, const std::uint32_t* d_efs,
                             std::uint32_t ef, vdx::dev::Neighbor* d_out, std::uint64_t out_stride,
                             std::uint32_t* d_count, std::uint32_t* d_queue, unsigned long long* d_stats) {
    vdx::require(nq <= 0xFFFFFFFFull, "too many queries in one call");
    vdx::dev::BeamArgs a{};
    a.vectors = h.vectors.as<float>();
    a.rows_u8 = plan.shape == vdx::dev::BeamShape::D128U8 ? h.rows_u8.as<unsigned char>() : nullptr;
    a.layer0 = h.layer0.as<std::uint32_t>();
    a.upper = h.upper.as<std::uint32_t>();
    a.upper_offset = h.upper_offset.as<std::uint32_t>();
    a.ld = h.split.ld;
    a.dim = static_cast<std::uint32_t>(h.dim);
    a.b0 = static_cast<std::uint32_t>(1 + 2 * h.M);
    a.bu = static_cast<std::uint32_t>(1 + h.M);
    a.n = static_cast<std::uint32_t>(h.n);
    a.max_level = static_cast<std::uint32_t>(h.max_level);
    a.entry_point = static_cast<std::uint32_t>(h.entry_point);
    a.l2 = h.metric == 0 ? 1u : 0u;
    a.nb = h.split.nb;
    a.n8 = h.split.n8;
    a.tail = h.split.tail;
    a.first_tail = h.split.first_tail;
    a.queries = d_queries;
    a.qld = static_cast<std::uint32_t>(h.dim);
    a.nq = static_cast<std::uint32_t>(nq);
    a.ks = d_ks;
    a.efs = d_efs;
    a.k = k;
    a.ef = ef;
    a.out = d_out;
    a.out_stride = out_stride;
    a.out_count = d_count;
    a.k_cap = static_cast<std::uint32_t>(std::min<std::uint64_t>(out_stride, plan.ef_cap));
    a.queue = d_queue;
    a.ef_cap = plan.ef_cap;
    a.visited_log2 = plan.visited_log2;
    a.lds_per_wave = static_cast<std::uint32_t>(plan.lds_per_wave);
    a.stats = d_stats;
    a.M2 = static_cast<std::uint32_t>(2 * h.M);
    a.max_exp = opts != nullptr && opts->max_expansions != 0 ? opts->max_expansions : 0xFFFFFFFFu;
    a.max_exps = opts != nullptr ? opts->max_expansions_per_query : nullptr;
    if (mode == vdx::dev::kBeamLearned) {
        const auto* model = static_cast<const vdx_gpu_mlp*>(opts->mlp);
        a.mlp = model->weights.as<void>();
        const float t = opts->mlp_threshold;
        a.stop_logit = std::log(t / (1.0f - t));
        a.min_steps = opts->min_steps;
        a.mlp_every = std::max<std::uint32_t>(1, opts->mlp_every);
    }
    return a;
}

void launch_beam(const vdx_gpu_hnsw& h, const BeamPlan& plan, int mode, const vdx::dev::BeamArgs& a,
                 hipStream_t stream, bool int_only) {
    if (plan.chunks > 0) {
        if (plan.shape == vdx::dev::BeamShape::D128U8 && a.qflag != nullptr) {
            // the integer kernel and the float kernel, in turn: the flag lets exactly one work
            check(vdx::dev::launch_hnsw_beam_reg(
                      a, plan.grid_int, plan.waves, plan.lds_bytes, stream, plan.shape,
                      mode == vdx::dev::kBeamLearned ? vdx::dev::kBeamU8IntLearned : vdx::dev::kBeamU8Int, plan.chunks),
                  "hnsw_beam_reg (integer) launch");
            if (int_only) {
                return;
            }
        }
        check(vdx::dev::launch_hnsw_beam_reg(a, plan.grid, plan.waves, plan.lds_bytes, stream,
                                             plan.shape, mode, plan.chunks),
              "hnsw_beam_reg launch");
    } else {
        check(vdx::dev::launch_hnsw_beam(a, plan.grid, plan.waves, plan.lds_bytes, stream,
                                         plan.shape, mode),
              "hnsw_beam launch");
    }
}

}  // namespace vdx::gpu
