// Host side of the MFMA brute force: operand preparation (cached per index and precision),
// per-chunk launch sequence, certification read-back and the exact fallback.

#include "bf_host.hpp"

#include <algorithm>
#include <cmath>
#include <vector>

#include "../host/capi_common.hpp"
#include "bf_gemm_args.hpp"

namespace vdx::gpu {
namespace {

constexpr std::uint32_t kMaxMfmaDim = 128;   // query operands held in registers
constexpr std::uint32_t kMaxShortlist = 1024;  // re-rank sorts in LDS: 8 KB per wave
constexpr std::uint64_t kQueryChunk = 16384;
constexpr int kGemmWaves = 8;  // wavefronts (32 queries each) sharing one staged base tile

// Operand row length: a multiple of the MFMA K step. BF16 rows of 97 to 128 elements are padded
// to 128 (at most a third more MFMA work), which selects the LDS-fed kernel.
std::uint32_t operand_kp(int prec, std::uint64_t dim) {
    if (prec == dev::kPrecBF16 && dim > 96 && dim <= 128) return 128;
    return prec == dev::kPrecF32 ? static_cast<std::uint32_t>((dim + 3) / 4 * 4)
                                 : static_cast<std::uint32_t>((dim + 15) / 16 * 16);
}

std::uint32_t operand_bytes(int prec) { return prec == dev::kPrecF32 ? 4 : (prec == dev::kPrecBF16 ? 2 : 1); }

std::uint32_t operand_steps(int prec, std::uint32_t kp) {
    return prec == dev::kPrecF32 ? kp / 4 : kp / 16;  // float2 pairs of K=2 steps, or K=16 steps
}

int rerank_shape(std::uint64_t dim) { return dim == 128 ? 0 : (dim == 100 ? 1 : 2); }

void prepare(const float* src, std::uint32_t ld_src, std::uint64_t rows, std::uint64_t dim,
             std::uint32_t kp, int prec, void* dst, float* n2, float* sc, float* qe, float* qn,
             std::uint32_t* integral, hipStream_t s, float* nhalf = nullptr) {
    if (rows == 0) return;
    dev::PrepArgs p{};
    p.src = src;
    p.ld_src = ld_src;
    p.rows = static_cast<std::uint32_t>(rows);
    p.dim = static_cast<std::uint32_t>(dim);
    p.kp = kp;
    p.prec = prec;
    p.dst = dst;
    p.norm2 = n2;
    p.nhalf = nhalf;
    p.scale = sc;
    p.qerr = qe;
    p.qnorm = qn;
    p.integral = integral;
    check(dev::launch_bf_prepare(p, s), "bf_prepare launch");
}

BfOperands& base_operands(vdx_gpu_bf& b, int prec, hipStream_t s) {
    const std::lock_guard<std::mutex> lock(b.prep_mutex);
    BfOperands& op = b.prep[prec];
    if (op.ready) return op;
    op.kp = operand_kp(prec, b.dim);
    op.rows = (b.n + 63) / 64 * 64;
    op.max_norm2 = op.max_qerr = op.max_qnorm = op.max_norm = 0.f;  // rows appended: made again
    op.xq.resize(op.rows * op.kp * operand_bytes(prec));
    op.norm2.resize(op.rows * 4);
    op.nhalf.resize(op.rows * 4);
    op.scale.resize(op.rows * 4);
    op.qerr.resize(op.rows * 4);
    op.qnorm.resize(op.rows * 4);
    op.integral.resize(op.rows * 4);
    check(hipMemsetAsync(op.xq.as<void>(), 0, op.xq.bytes(), s), "hipMemsetAsync");
    for (DeviceBuffer* d : {&op.norm2, &op.nhalf, &op.scale, &op.qerr, &op.qnorm, &op.integral}) {
        check(hipMemsetAsync(d->as<void>(), 0, d->bytes(), s), "hipMemsetAsync");
    }
    prepare(b.vectors.as<float>(), b.split.ld, b.n, b.dim, op.kp, prec, op.xq.as<void>(),
            op.norm2.as<float>(), op.scale.as<float>(), op.qerr.as<float>(), op.qnorm.as<float>(),
            op.integral.as<std::uint32_t>(), s, op.nhalf.as<float>());
    std::vector<float> n2(b.n), qe(b.n), qn(b.n);
    std::vector<std::uint32_t> integral(b.n);
    check(hipMemcpyAsync(integral.data(), op.integral.as<void>(), b.n * 4, hipMemcpyDeviceToHost, s), "copy");
    check(hipMemcpyAsync(n2.data(), op.norm2.as<void>(), b.n * 4, hipMemcpyDeviceToHost, s), "copy");
    check(hipMemcpyAsync(qe.data(), op.qerr.as<void>(), b.n * 4, hipMemcpyDeviceToHost, s), "copy");
    check(hipMemcpyAsync(qn.data(), op.qnorm.as<void>(), b.n * 4, hipMemcpyDeviceToHost, s), "copy");
    check(hipStreamSynchronize(s), "bf prepare");
    op.all_integral = true;
    for (std::uint64_t i = 0; i < b.n; ++i) {
        op.all_integral = op.all_integral && integral[i] != 0;
        op.max_norm2 = std::max(op.max_norm2, n2[i]);
        op.max_qerr = std::max(op.max_qerr, qe[i]);
        op.max_qnorm = std::max(op.max_qnorm, qn[i]);
    }
    // inflate the FP32-computed maxima slightly: they bound exact quantities from above
    op.max_norm2 *= 1.0f + 1e-5f;
    op.max_qerr *= 1.0f + 1e-5f;
    op.max_qnorm *= 1.0f + 1e-5f;
    op.max_norm = std::sqrt(op.max_norm2) * (1.0f + 1e-6f);
    op.ready = true;
    return op;
}

unsigned slabs_for(std::uint64_t qtiles, std::uint64_t row_tiles) {
    const std::uint64_t want = std::max<std::uint64_t>(1, (4096 + qtiles - 1) / qtiles);
    return static_cast<unsigned>(std::min<std::uint64_t>(want, std::max<std::uint64_t>(1, row_tiles)));
}

std::uint32_t next_pow2(std::uint32_t v) {
    std::uint32_t p = 1;
    while (p < v) p <<= 1;
    return p;
}

}  // namespace

void bf_mfma(vdx_gpu_bf& b, int prec, const float* d_queries, std::uint64_t nq, std::uint32_t k,
             vdx::dev::Neighbor* d_out, std::uint64_t out_stride, std::uint32_t* d_count,
             Workspace& ws, hipStream_t s, vdx_bf_stats* stats, std::uint32_t widen) {
    const std::uint32_t want = static_cast<std::uint32_t>(std::min<std::uint64_t>(k, b.n));
    const std::uint32_t mult = (prec == dev::kPrecF32 ? 2 : (prec == dev::kPrecBF16 ? 4 : 8)) * widen;
    std::uint64_t S = std::min<std::uint64_t>(std::max<std::uint64_t>(std::uint64_t{mult} * k, want), b.n);
    if (S > kMaxShortlist) S = std::max<std::uint64_t>(kMaxShortlist, want);
    if (b.dim > kMaxMfmaDim || want > kMaxShortlist) {
        // Outside the MFMA kernels' shapes: the exact scan answers (same results).
        bf_exact(b, d_queries, nq, k, d_out, out_stride, d_count, ws, s);
        if (stats) stats->exact_fallback += nq;
        return;
    }
    BfOperands& op = base_operands(b, prec, s);
    const std::uint32_t kp = op.kp;
    const std::uint32_t ks = operand_steps(prec, kp);
    const std::uint32_t sort_cap = next_pow2(static_cast<std::uint32_t>(S));
    const std::uint32_t rerank_R = rerank_shape(b.dim) == 2 ? 1 : 4;
    const std::size_t rerank_lds = static_cast<std::size_t>(sort_cap) * 8 + 8u * rerank_R * 8;
    const double u = std::ldexp(1.0, -24);
    const float gamma_k = static_cast<float>((kp + 3) * u / (1.0 - (kp + 3) * u));
    const float gamma_d = static_cast<float>((b.dim + 2) * u / (1.0 - (b.dim + 2) * u));

    // sample and candidate-buffer geometry
    const bool sample = b.n > 4096;
    std::uint32_t m = 0, stride = 1, rank = 0, cap = static_cast<std::uint32_t>(b.n);
    if (sample) {
        m = static_cast<std::uint32_t>(std::min<std::uint64_t>(8192, std::max<std::uint64_t>(512, b.n / 8)));
        stride = static_cast<std::uint32_t>(b.n / m);
        // The fraction of the base within tau is about Gamma(rank) / m. The shortlist and its
        // certification are the same for any tau that keeps more than S rows; with fewer, tau
        // itself bounds what was left out, which certifies unless barely k rows survive. rank
        // >= 4 and two above the rank of 2S survivors make fewer than k survivors about a 2e-6
        // event per query (the fraction within tau ~ Beta(rank, m); P ~ (m k / n)^rank / rank!),
        // while every survivor costs filter and shortlist work (the filter's staging of
        // survivors is about a quarter of its time at rank 6); the buffer holds 6x the expected
        // survivors, so overflows (redone exactly) are rarer still.
        rank = static_cast<std::uint32_t>(std::min<double>(m, std::max(4.0, std::ceil(2.0 * S * m / b.n) + 2.0)));
        const double expect = static_cast<double>(rank) * b.n / m;
        cap = static_cast<std::uint32_t>(std::clamp<double>(std::ceil(6.0 * expect / 256) * 256, 1024, 16384));
        cap = static_cast<std::uint32_t>(std::min<std::uint64_t>(cap, op.rows));
    }

    std::uint64_t fallbacks = 0;
    float max_err = 0.f;
    for (std::uint64_t c0 = 0; c0 < nq; c0 += kQueryChunk) {
        const std::uint64_t cq = std::min<std::uint64_t>(kQueryChunk, nq - c0);
        const std::uint64_t cq_pad = (cq + 255) / 256 * 256;
        const float* cq_src = d_queries + c0 * b.dim;
        DeviceBuffer& qq = ws.extra[0];
        DeviceBuffer& qn2 = ws.extra[1];
        DeviceBuffer& qsc = ws.extra[2];
        DeviceBuffer& qqe = ws.extra[3];
        DeviceBuffer& qqn = ws.extra[4];
        DeviceBuffer& tau = ws.extra[5];
        DeviceBuffer& dense = ws.extra[6];
        DeviceBuffer& buf = ws.extra[7];
        DeviceBuffer& cnt = ws.extra[8];
        DeviceBuffer& list = ws.extra[9];
        DeviceBuffer& approx = ws.extra[10];
        DeviceBuffer& list_count = ws.extra[11];
        DeviceBuffer& left = ws.extra[12];
        DeviceBuffer& strict = ws.extra[13];
        DeviceBuffer& fail = ws.extra[14];
        DeviceBuffer& errbuf = ws.extra[15];
        DeviceBuffer& qint = ws.aux2;
        qq.reserve(cq_pad * kp * operand_bytes(prec));
        for (DeviceBuffer* d : {&qn2, &qsc, &qqe, &qqn, &tau, &cnt, &list_count, &left, &strict, &fail, &qint}) {
            d->reserve(cq_pad * 4);
        }
        errbuf.reserve(256);
        check(hipMemsetAsync(qq.as<void>(), 0, cq_pad * kp * operand_bytes(prec), s), "hipMemsetAsync");
        check(hipMemsetAsync(errbuf.as<void>(), 0, 4, s), "hipMemsetAsync");
        prepare(cq_src, static_cast<std::uint32_t>(b.dim), cq, b.dim, kp, prec, qq.as<void>(),
                qn2.as<float>(), qsc.as<float>(), qqe.as<float>(), qqn.as<float>(),
                qint.as<std::uint32_t>(), s);

        dev::GemmArgs g{};
        g.xq = op.xq.as<void>();
        g.xn = op.norm2.as<float>();
        g.xn_max = op.max_norm2;
        g.xh = op.nhalf.as<float>();
        g.xs = prec == dev::kPrecFP8 ? op.scale.as<float>() : nullptr;
        g.qq = qq.as<void>();
        g.qn = qn2.as<float>();
        g.qs = prec == dev::kPrecFP8 ? qsc.as<float>() : nullptr;
        g.nq = static_cast<std::uint32_t>(cq);
        g.kp = kp;
        g.ks = ks;
        g.l2 = b.metric == 0 ? 1u : 0u;
        const unsigned qtiles = static_cast<unsigned>(cq_pad / (32 * kGemmWaves));
        if (sample) {
            dense.reserve(cq * static_cast<std::uint64_t>(m) * 4);
            g.rows = m;
            g.row_stride = stride;
            g.mode = 0;
            g.dense = dense.as<float>();
            g.dense_ld = m;
            check(dev::launch_bf_gemm(g, qtiles, slabs_for(qtiles, (m + 63) / 64), prec, kGemmWaves, s, b.gemm_shape.load()),
                  "bf_gemm(sample) launch");
            dev::ThresholdArgs t{};
            t.dense = dense.as<float>();
            t.dense_ld = m;
            t.count = m;
            t.rank = rank;
            t.tau = tau.as<float>();
            check(dev::launch_bf_threshold(t, static_cast<unsigned>(cq), s), "bf_threshold launch");
        }
        buf.reserve(cq * static_cast<std::uint64_t>(cap) * 8);
        check(hipMemsetAsync(cnt.as<void>(), 0, cq * 4, s), "hipMemsetAsync");
        g.rows = static_cast<std::uint32_t>(b.n);
        g.row_stride = 1;
        g.mode = 1;
        g.tau = sample ? tau.as<float>() : nullptr;
        g.buf = buf.as<std::uint64_t>();
        g.cap = cap;
        g.cnt = cnt.as<std::uint32_t>();
        check(dev::launch_bf_gemm(g, qtiles, slabs_for(qtiles, (b.n + 63) / 64), prec, kGemmWaves, s, b.gemm_shape.load()),
              "bf_gemm(filter) launch");

        list.reserve(cq * S * 4);
        approx.reserve(cq * S * 4);
        dev::ShortlistArgs sl{};
        sl.buf = buf.as<std::uint64_t>();
        sl.cnt = cnt.as<std::uint32_t>();
        sl.cap = cap;
        sl.S = static_cast<std::uint32_t>(S);
        sl.tau = sample ? tau.as<float>() : nullptr;
        sl.list = list.as<std::uint32_t>();
        sl.approx = approx.as<float>();
        sl.list_count = list_count.as<std::uint32_t>();
        sl.left = left.as<float>();
        sl.strict = strict.as<std::uint32_t>();
        check(dev::launch_bf_shortlist(sl, static_cast<unsigned>(cq), s), "bf_shortlist launch");

        dev::RerankArgs r{};
        r.vectors = b.vectors.as<float>();
        r.ld = b.split.ld;
        r.l2 = b.metric == 0 ? 1u : 0u;
        r.n = static_cast<std::uint32_t>(b.n);
        r.nq = static_cast<std::uint32_t>(cq);
        r.k = k;
        r.nb = b.split.nb;
        r.n8 = b.split.n8;
        r.tail = b.split.tail;
        r.first_tail = b.split.first_tail;
        r.queries = cq_src;
        r.qld = static_cast<std::uint32_t>(b.dim);
        r.S = static_cast<std::uint32_t>(S);
        r.sort_cap = sort_cap;
        r.lds_per_wave = static_cast<std::uint32_t>(rerank_lds);
        r.list = list.as<std::uint32_t>();
        r.approx = approx.as<float>();
        r.list_count = list_count.as<std::uint32_t>();
        r.left = left.as<float>();
        r.strict = strict.as<std::uint32_t>();
        r.q_norm2 = qn2.as<float>();
        r.q_qerr = qqe.as<float>();
        r.q_qnorm = qqn.as<float>();
        r.q_integral = qint.as<std::uint32_t>();
        r.x_integral = op.all_integral ? 1u : 0u;
        r.x_max_qerr = op.max_qerr;
        r.x_max_norm = op.max_norm;
        r.x_max_norm2 = op.max_norm2;
        r.x_max_qnorm = op.max_qnorm;
        r.gamma_k = gamma_k;
        r.gamma_d = gamma_d;
        r.out = d_out + c0 * out_stride;
        r.out_stride = out_stride;
        r.out_count = d_count + c0;
        r.fail = fail.as<std::uint32_t>();
        r.max_err = errbuf.as<float>();
        check(dev::launch_bf_rerank(r, static_cast<unsigned>(cq), rerank_lds, s, rerank_shape(b.dim)),
              "bf_rerank launch");

        std::vector<std::uint32_t> failed(cq);
        float chunk_err = 0.f;
        check(hipMemcpyAsync(failed.data(), fail.as<void>(), cq * 4, hipMemcpyDeviceToHost, s), "copy");
        check(hipMemcpyAsync(&chunk_err, errbuf.as<void>(), 4, hipMemcpyDeviceToHost, s), "copy");
        check(hipStreamSynchronize(s), "bf mfma");
        max_err = std::max(max_err, chunk_err);
        std::vector<std::uint32_t> redo;
        for (std::uint32_t i = 0; i < cq; ++i) {
            if (failed[i]) redo.push_back(i);
        }
        if (!redo.empty()) {
            const std::uint64_t nr = redo.size();
            // own buffers: the wider pass below reuses the workspace's
            DeviceBuffer idx(nr * 4), sub_q(nr * b.dim * 4), sub_out(nr * want * sizeof(dev::Neighbor)),
                sub_cnt(nr * 4);
            check(hipMemcpyAsync(idx.as<void>(), redo.data(), nr * 4, hipMemcpyHostToDevice, s), "copy");
            check(dev::launch_bf_gather_rows(cq_src, idx.as<std::uint32_t>(), static_cast<std::uint32_t>(nr),
                                             static_cast<std::uint32_t>(b.dim), sub_q.as<float>(), s),
                  "gather launch");
            if (widen == 1 && S < std::min<std::uint64_t>(kMaxShortlist, b.n)) {
                // a few uncertified queries: an 8x wider shortlist usually certifies them, at a
                // fraction of an exact scan (which a small batch leaves mostly idle)
                vdx_bf_stats inner{};
                bf_mfma(b, prec, sub_q.as<float>(), nr, k, sub_out.as<dev::Neighbor>(), want,
                        sub_cnt.as<std::uint32_t>(), ws, s, &inner, 8);
                if (stats) {
                    stats->widened += nr;
                    stats->exact_fallback += inner.exact_fallback;
                    stats->max_abs_error = std::max(stats->max_abs_error, inner.max_abs_error);
                }
            } else {
                bf_exact(b, sub_q.as<float>(), nr, k, sub_out.as<dev::Neighbor>(), want,
                         sub_cnt.as<std::uint32_t>(), ws, s);
                fallbacks += nr;
            }
            check(dev::launch_bf_scatter_results(sub_out.as<dev::Neighbor>(), sub_cnt.as<std::uint32_t>(),
                                                 idx.as<std::uint32_t>(), static_cast<std::uint32_t>(nr),
                                                 want, d_out + c0 * out_stride, out_stride, d_count + c0, s),
                  "scatter launch");
            check(hipStreamSynchronize(s), "bf fallback");
        }
    }
    if (stats) {
        stats->shortlist = S;
        stats->exact_fallback += fallbacks;
        stats->max_abs_error = std::max<double>(stats->max_abs_error, max_err);
    }
}

}  // namespace vdx::gpu
