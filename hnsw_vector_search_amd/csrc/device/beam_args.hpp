#pragma once

// Arguments of the HNSW beam kernel (hnsw_beam.hip), shared with its host-side launcher.

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace vdx::dev {

// Layout-identical to vecdex::Neighbor and vdx_neighbor.
struct Neighbor {
    float distance;
    std::uint32_t id;
};

struct BeamArgs {
    // the graph, in HBM
    const float* vectors;          // n rows of ld floats (dim used, the rest zero)
    const std::uint32_t* layer0;   // n blocks of b0 = 1 + 2M
    const std::uint32_t* upper;    // U blocks of bu = 1 + M, FORMAT order
    const std::uint32_t* upper_offset;  // n: first upper block of each node
    std::uint32_t ld, dim, b0, bu;
    std::uint32_t n, max_level, entry_point, l2;  // l2: 1 for L2, 0 for 1 - <a, b>
    // dimension split in the distance order: nb 32-blocks, n8 8-blocks, tail elements
    std::uint32_t nb, n8, tail, first_tail;
    // D128U8: the rows as bytes, 128 per row in the beam's lane order (pack_rows_u8)
    const unsigned char* rows_u8;
    // D128U8: a device word that launch_query_bytes_check sets to qflag_epoch when some query of
    // the launch has a component that is not an integer in [0, 255] (any other value: every
    // query is byte-valued, the integer kernel serves the launch); null: the float arithmetic
    const std::uint32_t* qflag;
    std::uint32_t qflag_epoch;
    // the queries (prepared), per-query k and ef
    const float* queries;
    // the integer kernel's queries as the host staged them when every one is byte-valued: 128
    // bytes each in element order (queries is then not read), or null
    const std::uint8_t* qbytes;
    std::uint32_t qld, nq;
    const std::uint32_t* ks;
    const std::uint32_t* efs;
    std::uint32_t k, ef;
    // results
    Neighbor* out;
    std::uint64_t out_stride;
    std::uint32_t* out_count;
    std::uint32_t k_cap;  // per-query k is clamped to this (<= out_stride, <= ef_cap)
    // launch state
    std::uint32_t* queue;  // next query to take; zero before the launch
    // optional per-query entry: entries[q] < n starts the beam there with no descent (an
    // insertion search, which starts from the layer above's result); other values descend
    const std::uint32_t* entries;
    // the layer the beam searches (0 for queries; an insertion searches each of its layers):
    // the descent stops above it and link blocks are the layer's (upper blocks for layer > 0)
    std::uint32_t layer;
    // optional work list: the launch serves queries order[range[0] .. range[1]) (device
    // values, written by an earlier launch on the stream); null: queries 0 .. nq
    const std::uint32_t* order;
    const std::uint32_t* range;
    std::uint32_t ef_cap;  // R capacity per wave (>= every max(ef, k)), a multiple of 64
    std::uint32_t visited_log2;
    std::uint32_t lds_per_wave;
    unsigned long long* stats;  // kBeamStats counters, or null
    // trace mode: per step features and top-10 hits against `truth` (10 ids per query)
    const std::uint32_t* truth;
    float* trace_feat;           // [nq][trace_cap][16]
    std::uint8_t* trace_hits;    // [nq][trace_cap]
    std::uint32_t* trace_steps;  // [nq] expansions taken
    std::uint32_t* trace_final;  // [nq] top-10 hits of the result
    std::uint32_t trace_cap;
    std::uint32_t M2;            // 2M, for feature scaling
    // learned mode: the stop model (device copy, staged in LDS) and its decision threshold
    const void* mlp;
    float stop_logit;
    std::uint32_t min_steps;
    std::uint32_t mlp_every;  // consult the model every this many expansions (>= 1)
    std::uint32_t max_exp;  // expansion budget: layer-0 link blocks per query at most (0xFFFFFFFF: none)
    const std::uint32_t* max_exps;  // per-query budgets replacing max_exp (0: none), or null
    // the visited set as an exact bitmap in HBM instead of the LDS cache (the LDS kernel): one
    // row of vbit_words words per resident wave, cleared by the wave at the start of each query
    std::uint32_t* vbits;
    std::uint32_t vbit_words;
};

constexpr int kBeamStats = 7;

enum BeamMode : int { kBeamPlain = 0, kBeamTrace = 1, kBeamLearned = 2,
                      // kernel selectors only: the plain and learned searches' integer kernels
                      kBeamU8Int = 3, kBeamU8IntLearned = 4 };

// D128U8: dimension 128 under squared L2 with the rows read from the 8-bit copy
// (BeamArgs::rows_u8); only the register-resident plain search has that variant, the other
// kernels take D128's.
enum class BeamShape { D128, D100, Generic, D128U8 };

// Bytes of LDS one wave uses for a given R capacity and visited-cache size.
// (bitmap: the visited set lives in HBM, BeamArgs::vbits)
inline std::size_t beam_lds_per_wave(std::uint32_t ef_cap, std::uint32_t visited_log2, bool bitmap = false) {
    return 8u * static_cast<std::size_t>(ef_cap) + 8u * 64 + 4u * 64 + 4u * 64 + 4u * 16 +
           (bitmap ? 0 : (std::size_t{4} << visited_log2));
}

// The register-resident beam (hnsw_beam_reg.hpp) holds R in registers in chunks of 64 keys;
// it serves ef' <= kRegMaxEf. Its LDS per wave: the merge buffer, fresh ids and distances,
// features (trace and learned modes only) and the visited cache. Without the features a
// 4-wave workgroup with capacity 64 and 2^10 visited slots takes 20,480 bytes, so eight of
// them share a CU's 160 KiB: 32 waves per CU, all the integer kernel's registers allow.
constexpr std::uint32_t kRegMaxEf = 512;
inline std::size_t beam_reg_lds_per_wave(std::uint32_t ef_cap, std::uint32_t visited_log2, bool features) {
    return 8u * static_cast<std::size_t>(ef_cap) + 4u * 64 + 4u * 64 + (features ? 4u * 16 : 0u) +
           (std::size_t{4} << visited_log2);
}

// LDS a workgroup needs besides its waves' regions (the stop model in learned mode).
std::size_t beam_lds_shared(int mode);

hipError_t launch_hnsw_beam(const BeamArgs& args, unsigned grid, unsigned waves_per_group,
                            std::size_t lds_bytes, hipStream_t stream, BeamShape shape, int mode);
const void* hnsw_beam_kernel_ptr(BeamShape shape, int mode);

// The register-resident kernel for R capacity 64 * chunks (chunks 1, 2, 4 or 8).
const void* hnsw_beam_reg_kernel_ptr(BeamShape shape, int mode, int chunks);
hipError_t launch_hnsw_beam_reg(const BeamArgs& args, unsigned grid, unsigned waves_per_group,
                                std::size_t lds_bytes, hipStream_t stream, BeamShape shape, int mode,
                                int chunks);

}  // namespace vdx::dev
