// K2/K3: brute force as an MFMA GEMM with a fused threshold filter, then an exact re-rank.
//
// BruteForceIndex::search (vecdex src/brute_force.cpp:20-84) returns the k smallest exact
// distances. Here, per batch of queries:
//
//  1. prepare: base rows and queries are quantized once into the MFMA operand format (FP32,
//     BF16, or FP8-e4m3 with a per-vector scale), with ||v||^2, the scale, and the exact
//     quantization error ||v~ - v|| and norm ||v~|| of every vector (bf_prepare_rows).
//  2. threshold: distances from every query to a strided sample of the base (the same GEMM,
//     dense output) give, by radix select, a per-query threshold tau whose expected number of
//     survivors over the whole base is a small multiple of the shortlist size.
//  3. filter GEMM: Q x X^T on MFMA, dist = |q|^2 + |x|^2 - 2 q.x (L2) or 1 - q.x; each
//     distance <= tau is appended (key = (distance, id)) to its query's candidate buffer.
//     (The BF16 kernel for 128-element rows accumulates q.x - |x|^2 / 2 from an accumulator
//     start of -|x|^2 / 2 and takes dist = |q|^2 - 2 acc.)
//  4. shortlist: per query, radix select of the S smallest keys, and the smallest distance
//     left out (L).
//  5. re-rank: exact-order distances of the S candidates (the library's arithmetic, bit-equal
//     to vecdex), sorted by (distance, id); the first k are the result.
//  6. certification: every vector left out has GEMM distance >= L, so its exact distance is
//     >= L - E, where E bounds |GEMM distance - exact distance| for this query over all rows
//     (Cauchy-Schwarz on the measured quantization errors, plus FP32 accumulation and rounding
//     terms). If L - E > (k-th exact distance) the result equals the exact scan; otherwise the
//     query is redone by the exact scan (bf_exact.hip). Every mode therefore returns exactly
//     the exact result.

#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>

#include "bf_gemm_args.hpp"
#include "device_common.hpp"
#include "gather_distances.hpp"

namespace vdx::dev {

namespace {

constexpr std::uint32_t kStageBytes = 64 * 8 + 64 * 4;  // per wave: 64 survivor keys and queries

// K steps of the query operand held in registers (F32: float2 pairs of K=2 steps; else K=16).
constexpr std::uint32_t bf_gemm_steps(int prec) { return prec == kPrecF32 ? 32 : 8; }

typedef float v16f __attribute__((ext_vector_type(16)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));

__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        v += __shfl_xor(v, off, kWave);
    }
    return v;
}

__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        v = fmaxf(v, __shfl_xor(v, off, kWave));
    }
    return v;
}

__device__ __forceinline__ unsigned char to_fp8(float x) {
    // OCP e4m3 on gfx950, round to nearest even, saturating to +-448.
    const int packed = __builtin_amdgcn_cvt_pk_fp8_f32(x, 0.0f, 0, false);
    return static_cast<unsigned char>(packed & 0xFF);
}

__device__ __forceinline__ float from_fp8(unsigned char b) {
    return __builtin_amdgcn_cvt_f32_fp8(static_cast<int>(b), 0);
}

}  // namespace

// One wave per row: quantizes `rows` rows of `dim` floats (row stride ld_src) into `dst`
// (rows of kp elements, zero padded), and writes ||v||^2, the scale (FP8; 1 otherwise), the
// quantization error ||v~ - v|| and ||v~||, where v~ is the dequantized vector.
__global__ __launch_bounds__(256) void bf_prepare_rows(PrepArgs a) {
    const int lane = lane_id();
    const std::uint32_t row = blockIdx.x * (blockDim.x / kWave) + (threadIdx.x / kWave);
    if (row >= a.rows) {
        return;
    }
    const float* src = a.src + static_cast<std::size_t>(row) * a.ld_src;
    float maxabs = 0.f;
    for (std::uint32_t e = lane; e < a.dim; e += kWave) {
        maxabs = fmaxf(maxabs, fabsf(src[e]));
    }
    maxabs = wave_max(maxabs);
    const float scale = (a.prec == kPrecFP8 && maxabs > 0.f) ? maxabs / 448.0f : 1.0f;
    const float inv = 1.0f / scale;
    float n2 = 0.f, err2 = 0.f, qn2 = 0.f;
    bool integral = true;
    for (std::uint32_t e = lane; e < a.kp; e += kWave) {
        const float x = e < a.dim ? src[e] : 0.f;
        float xq;
        if (a.prec == kPrecF32) {
            xq = x;
            static_cast<float*>(a.dst)[static_cast<std::size_t>(row) * a.kp + e] = x;
        } else if (a.prec == kPrecBF16) {
            const __bf16 b = static_cast<__bf16>(x);
            xq = static_cast<float>(b);
            static_cast<__bf16*>(a.dst)[static_cast<std::size_t>(row) * a.kp + e] = b;
        } else {
            const unsigned char b = to_fp8(x * inv);
            xq = from_fp8(b) * scale;
            static_cast<unsigned char*>(a.dst)[static_cast<std::size_t>(row) * a.kp + e] = b;
        }
        n2 += x * x;
        err2 += (xq - x) * (xq - x);
        qn2 += xq * xq;
        integral = integral && xq == x && rintf(x) == x;
    }
    integral = ballot(!integral) == 0;
    n2 = wave_sum(n2);
    err2 = wave_sum(err2);
    qn2 = wave_sum(qn2);
    if (lane == 0) {
        a.norm2[row] = n2;
        if (a.nhalf) a.nhalf[row] = -0.5f * n2;
        a.scale[row] = scale;
        a.qerr[row] = sqrtf(err2);
        a.qnorm[row] = sqrtf(qn2);
        a.integral[row] = integral ? 1u : 0u;
    }
}

// The filter's pre-test (L2, FP32 and BF16). A row passes the filter iff its GEMM distance
// d = fl(fl(qn + xn) - 2 dot) (one fma) is <= tau. The pre-test computes v = fl(xn - 2 dot),
// one fma cheaper per distance, and compares it with tq = fl(tau - qn) + 2^-20 (|tau| + qn + X),
// X >= every xn: d <= tau gives xn - 2 dot <= tau - qn + u (|tau| + qn + xn + |xn - 2 dot|)
// + O(u^2) (u = 2^-24), with |xn - 2 dot| <= 2 (qn + xn) up to the operands' quantization, so
// every row that passes the filter passes the pre-test. The exact test then runs only on the
// groups of four rows where some lane's pre-test passed, and the survivors are exactly those of
// the plain test. Padding queries get tq = -inf.
__device__ __forceinline__ float filter_pretest_bound(float tau, float qn, float xn_max, bool qok) {
    if (!qok) return -__int_as_float(0x7f800000);
    const float margin = 0x1p-20f * (fabsf(tau) + qn + xn_max);
    return (tau - qn) + margin;
}

// The epilogue of the GEMM kernels, over the accumulators acc[NH] of base tile tile_t, whose
// |x|^2 (and FP8 scales) sit in LDS at vals + tile_cb * NT (vals + (2 + tile_cb) * NT): rows
// 8 * g + 4 * half + (0..3) of each 32-row MFMA tile, four at a time. MODE 0 stores the dense
// distances; MODE 1 (always over consecutive rows, row_stride 1: a survivor's id is its row)
// stages the survivors (distance <= tau) for the wave's next flush, through the pre-test above
// where it applies (PRETEST). Padding queries never pass (tau = tq = -inf); the row bound is
// tested only in a tile that crosses the end of the rows.
#define VDX_EPILOGUE(CHECK_ROWS)                                                                  \
        _Pragma("unroll") for (int h = 0; h < NH; ++h) {                                          \
            _Pragma("unroll") for (int g = 0; g < 4; ++g) {                                       \
                constexpr bool PRETEST = MODE != 0 && L2 && PREC != kPrecFP8;                     \
                const int r0 = h * 32 + 8 * g + 4 * half;                                         \
                const float4 xn4 = *reinterpret_cast<const float4*>(vals + tile_cb * NT + r0);    \
                float4 xs4 = make_float4(1.f, 1.f, 1.f, 1.f);                                     \
                if constexpr (PREC == kPrecFP8) {                                                 \
                    xs4 = *reinterpret_cast<const float4*>(vals + (2 + tile_cb) * NT + r0);       \
                }                                                                                 \
                float d4[4];                                                                      \
                std::uint64_t m4[4];                                                              \
                _Pragma("unroll") for (int e = 0; e < 4; ++e) {                                   \
                    const float xnv = e == 0 ? xn4.x : (e == 1 ? xn4.y : (e == 2 ? xn4.z : xn4.w)); \
                    const float xsv = e == 0 ? xs4.x : (e == 1 ? xs4.y : (e == 2 ? xs4.z : xs4.w)); \
                    float dot = acc[h][4 * g + e];                                                \
                    if constexpr (PREC == kPrecFP8) {                                             \
                        dot = dot * qs * xsv;                                                     \
                    }                                                                             \
                    const std::uint32_t j = tile_t * NT + r0 + e;                                 \
                    if constexpr (PRETEST) {                                                      \
                        m4[e] = __builtin_amdgcn_fcmpf(__builtin_fmaf(-2.0f, dot, xnv), tq, 5);   \
                    } else {                                                                      \
                        d4[e] = L2 ? __builtin_fmaf(-2.0f, dot, qn + xnv) : 1.0f - dot;           \
                        if constexpr (MODE == 0) {                                                \
                            if (qok && (!(CHECK_ROWS) || j < a.rows)) {                           \
                                a.dense[static_cast<std::size_t>(q) * a.dense_ld + j] = d4[e];    \
                            }                                                                     \
                        } else {                                                                  \
                            /* the lanes whose distance passes, straight from the compare */      \
                            m4[e] = __builtin_amdgcn_fcmpf(d4[e], tau, 5 /* ordered <= */);       \
                            if (CHECK_ROWS) {                                                     \
                                m4[e] &= ballot(j < a.rows);                                      \
                            }                                                                     \
                        }                                                                         \
                    }                                                                             \
                }                                                                                 \
                if constexpr (MODE != 0) {                                                        \
                    if ((m4[0] | m4[1] | m4[2] | m4[3]) != 0) {                                   \
                        if constexpr (PRETEST) {                                                  \
                            _Pragma("unroll") for (int e = 0; e < 4; ++e) {                       \
                                const float xnv = e == 0 ? xn4.x : (e == 1 ? xn4.y : (e == 2 ? xn4.z : xn4.w)); \
                                d4[e] = __builtin_fmaf(-2.0f, acc[h][4 * g + e], qn + xnv);       \
                                m4[e] &= __builtin_amdgcn_fcmpf(d4[e], tau, 5 /* ordered <= */);  \
                                if (CHECK_ROWS) {                                                 \
                                    m4[e] &= ballot(tile_t * NT + r0 + e < a.rows);               \
                                }                                                                 \
                            }                                                                     \
                        }                                                                         \
                        _Pragma("unroll") for (int e = 0; e < 4; ++e) {                           \
                            const std::uint64_t m = m4[e];                                        \
                            if (m != 0) {                                                         \
                                const std::uint32_t np = static_cast<std::uint32_t>(__popcll(m)); \
                                if (staged + np > 64) {                                           \
                                    flush();                                                      \
                                }                                                                 \
                                if ((m >> lane) & 1u) {                                           \
                                    const std::uint32_t at = staged + lanes_below(m);             \
                                    skey[at] = make_key(d4[e], tile_t * NT + r0 + e);             \
                                    squery[at] = q;                                               \
                                }                                                                 \
                                staged += np;                                                     \
                            }                                                                     \
                        }                                                                         \
                    }                                                                             \
                }                                                                                 \
            }                                                                                     \
        }

// The GEMM. A workgroup of WAVES waves (8 at launch) takes 32 * WAVES queries (32 per wave, one
// MFMA column tile held in registers for the whole launch) and one slab of base rows. Base tiles
// of NT rows (64; 32 for FP32) are the MFMA row operand and go through LDS, double buffered: the
// workgroup issues the 16-byte global loads of tile t+2 before the MFMAs of tile t and stores
// tile t+1 to the other buffer afterwards, so one barrier per tile suffices and load latency
// hides behind compute. LDS rows are padded by 16 bytes so the 32 lanes of a fragment read hit
// distinct banks; the tile's |x|^2 and scales sit next to it in LDS.
//
// BF16 and FP8 are held to 128 VGPRs (4 waves per SIMD), so two 8-wave workgroups share a CU
// and one's epilogue and barrier overlap the other's MFMAs (1.3x over 3 waves per SIMD, which
// fits only one workgroup); FP32's larger fragments would spill at that bound.
//
// MFMA 32x32 output layout: lane l holds column n = l % 32 (its query) and rows
// m = 8 * (i / 4) + 4 * (l / 32) + (i % 4) (base rows) of its 16 accumulators i. Operand
// fragments: lane l holds row/column l % 32 and the K slice
// [kstep * K + (l / 32) * K / 2, ... + K / 2).
// MODE 0: dense distances to a sample of rows; MODE 1: the threshold filter. L2: squared L2,
// else 1 - <q, x>. FULL: the operand rows fill every K step (a.ks == KS), so the step loop has
// no run-time guard; a guarded step makes the compiler copy all accumulators after each MFMA.
template <int PREC, int WAVES, int MODE, bool L2, bool FULL>
__global__ __launch_bounds__(WAVES * 64, PREC == kPrecF32 || !FULL ? 2 : 4) void bf_gemm_kernel(GemmArgs a) {
    constexpr int EB = PREC == kPrecF32 ? 4 : (PREC == kPrecBF16 ? 2 : 1);  // bytes per element
    constexpr int NT = PREC == kPrecF32 ? 32 : 64;                           // rows per tile
    constexpr int NH = NT / 32;                                              // MFMA row tiles
    constexpr int KS = static_cast<int>(bf_gemm_steps(PREC));
    extern __shared__ __align__(16) unsigned char lds[];
    const int lane = lane_id();
    const int wave = threadIdx.x >> 6;
    const int col = lane & 31;
    const int half = lane >> 5;
    constexpr std::uint32_t kThreads = WAVES * 64;
    const std::uint32_t q = blockIdx.x * (32 * WAVES) + wave * 32 + col;  // this lane's query
    const std::uint32_t kp = a.kp;
    const std::uint32_t row_bytes = kp * EB;
    const std::uint32_t rs = row_bytes + 16;  // padded LDS row stride
    // buffer b: rows at lds + b * NT * rs, |x|^2 at vals + b * NT, scales at vals + (2 + b) * NT
    float* const vals = reinterpret_cast<float*>(lds + 2 * NT * rs);

    typedef typename std::conditional<PREC == kPrecF32, float2,
                                      typename std::conditional<PREC == kPrecBF16, v8bf, long>::type>::type Frag;
    Frag bq[KS];
    const std::uint32_t ks = a.ks;
    const unsigned char* qbase = static_cast<const unsigned char*>(a.qq) + static_cast<std::size_t>(q) * row_bytes;
#pragma unroll
    for (int s = 0; s < KS; ++s) {
        if (static_cast<std::uint32_t>(s) < ks) {
            if constexpr (PREC == kPrecF32) {
                bq[s] = *reinterpret_cast<const float2*>(qbase + (4 * s + 2 * half) * 4);
            } else if constexpr (PREC == kPrecBF16) {
                bq[s] = *reinterpret_cast<const v8bf*>(qbase + (16 * s + 8 * half) * 2);
            } else {
                bq[s] = *reinterpret_cast<const long*>(qbase + 16 * s + 8 * half);
            }
        }
    }
    const bool qok = q < a.nq;
    const float qn = qok ? a.qn[q] : 0.f;
    const float qs = qok && a.qs ? a.qs[q] : 1.f;
    // Padding lanes (q >= nq) get -inf, so no distance passes their filter: their candidate
    // rows would lie past the end of the buffer.
    const float tau = !qok ? -__int_as_float(0x7f800000) : (a.tau ? a.tau[q] : __int_as_float(0x7f800000));
    const float tq = filter_pretest_bound(tau, qn, a.xn_max, qok);

    const std::uint32_t tiles = (a.rows + NT - 1) / NT;
    const std::uint32_t per_slab = (tiles + gridDim.y - 1) / gridDim.y;
    const std::uint32_t t0 = blockIdx.y * per_slab;
    const std::uint32_t t1 = min(tiles, t0 + per_slab);
    if (t0 >= t1) {
        return;
    }
    // Cooperative tile loads: 16-byte pieces, piece p is row p / chunks, chunk p % chunks.
    const std::uint32_t chunks = row_bytes / 16;
    const std::uint32_t pieces = NT * chunks;
    // NT * row_bytes <= 16 KB: at most 1024 / kThreads pieces per thread. Two register stages
    // (a, b) hold the tiles fetched one and two steps ahead, so a tile's global loads have two
    // tiles of compute to land before they are stored to LDS.
    uint4 sa0 = make_uint4(0, 0, 0, 0), sa1 = sa0, sa2 = sa0, sa3 = sa0;
    uint4 sb0 = sa0, sb1 = sa0, sb2 = sa0, sb3 = sa0;
    float xa_n = 0.f, xa_s = 1.f, xb_n = 0.f, xb_s = 1.f;
    // Each thread's pieces sit at fixed positions within a tile: byte offsets in the operand
    // rows and in the LDS buffer are computed once. A tile past the end of the rows repeats the
    // last row (its distances are never used).
    std::uint32_t prow[4], pcol[4], plds[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const std::uint32_t p = threadIdx.x + kThreads * i;
        prow[i] = p / chunks;
        pcol[i] = (p % chunks) * 16;
        plds[i] = prow[i] * rs + pcol[i];
    }
    const unsigned char* const xq_bytes = static_cast<const unsigned char*>(a.xq);
#define VDX_FETCH_PIECE(T_, I, DST)                                                               \
    {                                                                                             \
        if (threadIdx.x + kThreads * (I) < pieces) {                                              \
            const std::uint32_t j = (T_) * NT + prow[I];                                          \
            const std::uint32_t xr = (j < a.rows ? j : a.rows - 1) * a.row_stride;               \
            DST = *reinterpret_cast<const uint4*>(xq_bytes + static_cast<std::size_t>(xr) * row_bytes + pcol[I]); \
        }                                                                                         \
    }
#define VDX_STORE_PIECE(B_, I, SRC)                                                               \
    {                                                                                             \
        if (threadIdx.x + kThreads * (I) < pieces) {                                              \
            *reinterpret_cast<uint4*>(lds + (B_) * NT * rs + plds[I]) = SRC;                     \
        }                                                                                         \
    }
#define VDX_FETCH_STAGE(T_, S0, S1, S2, S3, XN, XS)                                               \
    {                                                                                             \
        VDX_FETCH_PIECE(T_, 0, S0)                                                                \
        VDX_FETCH_PIECE(T_, 1, S1)                                                                \
        if constexpr (WAVES == 4) {                                                               \
            VDX_FETCH_PIECE(T_, 2, S2)                                                            \
            VDX_FETCH_PIECE(T_, 3, S3)                                                            \
        }                                                                                         \
        if (threadIdx.x < NT) {                                                                   \
            const std::uint32_t j = (T_) * NT + threadIdx.x;                                      \
            const std::uint32_t xr = (j < a.rows ? j : a.rows - 1) * a.row_stride;               \
            XN = a.xn[xr];                                                                        \
            if constexpr (PREC == kPrecFP8) {                                                     \
                XS = a.xs[xr];                                                                    \
            }                                                                                     \
        }                                                                                         \
    }
#define VDX_STORE_STAGE(B_, S0, S1, S2, S3, XN, XS)                                               \
    {                                                                                             \
        VDX_STORE_PIECE(B_, 0, S0)                                                                \
        VDX_STORE_PIECE(B_, 1, S1)                                                                \
        if constexpr (WAVES == 4) {                                                               \
            VDX_STORE_PIECE(B_, 2, S2)                                                            \
            VDX_STORE_PIECE(B_, 3, S3)                                                            \
        }                                                                                         \
        if (threadIdx.x < NT) {                                                                   \
            vals[(B_) * NT + threadIdx.x] = XN;                                                   \
            vals[(2 + (B_)) * NT + threadIdx.x] = XS;                                             \
        }                                                                                         \
    }
    // Filter survivors are staged per wave in LDS (64 keys with their queries, positions from a
    // ballot, no atomics) and flushed together: one instruction reserves the buffer slots of
    // all staged survivors, so the wave waits for one round trip per 64 survivors instead of one
    // per survivor.
    std::uint64_t* skey = reinterpret_cast<std::uint64_t*>(lds + 2 * NT * rs + 4 * NT * 4 + wave * kStageBytes);
    std::uint32_t* squery = reinterpret_cast<std::uint32_t*>(skey + 64);
    std::uint32_t staged = 0;  // wave-uniform
    const auto flush = [&]() {
        wave_lds_fence();
        if (static_cast<std::uint32_t>(lane) < staged) {
            const std::uint32_t sq = squery[lane];
            const std::uint32_t slot = atomicAdd(a.cnt + sq, 1u);
            if (slot < a.cap) {
                a.buf[static_cast<std::size_t>(sq) * a.cap + slot] = skey[lane];
            }
        }
        wave_lds_fence();
        staged = 0;
    };
    // Double-buffered LDS tiles and two register stages a and b, the loop unrolled by two so
    // the stages keep fixed roles (no register moves, so no wait right after a prefetch is
    // issued): an even step computes buffer 0 while tile t + 2 loads into a and stores b (tile
    // t + 1) to buffer 1; an odd step computes buffer 1, loads into b and stores a to buffer 0.
    // Loads past the last tile repeat it (never computed).
    const std::uint32_t last = t1 - 1;
    VDX_FETCH_STAGE(t0, sa0, sa1, sa2, sa3, xa_n, xa_s)
    VDX_STORE_STAGE(0, sa0, sa1, sa2, sa3, xa_n, xa_s)
    VDX_FETCH_STAGE(min(t0 + 1, last), sb0, sb1, sb2, sb3, xb_n, xb_s)
    // Every load so far (the query fragments above all) completes here, once: inside the loop
    // the compiler then waits only for the stage a step is about to store, never for vmcnt(0).
    __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0) only
    __syncthreads();
    for (std::uint32_t t = t0; t < t1; t += 2) {
        {
            const std::uint32_t tile_t = t;
            constexpr int tile_cb = 0;
            VDX_FETCH_STAGE(min(t + 2, last), sa0, sa1, sa2, sa3, xa_n, xa_s)
            v16f acc[NH];
#pragma unroll
            for (int h = 0; h < NH; ++h) {
                acc[h] = v16f{};
            }
#pragma unroll
            for (int s = 0; s < KS; ++s) {
                if (FULL || static_cast<std::uint32_t>(s) < ks) {
#pragma unroll
                    for (int h = 0; h < NH; ++h) {
                        const unsigned char* rowp = lds + tile_cb * NT * rs + (h * 32 + col) * rs;
                        if constexpr (PREC == kPrecF32) {
                            const float2 x = *reinterpret_cast<const float2*>(rowp + (4 * s + 2 * half) * 4);
                            acc[h] = __builtin_amdgcn_mfma_f32_32x32x2f32(x.x, bq[s].x, acc[h], 0, 0, 0);
                            acc[h] = __builtin_amdgcn_mfma_f32_32x32x2f32(x.y, bq[s].y, acc[h], 0, 0, 0);
                        } else if constexpr (PREC == kPrecBF16) {
                            const v8bf x = *reinterpret_cast<const v8bf*>(rowp + (16 * s + 8 * half) * 2);
                            acc[h] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(x, bq[s], acc[h], 0, 0, 0);
                        } else {
                            const long x = *reinterpret_cast<const long*>(rowp + 16 * s + 8 * half);
                            acc[h] = __builtin_amdgcn_mfma_f32_32x32x16_fp8_fp8(x, bq[s], acc[h], 0, 0, 0);
                        }
                    }
                }
            }

            // The filter tests exactly the distance it stores (2 dot is exact, so the fma equals
            // (|q|^2 + |x|^2) - 2 dot): a row is left out iff its GEMM distance exceeds tau, which
            // the certification relies on.
            if ((tile_t + 1) * NT <= a.rows) {
                VDX_EPILOGUE(false)
            } else {
                VDX_EPILOGUE(true)
            }

            VDX_STORE_STAGE(1, sb0, sb1, sb2, sb3, xb_n, xb_s)
            __syncthreads();
        }
        if (t + 1 >= t1) {
            break;
        }
        {
            const std::uint32_t tile_t = t + 1;
            constexpr int tile_cb = 1;
            VDX_FETCH_STAGE(min(t + 3, last), sb0, sb1, sb2, sb3, xb_n, xb_s)
            v16f acc[NH];
#pragma unroll
            for (int h = 0; h < NH; ++h) {
                acc[h] = v16f{};
            }
#pragma unroll
            for (int s = 0; s < KS; ++s) {
                if (FULL || static_cast<std::uint32_t>(s) < ks) {
#pragma unroll
                    for (int h = 0; h < NH; ++h) {
                        const unsigned char* rowp = lds + tile_cb * NT * rs + (h * 32 + col) * rs;
                        if constexpr (PREC == kPrecF32) {
                            const float2 x = *reinterpret_cast<const float2*>(rowp + (4 * s + 2 * half) * 4);
                            acc[h] = __builtin_amdgcn_mfma_f32_32x32x2f32(x.x, bq[s].x, acc[h], 0, 0, 0);
                            acc[h] = __builtin_amdgcn_mfma_f32_32x32x2f32(x.y, bq[s].y, acc[h], 0, 0, 0);
                        } else if constexpr (PREC == kPrecBF16) {
                            const v8bf x = *reinterpret_cast<const v8bf*>(rowp + (16 * s + 8 * half) * 2);
                            acc[h] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(x, bq[s], acc[h], 0, 0, 0);
                        } else {
                            const long x = *reinterpret_cast<const long*>(rowp + 16 * s + 8 * half);
                            acc[h] = __builtin_amdgcn_mfma_f32_32x32x16_fp8_fp8(x, bq[s], acc[h], 0, 0, 0);
                        }
                    }
                }
            }

            // The filter tests exactly the distance it stores (2 dot is exact, so the fma equals
            // (|q|^2 + |x|^2) - 2 dot): a row is left out iff its GEMM distance exceeds tau, which
            // the certification relies on.
            if ((tile_t + 1) * NT <= a.rows) {
                VDX_EPILOGUE(false)
            } else {
                VDX_EPILOGUE(true)
            }

            VDX_STORE_STAGE(0, sa0, sa1, sa2, sa3, xa_n, xa_s)
            __syncthreads();
        }
    }
    if (MODE != 0 && staged > 0) {
        flush();
    }
}

#undef VDX_FETCH_PIECE
#undef VDX_STORE_PIECE
#undef VDX_FETCH_STAGE
#undef VDX_STORE_STAGE

// BF16 with every K step filled (operand rows of 128 elements, dimensions 113 to 128). Base
// tiles go from global memory straight into LDS (global_load_lds_dwordx4: no staging registers,
// no LDS stores), triple buffered: step t issues the loads of tile t + 2 into the buffer that
// tile t - 1 used, computes tile t, waits until tile t + 1 has landed and meets the workgroup at
// one barrier. The registers this frees let four waves per SIMD keep MFMA operands in flight.
//
// A wave's load instruction fills 1 KB of LDS, 16 bytes per lane in lane order, so rows are not
// padded: row r keeps its 16-byte chunk c at slot c ^ (r % 16), which spreads the 32 rows of a
// fragment read over all banks. Lane (col, half) reads chunk 2s + half of row h * 32 + col at
// row * 256 + 16 * ((2s + half) ^ (col % 16)) = A0 ^ (32 s) + 8192 h, one XOR per step.
//
// The loads are issued from inline asm, which the compiler does not see writing LDS: its own
// bookkeeping for these loads would wait for every one in flight before any LDS read (the
// epilogue's |x|^2 reads included), which serializes the prefetch. The kernel orders them
// itself: a buffer is read only after wait_older() and the barrier that follow its loads, and
// is loaded again only after the barrier that ends the last step reading it.
__device__ __forceinline__ std::uint32_t lds_offset(const void* p) {
    return static_cast<std::uint32_t>(reinterpret_cast<std::uintptr_t>(p));  // LDS aperture offset
}

// 16 bytes per lane from `src` to LDS at m0 + 16 * lane (m0 wave-uniform).
__device__ __forceinline__ void lds_load16(const void* src, std::uint32_t m0) {
    asm volatile("global_load_lds_dwordx4 %0, off" ::"v"(src), "{m0}"(m0));
}

// 4 bytes per lane from `src` to LDS at m0 + 4 * lane (m0 wave-uniform).
__device__ __forceinline__ void lds_load4(const void* src, std::uint32_t m0) {
    asm volatile("global_load_lds_dword %0, off" ::"v"(src), "{m0}"(m0));
}

// The filter's accept bound for the LDS-fed kernel. Its accumulator starts at -|x|^2 / 2 under
// L2 (at 0 otherwise), so the MFMA chain leaves a = q~.x~ - |x|^2 / 2 and the GEMM distance is
// f(a) = fl(qn - 2a), one fma (fl(1 - a) for inner product and cosine). f never increases with
// a, so f(a) <= tau iff a >= a*, the smallest float with f(a*) <= tau: one compare per distance.
// a* comes from a bisection over the floats in order (32 steps, once per lane); NaN, which
// nothing passes, for tau = -inf (padding queries) and -inf for tau = +inf (no threshold).
__device__ float filter_accept_bound(float tau, float qn, bool l2) {
    if (!(tau > -__int_as_float(0x7f800000))) return __int_as_float(0x7fc00000);
    if (tau == __int_as_float(0x7f800000)) return -__int_as_float(0x7f800000);
    const auto f = [&](float x) { return l2 ? __builtin_fmaf(-2.0f, x, qn) : 1.0f - x; };
    std::uint32_t lo = sortable(-3.40282347e38f), hi = sortable(3.40282347e38f);
    if (f(unsortable(lo)) <= tau) return -__int_as_float(0x7f800000);
    if (f(unsortable(hi)) > tau) return __int_as_float(0x7fc00000);
    while (hi - lo > 1) {  // f(lo) > tau >= f(hi)
        const std::uint32_t mid = lo + (hi - lo) / 2;
        if (f(unsortable(mid)) <= tau) {
            hi = mid;
        } else {
            lo = mid;
        }
    }
    return unsortable(hi);
}

// The LDS-fed kernel's epilogue over acc[NH] of tile tile_t: MODE 0 stores f(acc); MODE 1
// stages the rows with acc >= a* (f(acc) <= tau) with key (f(acc), row). Lanes and rows as in
// VDX_EPILOGUE.
#define VDX_DMA_EPILOGUE(CHECK_ROWS)                                                              \
        _Pragma("unroll") for (int h = 0; h < NH; ++h) {                                          \
            _Pragma("unroll") for (int g = 0; g < 4; ++g) {                                       \
                const int r0 = h * 32 + 8 * g + 4 * half;                                         \
                if constexpr (MODE == 0) {                                                        \
                    _Pragma("unroll") for (int e = 0; e < 4; ++e) {                               \
                        const float acc_e = acc[h][4 * g + e];                                    \
                        const float d = L2 ? __builtin_fmaf(-2.0f, acc_e, qn) : 1.0f - acc_e;     \
                        const std::uint32_t j = tile_t * NT + r0 + e;                             \
                        if (qok && (!(CHECK_ROWS) || j < a.rows)) {                               \
                            a.dense[static_cast<std::size_t>(q) * a.dense_ld + j] = d;            \
                        }                                                                         \
                    }                                                                             \
                } else {                                                                          \
                    std::uint64_t m4[4];                                                          \
                    _Pragma("unroll") for (int e = 0; e < 4; ++e) {                               \
                        m4[e] = __builtin_amdgcn_fcmpf(acc[h][4 * g + e], astar, 3 /* ordered >= */); \
                        if (CHECK_ROWS) {                                                         \
                            m4[e] &= ballot(tile_t * NT + r0 + e < a.rows);                       \
                        }                                                                         \
                    }                                                                             \
                    if ((m4[0] | m4[1] | m4[2] | m4[3]) != 0) {                                   \
                        _Pragma("unroll") for (int e = 0; e < 4; ++e) {                           \
                            const std::uint64_t m = m4[e];                                        \
                            if (m != 0) {                                                         \
                                const std::uint32_t np = static_cast<std::uint32_t>(__popcll(m)); \
                                if (staged + np > 64) {                                           \
                                    flush();                                                      \
                                }                                                                 \
                                if ((m >> lane) & 1u) {                                           \
                                    const float acc_e = acc[h][4 * g + e];                        \
                                    const float d = L2 ? __builtin_fmaf(-2.0f, acc_e, qn) : 1.0f - acc_e; \
                                    const std::uint32_t at = staged + lanes_below(m);             \
                                    skey[at] = make_key(d, tile_t * NT + r0 + e);                 \
                                    squery[at] = q;                                               \
                                }                                                                 \
                                staged += np;                                                     \
                            }                                                                     \
                        }                                                                         \
                    }                                                                             \
                }                                                                                 \
            }                                                                                     \
        }

template <int MODE, bool L2>
__global__ __launch_bounds__(512, 4) void bf_gemm_bf16_dma_kernel(GemmArgs a) {
    constexpr int NT = 64;
    constexpr int NH = 2;
    constexpr int KS = 8;
    constexpr std::uint32_t kRowBytes = 256;
    constexpr std::uint32_t kTileBytes = NT * kRowBytes;
    extern __shared__ __align__(16) unsigned char lds[];
    const int lane = lane_id();
    const int wave = threadIdx.x >> 6;
    const int col = lane & 31;
    const int half = lane >> 5;
    const std::uint32_t q = blockIdx.x * 256 + wave * 32 + col;  // this lane's query
    // -|x|^2 / 2 of the three buffers (L2), then the survivor staging
    float* const vals = reinterpret_cast<float*>(lds + 3 * kTileBytes);

    v8bf bq[KS];
    const unsigned char* qbase = static_cast<const unsigned char*>(a.qq) + static_cast<std::size_t>(q) * kRowBytes;
#pragma unroll
    for (int s = 0; s < KS; ++s) {
        bq[s] = *reinterpret_cast<const v8bf*>(qbase + (16 * s + 8 * half) * 2);
    }
    const bool qok = q < a.nq;
    const float qn = qok ? a.qn[q] : 0.f;
    const float tau = !qok ? -__int_as_float(0x7f800000) : (a.tau ? a.tau[q] : __int_as_float(0x7f800000));
    const float astar = MODE == 0 ? 0.f : filter_accept_bound(tau, qn, L2);

    const std::uint32_t tiles = (a.rows + NT - 1) / NT;
    const std::uint32_t per_slab = (tiles + gridDim.y - 1) / gridDim.y;
    const std::uint32_t t0 = blockIdx.y * per_slab;
    const std::uint32_t t1 = min(tiles, t0 + per_slab);
    if (t0 >= t1) {
        return;
    }
    const std::uint32_t last = t1 - 1;
    // Wave w's two load instructions fill LDS bytes [2w KB, 2w KB + 2 KB) of a buffer: rows
    // 8w + lane / 16 (+4 for the second), slot lane % 16, holding chunk slot ^ (row % 16). The
    // per-lane source addresses are fixed; a tile adds tile * NT * row_stride rows (a scalar).
    // Under L2 wave 0 also loads the tile's -|x|^2 / 2.
    const std::uint32_t wave_u = __builtin_amdgcn_readfirstlane(wave);  // m0 must be uniform
    const std::uint32_t drow = 8 * wave + (lane >> 4);
    const std::uint32_t dcol0 = ((lane & 15) ^ (drow & 15)) * 16;
    const std::uint32_t dcol1 = ((lane & 15) ^ ((drow + 4) & 15)) * 16;
    const unsigned char* const xq_bytes = static_cast<const unsigned char*>(a.xq);
    const unsigned char* const src0 = xq_bytes + static_cast<std::size_t>(drow * a.row_stride) * kRowBytes + dcol0;
    const unsigned char* const src1 = xq_bytes + static_cast<std::size_t>((drow + 4) * a.row_stride) * kRowBytes + dcol1;
    const float* const xh_src = a.xh + static_cast<std::size_t>(lane) * a.row_stride;
    const std::uint32_t lds_base = lds_offset(lds);
    // A tile past the end of the rows repeats the last row (its distances are never used).
    const auto issue = [&](std::uint32_t tile, std::uint32_t b) {
        const std::uint32_t dst = lds_base + b * kTileBytes + wave_u * 2048;
        const std::size_t first = static_cast<std::size_t>(tile) * NT * a.row_stride;  // first row's index
        if ((tile + 1) * NT <= a.rows) {
            lds_load16(src0 + first * kRowBytes, dst);
            lds_load16(src1 + first * kRowBytes, dst + 1024);
            if (L2 && wave_u == 0) {
                lds_load4(xh_src + first, lds_base + 3 * kTileBytes + b * NT * 4);
            }
        } else {
            const std::uint32_t j0 = tile * NT + drow;
            const std::uint32_t j1 = j0 + 4;
            const std::uint32_t x0 = (j0 < a.rows ? j0 : a.rows - 1) * a.row_stride;
            const std::uint32_t x1 = (j1 < a.rows ? j1 : a.rows - 1) * a.row_stride;
            lds_load16(xq_bytes + static_cast<std::size_t>(x0) * kRowBytes + dcol0, dst);
            lds_load16(xq_bytes + static_cast<std::size_t>(x1) * kRowBytes + dcol1, dst + 1024);
            if (L2 && wave_u == 0) {
                const std::uint32_t j = tile * NT + lane;
                const std::uint32_t xr = (j < a.rows ? j : a.rows - 1) * a.row_stride;
                lds_load4(a.xh + xr, lds_base + 3 * kTileBytes + b * NT * 4);
            }
        }
    };
    // Loads complete in issue order: leaving the newest tile's loads (3 on wave 0 under L2, 2
    // otherwise) in flight means every older one has landed.
    const auto wait_older = [&]() {
        if (L2 && wave_u == 0) {
            __builtin_amdgcn_s_waitcnt(0x0F73);  // vmcnt(3)
        } else {
            __builtin_amdgcn_s_waitcnt(0x0F72);  // vmcnt(2)
        }
    };

    std::uint64_t* skey = reinterpret_cast<std::uint64_t*>(lds + 3 * kTileBytes + 3 * NT * 4 + wave * kStageBytes);
    std::uint32_t* squery = reinterpret_cast<std::uint32_t*>(skey + 64);
    std::uint32_t staged = 0;  // wave-uniform
    const auto flush = [&]() {
        wave_lds_fence();
        if (static_cast<std::uint32_t>(lane) < staged) {
            const std::uint32_t sq = squery[lane];
            const std::uint32_t slot = atomicAdd(a.cnt + sq, 1u);
            if (slot < a.cap) {
                a.buf[static_cast<std::size_t>(sq) * a.cap + slot] = skey[lane];
            }
        }
        wave_lds_fence();
        staged = 0;
    };

    // The query fragments, |q|^2 and tau land before the first tile load is issued, so the
    // compiler's own waits for them (it does not see the tile loads) never fall inside the loop.
    __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0)
    issue(t0, 0);
    issue(min(t0 + 1, last), 1);
    wait_older();
    __syncthreads();
    const std::uint32_t a0 = static_cast<std::uint32_t>(col) * kRowBytes +
                             16u * (static_cast<std::uint32_t>(half) ^ (static_cast<std::uint32_t>(col) & 15u));
    std::uint32_t tile_cb = 0;  // buffer of tile t
    for (std::uint32_t t = t0; t < t1; ++t) {
        const std::uint32_t tile_t = t;
        issue(min(t + 2, last), tile_cb == 0 ? 2 : tile_cb - 1);
        const std::uint32_t ab = a0 + tile_cb * kTileBytes;
        // accumulator element 4g + e of MFMA tile h is row h * 32 + 8g + 4 * half + e
        v16f acc[NH];
#pragma unroll
        for (int h = 0; h < NH; ++h) {
            if constexpr (L2) {
#pragma unroll
                for (int g = 0; g < 4; ++g) {
                    const float4 c = *reinterpret_cast<const float4*>(vals + tile_cb * NT + h * 32 + 8 * g + 4 * half);
                    acc[h][4 * g + 0] = c.x;
                    acc[h][4 * g + 1] = c.y;
                    acc[h][4 * g + 2] = c.z;
                    acc[h][4 * g + 3] = c.w;
                }
            } else {
                acc[h] = v16f{};
            }
        }
#pragma unroll
        for (int s = 0; s < KS; ++s) {
            const unsigned char* p = lds + (ab ^ (32u * s));
#pragma unroll
            for (int h = 0; h < NH; ++h) {
                const v8bf x = *reinterpret_cast<const v8bf*>(p + h * 32 * kRowBytes);
                acc[h] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(x, bq[s], acc[h], 0, 0, 0);
            }
        }
        if ((tile_t + 1) * NT <= a.rows) {
            VDX_DMA_EPILOGUE(false)
        } else {
            VDX_DMA_EPILOGUE(true)
        }
        wait_older();
        __syncthreads();
        tile_cb = tile_cb == 2 ? 0 : tile_cb + 1;
    }
    if (MODE != 0 && staged > 0) {
        flush();
    }
}

// The same filter with two query blocks per wave: each wave holds 64 queries (two 32-query
// register operands) and feeds every base fragment it reads from LDS to two MFMAs, so a tile
// costs half the LDS reads per MFMA of bf_gemm_bf16_dma_kernel (whose LDS reads about match
// the MFMA rate). 4-wave workgroups of 256 queries, three per CU (168 registers per lane: 64 of
// query fragments, 64 of accumulators). The tile layout, loads and epilogue are the single-block
// kernel's; a wave loads 16 of a tile's 64 rows.
#define VDX_DMA2_EPILOGUE(CHECK_ROWS)                                                             \
        _Pragma("unroll") for (int b = 0; b < QB; ++b) {                                          \
        _Pragma("unroll") for (int h = 0; h < NH; ++h) {                                          \
            _Pragma("unroll") for (int g = 0; g < 4; ++g) {                                       \
                const int r0 = h * 32 + 8 * g + 4 * half;                                         \
                if constexpr (MODE == 0) {                                                        \
                    _Pragma("unroll") for (int e = 0; e < 4; ++e) {                               \
                        const float acc_e = acc[b][h][4 * g + e];                                 \
                        const float d = L2 ? __builtin_fmaf(-2.0f, acc_e, qn[b]) : 1.0f - acc_e;  \
                        const std::uint32_t j = tile_t * NT + r0 + e;                             \
                        if (qok[b] && (!(CHECK_ROWS) || j < a.rows)) {                            \
                            a.dense[static_cast<std::size_t>(qv[b]) * a.dense_ld + j] = d;        \
                        }                                                                         \
                    }                                                                             \
                } else {                                                                          \
                    std::uint64_t m4[4];                                                          \
                    _Pragma("unroll") for (int e = 0; e < 4; ++e) {                               \
                        m4[e] = __builtin_amdgcn_fcmpf(acc[b][h][4 * g + e], astar[b], 3 /* ordered >= */); \
                        if (CHECK_ROWS) {                                                         \
                            m4[e] &= ballot(tile_t * NT + r0 + e < a.rows);                       \
                        }                                                                         \
                    }                                                                             \
                    if ((m4[0] | m4[1] | m4[2] | m4[3]) != 0) {                                   \
                        _Pragma("unroll") for (int e = 0; e < 4; ++e) {                           \
                            const std::uint64_t m = m4[e];                                        \
                            if (m != 0) {                                                         \
                                const std::uint32_t np = static_cast<std::uint32_t>(__popcll(m)); \
                                if (staged + np > 64) {                                           \
                                    flush();                                                      \
                                }                                                                 \
                                if ((m >> lane) & 1u) {                                           \
                                    const float acc_e = acc[b][h][4 * g + e];                     \
                                    const float d = L2 ? __builtin_fmaf(-2.0f, acc_e, qn[b]) : 1.0f - acc_e; \
                                    const std::uint32_t at = staged + lanes_below(m);             \
                                    skey[at] = make_key(d, tile_t * NT + r0 + e);                 \
                                    squery[at] = qv[b];                                           \
                                }                                                                 \
                                staged += np;                                                     \
                            }                                                                     \
                        }                                                                         \
                    }                                                                             \
                }                                                                                 \
            }                                                                                     \
        }                                                                                         \
        }

template <int MODE, bool L2>
__global__ __launch_bounds__(256, 2) void bf_gemm_bf16_dma2_kernel(GemmArgs a) {
    constexpr int NT = 64;
    constexpr int NH = 2;
    constexpr int KS = 8;
    constexpr int QB = 2;
    constexpr std::uint32_t kRowBytes = 256;
    constexpr std::uint32_t kTileBytes = NT * kRowBytes;
    extern __shared__ __align__(16) unsigned char lds[];
    const int lane = lane_id();
    const int wave = threadIdx.x >> 6;
    const int col = lane & 31;
    const int half = lane >> 5;
    float* const vals = reinterpret_cast<float*>(lds + 3 * kTileBytes);

    v8bf bq[QB][KS];
    std::uint32_t qv[QB];
    bool qok[QB];
    float qn[QB], astar[QB];
#pragma unroll
    for (int b = 0; b < QB; ++b) {
        qv[b] = blockIdx.x * 256 + wave * 64 + b * 32 + col;  // this lane's query of block b
        const unsigned char* qbase = static_cast<const unsigned char*>(a.qq) + static_cast<std::size_t>(qv[b]) * kRowBytes;
#pragma unroll
        for (int s = 0; s < KS; ++s) {
            bq[b][s] = *reinterpret_cast<const v8bf*>(qbase + (16 * s + 8 * half) * 2);
        }
        qok[b] = qv[b] < a.nq;
        qn[b] = qok[b] ? a.qn[qv[b]] : 0.f;
        const float tau = !qok[b] ? -__int_as_float(0x7f800000) : (a.tau ? a.tau[qv[b]] : __int_as_float(0x7f800000));
        astar[b] = MODE == 0 ? 0.f : filter_accept_bound(tau, qn[b], L2);
    }

    const std::uint32_t tiles = (a.rows + NT - 1) / NT;
    const std::uint32_t per_slab = (tiles + gridDim.y - 1) / gridDim.y;
    const std::uint32_t t0 = blockIdx.y * per_slab;
    const std::uint32_t t1 = min(tiles, t0 + per_slab);
    if (t0 >= t1) {
        return;
    }
    const std::uint32_t last = t1 - 1;
    // Wave w's four load instructions fill LDS bytes [4w KB, 4w KB + 4 KB) of a buffer: rows
    // 16w + 4i + lane / 16 (i = 0..3), slot lane % 16, holding chunk slot ^ (row % 16).
    const std::uint32_t wave_u = __builtin_amdgcn_readfirstlane(wave);  // m0 must be uniform
    const unsigned char* const xq_bytes = static_cast<const unsigned char*>(a.xq);
    std::uint32_t drow[4], dcol[4];
    const unsigned char* src[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        drow[i] = 16 * wave + 4 * i + (lane >> 4);
        dcol[i] = ((lane & 15) ^ (drow[i] & 15)) * 16;
        src[i] = xq_bytes + static_cast<std::size_t>(drow[i] * a.row_stride) * kRowBytes + dcol[i];
    }
    const float* const xh_src = a.xh + static_cast<std::size_t>(lane) * a.row_stride;
    const std::uint32_t lds_base = lds_offset(lds);
    const auto issue = [&](std::uint32_t tile, std::uint32_t b) {
        const std::uint32_t dst = lds_base + b * kTileBytes + wave_u * 4096;
        const std::size_t first = static_cast<std::size_t>(tile) * NT * a.row_stride;
        if ((tile + 1) * NT <= a.rows) {
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                lds_load16(src[i] + first * kRowBytes, dst + 1024 * i);
            }
            if (L2 && wave_u == 0) {
                lds_load4(xh_src + first, lds_base + 3 * kTileBytes + b * NT * 4);
            }
        } else {
            // a tile past the end of the rows repeats the last row (its distances are never used)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const std::uint32_t j = tile * NT + drow[i];
                const std::uint32_t x = (j < a.rows ? j : a.rows - 1) * a.row_stride;
                lds_load16(xq_bytes + static_cast<std::size_t>(x) * kRowBytes + dcol[i], dst + 1024 * i);
            }
            if (L2 && wave_u == 0) {
                const std::uint32_t j = tile * NT + lane;
                const std::uint32_t xr = (j < a.rows ? j : a.rows - 1) * a.row_stride;
                lds_load4(a.xh + xr, lds_base + 3 * kTileBytes + b * NT * 4);
            }
        }
    };
    // Loads complete in issue order: leaving the newest tile's loads (5 on wave 0 under L2, 4
    // otherwise) in flight means every older one has landed.
    const auto wait_older = [&]() {
        if (L2 && wave_u == 0) {
            __builtin_amdgcn_s_waitcnt(0x0F75);  // vmcnt(5)
        } else {
            __builtin_amdgcn_s_waitcnt(0x0F74);  // vmcnt(4)
        }
    };

    std::uint64_t* skey = reinterpret_cast<std::uint64_t*>(lds + 3 * kTileBytes + 3 * NT * 4 + wave * kStageBytes);
    std::uint32_t* squery = reinterpret_cast<std::uint32_t*>(skey + 64);
    std::uint32_t staged = 0;  // wave-uniform
    const auto flush = [&]() {
        wave_lds_fence();
        if (static_cast<std::uint32_t>(lane) < staged) {
            const std::uint32_t sq = squery[lane];
            const std::uint32_t slot = atomicAdd(a.cnt + sq, 1u);
            if (slot < a.cap) {
                a.buf[static_cast<std::size_t>(sq) * a.cap + slot] = skey[lane];
            }
        }
        wave_lds_fence();
        staged = 0;
    };

    __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0): query fragments, |q|^2, tau landed
    issue(t0, 0);
    issue(min(t0 + 1, last), 1);
    wait_older();
    __syncthreads();
    const std::uint32_t a0 = static_cast<std::uint32_t>(col) * kRowBytes +
                             16u * (static_cast<std::uint32_t>(half) ^ (static_cast<std::uint32_t>(col) & 15u));
    std::uint32_t tile_cb = 0;  // buffer of tile t
    for (std::uint32_t t = t0; t < t1; ++t) {
        const std::uint32_t tile_t = t;
        issue(min(t + 2, last), tile_cb == 0 ? 2 : tile_cb - 1);
        const std::uint32_t ab = a0 + tile_cb * kTileBytes;
        v16f acc[QB][NH];
#pragma unroll
        for (int h = 0; h < NH; ++h) {
            if constexpr (L2) {
#pragma unroll
                for (int b = 0; b < QB; ++b) {
#pragma unroll
                    for (int g = 0; g < 4; ++g) {
                        const float4 c = *reinterpret_cast<const float4*>(vals + tile_cb * NT + h * 32 + 8 * g + 4 * half);
                        acc[b][h][4 * g + 0] = c.x;
                        acc[b][h][4 * g + 1] = c.y;
                        acc[b][h][4 * g + 2] = c.z;
                        acc[b][h][4 * g + 3] = c.w;
                    }
                }
            } else {
#pragma unroll
                for (int b = 0; b < QB; ++b) {
                    acc[b][h] = v16f{};
                }
            }
        }
#pragma unroll
        for (int s = 0; s < KS; ++s) {
            const unsigned char* p = lds + (ab ^ (32u * s));
#pragma unroll
            for (int h = 0; h < NH; ++h) {
                const v8bf x = *reinterpret_cast<const v8bf*>(p + h * 32 * kRowBytes);
#pragma unroll
                for (int b = 0; b < QB; ++b) {
                    acc[b][h] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(x, bq[b][s], acc[b][h], 0, 0, 0);
                }
            }
        }
        if ((tile_t + 1) * NT <= a.rows) {
            VDX_DMA2_EPILOGUE(false)
        } else {
            VDX_DMA2_EPILOGUE(true)
        }
        wait_older();
        __syncthreads();
        tile_cb = tile_cb == 2 ? 0 : tile_cb + 1;
    }
    if (MODE != 0 && staged > 0) {
        flush();
    }
}

#undef VDX_DMA_EPILOGUE
#undef VDX_EPILOGUE

// Radix selection helpers for one 256-thread workgroup (4 waves). Each wave counts into its own
// 256-bin histogram (rows padded to 260 words so the waves' rows start on different banks); a
// digit that every active lane of the wave shares is added once, by one lane (the first passes
// over distances of one magnitude see mostly that case), others lane by lane.
constexpr int kSelWaves = 4;
constexpr int kSelRow = 260;

__device__ __forceinline__ void sel_count(std::uint32_t* hist, std::uint32_t digit) {
    const std::uint32_t d0 = __builtin_amdgcn_readfirstlane(digit);
    const std::uint64_t same = ballot(digit == d0);
    if (same == __builtin_amdgcn_read_exec()) {
        if (lanes_below(same) == 0) atomicAdd(&hist[d0], static_cast<std::uint32_t>(__popcll(same)));
    } else {
        atomicAdd(&hist[digit], 1u);
    }
}

// Wave 0 of the workgroup: the digit whose bin holds the want-th smallest counted key, and the
// rank within that bin, to sel[0] and sel[1] (want <= the number of keys counted).
__device__ __forceinline__ void sel_pick(const std::uint32_t* hist, std::uint32_t want, std::uint32_t* sel) {
    const int lane = lane_id();
    std::uint32_t c[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        c[j] = 0;
#pragma unroll
        for (int w = 0; w < kSelWaves; ++w) c[j] += hist[w * kSelRow + 4 * lane + j];
    }
    const std::uint32_t t = c[0] + c[1] + c[2] + c[3];
    std::uint32_t incl = t;
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) {
        const std::uint32_t v = __shfl_up(incl, off, kWave);
        if (lane >= off) incl += v;
    }
    const std::uint64_t reach = ballot(incl >= want);
    if (reach != 0 && lane == __builtin_ctzll(reach)) {
        std::uint32_t cum = incl - t;
        int j = 0;
        for (; j < 3; ++j) {
            if (cum + c[j] >= want) break;
            cum += c[j];
        }
        sel[0] = 4u * static_cast<std::uint32_t>(lane) + static_cast<std::uint32_t>(j);
        sel[1] = want - cum;
    }
}

// Per query (one workgroup): the `rank`-th smallest (1-based) of `count` (<= 8192) u32 sortable
// distances (dense sample row), by an MSB-first radix select, 8 bits per pass, over keys held in
// registers (32 per thread).
__global__ __launch_bounds__(256) void bf_threshold_kernel(ThresholdArgs a) {
    constexpr int kPer = 32;
    __shared__ std::uint32_t hist[kSelWaves * kSelRow];
    __shared__ std::uint32_t sel[2];
    const std::uint32_t q = blockIdx.x;
    const int wave = threadIdx.x >> 6;
    const float* row = a.dense + static_cast<std::size_t>(q) * a.dense_ld;
    std::uint32_t key[kPer];
#pragma unroll
    for (int i = 0; i < kPer; ++i) {
        const std::uint32_t idx = threadIdx.x + 256u * i;
        key[i] = idx < a.count ? sortable(row[idx]) : 0u;
    }
    std::uint32_t prefix = 0, mask = 0, want = a.rank;
    for (int shift = 24; shift >= 0; shift -= 8) {
        for (int i = threadIdx.x; i < kSelWaves * kSelRow; i += blockDim.x) hist[i] = 0;
        __syncthreads();
#pragma unroll
        for (int i = 0; i < kPer; ++i) {
            const std::uint32_t idx = threadIdx.x + 256u * i;
            if (idx < a.count && (key[i] & mask) == prefix) {
                sel_count(hist + wave * kSelRow, (key[i] >> shift) & 255u);
            }
        }
        __syncthreads();
        if (wave == 0) sel_pick(hist, want, sel);
        __syncthreads();
        prefix |= sel[0] << shift;
        mask |= 255u << shift;
        want = sel[1];
        __syncthreads();
    }
    if (threadIdx.x == 0) a.tau[q] = unsortable(prefix);
}

// Per query (one workgroup): from the candidate buffer keep the S smallest keys (exactly S,
// keys are unique) in the shortlist, and record the smallest GEMM distance left out: the
// (S+1)-th key's when the buffer held more than S, else tau (everything left out was > tau).
__global__ __launch_bounds__(256) void bf_shortlist_kernel(ShortlistArgs a) {
    __shared__ std::uint32_t hist[kSelWaves * kSelRow];
    __shared__ std::uint32_t sel[2], taken;
    const std::uint32_t q = blockIdx.x;
    const std::uint32_t total = a.cnt[q];
    const std::uint64_t* buf = a.buf + static_cast<std::size_t>(q) * a.cap;
    std::uint32_t* list = a.list + static_cast<std::size_t>(q) * a.S;
    if (total > a.cap) {  // overflowed: the exact scan redoes this query
        if (threadIdx.x == 0) {
            a.list_count[q] = 0;
            a.left[q] = -__int_as_float(0x7f800000);
            a.strict[q] = 1;
        }
        return;
    }
    if (total <= a.S) {
        for (std::uint32_t i = threadIdx.x; i < total; i += blockDim.x) {
            list[i] = key_id(buf[i]);
            a.approx[static_cast<std::size_t>(q) * a.S + i] = key_distance(buf[i]);
        }
        if (threadIdx.x == 0) {
            a.list_count[q] = total;
            a.left[q] = a.tau ? a.tau[q] : __int_as_float(0x7f800000);
            a.strict[q] = 0;  // left out means GEMM distance > tau
        }
        return;
    }
    std::uint64_t prefix = 0, mask = 0;
    std::uint32_t want = a.S;
    const int wave = threadIdx.x >> 6;
    for (int shift = 56; shift >= 0; shift -= 8) {
        for (int i = threadIdx.x; i < kSelWaves * kSelRow; i += blockDim.x) hist[i] = 0;
        __syncthreads();
        for (std::uint32_t i = threadIdx.x; i < total; i += blockDim.x) {
            const std::uint64_t v = buf[i];
            if ((v & mask) == prefix) sel_count(hist + wave * kSelRow, static_cast<std::uint32_t>(v >> shift) & 255u);
        }
        __syncthreads();
        if (wave == 0) sel_pick(hist, want, sel);
        __syncthreads();
        prefix |= static_cast<std::uint64_t>(sel[0]) << shift;
        mask |= 255ull << shift;
        want = sel[1];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        taken = 0;
    }
    __syncthreads();
    float my_left = __int_as_float(0x7f800000);
    for (std::uint32_t i = threadIdx.x; i < total; i += blockDim.x) {
        const std::uint64_t v = buf[i];
        if (v <= prefix) {
            const std::uint32_t at = atomicAdd(&taken, 1u);
            if (at < a.S) {  // exactly S keys qualify (keys are unique); the test keeps writes in bounds
                list[at] = key_id(v);
                a.approx[static_cast<std::size_t>(q) * a.S + at] = key_distance(v);
            }
        } else {
            my_left = fminf(my_left, key_distance(v));
        }
    }
    for (int off = 32; off > 0; off >>= 1) my_left = fminf(my_left, __shfl_xor(my_left, off, kWave));
    __shared__ float wave_left[4];
    if ((threadIdx.x & 63) == 0) wave_left[threadIdx.x >> 6] = my_left;
    __syncthreads();
    if (threadIdx.x == 0) {
        float m = wave_left[0];
        for (int w = 1; w < static_cast<int>(blockDim.x >> 6); ++w) m = fminf(m, wave_left[w]);
        a.list_count[q] = a.S;
        a.left[q] = m;
        a.strict[q] = 1;  // left out means GEMM distance >= left
    }
}

// One wave per query: exact-order distances of the shortlist, sorted by (distance, id); the
// first k are the result. Then the certification test of the file comment.
template <int NB, int N8, int TAIL, int R>
__global__ __launch_bounds__(256) void bf_rerank_kernel(RerankArgs a) {
    extern __shared__ __align__(16) unsigned char lds[];
    const int lane = lane_id();
    const int wave = threadIdx.x >> 6;
    const int m = lane & 7;
    const std::uint32_t q = blockIdx.x * (blockDim.x >> 6) + wave;
    if (q >= a.nq) {
        return;
    }
    unsigned char* base = lds + static_cast<std::size_t>(wave) * a.lds_per_wave;
    std::uint64_t* keys = reinterpret_cast<std::uint64_t*>(base);
    std::uint32_t* fid = reinterpret_cast<std::uint32_t*>(keys + a.sort_cap);
    float* fd = reinterpret_cast<float*>(fid + 8 * R);
    const int nb = static_cast<int>(a.nb), n8 = static_cast<int>(a.n8);
    const int tail = static_cast<int>(a.tail), first_tail = static_cast<int>(a.first_tail);
    const RowSource src{a.vectors, a.ld, a.l2};

    const float* qp = a.queries + static_cast<std::size_t>(q) * a.qld;
    float4 qb[NB > 0 ? NB : 1];
    float4 q8[N8 > 0 ? N8 : 1];
    float qt[8];
#pragma unroll
    for (int i = 0; i < NB; ++i) {
        qb[i] = make_float4(0.f, 0.f, 0.f, 0.f);
        if (i < nb) {
            const float* p = qp + 32 * i + 4 * m;
            qb[i] = make_float4(p[0], p[1], p[2], p[3]);
        }
    }
#pragma unroll
    for (int t = 0; t < N8; ++t) {
        q8[t] = make_float4(0.f, 0.f, 0.f, 0.f);
        if (t < n8 && m < 2) {
            const float* p = qp + 32 * nb + 8 * t + 4 * m;
            q8[t] = make_float4(p[0], p[1], p[2], p[3]);
        }
    }
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        qt[e] = (e < TAIL && e < tail) ? qp[first_tail + e] : 0.f;
    }

    const std::uint32_t cnt = a.list_count[q];
    const std::uint32_t* list = a.list + static_cast<std::size_t>(q) * a.S;
    const float* approx = a.approx + static_cast<std::size_t>(q) * a.S;
    float max_err = 0.f;
    for (std::uint32_t c0 = 0; c0 < cnt; c0 += 8 * R) {
        const std::uint32_t fc = min(8u * R, cnt - c0);
        if (static_cast<std::uint32_t>(lane) < fc) {
            fid[lane] = list[c0 + lane];
        }
        wave_lds_fence();
        distances<NB, N8, TAIL, R>(src, fid, fc, qb, q8, qt, nb, n8, tail, first_tail, fd);
        wave_lds_fence();
        if (static_cast<std::uint32_t>(lane) < fc) {
            keys[c0 + lane] = make_key(fd[lane], fid[lane]);
            max_err = fmaxf(max_err, fabsf(fd[lane] - approx[c0 + lane]));
        }
        wave_lds_fence();
    }
    for (std::uint32_t i = cnt + lane; i < a.sort_cap; i += kWave) {
        keys[i] = ~0ull;
    }
    wave_lds_fence();
    // Bitonic sort of sort_cap (a power of two) keys.
    for (std::uint32_t size = 2; size <= a.sort_cap; size <<= 1) {
        for (std::uint32_t stride = size >> 1; stride > 0; stride >>= 1) {
            for (std::uint32_t i = lane; i < a.sort_cap; i += kWave) {
                const std::uint32_t j = i ^ stride;
                if (j > i) {
                    const std::uint64_t x = keys[i], y = keys[j];
                    const bool ascending = (i & size) == 0;
                    if ((x > y) == ascending) {
                        keys[i] = y;
                        keys[j] = x;
                    }
                }
            }
            wave_lds_fence();
        }
    }
    const std::uint32_t want = min(a.k, a.n);
    const std::uint32_t got = min(want, cnt);
    Neighbor* out = a.out + static_cast<std::size_t>(q) * a.out_stride;
    for (std::uint32_t i = lane; i < got; i += kWave) {
        out[i] = Neighbor{key_distance(keys[i]), key_id(keys[i])};
    }
    // Certification (file comment, step 6).
    for (int off = 32; off > 0; off >>= 1) max_err = fmaxf(max_err, __shfl_xor(max_err, off, kWave));
    if (lane == 0) {
        a.out_count[q] = got;
        atomicMax(reinterpret_cast<unsigned int*>(a.max_err), __float_as_uint(max_err));
        bool ok = got == want;
        // Exact GEMM: integral operands and lossless quantization; every partial sum of
        // products is an integer of magnitude <= |q||x| < 2^24 (Cauchy-Schwarz), |q|^2 + |x|^2 <
        // 2^24, and the distance itself is then rounded once: exact when below 2^24, and >= 2^24
        // (so beyond a k-th distance below 2^24) otherwise. Under L2 the LDS-fed BF16 kernel
        // accumulates from -|x|^2 / 2: its partial sums are multiples of 1/2 of magnitude <=
        // |q||x| + |x|^2 / 2, exact below 2^23. The S smallest GEMM keys then contain the k
        // smallest exact ones: certified without a margin.
        const float two24 = 16777216.0f;
        const bool exact_gemm = a.x_integral && a.q_integral[q] && got == want && want > 0 &&
                                a.q_qnorm[q] * a.x_max_qnorm < two24 - 2.0f &&
                                (!a.l2 || (a.q_norm2[q] + a.x_max_norm2 < two24 &&
                                           a.q_qnorm[q] * a.x_max_qnorm + 0.5f * a.x_max_norm2 < 0.5f * two24 - 2.0f)) &&
                                fabsf(key_distance(keys[want - 1])) < two24;
        if (ok && !exact_gemm) {
            const float left = a.left[q];
            if (left != __int_as_float(0x7f800000)) {
                const float dk = key_distance(keys[want - 1]);
                const float qnorm = a.q_qnorm[q];   // |q~|
                const float qerr = a.q_qerr[q];     // |q~ - q|
                const float qlen = sqrtf(a.q_norm2[q]);
                const float e_dot = qnorm * a.x_max_qerr + qerr * a.x_max_norm +
                                    a.gamma_k * qnorm * a.x_max_qnorm;
                float lower;
                if (a.l2) {
                    // gamma_k |x|^2: the accumulation from -|x|^2 / 2 of the LDS-fed kernel
                    const float e = 2.0f * e_dot + a.gamma_k * a.x_max_norm2 + (a.gamma_d + 2.4e-7f) *
                                    (a.q_norm2[q] + a.x_max_norm2 + 2.0f * qnorm * a.x_max_qnorm);
                    lower = left * (1.0f - a.gamma_d) - e * (1.0f + a.gamma_d);
                } else {
                    const float e = e_dot + a.gamma_d * qlen * a.x_max_norm +
                                    2.4e-7f * (1.0f + fabsf(left) + e_dot);
                    lower = left - e;
                }
                lower = lower * (lower >= 0.f ? 0.99999f : 1.00001f) - 1e-30f;  // margin
                ok = a.strict[q] ? lower > dk : lower >= dk;
            }
        }
        a.fail[q] = ok ? 0 : 1;
    }
}

// Copies the rows listed in idx (count of them) from src to consecutive rows of dst.
__global__ void bf_gather_rows_kernel(const float* src, const std::uint32_t* idx, std::uint32_t count,
                                      std::uint32_t dim, float* dst) {
    const std::uint32_t r = blockIdx.x;
    if (r >= count) return;
    const float* from = src + static_cast<std::size_t>(idx[r]) * dim;
    for (std::uint32_t e = threadIdx.x; e < dim; e += blockDim.x) {
        dst[static_cast<std::size_t>(r) * dim + e] = from[e];
    }
}

// Writes result row r of `part` (rows of `width`) back to row idx[r] of out (rows of stride).
__global__ void bf_scatter_results_kernel(const Neighbor* part, const std::uint32_t* part_count,
                                          const std::uint32_t* idx, std::uint32_t count,
                                          std::uint32_t width, Neighbor* out,
                                          std::uint64_t out_stride, std::uint32_t* out_count) {
    const std::uint32_t r = blockIdx.x;
    if (r >= count) return;
    const std::uint32_t q = idx[r];
    const std::uint32_t c = part_count[r];
    for (std::uint32_t i = threadIdx.x; i < c; i += blockDim.x) {
        out[static_cast<std::size_t>(q) * out_stride + i] = part[static_cast<std::size_t>(r) * width + i];
    }
    if (threadIdx.x == 0) out_count[q] = c;
}

hipError_t launch_bf_gather_rows(const float* src, const std::uint32_t* idx, std::uint32_t count,
                                 std::uint32_t dim, float* dst, hipStream_t stream) {
    hipLaunchKernelGGL(bf_gather_rows_kernel, dim3(count), dim3(128), 0, stream, src, idx, count, dim, dst);
    return hipGetLastError();
}

hipError_t launch_bf_scatter_results(const Neighbor* part, const std::uint32_t* part_count,
                                     const std::uint32_t* idx, std::uint32_t count,
                                     std::uint32_t width, Neighbor* out, std::uint64_t out_stride,
                                     std::uint32_t* out_count, hipStream_t stream) {
    hipLaunchKernelGGL(bf_scatter_results_kernel, dim3(count), dim3(128), 0, stream, part, part_count,
                       idx, count, width, out, out_stride, out_count);
    return hipGetLastError();
}

hipError_t launch_bf_prepare(const PrepArgs& args, hipStream_t stream) {
    const unsigned blocks = (args.rows + 3) / 4;
    hipLaunchKernelGGL(bf_prepare_rows, dim3(blocks), dim3(256), 0, stream, args);
    return hipGetLastError();
}

std::size_t bf_gemm_dma_lds_bytes() { return 3 * 64 * 256 + 3 * 64 * sizeof(float) + 8 * kStageBytes; }
std::size_t bf_gemm_dma2_lds_bytes() { return 3 * 64 * 256 + 3 * 64 * sizeof(float) + 4 * kStageBytes; }

std::size_t bf_gemm_lds_bytes(int prec, std::uint32_t kp, int waves) {
    const std::size_t eb = prec == kPrecF32 ? 4 : (prec == kPrecBF16 ? 2 : 1);
    const std::size_t nt = prec == kPrecF32 ? 32 : 64;
    return 2 * nt * (kp * eb + 16) + 4 * nt * sizeof(float) + static_cast<std::size_t>(waves) * kStageBytes;
}

hipError_t launch_bf_gemm(const GemmArgs& args, unsigned qtiles, unsigned slabs, int prec, int waves,
                          hipStream_t stream, int shape) {
    if (waves != 8) return hipErrorInvalidValue;
    if (args.mode != 0 && args.row_stride != 1) return hipErrorInvalidValue;  // the filter's ids are rows
    const dim3 grid(qtiles, slabs);
    const std::size_t lds = bf_gemm_lds_bytes(prec, args.kp, waves);
    const dim3 block(64 * waves);
#define VDX_GEMM_MODES(P, F)                                                                      \
    if (args.mode == 0) {                                                                         \
        if (args.l2) {                                                                            \
            hipLaunchKernelGGL((bf_gemm_kernel<P, 8, 0, true, F>), grid, block, lds, stream, args);  \
        } else {                                                                                  \
            hipLaunchKernelGGL((bf_gemm_kernel<P, 8, 0, false, F>), grid, block, lds, stream, args); \
        }                                                                                         \
    } else {                                                                                      \
        if (args.l2) {                                                                            \
            hipLaunchKernelGGL((bf_gemm_kernel<P, 8, 1, true, F>), grid, block, lds, stream, args);  \
        } else {                                                                                  \
            hipLaunchKernelGGL((bf_gemm_kernel<P, 8, 1, false, F>), grid, block, lds, stream, args); \
        }                                                                                         \
    }
#define VDX_GEMM_LAUNCH(P)                                                                        \
    if (args.ks == bf_gemm_steps(P)) {                                                            \
        VDX_GEMM_MODES(P, true)                                                                   \
    } else {                                                                                      \
        VDX_GEMM_MODES(P, false)                                                                  \
    }
    if (prec == kPrecBF16 && args.ks == bf_gemm_steps(kPrecBF16) && args.kp == 128 && shape == 1) {
        const std::size_t lds2 = bf_gemm_dma2_lds_bytes();
        const dim3 block2(256);
        if (args.mode == 0) {
            if (args.l2) {
                hipLaunchKernelGGL((bf_gemm_bf16_dma2_kernel<0, true>), grid, block2, lds2, stream, args);
            } else {
                hipLaunchKernelGGL((bf_gemm_bf16_dma2_kernel<0, false>), grid, block2, lds2, stream, args);
            }
        } else if (args.l2) {
            hipLaunchKernelGGL((bf_gemm_bf16_dma2_kernel<1, true>), grid, block2, lds2, stream, args);
        } else {
            hipLaunchKernelGGL((bf_gemm_bf16_dma2_kernel<1, false>), grid, block2, lds2, stream, args);
        }
    } else if (prec == kPrecBF16 && args.ks == bf_gemm_steps(kPrecBF16) && args.kp == 128) {
        const std::size_t dma_lds = bf_gemm_dma_lds_bytes();
        if (args.mode == 0) {
            if (args.l2) {
                hipLaunchKernelGGL((bf_gemm_bf16_dma_kernel<0, true>), grid, block, dma_lds, stream, args);
            } else {
                hipLaunchKernelGGL((bf_gemm_bf16_dma_kernel<0, false>), grid, block, dma_lds, stream, args);
            }
        } else if (args.l2) {
            hipLaunchKernelGGL((bf_gemm_bf16_dma_kernel<1, true>), grid, block, dma_lds, stream, args);
        } else {
            hipLaunchKernelGGL((bf_gemm_bf16_dma_kernel<1, false>), grid, block, dma_lds, stream, args);
        }
    } else if (prec == kPrecF32) {
        VDX_GEMM_LAUNCH(kPrecF32)
    } else if (prec == kPrecBF16) {
        VDX_GEMM_LAUNCH(kPrecBF16)
    } else {
        VDX_GEMM_LAUNCH(kPrecFP8)
    }
#undef VDX_GEMM_MODES
#undef VDX_GEMM_LAUNCH
    return hipGetLastError();
}

hipError_t launch_bf_threshold(const ThresholdArgs& args, unsigned nq, hipStream_t stream) {
    if (args.count > 8192 || args.rank == 0 || args.rank > args.count) return hipErrorInvalidValue;  // keys in registers
    hipLaunchKernelGGL(bf_threshold_kernel, dim3(nq), dim3(256), 0, stream, args);
    return hipGetLastError();
}

hipError_t launch_bf_shortlist(const ShortlistArgs& args, unsigned nq, hipStream_t stream) {
    hipLaunchKernelGGL(bf_shortlist_kernel, dim3(nq), dim3(256), 0, stream, args);
    return hipGetLastError();
}

hipError_t launch_bf_rerank(const RerankArgs& args, unsigned nq, std::size_t lds_per_wave,
                            hipStream_t stream, int shape) {
    const unsigned waves = 4;
    const dim3 grid((nq + waves - 1) / waves);
    const std::size_t lds = lds_per_wave * waves;
    if (shape == 0) {
        hipLaunchKernelGGL((bf_rerank_kernel<4, 0, 0, 4>), grid, dim3(waves * 64), lds, stream, args);
    } else if (shape == 1) {
        hipLaunchKernelGGL((bf_rerank_kernel<3, 0, 4, 4>), grid, dim3(waves * 64), lds, stream, args);
    } else {
        hipLaunchKernelGGL((bf_rerank_kernel<8, 3, 7, 1>), grid, dim3(waves * 64), lds, stream, args);
    }
    return hipGetLastError();
}

}  // namespace vdx::dev
