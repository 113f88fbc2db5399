// Plain fill and copy kernels: clears that a captured HIP graph may replay (a memset node is
// not used there), and a device-to-device copy of known size against which HBM traffic
// counters are calibrated.

#include <hip/hip_runtime.h>

#include <algorithm>

#include <cstdint>

#include "fill_copy.hpp"

namespace vdx::dev {

__global__ __launch_bounds__(256) void fill_u32_kernel(std::uint32_t* dst, std::uint32_t value,
                                                       std::uint64_t count) {
    const std::uint64_t stride = static_cast<std::uint64_t>(gridDim.x) * blockDim.x;
    for (std::uint64_t i = static_cast<std::uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < count;
         i += stride) {
        dst[i] = value;
    }
}

// 16-byte copies for the aligned body, byte copies for the rest.
__global__ __launch_bounds__(256) void copy_bytes_kernel(unsigned char* dst, const unsigned char* src,
                                                         std::uint64_t bytes) {
    const std::uint64_t stride = static_cast<std::uint64_t>(gridDim.x) * blockDim.x;
    const std::uint64_t first = static_cast<std::uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    const std::uint64_t n16 = bytes / 16;
    for (std::uint64_t i = first; i < n16; i += stride) {
        reinterpret_cast<uint4*>(dst)[i] = reinterpret_cast<const uint4*>(src)[i];
    }
    for (std::uint64_t i = n16 * 16 + first; i < bytes; i += stride) {
        dst[i] = src[i];
    }
}

// dst[ids[r] * row_len + c] = src[r * row_len + c]: rows of u32 scattered by index.
__global__ __launch_bounds__(256) void scatter_rows_kernel(std::uint32_t* dst, const std::uint32_t* src,
                                                           const std::uint32_t* ids, std::uint64_t rows,
                                                           std::uint32_t row_len) {
    const std::uint64_t total = rows * row_len;
    const std::uint64_t stride = static_cast<std::uint64_t>(gridDim.x) * blockDim.x;
    for (std::uint64_t i = static_cast<std::uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < total;
         i += stride) {
        const std::uint64_t r = i / row_len;
        dst[static_cast<std::uint64_t>(ids[r]) * row_len + (i - r * row_len)] = src[i];
    }
}

__global__ __launch_bounds__(256) void gather_rows_kernel(float* dst, std::uint32_t dst_ld, const float* src,
                                                          std::uint32_t src_ld, const std::uint32_t* ids,
                                                          std::uint64_t rows, std::uint32_t width) {
    const std::uint64_t total = rows * width;
    const std::uint64_t stride = static_cast<std::uint64_t>(gridDim.x) * blockDim.x;
    for (std::uint64_t i = static_cast<std::uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < total;
         i += stride) {
        const std::uint64_t r = i / width;
        const std::uint64_t c = i - r * width;
        dst[r * dst_ld + c] = src[static_cast<std::uint64_t>(ids[r]) * src_ld + c];
    }
}

// Rows [first, first + rows) of `src` (128 floats used of each row of ld) as 128 bytes each in
// dst, in the beam's lane order: byte 16m + 4i + j of a row holds element 32i + 4m + j. One
// thread per 4 elements; *bad is set when an element is not an integer in [0, 255] (the copy is
// then of no use and is dropped by the caller).
__global__ __launch_bounds__(256) void pack_rows_u8_kernel(unsigned char* dst, const float* src, std::uint32_t ld,
                                                           std::uint64_t first, std::uint64_t rows,
                                                           std::uint32_t* bad) {
    const std::uint64_t total = rows * 32;
    const std::uint64_t stride = static_cast<std::uint64_t>(gridDim.x) * blockDim.x;
    for (std::uint64_t t = static_cast<std::uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; t < total;
         t += stride) {
        const std::uint64_t row = first + t / 32;
        const std::uint32_t quad = static_cast<std::uint32_t>(t % 32);
        const std::uint32_t m = quad / 4, i = quad % 4;
        const float* p = src + row * ld + 32 * i + 4 * m;
        std::uint32_t word = 0;
        bool ok = true;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const float v = p[j];
            ok = ok && v >= 0.0f && v <= 255.0f && v == __builtin_rintf(v);
            word |= (ok ? static_cast<std::uint32_t>(v) : 0u) << (8 * j);
        }
        if (!ok) atomicOr(bad, 1u);
        *reinterpret_cast<std::uint32_t*>(dst + row * 128 + 16 * m + 4 * i) = word;
    }
}

namespace {
unsigned blocks_for(std::uint64_t items) {
    const std::uint64_t b = (items + 255) / 256;
    return static_cast<unsigned>(b < 1 ? 1 : (b > 8192 ? 8192 : b));
}
}  // namespace

// A wave with any float that is not an integer in [0, 255] stamps the flag (one store per wave;
// every writer writes the same value, so the flag needs no clearing between launches).
__global__ __launch_bounds__(256) void query_bytes_check_kernel(const float* q, std::uint64_t count,
                                                                std::uint32_t* flag, std::uint32_t epoch,
                                                                std::uint32_t* zero) {
    if (zero != nullptr && blockIdx.x == 0 && threadIdx.x == 0) {
        *zero = 0u;
    }
    bool bad = false;
    const std::uint64_t step = static_cast<std::uint64_t>(gridDim.x) * blockDim.x * 4;
    for (std::uint64_t i = (static_cast<std::uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x) * 4; i < count;
         i += step) {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            if (i + j < count) {
                const float v = q[i + j];
                bad = bad || !(v >= 0.0f && v <= 255.0f && v == __builtin_floorf(v));
            }
        }
    }
    if (__ballot(bad) != 0 && __lane_id() == 0) {
        *flag = epoch;
    }
}

hipError_t launch_fill_u32(std::uint32_t* dst, std::uint32_t value, std::uint64_t count, hipStream_t s) {
    if (count == 0) return hipSuccess;
    hipLaunchKernelGGL(fill_u32_kernel, dim3(blocks_for(count)), dim3(256), 0, s, dst, value, count);
    return hipGetLastError();
}

// Random row gathers (the access of the beam's row and link-block reads, without the chain of
// dependent loads between them): each wave reads `per_wave` rows of row_bytes at ids drawn by a
// hash, row_bytes / 16 lanes per row (16 bytes a lane), four rows per lane group in flight, and
// leaves the XOR of everything it read in sink[wave] (so no load is dead).
__device__ __forceinline__ std::uint32_t probe_hash(std::uint32_t x) {
    x ^= x >> 16;
    x *= 0x7FEB352Du;
    x ^= x >> 15;
    x *= 0x846CA68Bu;
    x ^= x >> 16;
    return x;
}

__global__ __launch_bounds__(256) void gather_probe_kernel(const unsigned char* table, std::uint32_t rows,
                                                           std::uint32_t row_bytes, std::uint32_t per_wave,
                                                           std::uint32_t* sink) {
    constexpr std::uint32_t kRR = 4;
    const std::uint32_t lane = threadIdx.x & 63u;
    const std::uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const std::uint32_t lpr = row_bytes / 16;  // lanes per row, 1..64
    const std::uint32_t groups = 64 / lpr;      // rows per load instruction
    const std::uint32_t g = lane / lpr;
    const std::uint32_t m = lane % lpr;
    const bool active = g < groups;
    std::uint32_t acc = 0;
    for (std::uint32_t it = 0; it < per_wave; it += groups * kRR) {
        uint4 v[kRR];
#pragma unroll
        for (std::uint32_t r = 0; r < kRR; ++r) {
            const std::uint32_t h = probe_hash(wave * 0x9E3779B9u + (it + r * groups + g) * 0x85EBCA6Bu + 0x632BE5ABu);
            const std::uint32_t id = static_cast<std::uint32_t>((static_cast<std::uint64_t>(h) * rows) >> 32);
            v[r] = active ? *reinterpret_cast<const uint4*>(table + static_cast<std::size_t>(id) * row_bytes + 16 * m)
                          : make_uint4(0u, 0u, 0u, 0u);
        }
#pragma unroll
        for (std::uint32_t r = 0; r < kRR; ++r) {
            acc ^= v[r].x ^ v[r].y ^ v[r].z ^ v[r].w;
        }
    }
    for (int off = 32; off > 0; off >>= 1) {
        acc ^= static_cast<std::uint32_t>(__shfl_xor(static_cast<int>(acc), off, 64));
    }
    if (lane == 0) {
        sink[wave] = acc;
    }
}

hipError_t launch_gather_probe(const void* table, std::uint32_t rows, std::uint32_t row_bytes,
                               std::uint32_t per_wave, std::uint32_t waves, std::uint32_t* sink, hipStream_t s) {
    if (waves == 0 || waves % 4 != 0 || rows == 0 || row_bytes < 16 || row_bytes > 1024 || row_bytes % 16 != 0 ||
        (64 % (row_bytes / 16)) != 0 || reinterpret_cast<std::uintptr_t>(table) % 16 != 0) {
        return hipErrorInvalidValue;
    }
    hipLaunchKernelGGL(gather_probe_kernel, dim3(waves / 4), dim3(256), 0, s,
                       static_cast<const unsigned char*>(table), rows, row_bytes, per_wave, sink);
    return hipGetLastError();
}

hipError_t launch_scatter_rows(std::uint32_t* dst, const std::uint32_t* src, const std::uint32_t* ids,
                               std::uint64_t rows, std::uint32_t row_len, hipStream_t s) {
    if (rows == 0 || row_len == 0) return hipSuccess;
    hipLaunchKernelGGL(scatter_rows_kernel, dim3(blocks_for(rows * row_len)), dim3(256), 0, s, dst, src, ids,
                       rows, row_len);
    return hipGetLastError();
}

hipError_t launch_gather_rows(float* dst, std::uint32_t dst_ld, const float* src, std::uint32_t src_ld,
                              const std::uint32_t* ids, std::uint64_t rows, std::uint32_t width, hipStream_t s) {
    if (rows == 0 || width == 0) return hipSuccess;
    hipLaunchKernelGGL(gather_rows_kernel, dim3(blocks_for(rows * width)), dim3(256), 0, s, dst, dst_ld, src, src_ld,
                       ids, rows, width);
    return hipGetLastError();
}

hipError_t launch_pack_rows_u8(unsigned char* dst, const float* src, std::uint32_t ld, std::uint64_t first,
                               std::uint64_t rows, std::uint32_t* bad, hipStream_t s) {
    if (rows == 0) return hipSuccess;
    if (ld < 128 || ld % 4 != 0) return hipErrorInvalidValue;
    hipLaunchKernelGGL(pack_rows_u8_kernel, dim3(blocks_for(rows * 32)), dim3(256), 0, s, dst, src, ld, first, rows,
                       bad);
    return hipGetLastError();
}

hipError_t launch_copy_bytes(void* dst, const void* src, std::uint64_t bytes, hipStream_t s) {
    if (bytes == 0) return hipSuccess;
    const bool aligned = (reinterpret_cast<std::uintptr_t>(dst) % 16 == 0) &&
                         (reinterpret_cast<std::uintptr_t>(src) % 16 == 0);
    if (!aligned) return hipErrorInvalidValue;
    hipLaunchKernelGGL(copy_bytes_kernel, dim3(blocks_for(bytes / 16 + 1)), dim3(256), 0, s,
                       static_cast<unsigned char*>(dst), static_cast<const unsigned char*>(src), bytes);
    return hipGetLastError();
}

hipError_t launch_query_bytes_check(const float* queries, std::uint64_t count, std::uint32_t* flag,
                                    std::uint32_t epoch, std::uint32_t* zero, hipStream_t s) {
    const std::uint64_t blocks = std::max<std::uint64_t>(1, std::min<std::uint64_t>(1024, (count + 1023) / 1024));
    hipLaunchKernelGGL(query_bytes_check_kernel, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, s, queries, count,
                       flag, epoch, zero);
    return hipGetLastError();
}

}  // namespace vdx::dev
