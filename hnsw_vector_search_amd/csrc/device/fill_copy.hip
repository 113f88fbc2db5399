// Plain fill and copy kernels: clears that a captured HIP graph may replay (a memset node is
// not used there), and a device-to-device copy of known size against which HBM traffic
// counters are calibrated.

#include <hip/hip_runtime.h>

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

namespace {
unsigned blocks_for(std::uint64_t items) {
    const std::uint64_t b = (items + 255) / 256;
    return static_cast<unsigned>(b < 1 ? 1 : (b > 8192 ? 8192 : b));
}
}  // namespace

hipError_t launch_fill_u32(std::uint32_t* dst, std::uint32_t value, std::uint64_t count, hipStream_t s) {
    if (count == 0) return hipSuccess;
    hipLaunchKernelGGL(fill_u32_kernel, dim3(blocks_for(count)), dim3(256), 0, s, dst, value, count);
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

hipError_t launch_copy_bytes(void* dst, const void* src, std::uint64_t bytes, hipStream_t s) {
    if (bytes == 0) return hipSuccess;
    const bool aligned = (reinterpret_cast<std::uintptr_t>(dst) % 16 == 0) &&
                         (reinterpret_cast<std::uintptr_t>(src) % 16 == 0);
    if (!aligned) return hipErrorInvalidValue;
    hipLaunchKernelGGL(copy_bytes_kernel, dim3(blocks_for(bytes / 16 + 1)), dim3(256), 0, s,
                       static_cast<unsigned char*>(dst), static_cast<const unsigned char*>(src), bytes);
    return hipGetLastError();
}

}  // namespace vdx::dev
