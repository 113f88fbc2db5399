#pragma once

// Fill and copy kernels (fill_copy.hip).

#include <hip/hip_runtime.h>

#include <cstdint>

namespace vdx::dev {

hipError_t launch_fill_u32(std::uint32_t* dst, std::uint32_t value, std::uint64_t count, hipStream_t s);
// Rows of row_len u32 from src scattered to dst rows ids[0..rows) (ids valid row indices of dst).
hipError_t launch_scatter_rows(std::uint32_t* dst, const std::uint32_t* src, const std::uint32_t* ids,
                               std::uint64_t rows, std::uint32_t row_len, hipStream_t s);
// dst row r (width floats, row stride dst_ld) = src row ids[r] (row stride src_ld); ids valid
// row indices of src.
hipError_t launch_gather_rows(float* dst, std::uint32_t dst_ld, const float* src, std::uint32_t src_ld,
                              const std::uint32_t* ids, std::uint64_t rows, std::uint32_t width, hipStream_t s);
// Rows [first, first + rows) of src (the first 128 of each row's ld floats) packed as 128 bytes
// each into dst (rows at dst + row * 128, the beam's lane order); *bad becomes nonzero if an
// element is not an integer in [0, 255]. ld >= 128, a multiple of 4.
hipError_t launch_pack_rows_u8(unsigned char* dst, const float* src, std::uint32_t ld, std::uint64_t first,
                               std::uint64_t rows, std::uint32_t* bad, hipStream_t s);
// Random gathers of `per_wave` rows (rounded up to whole rounds) of row_bytes (16..1024, a
// multiple of 16 dividing 1024) from rows [0, rows) of `table` (16-byte aligned) by each of
// `waves` waves (a multiple of 4); sink[waves] receives one word per wave.
hipError_t launch_gather_probe(const void* table, std::uint32_t rows, std::uint32_t row_bytes,
                               std::uint32_t per_wave, std::uint32_t waves, std::uint32_t* sink, hipStream_t s);
// dst and src 16-byte aligned (device allocations are).
hipError_t launch_copy_bytes(void* dst, const void* src, std::uint64_t bytes, hipStream_t s);
// Writes `epoch` to *flag if any of the `count` floats at `queries` is not an integer in
// [0, 255] (the exactness condition of the byte rows' integer arithmetic), leaving it alone
// otherwise; with `zero` non-null it also clears that word (a work counter of the launches after
// it).
hipError_t launch_query_bytes_check(const float* queries, std::uint64_t count, std::uint32_t* flag,
                                    std::uint32_t epoch, std::uint32_t* zero, hipStream_t s);

}  // namespace vdx::dev
