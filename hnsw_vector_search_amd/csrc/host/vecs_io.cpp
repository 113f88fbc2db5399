// TEXMEX .fvecs / .ivecs readers (vecdex src/io.cpp:30-87, include/vecdex/io.hpp:10-20): each
// record is a little-endian int32 dimension d and d 4-byte values; every record has the same d.
// Nothing in the file is trusted: the row count comes from the file size, so a corrupted file
// cannot ask for an allocation larger than itself, and every record's dimension is checked.
// Errors are std::runtime_error with the path first, as vecdex reports them.

#include <cstdio>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <system_error>

#include "capi_common.hpp"

namespace vdx {
namespace {

static_assert(sizeof(float) == 4 && sizeof(std::int32_t) == 4, "4-byte elements");

[[noreturn]] void fail(const std::string& path, const std::string& what) {
    throw std::runtime_error(path + ": " + what);
}

void read_exact(std::ifstream& in, void* dst, std::size_t bytes, const std::string& path) {
    in.read(static_cast<char*>(dst), static_cast<std::streamsize>(bytes));
    if (in.gcount() != static_cast<std::streamsize>(bytes)) {
        fail(path, "unexpected end of file");
    }
}

struct VecsShape {
    std::uint64_t rows = 0, dim = 0;
};

VecsShape vecs_shape(const std::string& path, std::ifstream& in) {
    std::error_code ec;
    const std::uintmax_t file_size = std::filesystem::file_size(path, ec);
    if (ec) {
        fail(path, ec.message());
    }
    if (file_size == 0) {
        fail(path, "file is empty");
    }
    in.open(path, std::ios::binary);
    if (!in) {
        fail(path, "cannot open file");
    }
    std::int32_t dim = 0;
    read_exact(in, &dim, sizeof dim, path);
    if (dim <= 0) {
        fail(path, fmt("invalid dimension %d in the first record", dim));
    }
    const std::uintmax_t record = 4u + static_cast<std::uintmax_t>(dim) * 4u;
    if (file_size % record != 0) {
        fail(path, fmt("size %llu is not a multiple of the record size %llu (dimension %d); "
                       "the file is truncated or not in this format",
                       static_cast<unsigned long long>(file_size), static_cast<unsigned long long>(record),
                       dim));
    }
    return {static_cast<std::uint64_t>(file_size / record), static_cast<std::uint64_t>(dim)};
}

}  // namespace
}  // namespace vdx

using vdx::guarded;
using vdx::require;

extern "C" {

int vdx_vecs_info(const char* path, uint64_t* rows, uint64_t* dim) {
    return guarded([&] {
        require(path != nullptr && rows != nullptr && dim != nullptr, "vdx_vecs_info: null argument");
        std::ifstream in;
        const vdx::VecsShape s = vdx::vecs_shape(path, in);
        *rows = s.rows;
        *dim = s.dim;
    });
}

int vdx_vecs_read(const char* path, void* dst, uint64_t rows, uint64_t dim) {
    return guarded([&] {
        require(path != nullptr && (rows == 0 || dst != nullptr), "vdx_vecs_read: null argument");
        std::ifstream in;
        const vdx::VecsShape s = vdx::vecs_shape(path, in);
        require(s.rows == rows && s.dim == dim, "vdx_vecs_read: the file's shape differs from the buffer's");
        in.seekg(0);
        auto* out = static_cast<unsigned char*>(dst);
        for (std::uint64_t i = 0; i < rows; ++i) {
            std::int32_t record_dim = 0;
            vdx::read_exact(in, &record_dim, sizeof record_dim, path);
            if (record_dim < 0 || static_cast<std::uint64_t>(record_dim) != dim) {
                vdx::fail(path, vdx::fmt("record %llu has dimension %d, expected %llu",
                                         static_cast<unsigned long long>(i), record_dim,
                                         static_cast<unsigned long long>(dim)));
            }
            vdx::read_exact(in, out + i * dim * 4, dim * 4, path);
        }
    });
}

}  // extern "C"
