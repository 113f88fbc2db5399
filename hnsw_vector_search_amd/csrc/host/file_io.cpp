// File helpers shared by FORMAT v1 and v2 (docs/FORMAT.md "Saving", src/hnsw_io.cpp,
// src/file_sync.cpp): the replace is atomic (rename within one directory) and durable (the file,
// then its directory, flushed with fsync).

#include "file_io.hpp"

#include <fcntl.h>
#include <unistd.h>

#include <filesystem>
#include <limits>
#include <random>
#include <stdexcept>
#include <system_error>

#include "validate.hpp"

namespace vdx {
namespace {

namespace fs = std::filesystem;

void sync_path(const std::string& path, bool directory) {
    const int fd = ::open(path.c_str(), directory ? (O_RDONLY | O_DIRECTORY) : O_RDONLY);
    if (fd < 0) {
        throw std::system_error(errno, std::generic_category());
    }
    const int rc = ::fsync(fd);
    const int err = errno;
    ::close(fd);
    if (rc != 0) {
        throw std::system_error(err, std::generic_category());
    }
}

}  // namespace

void fail_path(const std::string& path, const std::string& what) {
    throw std::runtime_error(path + ": " + what);
}

std::uint64_t mul_checked(std::uint64_t a, std::uint64_t b, const std::string& path) {
    if (b != 0 && a > std::numeric_limits<std::uint64_t>::max() / b) {
        fail_path(path, "the header's sizes overflow 64 bits");
    }
    return a * b;
}

std::uint64_t add_checked(std::uint64_t a, std::uint64_t b, const std::string& path) {
    if (a > std::numeric_limits<std::uint64_t>::max() - b) {
        fail_path(path, "the header's sizes overflow 64 bits");
    }
    return a + b;
}

void save_atomically(const std::string& path, const std::function<void(std::ofstream&)>& body) {
    std::random_device random;
    const std::uint64_t suffix = (std::uint64_t{random()} << 32U) | random();
    const std::string temp = path + fmt(".tmp-%016llx", static_cast<unsigned long long>(suffix));
    struct Cleanup {
        std::string path;
        ~Cleanup() {
            if (!path.empty()) {
                std::error_code ignored;
                fs::remove(path, ignored);
            }
        }
    } cleanup{temp};
    {
        std::ofstream out(temp, std::ios::binary | std::ios::trunc);
        if (!out) {
            fail_path(path, "cannot create the temporary file " + temp);
        }
        body(out);
        out.close();
        if (!out) {
            fail_path(path, "cannot write the temporary file " + temp);
        }
    }
    try {
        sync_path(temp, false);
    } catch (const std::system_error& e) {
        fail_path(path, "cannot flush the temporary file " + temp + " to storage: " + e.what());
    }
    std::error_code ec;
    fs::rename(temp, path, ec);
    if (ec) {
        fail_path(path, "cannot replace it with " + temp + ": " + ec.message());
    }
    cleanup.path.clear();
    const fs::path parent = fs::path(path).parent_path();
    const std::string dir = parent.empty() ? std::string(".") : parent.string();
    try {
        sync_path(dir, true);
    } catch (const std::system_error& e) {
        fail_path(path, "saved, but cannot flush the directory " + dir + " to storage: " + e.what());
    }
}

}  // namespace vdx
