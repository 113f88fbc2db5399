#pragma once

// File helpers shared by the index formats: errors that start with the path, overflow-checked
// size arithmetic, and the atomic, durable replace of docs/FORMAT.md's "Saving" section
// (temporary file, flush, rename, flush of the directory).

#include <cstdint>
#include <fstream>
#include <functional>
#include <string>

namespace vdx {

[[noreturn]] void fail_path(const std::string& path, const std::string& what);
std::uint64_t mul_checked(std::uint64_t a, std::uint64_t b, const std::string& path);
std::uint64_t add_checked(std::uint64_t a, std::uint64_t b, const std::string& path);

// Writes `path` through `body` into `<path>.tmp-<16 hex>`, flushes it, renames it over `path`
// and flushes the directory. On failure the temporary file is removed and `path` is untouched.
void save_atomically(const std::string& path, const std::function<void(std::ofstream&)>& body);

}  // namespace vdx
