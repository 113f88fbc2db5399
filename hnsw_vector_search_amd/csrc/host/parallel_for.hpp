#pragma once

#include <algorithm>
#include <atomic>
#include <exception>
#include <mutex>
#include <thread>
#include <vector>

namespace vdx {

inline std::size_t resolve_threads(std::size_t num_threads) noexcept {
    return num_threads != 0 ? num_threads : std::max(1U, std::thread::hardware_concurrency());
}

template <typename Fn>
void parallel_for(std::size_t count, std::size_t threads, const Fn& fn) {
    if (count == 0) {
        return;
    }
    threads = std::max<std::size_t>(1, std::min(threads, count));
    std::atomic<std::size_t> next{0};
    std::exception_ptr error;
    std::mutex error_mutex;
    const auto worker = [&] {
        try {
            for (std::size_t i = next.fetch_add(1, std::memory_order_relaxed); i < count;
                 i = next.fetch_add(1, std::memory_order_relaxed)) {
                fn(i);
            }
        } catch (...) {
            next.store(count, std::memory_order_relaxed);
            const std::lock_guard<std::mutex> lock(error_mutex);
            if (!error) {
                error = std::current_exception();
            }
        }
    };
    std::vector<std::thread> helpers;
    helpers.reserve(threads - 1);
    for (std::size_t t = 1; t < threads; ++t) {
        try {
            helpers.emplace_back(worker);
        } catch (...) {
            break;  // no more threads: the ones running still claim every item
        }
    }
    worker();
    for (std::thread& h : helpers) {
        h.join();
    }
    if (error) {
        std::rethrow_exception(error);
    }
}

}  // namespace vdx
