#pragma once

// A persistent pool of host threads for the per-call host work of the search entry points
// (query validation and staging, result unpacking): spawning threads on every call would cost
// more than the work on a 10,000-query batch.

#include <condition_variable>
#include <cstddef>
#include <exception>
#include <functional>
#include <mutex>
#include <thread>
#include <vector>

namespace vdx {

class HostPool {
public:
    // `threads` counts the calling thread: threads - 1 workers are started.
    explicit HostPool(std::size_t threads);
    ~HostPool();
    HostPool(const HostPool&) = delete;
    HostPool& operator=(const HostPool&) = delete;

    std::size_t threads() const noexcept { return workers_.size() + 1; }

    // Runs fn(i) for every i in [0, count) on the workers and the calling thread. If any call
    // throws, the exception of the lowest failing index is rethrown (what a sequential loop
    // would have thrown first); items above a failed one may be skipped. A pool busy with
    // another caller's items runs these on the calling thread alone.
    void run(std::size_t count, const std::function<void(std::size_t)>& fn);

private:
    void work(std::size_t generation_seen);
    void drain();

    std::vector<std::thread> workers_;
    std::mutex busy_;  // one caller's items at a time
    std::mutex m_;
    std::condition_variable wake_, done_;
    std::size_t generation_ = 0;
    bool stop_ = false;
    const std::function<void(std::size_t)>* fn_ = nullptr;
    std::size_t count_ = 0;
    std::size_t next_ = 0;
    std::size_t active_ = 0;
    std::size_t error_index_ = 0;
    std::exception_ptr error_;
};

}  // namespace vdx
