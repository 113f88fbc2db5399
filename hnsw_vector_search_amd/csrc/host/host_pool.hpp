#pragma once

// A persistent pool of host threads for the per-call host work of the search entry points
// (query validation and staging, result unpacking): spawning threads on every call would cost
// more than the work on a 10,000-query batch. A thread that finds no work spins for up to
// kSpin before it sleeps (and a caller spins as long for the last items to finish): a sleeping
// thread takes tens of microseconds to wake, about as long as a 10,000-query batch's staging,
// while back-to-back calls come within a kernel's time of each other.

#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstddef>
#include <cstdint>
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

    static constexpr std::chrono::microseconds kSpin{1000};

private:
    void work();
    // Runs items of batch `gen` (fn, count) until none is left to claim.
    void drain(std::uint32_t gen, const std::function<void(std::size_t)>& fn, std::size_t count);

    std::vector<std::thread> workers_;
    std::mutex busy_;  // one caller's items at a time
    std::mutex m_;
    std::condition_variable wake_, done_;
    // Under m_: the current batch (its number, function and item count), the workers still in
    // it, the first failure.
    std::uint32_t generation_ = 0;
    bool stop_ = false;
    const std::function<void(std::size_t)>* fn_ = nullptr;
    std::size_t count_ = 0;
    std::size_t active_ = 0;
    std::size_t error_index_ = 0;
    std::exception_ptr error_;
    // Items are claimed without the lock: the batch number (high 32 bits) and the next item
    // (low 32 bits) in one word, so a claim can only take an item of the batch its thread
    // joined. `limit_` drops to a failed item + 1 (items above it need not run).
    std::atomic<std::uint64_t> claim_{0};
    std::atomic<std::size_t> limit_{0};
    // mirrors of generation_, stop_ and active_ (written under m_) for the spins, which read
    // them without the lock and decide nothing but when to take it
    std::atomic<std::uint32_t> generation_seen_{0};
    std::atomic<bool> stopping_{false};
    std::atomic<std::size_t> active_seen_{0};
};

}  // namespace vdx
