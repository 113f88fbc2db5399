#include "host_pool.hpp"

#include <algorithm>
#include <immintrin.h>

namespace vdx {

namespace {

// Spins until `ready()` or kSpin has passed.
template <typename Ready>
void spin_until(Ready ready) {
    const auto until = std::chrono::steady_clock::now() + HostPool::kSpin;
    while (!ready()) {
        for (int i = 0; i < 64; ++i) {
            _mm_pause();
        }
        if (std::chrono::steady_clock::now() >= until) {
            return;
        }
    }
}

}  // namespace

HostPool::HostPool(std::size_t threads) {
    const std::size_t n = threads > 1 ? threads - 1 : 0;
    workers_.reserve(n);
    for (std::size_t t = 0; t < n; ++t) {
        try {
            workers_.emplace_back([this] { work(); });
        } catch (...) {
            break;  // fewer workers: the caller still runs every item
        }
    }
}

HostPool::~HostPool() {
    {
        const std::lock_guard<std::mutex> lock(m_);
        stop_ = true;
        stopping_.store(true, std::memory_order_release);
    }
    wake_.notify_all();
    for (std::thread& t : workers_) {
        t.join();
    }
}

void HostPool::drain(std::uint32_t gen, const std::function<void(std::size_t)>& fn, std::size_t count) {
    for (;;) {
        std::uint64_t v = claim_.load(std::memory_order_acquire);
        std::size_t i;
        for (;;) {
            i = static_cast<std::size_t>(v & 0xFFFFFFFFu);
            if (static_cast<std::uint32_t>(v >> 32) != gen || i >= count ||
                i >= limit_.load(std::memory_order_acquire)) {
                return;
            }
            if (claim_.compare_exchange_weak(v, v + 1, std::memory_order_acq_rel)) {
                break;
            }
        }
        try {
            fn(i);
        } catch (...) {
            const std::lock_guard<std::mutex> lock(m_);
            if (!error_ || i < error_index_) {
                error_ = std::current_exception();
                error_index_ = i;
            }
            // items above a failure need not run (every item below it is claimed already)
            if (limit_.load(std::memory_order_relaxed) > i + 1) {
                limit_.store(i + 1, std::memory_order_release);
            }
        }
    }
}

void HostPool::work() {
    std::uint32_t seen = 0;
    for (;;) {
        spin_until([&] {
            return stopping_.load(std::memory_order_acquire) ||
                   generation_seen_.load(std::memory_order_acquire) != seen;
        });
        const std::function<void(std::size_t)>* fn;
        std::size_t count;
        {
            std::unique_lock<std::mutex> lock(m_);
            wake_.wait(lock, [&] { return stop_ || generation_ != seen; });
            if (stop_) {
                return;
            }
            // joins the current batch: run() waits for it to leave before the batch ends
            seen = generation_;
            fn = fn_;
            count = count_;
            active_seen_.store(++active_, std::memory_order_release);
        }
        if (fn != nullptr) {
            drain(seen, *fn, count);
        }
        {
            const std::lock_guard<std::mutex> lock(m_);
            active_seen_.store(--active_, std::memory_order_release);
            if (active_ == 0) {
                done_.notify_all();
            }
        }
    }
}

void HostPool::run(std::size_t count, const std::function<void(std::size_t)>& fn) {
    if (count == 0) {
        return;
    }
    std::unique_lock<std::mutex> busy(busy_, std::try_to_lock);
    if (!busy.owns_lock() || workers_.empty() || count == 1 || count > 0xFFFFFFFFu) {
        for (std::size_t i = 0; i < count; ++i) {
            fn(i);
        }
        return;
    }
    std::uint32_t gen;
    {
        const std::lock_guard<std::mutex> lock(m_);
        fn_ = &fn;
        count_ = count;
        error_ = nullptr;
        error_index_ = 0;
        gen = ++generation_;
        limit_.store(count, std::memory_order_release);
        claim_.store(static_cast<std::uint64_t>(gen) << 32, std::memory_order_release);
        generation_seen_.store(gen, std::memory_order_release);
    }
    wake_.notify_all();
    drain(gen, fn, count);
    spin_until([&] { return active_seen_.load(std::memory_order_acquire) == 0; });
    std::exception_ptr error;
    {
        std::unique_lock<std::mutex> lock(m_);
        // every worker that joined this batch has left it; one that joins later finds no item
        // of it to claim (the batch number in the claim word) and touches nothing of it
        done_.wait(lock, [&] { return active_ == 0; });
        error = error_;
        error_ = nullptr;
        fn_ = nullptr;
        count_ = 0;
    }
    if (error) {
        std::rethrow_exception(error);
    }
}

}  // namespace vdx
