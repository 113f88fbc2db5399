#include "host_pool.hpp"

#include <algorithm>

namespace vdx {

HostPool::HostPool(std::size_t threads) {
    const std::size_t n = threads > 1 ? threads - 1 : 0;
    workers_.reserve(n);
    for (std::size_t t = 0; t < n; ++t) {
        try {
            workers_.emplace_back([this] { work(0); });
        } catch (...) {
            break;  // fewer workers: the caller still runs every item
        }
    }
}

HostPool::~HostPool() {
    {
        const std::lock_guard<std::mutex> lock(m_);
        stop_ = true;
    }
    wake_.notify_all();
    for (std::thread& t : workers_) {
        t.join();
    }
}

// Claims and runs items of the current batch until none is left.
void HostPool::drain() {
    for (;;) {
        std::size_t i;
        const std::function<void(std::size_t)>* fn;
        {
            const std::lock_guard<std::mutex> lock(m_);
            if (next_ >= count_) {
                return;
            }
            i = next_++;
            fn = fn_;
        }
        try {
            (*fn)(i);
        } catch (...) {
            const std::lock_guard<std::mutex> lock(m_);
            if (!error_ || i < error_index_) {
                error_ = std::current_exception();
                error_index_ = i;
            }
            // items above a failure need not run
            count_ = std::min(count_, std::max(next_, i + 1));
        }
    }
}

void HostPool::work(std::size_t generation_seen) {
    for (;;) {
        {
            std::unique_lock<std::mutex> lock(m_);
            wake_.wait(lock, [&] { return stop_ || generation_ != generation_seen; });
            if (stop_) {
                return;
            }
            generation_seen = generation_;
            ++active_;
        }
        drain();
        {
            const std::lock_guard<std::mutex> lock(m_);
            if (--active_ == 0) {
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
    if (!busy.owns_lock() || workers_.empty() || count == 1) {
        for (std::size_t i = 0; i < count; ++i) {
            fn(i);
        }
        return;
    }
    {
        const std::lock_guard<std::mutex> lock(m_);
        fn_ = &fn;
        count_ = count;
        next_ = 0;
        error_ = nullptr;
        error_index_ = 0;
        ++generation_;
    }
    wake_.notify_all();
    drain();
    std::exception_ptr error;
    {
        std::unique_lock<std::mutex> lock(m_);
        // every worker that woke for this batch has left it (a worker that had not woken yet
        // finds no item left and touches nothing of it)
        done_.wait(lock, [&] { return active_ == 0; });
        error = error_;
        error_ = nullptr;
        fn_ = nullptr;
        count_ = 0;
        next_ = 0;
    }
    if (error) {
        std::rethrow_exception(error);
    }
}

}  // namespace vdx
