// step_gate.hpp -- the scheduling of batched steps, without HIP (so it also
// builds and is tested on a CPU-only host, under the thread sanitizer).
//
// Jobs that run in step -- each an iterative loop on its own host thread
// whose every iteration queues one step of device work -- are `active`
// between join() and leave().  run(job) queues the job's next step and
// returns once it is launched: a batch of queued steps (at most max_batch)
// is handed to the launch function when every active job has queued one,
// or when a queued step has waited `gather`, so jobs that share a batch stay
// in step.  Launches happen under the gate's lock, in queue order.
#pragma once

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdint>
#include <exception>
#include <functional>
#include <mutex>
#include <vector>

namespace tiplhip {

// The gate's per-job state (a job type derives from it).
struct GateSlot {
    std::atomic<bool> launched{false};
    std::exception_ptr error;
};

template <class Job>
class StepGate {
public:
    // launch(jobs, n): queues the steps of n jobs; an exception fails all n
    using Launch = std::function<void(Job* const*, int)>;

    StepGate(int max_batch, std::chrono::microseconds gather, Launch launch)
        : max_(max_batch), gather_(gather), launch_(std::move(launch))
    {
    }
    StepGate(const StepGate&) = delete;
    StepGate& operator=(const StepGate&) = delete;

    void join()
    {
        std::lock_guard<std::mutex> lk(mu_);
        ++active_;
    }
    // Ends a job's membership; f() runs under the lock, after every step the
    // job queued has been launched.  Steps waiting for this job go out now.
    template <class F>
    void leave(F&& f)
    {
        std::lock_guard<std::mutex> lk(mu_);
        --active_;
        f();
        if (!pending_.empty() && int(pending_.size()) >= threshold()) launch_pending();
    }
    // Queues j's step and returns once it is launched; rethrows the error
    // of a failed launch.
    void run(Job& j)
    {
        GateSlot& s = j;
        const auto t0 = std::chrono::steady_clock::now();
        {
            std::lock_guard<std::mutex> lk(mu_);
            s.launched.store(false, std::memory_order_relaxed);
            s.error = nullptr;
            pending_.push_back(&j);
            if (int(pending_.size()) >= threshold()) launch_pending();
        }
        while (!s.launched.load(std::memory_order_acquire)) {
            if (std::chrono::steady_clock::now() - t0 > gather_) {
                std::lock_guard<std::mutex> lk(mu_);
                if (!s.launched.load(std::memory_order_relaxed)) launch_pending();
            } else {
#if defined(__x86_64__) || defined(__i386__)
                __builtin_ia32_pause();
#endif
            }
        }
        if (s.error) std::rethrow_exception(s.error);
    }
    // batches launched, and job steps over them
    int64_t batches() const { return batches_.load(); }
    int64_t steps() const { return steps_.load(); }

private:
    int threshold() const { return std::max(1, std::min(active_, max_)); }

    // Launches every queued step, max_ at a time (mu_ held).
    void launch_pending()
    {
        while (!pending_.empty()) {
            const int m = std::min(int(pending_.size()), max_);
            std::vector<Job*> take(pending_.begin(), pending_.begin() + m);
            pending_.erase(pending_.begin(), pending_.begin() + m);
            std::exception_ptr err;
            try {
                launch_(take.data(), m);
            } catch (...) {
                err = std::current_exception();
            }
            batches_.fetch_add(1);
            steps_.fetch_add(m);
            for (Job* j : take) {
                GateSlot& s = *j;
                s.error = err;
                s.launched.store(true, std::memory_order_release);
            }
        }
    }

    const int max_;
    const std::chrono::microseconds gather_;
    const Launch launch_;
    std::mutex mu_;
    std::vector<Job*> pending_;
    int active_ = 0;
    std::atomic<int64_t> batches_{0}, steps_{0};
};

}  // namespace tiplhip
