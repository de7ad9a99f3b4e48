// server_core.hpp -- the serving loop's scheduling: admission by a device
// memory budget, worker slots, tickets, latency statistics and the batcher
// that merges concurrent candidate sets into multi-pair cost launches.
//
// Plain C++ (no HIP).  The device work comes from a backend type B:
//   B::Context                              one stream; B::make(device)
//   B::register_pair(ctx, req, eval, res)   one registration (throws Error)
//   B::direct(ctx)                          a cost evaluator on ctx's stream
//   B::dims(volume, int[3])                 a volume's shape
//   B::transform(param, fd, fvs, td, tvs, T)  affine_param -> matrix
//   B::launch(ctx, pairs, n, cost, bins)    one multi-pair cost launch (throws Error)
// server.cpp instantiates it with the HIP backend; tests build it on the CPU
// with a stub backend under AddressSanitizer and ThreadSanitizer.
#pragma once

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <exception>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "errors.hpp"

namespace tiplhip {

using steady = std::chrono::steady_clock;

inline double ms_between(steady::time_point a, steady::time_point b)
{
    return std::chrono::duration<double, std::milli>(b - a).count();
}

inline size_t vox3(const int* d) { return size_t(d[0]) * size_t(d[1]) * size_t(d[2]); }

// The optimiser's batched cost callback (the same signature as
// tipl::reg::detail::cost_evaluator).
using CostEvaluator =
    std::function<void(const tiplhip_volume* from, const float* from_vs, const tiplhip_volume* to,
                       const float* to_vs, const float* params, int K, int cost, int bins,
                       double* costs)>;

inline void check_request(const tiplhip_reg_request* r)
{
    TIPLHIP_REQUIRE(r != nullptr, "null request");
    TIPLHIP_REQUIRE(r->fixed != nullptr && r->moving != nullptr, "null image");
    for (int i = 0; i < 3; ++i)
        TIPLHIP_REQUIRE(r->fixed_dim[i] > 0 && r->moving_dim[i] > 0, "dims must be positive");
    for (int dt : {r->fixed_dtype, r->moving_dtype})
        TIPLHIP_REQUIRE(dt == TIPLHIP_U8 || dt == TIPLHIP_I16 || dt == TIPLHIP_F32, "unknown dtype");
    TIPLHIP_REQUIRE(r->cost == TIPLHIP_COST_MI || r->cost == TIPLHIP_COST_CORR, "unknown cost type");
    TIPLHIP_REQUIRE(r->bins == 256 || r->bins == 64, "bins must be 256 or 64");
    TIPLHIP_REQUIRE(r->preprocess == 0 || r->preprocess == 1, "preprocess must be 0 or 1");
    TIPLHIP_REQUIRE(!(r->f2t || r->from2to) || (r->nonlinear && r->reverse),
                    "f2t and from2to need a nonlinear request with reverse set");
    if (r->skip_linear)
        for (float v : r->arg) TIPLHIP_REQUIRE(std::isfinite(v), "arg must be finite");
}

// Estimated device bytes a request holds while it runs (what admission counts).
inline size_t request_bytes(const tiplhip_reg_request& r)
{
    const double nf = double(vox3(r.fixed_dim)), nm = double(vox3(r.moving_dim));
    const double pyramid = 8.0 / 7.0;  // 1 + 1/8 + 1/64 + ...
    double b = (nf + nm) * pyramid;                    // uint8 volumes and their pyramids
    b += 8.0 * std::max(nf, nm) * pyramid;            // 8-tap packed copy of the sampled side
    auto es = [](int dt) { return dt == TIPLHIP_F32 ? 4.0 : dt == TIPLHIP_I16 ? 2.0 : 0.0; };
    b += std::max(es(r.fixed_dtype) * nf, es(r.moving_dtype) * nm);  // quantisation staging
    b += 2.0 * nf;                                     // J and its correlation with the fixed image
    if (r.nonlinear) {
        b += nf * 75.0 * pyramid;                      // best field, 4 work fields, Js per direction
        if (r.reverse) b += nf * 75.0 * pyramid;
        b += 25.0 * nf + (r.from2to ? 12.0 * nm : 0.0);  // mappings and the warped image
    }
    if (r.preprocess == 1) b += 12.0 * nm;            // subject preprocessing: 3 float images
    return size_t(b) + (size_t(4) << 20);
}

// ---------------------------------------------------------------------------
// Batcher: merges concurrent candidate sets into multi-pair launches.
// ---------------------------------------------------------------------------
template <class B>
class Batcher {
public:
    // `launchers` threads, each with its own context (stream): while one
    // batch runs on the device the next one is formed and launched.
    Batcher(int device, int launchers)
    {
        for (int i = 0; i < launchers; ++i) ctx_.push_back(B::make(device));
        for (int i = 0; i < launchers; ++i) threads_.emplace_back([this, i] { run(*ctx_[size_t(i)]); });
    }
    ~Batcher()
    {
        {
            std::lock_guard<std::mutex> lk(mu_);
            stop_ = true;
        }
        cv_.notify_all();
        for (auto& t : threads_) t.join();
    }
    Batcher(const Batcher&) = delete;
    Batcher& operator=(const Batcher&) = delete;

    // The cost evaluator of a served registration: blocks until the
    // candidates have been evaluated in some launch.
    void evaluate(const tiplhip_volume* from, const float* from_vs, const tiplhip_volume* to,
                  const float* to_vs, const float* params, int K, int cost, int bins, double* costs)
    {
        int fd[3], td[3];
        B::dims(from, fd);
        B::dims(to, td);
        for (int k0 = 0; k0 < K; k0 += kMaxLaunch) {
            Job j;
            j.from = from;
            j.to = to;
            j.K = std::min(kMaxLaunch, K - k0);
            j.cost = cost;
            j.bins = bins;
            j.costs = costs + k0;
            j.T.resize(12 * size_t(j.K));
            for (int k = 0; k < j.K; ++k)
                B::transform(params + 12 * size_t(k0 + k), fd, from_vs, td, to_vs, &j.T[12 * size_t(k)]);
            std::unique_lock<std::mutex> lk(mu_);
            pending_.push_back(&j);
            cv_.notify_all();
            done_cv_.wait(lk, [&] { return j.done; });
            if (j.rc != TIPLHIP_OK) throw Error(j.rc, j.err);
        }
    }

    void counts(int64_t& launches, int64_t& pairs)
    {
        std::lock_guard<std::mutex> lk(mu_);
        launches = launches_;
        pairs = pairs_;
    }

private:
    static constexpr int kMaxLaunch = 4096;
    struct Job {
        const tiplhip_volume* from = nullptr;
        const tiplhip_volume* to = nullptr;
        std::vector<float> T;
        int K = 0, cost = 0, bins = 0;
        double* costs = nullptr;
        bool done = false;
        int rc = TIPLHIP_OK;
        std::string err;
    };

    // One launch over `batch`; rc and message of its failure.
    static int launch(typename B::Context& ctx, const std::vector<Job*>& batch, int cost, int bins,
                      std::string& err)
    {
        std::vector<tiplhip_pair_costs> pc;
        for (Job* j : batch) pc.push_back({j->from, j->to, j->T.data(), j->K, j->costs});
        try {
            B::launch(ctx, pc.data(), int(pc.size()), cost, bins);
            return TIPLHIP_OK;
        } catch (const Error& e) {
            err = e.what();
            return e.code;
        } catch (const std::exception& e) {
            err = e.what();
            return TIPLHIP_ERR_INVALID;
        }
    }

    void run(typename B::Context& ctx)
    {
        for (;;) {
            std::vector<Job*> batch;
            int cost = 0, bins = 0;
            {
                std::unique_lock<std::mutex> lk(mu_);
                cv_.wait(lk, [&] { return stop_ || !pending_.empty(); });
                if (pending_.empty()) return;  // stop_ with nothing left
                cost = pending_.front()->cost;
                bins = pending_.front()->bins;
                int K = 0;
                for (auto it = pending_.begin(); it != pending_.end();) {
                    Job* j = *it;
                    const bool same = j->cost == cost && (cost == TIPLHIP_COST_CORR || j->bins == bins);
                    if (same && K + j->K <= kMaxLaunch) {
                        K += j->K;
                        batch.push_back(j);
                        it = pending_.erase(it);
                    } else {
                        ++it;
                    }
                }
            }
            std::vector<int> rc(batch.size(), TIPLHIP_OK);
            std::vector<std::string> err(batch.size());
            std::string e0;
            const int r0 = launch(ctx, batch, cost, bins, e0);
            int64_t n_launches = 1;
            if (r0 != TIPLHIP_OK && batch.size() > 1) {
                // One pair's failure (a volume the kernels do not take, a
                // device error) must not fail the registrations it happened
                // to share the launch with: each job again on its own.
                for (size_t i = 0; i < batch.size(); ++i) {
                    rc[i] = launch(ctx, {batch[i]}, cost, bins, err[i]);
                    ++n_launches;
                }
            } else {
                std::fill(rc.begin(), rc.end(), r0);
                std::fill(err.begin(), err.end(), e0);
            }
            {
                std::lock_guard<std::mutex> lk(mu_);
                for (size_t i = 0; i < batch.size(); ++i) {
                    batch[i]->rc = rc[i];
                    batch[i]->err = err[i];
                    batch[i]->done = true;
                }
                launches_ += n_launches;
                pairs_ += int64_t(batch.size());
            }
            done_cv_.notify_all();
        }
    }

    std::vector<std::unique_ptr<typename B::Context>> ctx_;
    std::mutex mu_;
    std::condition_variable cv_, done_cv_;
    std::deque<Job*> pending_;
    bool stop_ = false;
    int64_t launches_ = 0, pairs_ = 0;
    std::vector<std::thread> threads_;
};

struct Ticket {
    tiplhip_reg_request req;
    tiplhip_reg_result res;
    size_t bytes = 0;
    std::string err;
    steady::time_point submitted, started, finished;
    bool done = false;
    bool claimed = false;  // a caller is waiting for it
};

// ---------------------------------------------------------------------------
// The server: `slots` worker threads, each with its own context (stream),
// take requests FIFO when their estimated device memory fits the budget.
// ---------------------------------------------------------------------------
template <class B>
class ServerCore {
public:
    // launchers: batch launcher threads (-1: one per 4 slots, 1..4; 0: no
    // batcher, each slot launches its own candidate sets on its stream).
    ServerCore(int device, int slots, int launchers, size_t budget) : budget_(budget)
    {
        TIPLHIP_REQUIRE(slots >= 1 && slots <= 64, "slots must be in [1, 64]");
        TIPLHIP_REQUIRE(launchers >= -1 && launchers <= 16, "launchers must be in [-1, 16]");
        TIPLHIP_REQUIRE(budget > 0, "zero memory budget");
        if (launchers < 0) launchers = std::min(4, std::max(1, slots / 4));
        if (launchers > 0) batcher_ = std::make_unique<Batcher<B>>(device, launchers);
        for (int i = 0; i < slots; ++i) contexts_.push_back(B::make(device));
        try {
            for (int i = 0; i < slots; ++i)
                workers_.emplace_back([this, c = contexts_[size_t(i)].get()] { worker(*c); });
        } catch (...) {
            stop_and_join();
            throw;
        }
    }
    // Finishes every submitted request, then stops the workers.
    ~ServerCore() { stop_and_join(); }
    ServerCore(const ServerCore&) = delete;
    ServerCore& operator=(const ServerCore&) = delete;

    uint64_t submit(const tiplhip_reg_request* req)
    {
        check_request(req);
        auto t = std::make_unique<Ticket>();
        t->req = *req;
        std::memset(&t->res, 0, sizeof(t->res));
        t->bytes = request_bytes(*req);
        TIPLHIP_REQUIRE(t->bytes <= budget_, "request needs more device memory than the budget");
        uint64_t id = 0;
        {
            std::lock_guard<std::mutex> lk(mu_);
            TIPLHIP_REQUIRE(!stopping_, "server is stopping");
            t->submitted = steady::now();
            if (submitted_ == 0) first_submit_ = t->submitted;
            ++submitted_;
            id = next_ticket_++;
            queue_.push_back(id);
            tickets_[id] = std::move(t);
        }
        work_cv_.notify_all();
        return id;
    }

    // Blocks until the ticket is done; copies its result out and forgets it.
    void wait(uint64_t ticket, tiplhip_reg_result* res)
    {
        std::unique_lock<std::mutex> lk(mu_);
        auto it = tickets_.find(ticket);
        TIPLHIP_REQUIRE(it != tickets_.end(), "unknown ticket");
        Ticket* t = it->second.get();
        TIPLHIP_REQUIRE(!t->claimed, "another caller is already waiting for this ticket");
        t->claimed = true;
        done_cv_.wait(lk, [&] { return t->done; });
        *res = t->res;
        const int status = t->res.status;
        const std::string err = t->err;
        tickets_.erase(it);
        if (status != TIPLHIP_OK) throw Error(status, err);
    }

    void stats(tiplhip_server_stats* st)
    {
        std::memset(st, 0, sizeof(*st));
        std::vector<double> lat;
        {
            std::lock_guard<std::mutex> lk(mu_);
            st->submitted = submitted_;
            st->completed = completed_;
            st->failed = failed_;
            st->running = running_;
            st->queued = int64_t(queue_.size());
            st->hbm_in_use = in_use_;
            st->hbm_peak = peak_;
            lat = latencies_;
            if (completed_ > 0) st->wall_s = ms_between(first_submit_, last_done_) * 1e-3;
        }
        if (batcher_) batcher_->counts(st->batch_launches, st->batched_pairs);
        if (!lat.empty()) {
            std::sort(lat.begin(), lat.end());
            // nearest-rank percentiles
            auto pct = [&](double q) {
                size_t r = size_t(std::ceil(q * double(lat.size())));
                return lat[std::min(lat.size(), std::max<size_t>(r, 1)) - 1];
            };
            st->latency_p50_ms = pct(0.50);
            st->latency_p95_ms = pct(0.95);
            st->latency_p99_ms = pct(0.99);
        }
        if (st->wall_s > 0.0) st->regs_per_s = double(st->completed) / st->wall_s;
    }

private:
    void stop_and_join()
    {
        {
            std::lock_guard<std::mutex> lk(mu_);
            stopping_ = true;
        }
        work_cv_.notify_all();
        for (auto& w : workers_)
            if (w.joinable()) w.join();
        batcher_.reset();
    }

    void worker(typename B::Context& ctx)
    {
        // with a batcher, candidate sets of all slots are merged into shared
        // launches; without one, each slot launches its own on its stream
        CostEvaluator eval =
            batcher_ ? CostEvaluator([this](const tiplhip_volume* from, const float* fvs,
                                            const tiplhip_volume* to, const float* tvs,
                                            const float* params, int K, int cost, int bins,
                                            double* costs) {
                batcher_->evaluate(from, fvs, to, tvs, params, K, cost, bins, costs);
            })
                     : B::direct(ctx);
        for (;;) {
            Ticket* t = nullptr;
            {
                std::unique_lock<std::mutex> lk(mu_);
                work_cv_.wait(lk, [&] {
                    if (queue_.empty()) return stopping_;
                    return in_use_ + tickets_.at(queue_.front())->bytes <= budget_;
                });
                if (queue_.empty()) return;
                t = tickets_.at(queue_.front()).get();
                queue_.pop_front();
                in_use_ += t->bytes;
                peak_ = std::max(peak_, in_use_);
                ++running_;
                t->started = steady::now();
            }
            int status = TIPLHIP_OK;
            std::string err;
            try {
                B::register_pair(ctx, t->req, eval, t->res);
            } catch (const Error& e) {
                status = e.code;
                err = e.what();
            } catch (const std::bad_alloc&) {
                status = TIPLHIP_ERR_MEMORY;
                err = "host allocation failed";
            } catch (const std::exception& e) {
                status = TIPLHIP_ERR_INVALID;
                err = e.what();
            }
            {
                std::lock_guard<std::mutex> lk(mu_);
                t->finished = steady::now();
                t->res.status = status;
                t->err = err;
                t->res.queue_ms = ms_between(t->submitted, t->started);
                t->res.run_ms = ms_between(t->started, t->finished);
                t->done = true;
                in_use_ -= t->bytes;
                --running_;
                if (status == TIPLHIP_OK) {
                    ++completed_;
                    latencies_.push_back(ms_between(t->submitted, t->finished));
                    last_done_ = t->finished;
                } else {
                    ++failed_;
                }
            }
            done_cv_.notify_all();
            work_cv_.notify_all();
        }
    }

    size_t budget_ = 0;
    std::unique_ptr<Batcher<B>> batcher_;
    std::mutex mu_;
    std::condition_variable work_cv_, done_cv_;
    std::deque<uint64_t> queue_;
    std::map<uint64_t, std::unique_ptr<Ticket>> tickets_;
    uint64_t next_ticket_ = 1;
    size_t in_use_ = 0, peak_ = 0;
    bool stopping_ = false;
    int64_t submitted_ = 0, completed_ = 0, failed_ = 0, running_ = 0;
    std::vector<double> latencies_;
    steady::time_point first_submit_, last_done_;
    std::vector<std::unique_ptr<typename B::Context>> contexts_;
    std::vector<std::thread> workers_;
};

}  // namespace tiplhip
