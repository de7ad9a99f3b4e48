// server.cpp -- one-pair registration pipeline and the serving loop.
//
// tiplhip_register runs the reference's mm_reg order for one image pair
// (reg/mm_reg.hpp:348-416): tipl::reg::linear(fixed, moving) -- the body of
// the header-only mirror in include/tipl/reg/linear.hpp, so the optimiser is
// the one C++ callers get -- then, for nonlinear requests, the moving image
// resampled onto the fixed grid and cdm(fixed, J) (reg/cdm.hpp:373-474).
//
// The server runs many of these at once on one GPU:
//   * `slots` worker threads, each with its own context (HIP stream), take
//     requests FIFO when their estimated device memory fits the budget;
//   * one batcher thread owns the affine cost launches: every optimiser
//     candidate set of every running registration is queued to it, and each
//     time it is free it launches everything queued (up to 4096 candidates,
//     one cost type) as one multi-pair launch -- registrations join and
//     leave the batch at every optimiser step;
//   * CDM iterations of different requests run on their workers' streams
//     and overlap on the device.
// Batched costs are bit-identical to single-pair costs and the optimiser is
// deterministic, so a served result equals tiplhip_register's for the same
// request.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <deque>
#include <exception>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/tipl/reg/linear.hpp"
#include "field.hpp"
#include "internal.hpp"
#include "warp.hpp"

namespace tiplhip {
namespace {

using steady = std::chrono::steady_clock;

double ms_between(steady::time_point a, steady::time_point b)
{
    return std::chrono::duration<double, std::milli>(b - a).count();
}

void rc_check(int rc)
{
    if (rc != TIPLHIP_OK) throw Error(rc, tiplhip_last_error());
}

void check_request(const tiplhip_reg_request* r)
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

size_t request_bytes(const tiplhip_reg_request& r)
{
    const double nf = double(vox(r.fixed_dim)), nm = double(vox(r.moving_dim));
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

// Joins a thread when the scope unwinds (the reverse CDM must not outlive
// the buffers it reads if the forward direction throws).
struct Joiner {
    std::thread t;
    ~Joiner()
    {
        if (t.joinable()) t.join();
    }
};

// When the scope unwinds on an exception, waits for ctx's stream: buffers
// declared before it (some owned by another stream, whose frees are not
// ordered after this stream's work) may still be read by kernels queued here.
struct DrainOnThrow {
    tiplhip_ctx* ctx;
    ~DrainOnThrow()
    {
        if (std::uncaught_exceptions() > 0) (void)hipStreamSynchronize(ctx->stream);
    }
};

void to_host(tiplhip_ctx* ctx, void* dst, const void* src, size_t bytes)
{
    TIPLHIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
}

// One registration on ctx; costs go through `eval`.
void register_pair(tipl::hip::context& ctx, const tiplhip_reg_request& req,
                   const tipl::reg::detail::cost_evaluator& eval, tiplhip_reg_result& res)
{
    const auto t0 = steady::now();
    tiplhip_ctx* c = ctx.get();
    tiplhip_volume* raw = nullptr;
    const bool mm = req.preprocess == 1;
    rc_check(tiplhip_volume_create(c, req.fixed, req.fixed_dtype, req.fixed_dim,
                                   mm ? 255.999f : 255.99f, &raw));
    auto df = std::make_shared<tipl::hip::volume>(ctx, raw);
    raw = nullptr;
    if (mm)
        rc_check(tiplhip_volume_create_subject(c, req.moving, req.moving_dtype, req.moving_dim, &raw));
    else
        rc_check(tiplhip_volume_create(c, req.moving, req.moving_dtype, req.moving_dim, 255.99f, &raw));
    auto dm = std::make_shared<tipl::hip::volume>(ctx, raw);
    const tipl::vector<3> fvs(req.fixed_vs[0], req.fixed_vs[1], req.fixed_vs[2]);
    const tipl::vector<3> mvs(req.moving_vs[0], req.moving_vs[1], req.moving_vs[2]);
    const tipl::shape<3> fs(req.fixed_dim[0], req.fixed_dim[1], req.fixed_dim[2]);
    const tipl::shape<3> ms(req.moving_dim[0], req.moving_dim[1], req.moving_dim[2]);

    // linear_reg (reg/mm_reg.hpp:348-368)
    tipl::affine_param<float, 3> arg;
    float cost = 0.0f;
    size_t evals = 0;
    auto t1 = steady::now();
    if (req.skip_linear) {
        std::copy_n(req.arg, 12, arg.data());
    } else {
        tipl::reg::linear_reg_param p;
        p.cuda = true;
        p.cost_type = req.cost == TIPLHIP_COST_MI ? tipl::reg::mutual_info : tipl::reg::corr;
        p.mi_bins = req.bins;
        bool terminated = false;
        // has_mask and the moment estimate read the quantised images on the host
        const tipl::image<3, uint8_t> hf = df->read(), hm = dm->read();
        t1 = steady::now();
        cost = tipl::reg::detail::linear_impl<void>(
            {tipl::const_pointer_image<3, unsigned char>(hf)}, fvs,
            {tipl::const_pointer_image<3, unsigned char>(hm)}, mvs, {df}, {dm}, arg, p, terminated,
            eval, &evals);
    }
    const tipl::transformation_matrix<float, 3> T(arg, fs, fvs, ms, mvs);
    std::copy_n(arg.data(), 12, res.affine);
    std::copy_n(T.data(), 12, res.T);
    res.cost = cost;
    res.evaluations = int64_t(evals);
    res.cdm_iterations = 0;
    const auto t2 = steady::now();
    res.prep_ms = ms_between(t0, t1);
    res.affine_ms = ms_between(t1, t2);

    // calculate_linear_r (reg/mm_reg.hpp:285-311): J = resample(I, Its, T)
    const size_t nf = vox(req.fixed_dim), nm = vox(req.moving_dim);
    const bool masked = req.masked_r != 0;
    raw = nullptr;
    rc_check(tiplhip_volume_resample(c, dm->get(), req.fixed_dim, T.data(), 1, &raw));
    tipl::hip::volume dj(ctx, raw);
    tiplhip_volume* It = df->get();
    res.r_linear = float(device_correlation(c, dj.get()->d, It->d, nf, masked));
    if (!req.nonlinear) {
        if (req.warped) {
            to_host(c, req.warped, dj.get()->d, nf);
            TIPLHIP_CHECK(hipStreamSynchronize(c->stream));
        }
        res.cdm_ms = ms_between(t2, steady::now());
        return;
    }
    if (req.match_fov) {
        launch_match_fov(c, It->d, It->dim, T.data(), req.moving_dim);
        volume_invalidate(c, It);
    }
    TIPLHIP_CHECK(hipStreamSynchronize(c->stream));

    // nonlinear_reg (reg/mm_reg.hpp:370-410): the reverse cdm(J, It) on the
    // auxiliary stream while the forward cdm(It, J) runs on this one
    const size_t fb = sizeof(float) * 3 * nf;
    DevMem t2f(c, fb);
    tiplhip_ctx* aux = req.reverse ? aux_context(c) : nullptr;
    std::unique_ptr<DevMem> f2t;  // on the auxiliary stream, read on this one below
    DrainOnThrow drain{c};
    int rev_rc = TIPLHIP_OK;
    std::string rev_err;
    {
        Joiner rev;
        if (aux) {
            f2t.reset(new DevMem(aux, fb));
            rev.t = std::thread([&] {
                try {
                    DeviceGuard dg(aux->device);
                    const tiplhip_volume* a = dj.get();
                    const tiplhip_volume* b = It;
                    const int rc = cdm_device(aux, &a, &b, 1, &req.cdm, f2t->as<float>(), nullptr, nullptr);
                    if (rc == TIPLHIP_ERR_FAILED) throw Error(rc, "reverse cdm returned false (zero cost)");
                    TIPLHIP_CHECK(hipStreamSynchronize(aux->stream));
                } catch (const Error& e) {
                    rev_rc = e.code;
                    rev_err = e.what();
                } catch (const std::exception& e) {
                    rev_rc = TIPLHIP_ERR_INVALID;
                    rev_err = e.what();
                }
            });
        }
        const tiplhip_volume* a = It;
        const tiplhip_volume* b = dj.get();
        int iterations = 0;
        const int rc = cdm_device(c, &a, &b, 1, &req.cdm, t2f.as<float>(), nullptr, &iterations);
        res.cdm_iterations = iterations;
        if (rc == TIPLHIP_ERR_FAILED) throw Error(rc, "cdm returned false (zero cost)");
    }
    if (rev_rc != TIPLHIP_OK) throw Error(rev_rc, rev_err);

    // every buffer is allocated before the first launch that reads f2t (an
    // allocation failure then cannot release f2t under a running kernel)
    DevMem to2from(c, fb), jj(c, nf);
    std::unique_ptr<DevMem> from2to;
    if (aux && req.from2to) from2to.reset(new DevMem(c, sizeof(float) * 3 * nm));
    // compute_mapping_from_displacement (reg/mm_reg.hpp:271-284)
    launch_mapping(c, t2f.as<float>(), req.fixed_dim, T.data(), to2from.as<float>());
    if (from2to)
        launch_inv_mapping(c, f2t->as<float>(), req.fixed_dim, T.data(), req.moving_dim,
                           from2to->as<float>());
    // calculate_nonlinear_r (reg/mm_reg.hpp:330-346): J = compose_mapping(I, to2from)
    launch_compose_mapping_u8(c, dm->get()->d, req.moving_dim, to2from.as<float>(), req.fixed_dim,
                              jj.as<uint8_t>());
    res.r_nonlinear = float(device_correlation(c, jj.as<uint8_t>(), It->d, nf, masked));
    if (req.field) to_host(c, req.field, t2f.p, fb);
    if (aux && req.f2t) to_host(c, req.f2t, f2t->p, fb);
    if (req.to2from) to_host(c, req.to2from, to2from.p, fb);
    if (from2to) to_host(c, req.from2to, from2to->p, sizeof(float) * 3 * nm);
    if (req.warped) to_host(c, req.warped, jj.p, nf);
    TIPLHIP_CHECK(hipStreamSynchronize(c->stream));
    res.cdm_ms = ms_between(t2, steady::now());
}

// ---------------------------------------------------------------------------
// Batcher: merges concurrent candidate sets into multi-pair launches.
// ---------------------------------------------------------------------------
class Batcher {
public:
    // `launchers` threads, each with its own context (stream): while one
    // batch runs on the device the next one is formed and launched.
    Batcher(int device, int launchers)
    {
        for (int i = 0; i < launchers; ++i) ctx_.push_back(std::make_unique<tipl::hip::context>(device));
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

    // The cost_evaluator of a served registration: blocks until the
    // candidates have been evaluated in some launch.
    void evaluate(const tiplhip_volume* from, const float* from_vs, const tiplhip_volume* to,
                  const float* to_vs, const float* params, int K, int cost, int bins, double* costs)
    {
        int fd[3], td[3];
        rc_check(tiplhip_volume_dims(from, fd));
        rc_check(tiplhip_volume_dims(to, td));
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
                rc_check(tiplhip_affine_matrix(params + 12 * size_t(k0 + k), fd, from_vs, td, to_vs,
                                               &j.T[12 * size_t(k)]));
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

    void run(tipl::hip::context& ctx)
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
            std::vector<tiplhip_pair_costs> pc;
            for (Job* j : batch) pc.push_back({j->from, j->to, j->T.data(), j->K, j->costs});
            const int rc = tiplhip_cost_batch_pairs(ctx.get(), pc.data(), int(pc.size()), cost, bins);
            const std::string err = rc == TIPLHIP_OK ? std::string() : std::string(tiplhip_last_error());
            {
                std::lock_guard<std::mutex> lk(mu_);
                for (Job* j : batch) {
                    j->rc = rc;
                    j->err = err;
                    j->done = true;
                }
                ++launches_;
                pairs_ += int64_t(batch.size());
            }
            done_cv_.notify_all();
        }
    }

    std::vector<std::unique_ptr<tipl::hip::context>> ctx_;
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
};

}  // namespace
}  // namespace tiplhip

struct tiplhip_server {
    int device = 0;
    size_t budget = 0;
    std::unique_ptr<tiplhip::Batcher> batcher;
    std::mutex mu;
    std::condition_variable work_cv, done_cv;
    std::deque<uint64_t> queue;
    std::map<uint64_t, std::unique_ptr<tiplhip::Ticket>> tickets;
    uint64_t next_ticket = 1;
    size_t in_use = 0, peak = 0;
    bool stopping = false;
    int64_t submitted = 0, completed = 0, failed = 0, running = 0;
    std::vector<double> latencies;
    tiplhip::steady::time_point first_submit, last_done;
    std::vector<std::unique_ptr<tipl::hip::context>> contexts;  // one stream per slot
    std::vector<std::thread> workers;

    void worker(tipl::hip::context& ctx)
    {
        using namespace tiplhip;
        // with a batcher, candidate sets of all slots are merged into shared
        // launches; without one, each slot launches its own on its stream
        tipl::reg::detail::cost_evaluator eval =
            batcher ? tipl::reg::detail::cost_evaluator(
                          [this](const tiplhip_volume* from, const float* fvs, const tiplhip_volume* to,
                                 const float* tvs, const float* params, int K, int cost, int bins,
                                 double* costs) {
                              batcher->evaluate(from, fvs, to, tvs, params, K, cost, bins, costs);
                          })
                    : tipl::reg::detail::direct_evaluator(ctx);
        for (;;) {
            Ticket* t = nullptr;
            {
                std::unique_lock<std::mutex> lk(mu);
                work_cv.wait(lk, [&] {
                    if (queue.empty()) return stopping;
                    return in_use + tickets[queue.front()]->bytes <= budget;
                });
                if (queue.empty()) return;
                t = tickets[queue.front()].get();
                queue.pop_front();
                in_use += t->bytes;
                peak = std::max(peak, in_use);
                ++running;
                t->started = steady::now();
            }
            int status = TIPLHIP_OK;
            std::string err;
            try {
                register_pair(ctx, t->req, eval, t->res);
            } catch (const Error& e) {
                status = e.code;
                err = e.what();
            } catch (const std::exception& e) {
                status = TIPLHIP_ERR_INVALID;
                err = e.what();
            }
            {
                std::lock_guard<std::mutex> lk(mu);
                t->finished = steady::now();
                t->res.status = status;
                t->err = err;
                t->res.queue_ms = ms_between(t->submitted, t->started);
                t->res.run_ms = ms_between(t->started, t->finished);
                t->done = true;
                in_use -= t->bytes;
                --running;
                if (status == TIPLHIP_OK) {
                    ++completed;
                    latencies.push_back(ms_between(t->submitted, t->finished));
                    last_done = t->finished;
                } else {
                    ++failed;
                }
            }
            done_cv.notify_all();
            work_cv.notify_all();
        }
    }
};

using namespace tiplhip;

extern "C" {

void tiplhip_reg_request_default(tiplhip_reg_request* r)
{
    if (!r) return;
    std::memset(r, 0, sizeof(*r));
    r->fixed_dtype = r->moving_dtype = TIPLHIP_U8;
    for (int i = 0; i < 3; ++i) r->fixed_vs[i] = r->moving_vs[i] = 1.0f;
    r->cost = TIPLHIP_COST_MI;
    r->bins = 256;
    tiplhip_cdm_param_default(&r->cdm);
    r->match_fov = 1;
    r->masked_r = 1;
    for (int i = 6; i < 9; ++i) r->arg[i] = 1.0f;  // affine_param's identity scaling
}

size_t tiplhip_request_bytes(const tiplhip_reg_request* req)
{
    return req ? request_bytes(*req) : 0;
}

int tiplhip_register(tiplhip_ctx* ctx, const tiplhip_reg_request* req, tiplhip_reg_result* res)
{
    return guarded([&] {
        TIPLHIP_REQUIRE(ctx != nullptr && res != nullptr, "null argument");
        std::memset(res, 0, sizeof(*res));
        res->status = TIPLHIP_ERR_INVALID;
        check_request(req);
        tipl::hip::context c(ctx);
        const auto t0 = steady::now();
        try {
            register_pair(c, *req, tipl::reg::detail::direct_evaluator(c), *res);
        } catch (const Error& e) {
            res->status = e.code;
            throw;
        }
        res->run_ms = ms_between(t0, steady::now());
        res->status = TIPLHIP_OK;
    });
}

int tiplhip_server_create(int device, int slots, size_t hbm_budget, tiplhip_server** out)
{
    return tiplhip_server_create_ex(device, slots, -1, hbm_budget, out);
}

int tiplhip_server_create_ex(int device, int slots, int launchers, size_t hbm_budget,
                             tiplhip_server** out)
{
    return guarded([&] {
        TIPLHIP_REQUIRE(out != nullptr, "null output");
        *out = nullptr;
        TIPLHIP_REQUIRE(slots >= 1 && slots <= 64, "slots must be in [1, 64]");
        TIPLHIP_REQUIRE(launchers >= -1 && launchers <= 16, "launchers must be in [-1, 16]");
        TIPLHIP_REQUIRE(hbm_budget > 0, "zero memory budget");
        int n = 0;
        rc_check(tiplhip_device_count(&n));
        TIPLHIP_REQUIRE(device >= 0 && device < n, "no such device");
        auto s = std::make_unique<tiplhip_server>();
        s->device = device;
        s->budget = hbm_budget;
        // default: one launcher per 4 slots (at least 1, at most 4); 0: no batcher
        if (launchers < 0) launchers = std::min(4, std::max(1, slots / 4));
        if (launchers > 0) s->batcher = std::make_unique<Batcher>(device, launchers);
        for (int i = 0; i < slots; ++i) s->contexts.push_back(std::make_unique<tipl::hip::context>(device));
        for (int i = 0; i < slots; ++i)
            s->workers.emplace_back([p = s.get(), c = s->contexts[size_t(i)].get()] { p->worker(*c); });
        *out = s.release();
    });
}

int tiplhip_server_submit(tiplhip_server* s, const tiplhip_reg_request* req, uint64_t* ticket)
{
    return guarded([&] {
        TIPLHIP_REQUIRE(s != nullptr && ticket != nullptr, "null argument");
        check_request(req);
        auto t = std::make_unique<Ticket>();
        t->req = *req;
        std::memset(&t->res, 0, sizeof(t->res));
        t->bytes = request_bytes(*req);
        TIPLHIP_REQUIRE(t->bytes <= s->budget, "request needs more device memory than the budget");
        {
            std::lock_guard<std::mutex> lk(s->mu);
            TIPLHIP_REQUIRE(!s->stopping, "server is stopping");
            t->submitted = steady::now();
            if (s->submitted == 0) s->first_submit = t->submitted;
            ++s->submitted;
            *ticket = s->next_ticket++;
            s->queue.push_back(*ticket);
            s->tickets[*ticket] = std::move(t);
        }
        s->work_cv.notify_all();
    });
}

int tiplhip_server_wait(tiplhip_server* s, uint64_t ticket, tiplhip_reg_result* res)
{
    return guarded([&] {
        TIPLHIP_REQUIRE(s != nullptr && res != nullptr, "null argument");
        std::unique_lock<std::mutex> lk(s->mu);
        auto it = s->tickets.find(ticket);
        TIPLHIP_REQUIRE(it != s->tickets.end(), "unknown ticket");
        Ticket* t = it->second.get();
        s->done_cv.wait(lk, [&] { return t->done; });
        *res = t->res;
        const int status = t->res.status;
        const std::string err = t->err;
        s->tickets.erase(it);
        if (status != TIPLHIP_OK) throw Error(status, err);
    });
}

int tiplhip_server_stats_get(tiplhip_server* s, tiplhip_server_stats* st)
{
    return guarded([&] {
        TIPLHIP_REQUIRE(s != nullptr && st != nullptr, "null argument");
        std::memset(st, 0, sizeof(*st));
        std::vector<double> lat;
        {
            std::lock_guard<std::mutex> lk(s->mu);
            st->submitted = s->submitted;
            st->completed = s->completed;
            st->failed = s->failed;
            st->running = s->running;
            st->queued = int64_t(s->queue.size());
            st->hbm_in_use = s->in_use;
            st->hbm_peak = s->peak;
            lat = s->latencies;
            if (s->completed > 0) st->wall_s = ms_between(s->first_submit, s->last_done) * 1e-3;
        }
        if (s->batcher) s->batcher->counts(st->batch_launches, st->batched_pairs);
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
    });
}

int tiplhip_server_destroy(tiplhip_server* s)
{
    return guarded([&] {
        if (!s) return;
        {
            std::lock_guard<std::mutex> lk(s->mu);
            s->stopping = true;
        }
        s->work_cv.notify_all();
        for (auto& w : s->workers) w.join();
        s->batcher.reset();
        delete s;
    });
}

}  // extern "C"
