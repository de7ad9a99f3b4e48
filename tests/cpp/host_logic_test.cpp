// host_logic_test.cpp -- the library's host-only logic on a CPU-only host,
// built by tests/test_host_sanitizers.py under AddressSanitizer +
// UndefinedBehaviorSanitizer and under ThreadSanitizer:
//   * the serving loop (tipl_amd/csrc/server_core.hpp) over a stub backend
//     whose "registration" reads every input voxel, evaluates candidate sets
//     through the server's cost evaluator and writes every output element:
//     admission, concurrent submit / wait, batching across slots, failure
//     isolation (a throwing registration, a pair the cost launch refuses),
//     ticket misuse, shutdown with requests still queued;
//   * the host's wait for results that kernels store into host memory
//     (tipl_amd/csrc/host_wait.hpp): spin, timeout, fallback, drain on unwind;
//   * the gate of batched CDM steps (tipl_amd/csrc/step_gate.hpp): jobs in
//     step share batches, every step is launched once and in order, a
//     leaving job releases the steps waiting for it, a failed launch fails
//     exactly the jobs of its batch.
// Exits 0 when every check holds; prints the failed check otherwise.
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <mutex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <thread>
#include <vector>

#include "../../tipl_amd/csrc/host_wait.hpp"
#include "../../tipl_amd/csrc/server_core.hpp"
#include "../../tipl_amd/csrc/step_gate.hpp"

// errors.hpp declares these for guarded(); the core does not call them.
namespace tiplhip {
void set_last_error(const std::string&) {}
void clear_last_error() {}
}  // namespace tiplhip

struct tiplhip_volume {
    int dim[3];
    bool refused;  // the cost launch rejects this volume
};

#define CHECK(cond)                                                                          \
    do {                                                                                     \
        if (!(cond)) {                                                                       \
            std::fprintf(stderr, "FAILED %s:%d: %s\n", __FILE__, __LINE__, #cond);           \
            std::exit(1);                                                                    \
        }                                                                                    \
    } while (0)

namespace {

using tiplhip::CostEvaluator;
using tiplhip::Error;

thread_local std::mt19937 jitter_rng{std::random_device{}()};

void jitter(int max_us)
{
    std::uniform_int_distribution<int> d(0, max_us);
    std::this_thread::sleep_for(std::chrono::microseconds(d(jitter_rng)));
}

// A pure function of the transform and the pair's shapes: the same wherever
// and with whatever else a candidate is evaluated.
double stub_cost(const float* T, const int* fd, const int* td)
{
    double c = 0.0;
    for (int i = 0; i < 12; ++i) c += double(T[i]) * double(i + 1);
    return c + 1e-3 * double(fd[0] + 2 * td[1] + 3 * fd[2]);
}

struct StubBackend {
    struct Context {
        int device;
    };
    static std::atomic<int> launches;

    static std::unique_ptr<Context> make(int device) { return std::make_unique<Context>(Context{device}); }
    static void dims(const tiplhip_volume* v, int* d) { std::memcpy(d, v->dim, sizeof(v->dim)); }
    static void transform(const float* p, const int* fd, const float* fvs, const int* td,
                          const float* tvs, float* T)
    {
        for (int i = 0; i < 12; ++i) T[i] = p[i] * fvs[i % 3] + float(td[i % 3]) / tvs[i % 3] - float(fd[i % 3]);
    }
    static void launch(Context&, tiplhip_pair_costs* pairs, int n, int, int)
    {
        ++launches;
        jitter(150);
        for (int j = 0; j < n; ++j)
            if (pairs[j].to->refused) throw Error(TIPLHIP_ERR_INVALID, "volume refused by the cost launch");
        for (int j = 0; j < n; ++j)
            for (int k = 0; k < pairs[j].K; ++k)
                pairs[j].costs[k] = stub_cost(pairs[j].T + 12 * k, pairs[j].from->dim, pairs[j].to->dim);
    }
    static CostEvaluator direct(Context& ctx)
    {
        return [&ctx](const tiplhip_volume* from, const float* fvs, const tiplhip_volume* to,
                      const float* tvs, const float* params, int K, int cost, int bins, double* costs) {
            std::vector<float> T(12 * size_t(K));
            for (int k = 0; k < K; ++k) transform(params + 12 * k, from->dim, fvs, to->dim, tvs, &T[12 * size_t(k)]);
            tiplhip_pair_costs pc{from, to, T.data(), K, costs};
            launch(ctx, &pc, 1, cost, bins);
        };
    }

    // Reads every input voxel, runs a few optimiser-like rounds of candidate
    // sets through eval, writes every requested output element.
    static void register_pair(Context&, const tiplhip_reg_request& req, const CostEvaluator& eval,
                              tiplhip_reg_result& res)
    {
        const size_t nf = tiplhip::vox3(req.fixed_dim), nm = tiplhip::vox3(req.moving_dim);
        auto checksum = [](const void* p, int dt, size_t n) {
            double s = 0.0;
            if (dt == TIPLHIP_U8)
                for (size_t i = 0; i < n; ++i) s += static_cast<const uint8_t*>(p)[i];
            else if (dt == TIPLHIP_I16)
                for (size_t i = 0; i < n; ++i) s += static_cast<const int16_t*>(p)[i];
            else
                for (size_t i = 0; i < n; ++i) s += double(static_cast<const float*>(p)[i]);
            return s;
        };
        const double sf = checksum(req.fixed, req.fixed_dtype, nf);
        const double sm = checksum(req.moving, req.moving_dtype, nm);
        auto from = std::make_unique<tiplhip_volume>(tiplhip_volume{{req.fixed_dim[0], req.fixed_dim[1], req.fixed_dim[2]}, false});
        auto to = std::make_unique<tiplhip_volume>(tiplhip_volume{{req.moving_dim[0], req.moving_dim[1], req.moving_dim[2]},
                                                                  req.moving_vs[0] == 13.0f});
        double best = 1e300;
        int64_t evals = 0;
        float best_p[12] = {};
        for (int round = 0; round < 6; ++round) {
            if (req.fixed_vs[1] == 7.0f && round == 2) throw Error(TIPLHIP_ERR_DEVICE, "stub registration failure");
            const int K = 1 + int(std::fmod(sf + 31.0 * round, 23.0));
            std::vector<float> P(12 * size_t(K));
            for (size_t i = 0; i < P.size(); ++i) P[i] = float(std::fmod(sm * 0.001 + double(i) * 0.37 + round, 5.0));
            std::vector<double> c(static_cast<size_t>(K));
            eval(from.get(), req.fixed_vs, to.get(), req.moving_vs, P.data(), K, req.cost, req.bins, c.data());
            evals += K;
            for (int k = 0; k < K; ++k)
                if (c[size_t(k)] < best) {
                    best = c[size_t(k)];
                    std::memcpy(best_p, &P[12 * size_t(k)], sizeof(best_p));
                }
            jitter(100);
        }
        std::memcpy(res.affine, best_p, sizeof(best_p));
        res.cost = float(best);
        res.evaluations = evals;
        res.cdm_iterations = req.nonlinear ? int(nf % 97) : 0;
        if (req.nonlinear && req.field)
            for (size_t i = 0; i < 3 * nf; ++i) req.field[i] = float(i % 13) * float(best);
        if (req.warped)
            for (size_t i = 0; i < nf; ++i) req.warped[i] = uint8_t((i + evals) & 0xff);
    }
};
std::atomic<int> StubBackend::launches{0};

using Server = tiplhip::ServerCore<StubBackend>;

struct Req {
    std::vector<uint8_t> fixed, moving;  // raw bytes of the request's dtype
    std::vector<float> field;
    std::vector<uint8_t> warped;
    tiplhip_reg_request r{};
};

std::unique_ptr<Req> make_req(int i, bool nonlinear, bool refused = false, bool throws = false)
{
    auto q = std::make_unique<Req>();
    std::memset(&q->r, 0, sizeof(q->r));
    const int dts[3] = {TIPLHIP_U8, TIPLHIP_I16, TIPLHIP_F32};
    const int es[3] = {1, 2, 4};
    const int a = 3 + i % 7, b = 4 + (i * 5) % 9, c = 2 + (i * 3) % 6;
    const int fdt = i % 3, mdt = (i / 3) % 3;
    q->r.fixed_dim[0] = a, q->r.fixed_dim[1] = b, q->r.fixed_dim[2] = c;
    q->r.moving_dim[0] = b, q->r.moving_dim[1] = c + 1, q->r.moving_dim[2] = a;
    q->fixed.resize(size_t(es[fdt]) * tiplhip::vox3(q->r.fixed_dim));
    q->moving.resize(size_t(es[mdt]) * tiplhip::vox3(q->r.moving_dim));
    std::mt19937 g(uint32_t(1000 + i));
    for (auto& v : q->fixed) v = uint8_t(g() & 0x3f);
    for (auto& v : q->moving) v = uint8_t(g() & 0x3f);
    q->r.fixed = q->fixed.data();
    q->r.moving = q->moving.data();
    q->r.fixed_dtype = dts[fdt];
    q->r.moving_dtype = dts[mdt];
    for (int k = 0; k < 3; ++k) q->r.fixed_vs[k] = q->r.moving_vs[k] = 1.0f;
    if (refused) q->r.moving_vs[0] = 13.0f;
    if (throws) q->r.fixed_vs[1] = 7.0f;
    q->r.cost = i % 4 == 0 ? TIPLHIP_COST_CORR : TIPLHIP_COST_MI;
    q->r.bins = i % 5 == 0 ? 64 : 256;
    q->r.nonlinear = nonlinear ? 1 : 0;
    if (nonlinear) {
        q->field.resize(3 * tiplhip::vox3(q->r.fixed_dim));
        q->r.field = q->field.data();
    }
    q->warped.resize(tiplhip::vox3(q->r.fixed_dim));
    q->r.warped = q->warped.data();
    return q;
}

// The plain path: the same registration with costs evaluated directly.
tiplhip_reg_result plain(Req& q, std::vector<float>& field, std::vector<uint8_t>& warped)
{
    tiplhip_reg_result res;
    std::memset(&res, 0, sizeof(res));
    StubBackend::Context ctx{0};
    tiplhip_reg_request r = q.r;
    field.assign(q.field.size(), 0.0f);
    warped.assign(q.warped.size(), 0);
    r.field = q.field.empty() ? nullptr : field.data();
    r.warped = warped.data();
    StubBackend::register_pair(ctx, r, StubBackend::direct(ctx), res);
    return res;
}

void same(const tiplhip_reg_result& a, const tiplhip_reg_result& b)
{
    CHECK(a.status == TIPLHIP_OK);
    CHECK(std::memcmp(a.affine, b.affine, sizeof(a.affine)) == 0);
    CHECK(a.cost == b.cost);
    CHECK(a.evaluations == b.evaluations);
    CHECK(a.cdm_iterations == b.cdm_iterations);
}

void test_concurrent_clients(int launchers)
{
    const int kClients = 4, kPer = 12;
    std::vector<std::unique_ptr<Req>> reqs;
    for (int i = 0; i < kClients * kPer; ++i) reqs.push_back(make_req(i, i % 3 == 1));
    size_t largest = 0;
    for (auto& q : reqs) largest = std::max(largest, tiplhip::request_bytes(q->r));
    // room for about three requests at once: admission has to hold some back
    const size_t budget = 3 * largest;
    Server srv(0, 8, launchers, budget);
    std::vector<tiplhip_reg_result> got(reqs.size());
    std::vector<std::thread> clients;
    for (int c = 0; c < kClients; ++c)
        clients.emplace_back([&, c] {
            std::vector<uint64_t> tickets;
            for (int i = c * kPer; i < (c + 1) * kPer; ++i) {
                tickets.push_back(srv.submit(&reqs[size_t(i)]->r));
                jitter(200);
            }
            // wait in reverse order of submission
            for (int i = kPer - 1; i >= 0; --i) srv.wait(tickets[size_t(i)], &got[size_t(c * kPer + i)]);
        });
    for (auto& t : clients) t.join();
    tiplhip_server_stats st;
    srv.stats(&st);
    CHECK(st.completed == int64_t(reqs.size()) && st.failed == 0);
    CHECK(st.running == 0 && st.queued == 0 && st.hbm_in_use == 0);
    CHECK(st.hbm_peak <= budget);
    CHECK(st.latency_p50_ms <= st.latency_p95_ms && st.latency_p95_ms <= st.latency_p99_ms);
    if (launchers > 0) CHECK(st.batched_pairs > 0 && st.batch_launches > 0);
    for (size_t i = 0; i < reqs.size(); ++i) {
        std::vector<float> f;
        std::vector<uint8_t> w;
        same(got[i], plain(*reqs[i], f, w));
        CHECK(f == reqs[i]->field);
        CHECK(w == reqs[i]->warped);
    }
}

void test_failure_isolation()
{
    std::vector<std::unique_ptr<Req>> reqs;
    for (int i = 0; i < 16; ++i) reqs.push_back(make_req(100 + i, i % 4 == 0, i == 5, i == 9));
    Server srv(0, 6, 1, size_t(1) << 40);
    std::vector<uint64_t> tickets;
    for (auto& q : reqs) tickets.push_back(srv.submit(&q->r));
    for (size_t i = 0; i < reqs.size(); ++i) {
        tiplhip_reg_result res;
        try {
            srv.wait(tickets[i], &res);
            CHECK(i != 5 && i != 9);
            std::vector<float> f;
            std::vector<uint8_t> w;
            same(res, plain(*reqs[i], f, w));
        } catch (const Error& e) {
            CHECK(i == 5 || i == 9);
            CHECK(e.code == (i == 5 ? TIPLHIP_ERR_INVALID : TIPLHIP_ERR_DEVICE));
        }
    }
    tiplhip_server_stats st;
    srv.stats(&st);
    CHECK(st.completed == 14 && st.failed == 2);
}

void test_ticket_misuse_and_admission()
{
    Server srv(0, 2, 1, size_t(64) << 20);
    auto q = make_req(7, false);
    auto big = make_req(8, true);
    big->r.fixed_dim[0] = big->r.fixed_dim[1] = big->r.fixed_dim[2] = 4096;  // never read: rejected
    bool threw = false;
    try {
        srv.submit(&big->r);
    } catch (const Error& e) {
        threw = e.code == TIPLHIP_ERR_INVALID;
    }
    CHECK(threw);
    tiplhip_reg_request bad = q->r;
    bad.bins = 100;
    threw = false;
    try {
        srv.submit(&bad);
    } catch (const Error&) {
        threw = true;
    }
    CHECK(threw);
    const uint64_t t = srv.submit(&q->r);
    std::atomic<bool> second_rejected{false};
    std::thread waiter([&] {
        tiplhip_reg_result res;
        srv.wait(t, &res);
        CHECK(res.status == TIPLHIP_OK);
    });
    // a second wait on the same ticket is rejected, whether it comes while
    // the first caller waits or after the ticket was returned and forgotten
    std::this_thread::sleep_for(std::chrono::milliseconds(1));
    try {
        tiplhip_reg_result res;
        srv.wait(t, &res);
    } catch (const Error&) {
        second_rejected = true;
    }
    waiter.join();
    CHECK(second_rejected);
    threw = false;
    try {
        tiplhip_reg_result res;
        srv.wait(12345, &res);
    } catch (const Error&) {
        threw = true;
    }
    CHECK(threw);
}

void test_shutdown_with_queue()
{
    std::vector<std::unique_ptr<Req>> reqs;
    for (int i = 0; i < 10; ++i) reqs.push_back(make_req(200 + i, i % 2 == 0));
    {
        Server srv(0, 2, 2, size_t(1) << 40);
        for (auto& q : reqs) srv.submit(&q->r);
        // destroyed with requests queued and never waited for: it finishes them
    }
    for (auto& q : reqs) {
        std::vector<float> f;
        std::vector<uint8_t> w;
        plain(*q, f, w);
        CHECK(w == q->warped);
    }
}

void test_wait_for_slots()
{
    using tiplhip::wait_for_slots;
    constexpr uint32_t kPending = 0x7FC0C057u;
    // already written: no fallback
    {
        uint32_t s[4] = {1, 2, 3, 4};
        int calls = 0;
        CHECK(wait_for_slots<uint32_t>(s, 4, kPending, std::chrono::microseconds(100), [&] { ++calls; }) == 4);
        CHECK(calls == 0);
    }
    // never written: the fallback runs once, after the spin budget
    {
        uint32_t s[3] = {5, kPending, kPending};
        int calls = 0;
        const auto t0 = std::chrono::steady_clock::now();
        const size_t r = wait_for_slots<uint32_t>(s, 3, kPending, std::chrono::microseconds(3000), [&] { ++calls; });
        const auto dt = std::chrono::steady_clock::now() - t0;
        CHECK(r == 1 && calls == 1);
        CHECK(dt >= std::chrono::microseconds(3000));
    }
    // written by another thread while the host spins: no fallback
    {
        std::vector<uint32_t> s(8, kPending);
        std::atomic<bool> go{false};
        std::thread w([&] {
            while (!go) std::this_thread::yield();
            for (size_t i = 0; i < s.size(); ++i) __atomic_store_n(&s[i], uint32_t(i), __ATOMIC_RELAXED);
        });
        int calls = 0;
        go = true;
        const size_t r = wait_for_slots<uint32_t>(reinterpret_cast<volatile uint32_t*>(s.data()), s.size(),
                                                  kPending, std::chrono::seconds(5), [&] { ++calls; });
        w.join();
        CHECK(r == s.size() && calls == 0);
    }
    // a fallback that throws (a failed launch reported by the synchronisation)
    {
        uint32_t s[1] = {kPending};
        bool caught = false;
        try {
            wait_for_slots<uint32_t>(s, 1, kPending, std::chrono::microseconds(10),
                                     [] { throw Error(TIPLHIP_ERR_DEVICE, "launch failed"); });
        } catch (const Error& e) {
            caught = e.code == TIPLHIP_ERR_DEVICE;
        }
        CHECK(caught);
    }
}

// ---------------------------------------------------------------------------
// StepGate
// ---------------------------------------------------------------------------
struct GateJob : tiplhip::GateSlot {
    int id = 0;
    int next = 0;        // the step the owner queues (set before run)
    int launched_steps = 0;  // written by the launch (under the gate's lock)
    bool poison = false; // the launch fails the batch holding this step
};

struct GateLog {
    std::vector<int> sizes;  // batch sizes, in launch order (gate lock held)
};

// Start barrier for the gate's threads.
struct Barrier {
    std::mutex mu;
    std::condition_variable cv;
    int waiting = 0, n;
    explicit Barrier(int count) : n(count) {}
    void arrive()
    {
        std::unique_lock<std::mutex> lk(mu);
        if (++waiting == n) {
            cv.notify_all();
            return;
        }
        cv.wait(lk, [&] { return waiting >= n; });
    }
};

void test_step_gate_lockstep(int threads, int max_batch)
{
    GateLog log;
    tiplhip::StepGate<GateJob> gate(max_batch, std::chrono::seconds(10), [&](GateJob* const* J, int n) {
        CHECK(n >= 1 && n <= max_batch);
        for (int i = 0; i < n; ++i) {
            for (int k = 0; k < i; ++k) CHECK(J[k] != J[i]);
            CHECK(J[i]->next == J[i]->launched_steps);  // in order, once each
            ++J[i]->launched_steps;
        }
        log.sizes.push_back(n);
    });
    const int steps = 40;
    std::vector<GateJob> jobs(static_cast<size_t>(threads));
    Barrier start(threads);
    std::vector<std::thread> ts;
    const auto t0 = std::chrono::steady_clock::now();
    for (int t = 0; t < threads; ++t) {
        ts.emplace_back([&, t] {
            GateJob& j = jobs[size_t(t)];
            j.id = t;
            gate.join();
            start.arrive();
            for (int i = 0; i < steps; ++i) {
                j.next = i;
                gate.run(j);
                jitter(30);  // the device step and the host's decision
            }
            gate.leave([] {});
        });
    }
    for (auto& th : ts) th.join();
    CHECK(std::chrono::steady_clock::now() - t0 < std::chrono::seconds(5));  // no gather timeout
    for (const GateJob& j : jobs) CHECK(j.launched_steps == steps);
    CHECK(gate.steps() == int64_t(threads) * steps);
    CHECK(gate.batches() == int64_t(log.sizes.size()));
    if (threads <= max_batch) {
        // every job joined before the first step: each batch holds all of them
        CHECK(gate.batches() == steps);
        for (int n : log.sizes) CHECK(n == threads);
    } else {
        for (int n : log.sizes) CHECK(n <= max_batch);
        CHECK(gate.batches() < gate.steps());
    }
}

void test_step_gate_leave_releases()
{
    // A and B wait for C, which leaves without queueing another step: its
    // leave() launches their steps (the gather time is far longer than the test).
    tiplhip::StepGate<GateJob> gate(8, std::chrono::seconds(30), [](GateJob* const* J, int n) {
        for (int i = 0; i < n; ++i) ++J[i]->launched_steps;
    });
    GateJob a, b, c;
    gate.join();
    gate.join();
    gate.join();
    const auto t0 = std::chrono::steady_clock::now();
    std::thread ta([&] { gate.run(a); });
    std::thread tb([&] { gate.run(b); });
    std::this_thread::sleep_for(std::chrono::milliseconds(20));
    CHECK(!a.launched.load() && !b.launched.load());
    gate.leave([] {});
    ta.join();
    tb.join();
    CHECK(a.launched_steps == 1 && b.launched_steps == 1 && c.launched_steps == 0);
    CHECK(std::chrono::steady_clock::now() - t0 < std::chrono::seconds(5));
    gate.leave([] {});
    gate.leave([] {});
    // alone: a step goes out at once
    gate.join();
    gate.run(a);
    CHECK(a.launched_steps == 2);
    gate.leave([] {});
    // gather timeout: two active, one queues; it is launched after ~gather
    tiplhip::StepGate<GateJob> quick(8, std::chrono::milliseconds(2), [](GateJob* const* J, int n) {
        for (int i = 0; i < n; ++i) ++J[i]->launched_steps;
    });
    quick.join();
    quick.join();
    const auto t1 = std::chrono::steady_clock::now();
    quick.run(c);
    CHECK(c.launched_steps == 1);
    CHECK(std::chrono::steady_clock::now() - t1 >= std::chrono::milliseconds(2));
    quick.leave([] {});
    quick.leave([] {});
}

void test_step_gate_failure()
{
    // Four jobs in step; at step 3 job 2's launch throws: exactly the jobs of
    // that batch see the error, and later steps launch normally.
    const int threads = 4, steps = 8;
    tiplhip::StepGate<GateJob> gate(8, std::chrono::seconds(10), [](GateJob* const* J, int n) {
        for (int i = 0; i < n; ++i)
            if (J[i]->poison) throw Error(TIPLHIP_ERR_DEVICE, "launch failed");
        for (int i = 0; i < n; ++i) ++J[i]->launched_steps;
    });
    std::vector<GateJob> jobs(static_cast<size_t>(threads));
    std::vector<int> errors(static_cast<size_t>(threads), 0);
    Barrier start(threads);
    std::vector<std::thread> ts;
    for (int t = 0; t < threads; ++t) {
        ts.emplace_back([&, t] {
            GateJob& j = jobs[size_t(t)];
            gate.join();
            start.arrive();
            for (int i = 0; i < steps; ++i) {
                j.poison = t == 2 && i == 3;
                try {
                    gate.run(j);
                } catch (const Error& e) {
                    CHECK(e.code == TIPLHIP_ERR_DEVICE);
                    ++errors[size_t(t)];
                }
            }
            gate.leave([] {});
        });
    }
    for (auto& th : ts) th.join();
    // all four were in the failed batch (each batch holds every job)
    for (int t = 0; t < threads; ++t) {
        CHECK(errors[size_t(t)] == 1);
        CHECK(jobs[size_t(t)].launched_steps == steps - 1);
    }
}

void test_drain_on_unwind()
{
    int drained = 0;
    auto d = [&] { ++drained; };
    {
        tiplhip::DrainOnUnwind<decltype(d)> g(d);
    }
    CHECK(drained == 0);
    try {
        tiplhip::DrainOnUnwind<decltype(d)> g(d);
        throw Error(TIPLHIP_ERR_DEVICE, "x");
    } catch (const Error&) {
    }
    CHECK(drained == 1);
    // constructed while another exception unwinds, left normally: no drain
    struct Probe {
        decltype(d)& dr;
        ~Probe()
        {
            tiplhip::DrainOnUnwind<decltype(d)> g(dr);
        }
    };
    try {
        Probe p{d};
        throw Error(TIPLHIP_ERR_DEVICE, "y");
    } catch (const Error&) {
    }
    CHECK(drained == 1);
}

}  // namespace

int main()
{
    test_wait_for_slots();
    test_drain_on_unwind();
    test_concurrent_clients(2);
    test_concurrent_clients(0);
    test_failure_isolation();
    test_ticket_misuse_and_admission();
    test_shutdown_with_queue();
    test_step_gate_lockstep(6, 8);
    test_step_gate_lockstep(12, 8);
    test_step_gate_leave_releases();
    test_step_gate_failure();
    std::printf("host logic ok (%d stub launches)\n", StubBackend::launches.load());
    return 0;
}
