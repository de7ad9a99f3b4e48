// server.cpp -- one-pair registration pipeline and the serving loop.
//
// tiplhip_register runs the reference's mm_reg order for one image pair
// (reg/mm_reg.hpp:348-416): tipl::reg::linear(fixed, moving) -- the body of
// the header-only mirror in include/tipl/reg/linear.hpp, so the optimiser is
// the one C++ callers get -- then, for nonlinear requests, the moving image
// resampled onto the fixed grid and cdm(fixed, J) (reg/cdm.hpp:373-474).
//
// The server (scheduling in server_core.hpp) runs many of these at once on
// one GPU:
//   * `slots` worker threads, each with its own context (HIP stream), take
//     requests FIFO when their estimated device memory fits the budget;
//   * batch launcher threads own the affine cost launches: every optimiser
//     candidate set of every running registration is queued to them, and a
//     free launcher launches everything queued (up to 4096 candidates, one
//     cost type) as one multi-pair launch -- registrations join and leave
//     the batch at every optimiser step;
//   * the CDM iterations of all running registrations (both directions)
//     are launched together: each iteration's device work is one step for
//     the device's CdmBatcher (cdm.hip), which merges the steps of up to 8
//     fields into one launch per kernel.
// Batched costs are bit-identical to single-pair costs and the optimiser is
// deterministic, so a served result equals tiplhip_register's for the same
// request.
#include <algorithm>
#include <memory>
#include <string>
#include <thread>

#include "../../include/tipl/reg/linear.hpp"
#include "field.hpp"
#include "host_wait.hpp"
#include "internal.hpp"
#include "server_core.hpp"
#include "warp.hpp"

namespace tiplhip {
namespace {

void rc_check(int rc)
{
    if (rc != TIPLHIP_OK) throw Error(rc, tiplhip_last_error());
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
    // on an exception, wait for this stream: buffers declared before this
    // point (some owned by another stream, whose frees are not ordered after
    // this stream's work) may still be read by kernels queued here
    auto drain_stream = [c] { (void)hipStreamSynchronize(c->stream); };
    DrainOnUnwind<decltype(drain_stream)> drain(drain_stream);
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

// The device side of the serving loop (server_core.hpp): one context
// (stream) per worker slot and per batch launcher.
struct HipBackend {
    using Context = tipl::hip::context;
    // Every worker context shares the device's CDM batcher, so the cdm()
    // iterations of all running registrations go out as batched steps
    // (TIPLHIP_SERVER_CDM_BATCH=0 at server creation turns this off).
    static std::unique_ptr<Context> make(int device)
    {
        auto c = std::make_unique<Context>(device);
        if (cdm_batching_enabled()) c->get()->cdm_batch = shared_cdm_batcher(device);
        return c;
    }
    static void register_pair(Context& ctx, const tiplhip_reg_request& req, const CostEvaluator& eval,
                              tiplhip_reg_result& res)
    {
        tiplhip::register_pair(ctx, req, eval, res);
    }
    static CostEvaluator direct(Context& ctx) { return tipl::reg::detail::direct_evaluator(ctx); }
    static void dims(const tiplhip_volume* v, int* d) { rc_check(tiplhip_volume_dims(v, d)); }
    static void transform(const float* p, const int* fd, const float* fvs, const int* td,
                          const float* tvs, float* T)
    {
        rc_check(tiplhip_affine_matrix(p, fd, fvs, td, tvs, T));
    }
    static void launch(Context& ctx, tiplhip_pair_costs* pairs, int n, int cost, int bins)
    {
        rc_check(tiplhip_cost_batch_pairs(ctx.get(), pairs, n, cost, bins));
    }
};

}  // namespace
}  // namespace tiplhip

struct tiplhip_server {
    tiplhip::ServerCore<tiplhip::HipBackend> core;
    // the device's CDM batcher (shared with any other server on the device)
    // and its counts when this server started
    std::shared_ptr<tiplhip::CdmBatcher> cdm;
    int64_t cdm_steps0 = 0, cdm_fields0 = 0;
    tiplhip_server(int device, int slots, int launchers, size_t budget)
        : core(device, slots, launchers, budget)
    {
        if (tiplhip::cdm_batching_enabled()) {
            cdm = tiplhip::shared_cdm_batcher(device);
            tiplhip::cdm_batcher_counts(*cdm, &cdm_steps0, &cdm_fields0);
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
        int n = 0;
        rc_check(tiplhip_device_count(&n));
        TIPLHIP_REQUIRE(device >= 0 && device < n, "no such device");
        *out = new tiplhip_server(device, slots, launchers, hbm_budget);
    });
}

int tiplhip_server_submit(tiplhip_server* s, const tiplhip_reg_request* req, uint64_t* ticket)
{
    return guarded([&] {
        TIPLHIP_REQUIRE(s != nullptr && ticket != nullptr, "null argument");
        *ticket = s->core.submit(req);
    });
}

int tiplhip_server_wait(tiplhip_server* s, uint64_t ticket, tiplhip_reg_result* res)
{
    return guarded([&] {
        TIPLHIP_REQUIRE(s != nullptr && res != nullptr, "null argument");
        s->core.wait(ticket, res);
    });
}

int tiplhip_server_stats_get(tiplhip_server* s, tiplhip_server_stats* st)
{
    return guarded([&] {
        TIPLHIP_REQUIRE(s != nullptr && st != nullptr, "null argument");
        s->core.stats(st);
        st->cdm_steps = st->cdm_fields = 0;
        if (s->cdm) {
            tiplhip::cdm_batcher_counts(*s->cdm, &st->cdm_steps, &st->cdm_fields);
            st->cdm_steps -= s->cdm_steps0;
            st->cdm_fields -= s->cdm_fields0;
        }
    });
}

int tiplhip_server_destroy(tiplhip_server* s)
{
    return guarded([&] { delete s; });
}

}  // extern "C"
