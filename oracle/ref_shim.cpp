// ref_shim.cpp -- C entry points into the reference TIPL headers.
//
// TEST INFRASTRUCTURE ONLY.  Built by oracle/Makefile against the reference
// headers where they lie (-I <reference root>) into oracle/_ref/libtiplref.so.
// It contains no reference code: every function below only instantiates a
// reference template on plain buffers so that tests can check the C oracle
// (oracle/oracle.c) against the real implementation, and bench.py can time the
// reference's CPU path as its baseline.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdint>
#include <cstring>
#include <vector>

#include <map>     // io/gz_stream.hpp uses std::map without including it
#include <zlib.h>  // before io/gz_stream.hpp, which is compiled only under ZLIB_H

#include "filter/gaussian.hpp"
#include "io/gz_stream.hpp"
#include "io/nifti.hpp"
#include "numerical/otsu.hpp"
#include "reg/cdm.hpp"
#include "reg/mm_reg.hpp"
#include "reg/linear.hpp"

namespace {

tipl::shape<3> shp(const int* d) { return tipl::shape<3>(d[0], d[1], d[2]); }

tipl::transformation_matrix<float, 3> mat(const float* T)
{
    tipl::transformation_matrix<float, 3> M;
    std::copy_n(T, 12, M.begin());
    return M;
}

using field_t = tipl::image<3, tipl::vector<3>>;
using field2_t = tipl::image<2, tipl::vector<2>>;

tipl::shape<2> shp2(const int* d) { return tipl::shape<2>(d[0], d[1]); }

field2_t to_field2(const float* f, const int* d)
{
    field2_t F(shp2(d));
    std::memcpy(&F[0][0], f, sizeof(float) * 2 * F.size());
    return F;
}

void from_field2(const field2_t& F, float* f)
{
    std::memcpy(f, reinterpret_cast<const float*>(F.data()), sizeof(float) * 2 * F.size());
}

// Log sink for mm_reg<out_type>, whose show_r writes unconditionally.
struct null_out {
    template <class T>
    null_out& operator<<(const T&) { return *this; }
};

field_t to_field(const float* f, const int* d)
{
    field_t F(shp(d));
    std::memcpy(&F[0][0], f, sizeof(float) * 3 * F.size());
    return F;
}

void from_field(const field_t& F, float* f)
{
    std::memcpy(f, &F[0][0], sizeof(float) * 3 * F.size());
}

// Cost callbacks supplied by the tests: a candidate set of K 12-float
// transforms (sr[9], shift[3]) on one image pair, and a plain function of
// parameter vectors (the quadratic-bowl checks of the optimisers).
typedef void (*tref_cost_cb)(const uint8_t* from, const int* fdim, const uint8_t* to,
                             const int* tdim, const float* T, int K, int cost, int bins,
                             double* costs, void* user);
typedef void (*tref_param_cb)(const float* x, int n, int K, double* costs, void* user);

// reg::linear with line search and L-BFGS run one after the other instead of
// racing on two threads (reg/linear.hpp:446-499): line_search to completion,
// then lbfgs from the better of the two points.  Every piece is the
// reference's own -- linear_reg_runner's bounds and stage list, line_search,
// lbfgs, transformation_matrix, downsample_with_padding, has_mask,
// estimate_affine_param -- only the schedule is fixed, so the run is
// deterministic; costs come from the callback.
struct sequential_linear {
    using runner_t = tipl::reg::linear_reg_runner<3, uint8_t, void>;
    tref_cost_cb cb;
    void* user;
    int cost;
    unsigned count = 0;

    double fun(runner_t& r, const float* x)
    {
        tipl::affine_param<float, 3> p;
        std::copy_n(x, 12, p.begin());
        ++count;
        double c = 0.0f;
        for (size_t i = 0; i < r.from.size(); ++i) {
            tipl::transformation_matrix<float, 3> trans(p, r.from[i].shape(), r.from_vs,
                                                       r.to[i].shape(), r.to_vs);
            const auto& fs = r.from[i].shape();
            const auto& ts = r.to[i].shape();
            const int fd[3] = {int(fs[0]), int(fs[1]), int(fs[2])};
            const int td[3] = {int(ts[0]), int(ts[1]), int(ts[2])};
            double ci = 0.0;
            cb(r.from[i].data(), fd, r.to[i].data(), td, &trans[0], 1, cost, 256, &ci, user);
            c += ci;
        }
        c /= r.from.size() * 2.0f;
        return c;
    }

    float run_optimize(runner_t& r, unsigned cur_search_count)
    {
        auto f = [&](const auto& x) { return fun(r, &*x.begin()); };
        auto never = [] { return false; };
        double optimal_value = f(r.arg_min);
        if (cur_search_count) {
            for (size_t iter = 0; iter < r.reg_list.size(); ++iter) {
                auto cur_bound = r.get_current_bound(r.reg_list[iter]);
                auto arg_min2 = r.arg_min;
                double optimal_value2 = optimal_value;
                tipl::optimization::line_search(arg_min2.begin(), arg_min2.end(),
                                                cur_bound.first.begin(), cur_bound.second.begin(),
                                                {0, 0, 0, 1, 1, 1, 1, 1, 1, 1, 1, 1}, f,
                                                optimal_value2, int(cur_search_count), never);
                if (optimal_value2 < optimal_value) {
                    optimal_value = optimal_value2;
                    r.arg_min = arg_min2;
                }
                tipl::optimization::lbfgs(r.arg_min.begin(), r.arg_min.end(), cur_bound.first.begin(),
                                          cur_bound.second.begin(), f, optimal_value, never,
                                          r.precision, int(r.max_iterations));
            }
        } else {
            auto cur_bound = r.get_current_bound(r.reg_list.back());
            tipl::optimization::lbfgs(r.arg_min.begin(), r.arg_min.end(), cur_bound.first.begin(),
                                      cur_bound.second.begin(), f, optimal_value, never,
                                      r.precision * 0.2f, int(r.max_iterations));
        }
        return optimal_value;
    }

    // run_optimize_mr (reg/linear.hpp:527-550); the private down_sampling is
    // the same downsample_with_padding calls on the runner's images.
    float run_optimize_mr(runner_t& r)
    {
        const size_t sz = r.from[0].size();
        const bool is_large = sz > std::pow(96, 3);
        if (is_large) {
            runner_t low(r);
            std::vector<tipl::image<3, uint8_t>> buf(low.from.size() + low.to.size());
            for (size_t i = 0; i < buf.size(); ++i)
                tipl::downsample_with_padding(i < low.from.size() ? low.from[i]
                                                                  : low.to[i - low.from.size()],
                                              buf[i]);
            for (size_t i = 0; i < low.from.size(); ++i)
                low.from[i] = tipl::const_pointer_image<3, uint8_t>(&buf[i][0], buf[i].shape());
            for (size_t i = 0; i < low.to.size(); ++i) {
                auto& b = buf[low.from.size() + i];
                low.to[i] = tipl::const_pointer_image<3, uint8_t>(&b[0], b.shape());
            }
            low.from_vs *= 2.0f;
            low.to_vs *= 2.0f;
            const float v = run_optimize_mr(low);
            if (sz > std::pow(256, 3)) return v;
        }
        return run_optimize(r, is_large ? 0 : r.search_count);
    }
};

}  // namespace

extern "C" {

void tref_set_threads(int n) { tipl::max_thread_count = std::max(1, n); }
int tref_get_threads(void) { return tipl::max_thread_count; }

void tref_affine_matrix(const float* p, const int* fdim, const float* fvs, const int* tdim,
                        const float* tvs, float* T)
{
    tipl::affine_param<float, 3> arg(p);
    tipl::transformation_matrix<float, 3> M(arg, shp(fdim), tipl::vector<3>(fvs), shp(tdim),
                                            tipl::vector<3>(tvs));
    std::copy_n(M.begin(), 12, T);
}

int tref_matrix_inverse(float* T)
{
    auto M = mat(T);
    bool ok = M.inverse();
    std::copy_n(M.begin(), 12, T);
    return ok ? 1 : 0;
}

void tref_resample_u8(const uint8_t* from, const int* fdim, uint8_t* to, const int* tdim,
                      const float* T)
{
    tipl::const_pointer_image<3, uint8_t> F(from, shp(fdim));
    tipl::image<3, uint8_t> out(shp(tdim));
    tipl::resample(F, out, mat(T));
    std::memcpy(to, out.data(), out.size());
}

uint8_t tref_sample_u8(const uint8_t* I, const int* dim, float x, float y, float z)
{
    tipl::const_pointer_image<3, uint8_t> F(I, shp(dim));
    return F[tipl::vector<3>(x, y, z)];
}

double tref_corr_cost(const uint8_t* from, const int* fdim, const uint8_t* to, const int* tdim,
                      const float* T)
{
    tipl::const_pointer_image<3, uint8_t> F(from, shp(fdim)), G(to, shp(tdim));
    return tipl::reg::correlation()(F, G, mat(T));
}

void tref_mi_hist(const uint8_t* from, const int* fdim, const uint8_t* to, const int* tdim,
                  const float* T, int32_t* hist)
{
    tipl::const_pointer_image<3, uint8_t> F(from, shp(fdim)), G(to, shp(tdim));
    auto h = tipl::reg::get_mutual_info(G, F, mat(T));
    std::memcpy(hist, h.data(), sizeof(int32_t) * h.size());
}

double tref_mi_cost(const uint8_t* from, const int* fdim, const uint8_t* to, const int* tdim,
                    const float* T)
{
    tipl::const_pointer_image<3, uint8_t> F(from, shp(fdim)), G(to, shp(tdim));
    tipl::reg::mutual_information<3> mi;
    return mi(F, G, mat(T));
}

// K cost evaluations the way the reference optimisers issue them: one
// shared cost functor, tipl::par_for over the candidates
// (numerical/optimization.hpp:21-28, 82-87, 356-404), so each evaluation runs
// on one thread (nested par_for is serial, mt.hpp:126-129).
void tref_cost_batch(int kind, const uint8_t* from, const int* fdim, const uint8_t* to,
                     const int* tdim, const float* T, int K, double* costs)
{
    tipl::const_pointer_image<3, uint8_t> F(from, shp(fdim)), G(to, shp(tdim));
    if (kind == 1) {
        tipl::reg::mutual_information<3> mi;
        tipl::par_for(K, [&](int k) { costs[k] = mi(F, G, mat(T + 12 * k)); });
    } else {
        tipl::reg::correlation corr;
        tipl::par_for(K, [&](int k) { costs[k] = corr(F, G, mat(T + 12 * k)); });
    }
}

int tref_normalize_f32(const float* in, size_t n, const int* dim, uint8_t* out, float upper)
{
    tipl::const_pointer_image<3, float> I(in, shp(dim));
    tipl::image<3, uint8_t> O;
    tipl::normalize_upper_lower2(I, O, upper);
    if (O.empty()) return 0;
    std::memcpy(out, O.data(), n);
    return 1;
}

void tref_downsample_u8(const uint8_t* in, const int* dim, uint8_t* out, int* odim)
{
    tipl::const_pointer_image<3, uint8_t> I(in, shp(dim));
    tipl::image<3, uint8_t> O;
    tipl::downsample_with_padding(I, O);
    odim[0] = O.width();
    odim[1] = O.height();
    odim[2] = O.depth();
    std::memcpy(out, O.data(), O.size());
}

void tref_upsample_f3(const float* in, const int* idim, float* out, const int* odim)
{
    field_t F = to_field(in, idim);
    tipl::upsample_with_padding(F, shp(odim));
    from_field(F, out);
}

void tref_compose_displacement(const uint8_t* Is, const int* dim, const float* d, uint8_t* Js)
{
    tipl::const_pointer_image<3, uint8_t> I(Is, shp(dim));
    field_t D = to_field(d, dim);
    auto J = tipl::compose_displacement(I, D);
    std::memcpy(Js, J.data(), J.size());
}

float tref_cdm_gradient(const uint8_t* Js, const uint8_t* It, const int* dim, char type,
                        float* new_d)
{
    tipl::image<3, uint8_t> J(Js, shp(dim));
    tipl::const_pointer_image<3, uint8_t> T(It, shp(dim));
    field_t D;
    float c = tipl::reg::cdm_get_gradient(J, T, D, (unsigned char)type);
    from_field(D, new_d);
    return c;
}

void tref_cdm_poisson(float* new_d, const int* dim)
{
    field_t D = to_field(new_d, dim);
    bool terminated = false;
    tipl::reg::cdm_solve_poisson(D, terminated);
    from_field(D, new_d);
}

float tref_cdm_max_length(const float* d, const int* dim)
{
    field_t D = to_field(d, dim);
    return tipl::reg::cdm_max_displacement_length(D);
}

void tref_accumulate_displacement(const float* cur, float* nd, const int* dim)
{
    field_t C = to_field(cur, dim), N = to_field(nd, dim);
    tipl::accumulate_displacement(C, N);
    from_field(N, nd);
}

void tref_cdm_smooth(const float* d, float* dd, const int* dim, float smoothing)
{
    field_t D = to_field(d, dim), DD = to_field(dd, dim);
    tipl::reg::cdm_smooth(D, DD, smoothing);
    from_field(DD, dd);
}

void tref_gaussian_f32(float* img, const int* dim, int ch, unsigned bandwidth)
{
    if (ch == 3) {
        field_t F = to_field(img, dim);
        tipl::filter::gaussian(F, bandwidth);
        from_field(F, img);
    } else {
        tipl::image<3, float> I(shp(dim));
        std::memcpy(&I[0], img, sizeof(float) * I.size());
        tipl::filter::gaussian(I, bandwidth);
        std::memcpy(img, &I[0], sizeof(float) * I.size());
    }
}

int tref_cdm_improved(float* cost, int* n)
{
    std::vector<float> c(cost, cost + *n);
    bool r = tipl::reg::cdm_improved(c);
    std::copy(c.begin(), c.end(), cost);
    *n = int(c.size());
    return r ? 1 : 0;
}

int tref_cdm(const uint8_t* const* It, const uint8_t* const* Is, int nmod, const int* dim,
             float resolution, float speed, float smoothing, int min_dimension,
             char gradient_type, float* field)
{
    std::vector<tipl::const_pointer_image<3, uint8_t>> vt, vs;
    for (int i = 0; i < nmod; ++i) {
        vt.emplace_back(It[i], shp(dim));
        vs.emplace_back(Is[i], shp(dim));
    }
    tipl::reg::cdm_param p;
    p.resolution = resolution;
    p.speed = speed;
    p.smoothing = smoothing;
    p.min_dimension = (unsigned int)min_dimension;
    p.gradient_type = std::string(size_t(nmod), gradient_type);
    field_t D;
    bool terminated = false;
    bool ok = tipl::reg::cdm(vt, vs, D, terminated, p);
    if (ok) from_field(D, field);
    return ok ? 1 : 0;
}

// ---- the 2-D instantiations (cdm<out,uint8,2>) ----------------------------

uint8_t tref2_sample_u8(const uint8_t* I, const int* dim, float x, float y)
{
    tipl::const_pointer_image<2, uint8_t> P(I, shp2(dim));
    uint8_t v = 0;
    tipl::estimate<tipl::interpolation::linear>(P, tipl::vector<2>(x, y), v);
    return v;
}

void tref2_downsample_u8(const uint8_t* in, const int* dim, uint8_t* out, int* odim)
{
    tipl::const_pointer_image<2, uint8_t> I(in, shp2(dim));
    tipl::image<2, uint8_t> O;
    tipl::downsample_with_padding(I, O);
    odim[0] = O.width();
    odim[1] = O.height();
    std::memcpy(out, O.data(), O.size());
}

void tref2_upsample_f2(const float* in, const int* idim, float* out, const int* odim)
{
    field2_t F = to_field2(in, idim);
    tipl::upsample_with_padding(F, shp2(odim));
    from_field2(F, out);
}

void tref2_compose_displacement(const uint8_t* Is, const int* dim, const float* d, uint8_t* Js)
{
    tipl::const_pointer_image<2, uint8_t> I(Is, shp2(dim));
    field2_t D = to_field2(d, dim);
    auto J = tipl::compose_displacement(I, D);
    std::memcpy(Js, J.data(), J.size());
}

float tref2_cdm_gradient(const uint8_t* Js, const uint8_t* It, const int* dim, char type,
                         float* new_d)
{
    tipl::image<2, uint8_t> J(Js, shp2(dim));
    tipl::const_pointer_image<2, uint8_t> T(It, shp2(dim));
    field2_t D;
    float c = tipl::reg::cdm_get_gradient(J, T, D, (unsigned char)type);
    from_field2(D, new_d);
    return c;
}

void tref2_cdm_poisson(float* new_d, const int* dim)
{
    field2_t D = to_field2(new_d, dim);
    bool terminated = false;
    tipl::reg::cdm_solve_poisson(D, terminated);
    from_field2(D, new_d);
}

float tref2_cdm_max_length(const float* d, const int* dim)
{
    field2_t D = to_field2(d, dim);
    return tipl::reg::cdm_max_displacement_length(D);
}

void tref2_accumulate_displacement(const float* cur, float* nd, const int* dim)
{
    field2_t C = to_field2(cur, dim), N = to_field2(nd, dim);
    tipl::accumulate_displacement(C, N);
    from_field2(N, nd);
}

void tref2_cdm_smooth(const float* d, float* dd, const int* dim, float smoothing)
{
    field2_t D = to_field2(d, dim), DD = to_field2(dd, dim);
    tipl::reg::cdm_smooth(D, DD, smoothing);
    from_field2(DD, dd);
}

int tref2_cdm(const uint8_t* const* It, const uint8_t* const* Is, int nmod, const int* dim,
              float resolution, float speed, float smoothing, int min_dimension,
              char gradient_type, float* field)
{
    std::vector<tipl::const_pointer_image<2, uint8_t>> vt, vs;
    for (int i = 0; i < nmod; ++i) {
        vt.emplace_back(It[i], shp2(dim));
        vs.emplace_back(Is[i], shp2(dim));
    }
    tipl::reg::cdm_param p;
    p.resolution = resolution;
    p.speed = speed;
    p.smoothing = smoothing;
    p.min_dimension = (unsigned int)min_dimension;
    p.gradient_type = std::string(size_t(nmod), gradient_type);
    field2_t D;
    bool terminated = false;
    bool ok = tipl::reg::cdm(vt, vs, D, terminated, p);
    if (ok) from_field2(D, field);
    return ok ? 1 : 0;
}

// `iters` iterations of the body of the reference's cdm() loop
// (reg/cdm.hpp:418-470) at one resolution level, from a zero field, with
// the reference's own step functions and par_for threading; no early exit.
// Returns the wall time in seconds; the field is written to `field`.
double tref_cdm_iterations(const uint8_t* It, const uint8_t* Is, const int* dim, int iters,
                           float* field)
{
    tipl::const_pointer_image<3, uint8_t> T(It, shp(dim)), S(Is, shp(dim));
    field_t cur(shp(dim));
    float theta = 0.0f, cur_smoothing = 0.8f;
    const float speed = 0.3f;
    bool terminated = false;
    auto t0 = std::chrono::steady_clock::now();
    for (int i = 0; i < iters; ++i) {
        field_t new_d;
        (void)tipl::reg::cdm_get_gradient(tipl::compose_displacement(S, cur), T, new_d, 'r');
        tipl::reg::cdm_solve_poisson(new_d, terminated);
        if (theta == 0.0f) theta = tipl::reg::cdm_max_displacement_length(new_d);
        if (theta == 0.0f) break;
        new_d *= speed / theta;
        tipl::accumulate_displacement(cur, new_d);
        tipl::reg::cdm_smooth(new_d, cur, cur_smoothing);
    }
    auto t1 = std::chrono::steady_clock::now();
    if (field) from_field(cur, field);
    return std::chrono::duration<double>(t1 - t0).count();
}

// The fixed-schedule CDM extension composed from the reference's own
// routines: its pyramid (downsample_with_padding, best_d *= 2,
// upsample_with_padding, reg/cdm.hpp:385-405), its loop body as above for
// exactly `iters` iterations per level, and filter::gaussian(new_d, bw) in
// place of cdm_smooth when bw > 0.  Checks the oracle's restatement.
static void ref_cdm_fixed(const tipl::const_pointer_image<3, uint8_t>& T,
                          const tipl::const_pointer_image<3, uint8_t>& S, field_t& cur, int levels,
                          int iters, float smoothing, unsigned bw)
{
    cur.resize(T.shape());
    std::fill(cur.begin(), cur.end(), tipl::vector<3>(0.0f, 0.0f, 0.0f));
    if (levels > 1) {
        tipl::image<3, uint8_t> rT, rS;
        tipl::downsample_with_padding(T, rT);
        tipl::downsample_with_padding(S, rS);
        field_t rd;
        ref_cdm_fixed(tipl::const_pointer_image<3, uint8_t>(rT.data(), rT.shape()),
                      tipl::const_pointer_image<3, uint8_t>(rS.data(), rS.shape()), rd, levels - 1,
                      iters, smoothing, bw);
        rd *= 2.0f;
        tipl::upsample_with_padding(rd, T.shape());
        cur = rd;
    }
    float theta = 0.0f;
    bool terminated = false;
    for (int i = 0; i < iters; ++i) {
        field_t new_d;
        (void)tipl::reg::cdm_get_gradient(tipl::compose_displacement(S, cur), T, new_d, 'r');
        tipl::reg::cdm_solve_poisson(new_d, terminated);
        if (theta == 0.0f) theta = tipl::reg::cdm_max_displacement_length(new_d);
        if (theta == 0.0f) break;
        new_d *= 0.3f / theta;
        tipl::accumulate_displacement(cur, new_d);
        if (bw) {
            tipl::filter::gaussian(new_d, bw);
            cur = new_d;
        } else {
            tipl::reg::cdm_smooth(new_d, cur, smoothing);
        }
    }
}

void tref_cdm_fixed(const uint8_t* It, const uint8_t* Is, const int* dim, int levels, int iters,
                    float smoothing, unsigned bw, float* field)
{
    tipl::const_pointer_image<3, uint8_t> T(It, shp(dim)), S(Is, shp(dim));
    field_t cur;
    ref_cdm_fixed(T, S, cur, levels, iters, smoothing, bw);
    from_field(cur, field);
}

// reg::subject_image_pre (gaussian, normalize_otsu_median, uint8 conversion).
void tref_subject_pre(const float* in, const int* dim, uint8_t* out)
{
    tipl::image<3> I(shp(dim));
    std::memcpy(&I[0], in, sizeof(float) * I.size());
    auto J = tipl::reg::subject_image_pre(std::move(I));
    std::memcpy(out, &J[0], J.size());
}

float tref_otsu_median(const float* in, size_t n)
{
    tipl::image<3> I(tipl::shape<3>(unsigned(n), 1u, 1u));
    std::memcpy(&I[0], in, sizeof(float) * n);
    return tipl::segmentation::otsu_median(I);
}

// segmentation::otsu_threshold (numerical/otsu.hpp:66-80) of a float image.
float tref_otsu_threshold(const float* in, size_t n)
{
    tipl::image<3> I(tipl::shape<3>(unsigned(n), 1u, 1u));
    std::memcpy(&I[0], in, sizeof(float) * n);
    return tipl::segmentation::otsu_threshold(I);
}

// normalize_upper_lower2 (numerical/numerical.hpp:909-927) of an int16 image.
int tref_normalize_i16(const int16_t* in, size_t n, const int* dim, uint8_t* out, float upper)
{
    tipl::const_pointer_image<3, int16_t> I(in, shp(dim));
    tipl::image<3, uint8_t> O;
    tipl::normalize_upper_lower2(I, O, upper);
    if (O.empty()) return 0;
    std::memcpy(out, O.data(), n);
    return 1;
}

float tref_median(const float* in, size_t n)
{
    std::vector<float> v(in, in + n);
    return tipl::median(v);
}

// tipl::reg::linear on the CPU; arg (12 floats) is in/out.
float tref_linear(const uint8_t* from, const int* fdim, const float* fvs, const uint8_t* to,
                  const int* tdim, const float* tvs, int cost_type, int reg_type, float* arg)
{
    std::vector<tipl::const_pointer_image<3, uint8_t>> f{tipl::const_pointer_image<3, uint8_t>(from, shp(fdim))};
    std::vector<tipl::const_pointer_image<3, uint8_t>> t{tipl::const_pointer_image<3, uint8_t>(to, shp(tdim))};
    tipl::affine_param<float, 3> a(arg);
    tipl::reg::linear_reg_param param;
    param.cost_type = cost_type ? tipl::reg::mutual_info : tipl::reg::corr;
    param.reg_type = tipl::reg::reg_type(reg_type);
    param.cuda = false;
    bool terminated = false;
    float r = tipl::reg::linear<void>(f, tipl::vector<3>(fvs), t, tipl::vector<3>(tvs), a, param,
                                      terminated);
    std::copy_n(a.begin(), 12, arg);
    return r;
}

// reg::linear (reg/linear.hpp:622-700) with the sequential schedule above;
// arg (12 floats) is in/out, *count receives the number of cost evaluations.
float tref_linear_sequential(const uint8_t* from, const int* fdim, const float* fvs,
                             const uint8_t* to, const int* tdim, const float* tvs, int cost_type,
                             int reg_type, float* arg, tref_cost_cb cb, void* user,
                             unsigned* count)
{
    std::vector<tipl::const_pointer_image<3, uint8_t>> f{tipl::const_pointer_image<3, uint8_t>(from, shp(fdim))};
    std::vector<tipl::const_pointer_image<3, uint8_t>> t{tipl::const_pointer_image<3, uint8_t>(to, shp(tdim))};
    const tipl::vector<3> from_vs(fvs), to_vs(tvs);
    tipl::affine_param<float, 3> a(arg);
    tipl::reg::linear_reg_param param;
    param.cost_type = cost_type ? tipl::reg::mutual_info : tipl::reg::corr;
    param.reg_type = tipl::reg::reg_type(reg_type);
    sequential_linear seq{cb, user, cost_type};
    tipl::vector<3> new_to_vs = to_vs;
    tipl::affine_param<float, 3> surrogate = a;
    bool adjust_vs = false;
    if (param.reg_type == tipl::reg::affine && tipl::has_mask(f[0]) && tipl::has_mask(t[0])) {
        tipl::estimate_affine_param(f[0], from_vs, t[0], to_vs, surrogate);
        a = surrogate;
        for (int i = 0; i < 3; ++i) {
            new_to_vs[i] = to_vs[i] / std::max<float>(surrogate.scaling[i], 0.01f);
            surrogate.scaling[i] = 1.0f;
        }
        adjust_vs = true;
    }
    param.absolute_bound = true;
    sequential_linear::runner_t reg(f, t, adjust_vs ? surrogate : a);
    reg.from_vs = from_vs;
    reg.to_vs = new_to_vs;
    reg.set(param);
    // one thread: the optimisers' par_for loops evaluate in index order
    const int saved = tipl::max_thread_count;
    tipl::max_thread_count = 1;
    const float result = seq.run_optimize_mr(reg);
    tipl::max_thread_count = saved;
    if (adjust_vs)
        a = tipl::transformation_matrix<float, 3>(surrogate, f[0], from_vs, t[0], new_to_vs)
                .to_affine_param(f[0], from_vs, t[0], to_vs);
    std::copy_n(a.begin(), 12, arg);
    if (count) *count = seq.count;
    return result;
}

// has_mask (numerical/basic_op.hpp:811-830).
int tref_has_mask(const uint8_t* I, const int* dim)
{
    return tipl::has_mask(tipl::const_pointer_image<3, uint8_t>(I, shp(dim))) ? 1 : 0;
}

// estimate_affine_param (numerical/transformation.hpp:933-1021).
void tref_estimate_affine_param(const uint8_t* src, const int* sdim, const float* svs,
                                const uint8_t* tgt, const int* tdim, const float* tvs, float* arg)
{
    tipl::const_pointer_image<3, uint8_t> S(src, shp(sdim)), T(tgt, shp(tdim));
    tipl::affine_param<float, 3> a;
    tipl::estimate_affine_param(S, tipl::vector<3>(svs), T, tipl::vector<3>(tvs), a);
    std::copy_n(a.begin(), 12, arg);
}

// optimization::line_search / lbfgs (numerical/optimization.hpp:137-512) on
// n parameters with one thread (their par_for loops then run in index
// order); x is in/out, the optimal value is returned.
double tref_line_search(float* x, int n, const float* upper, const float* lower,
                        const int* strategy, int grid, double optimal, tref_param_cb cb, void* user)
{
    const int saved = tipl::max_thread_count;
    tipl::max_thread_count = 1;
    std::vector<float> X(x, x + n), U(upper, upper + n), L(lower, lower + n);
    std::vector<int> S(strategy, strategy + n);
    auto fun = [&](const std::vector<float>& p) {
        double c = 0.0;
        cb(p.data(), n, 1, &c, user);
        return c;
    };
    tipl::optimization::line_search(X.begin(), X.end(), U.begin(), L.begin(), S, fun, optimal, grid,
                                    [] { return false; });
    tipl::max_thread_count = saved;
    std::copy(X.begin(), X.end(), x);
    return optimal;
}

double tref_lbfgs(float* x, int n, const float* upper, const float* lower, double fun_x,
                  double precision, int max_iteration, tref_param_cb cb, void* user)
{
    const int saved = tipl::max_thread_count;
    tipl::max_thread_count = 1;
    std::vector<float> X(x, x + n), U(upper, upper + n), L(lower, lower + n);
    auto fun = [&](const std::vector<float>& p) {
        double c = 0.0;
        cb(p.data(), n, 1, &c, user);
        return c;
    };
    auto never = [] { return false; };
    tipl::optimization::lbfgs(X.begin(), X.end(), U.begin(), L.begin(), fun, fun_x, never, precision,
                              max_iteration);
    tipl::max_thread_count = saved;
    std::copy(X.begin(), X.end(), x);
    return fun_x;
}

// Post-registration warping (numerical/dif.hpp) and mm_reg's r values.
void tref_disp_to_mapping(const float* dis, const int* dim, const float* T, float* out)
{
    field_t D = to_field(dis, dim), M;
    if (T)
        tipl::displacement_to_mapping(D, M, mat(T));
    else
        tipl::displacement_to_mapping(D, M);
    from_field(M, out);
}

void tref_mapping_to_disp(const float* mapping, const int* dim, float* out)
{
    field_t M = to_field(mapping, dim);
    tipl::mapping_to_displacement(M);
    from_field(M, out);
}

void tref_inv_disp_to_mapping(const float* inv_dis, const int* ddim, const float* T,
                              const int* odim, float* out)
{
    field_t D = to_field(inv_dis, ddim), M(shp(odim));
    tipl::inv_displacement_to_mapping(D, M, mat(T));
    from_field(M, out);
}

void tref_compose_mapping_u8(const uint8_t* from, const int* fdim, const float* mapping,
                             const int* dim, uint8_t* out)
{
    tipl::const_pointer_image<3, uint8_t> I(from, shp(fdim));
    field_t M = to_field(mapping, dim);
    tipl::image<3, uint8_t> J;
    tipl::compose_mapping(I, M, J);
    std::memcpy(out, J.data(), J.size());
}

void tref_compose_mapping_f32(const float* from, const int* fdim, const float* mapping,
                              const int* dim, float* out)
{
    tipl::const_pointer_image<3, float> I(from, shp(fdim));
    field_t M = to_field(mapping, dim);
    tipl::image<3, float> J;
    tipl::compose_mapping(I, M, J);
    std::memcpy(out, J.data(), sizeof(float) * J.size());
}

void tref_jacobian_det_dis(const float* dis, const int* dim, float* out)
{
    field_t D = to_field(dis, dim);
    tipl::image<3, float> J;
    tipl::jacobian_determinant_dis(D, J);
    std::memcpy(out, J.data(), sizeof(float) * J.size());
}

void tref_jacobian_det(const float* mapping, const int* dim, float* out)
{
    field_t M = to_field(mapping, dim);
    auto J = tipl::jacobian_determinant(M);
    std::memcpy(out, J.data(), sizeof(float) * J.size());
}

double tref_correlation(const uint8_t* x, const uint8_t* y, size_t n, int masked)
{
    return masked ? tipl::correlation_ygz(x, x + n, y) : tipl::correlation(x, x + n, y);
}

// mm_reg::It_match_fov on a one-modality mm_reg (reg/mm_reg.hpp:313-328);
// arg is the affine parameter vector, as mm_reg holds it.
void tref_match_fov(uint8_t* It, const int* tdim, const float* tvs, const int* sdim,
                    const float* svs, const float* arg)
{
    tipl::reg::mm_reg<null_out> m;
    m.It[0] = tipl::image<3, uint8_t>(It, shp(tdim));
    m.Its = shp(tdim);
    m.Is = shp(sdim);
    m.Itvs = tipl::vector<3>(tvs);
    m.Ivs = tipl::vector<3>(svs);
    m.arg = tipl::affine_param<float, 3>(arg);
    m.It_match_fov();
    std::memcpy(It, m.It[0].data(), m.It[0].size());
}

// The nonlinear half of mm_reg with a given affine (skip_linear): linear_reg
// -> calculate_linear_r (J, r, It_match_fov), then nonlinear_reg (reverse CDM
// on its own thread, forward CDM, mappings, calculate_nonlinear_r), all on
// the CPU (reg/mm_reg.hpp:271-311, 330-416).  r[0] / r[1] are the linear and
// nonlinear r; the images and fields are copied out after each stage.
void tref_mm_reg(const uint8_t* It, const int* tdim, const float* tvs, const uint8_t* I,
                 const int* sdim, const float* svs, const float* arg, int reverse, int match_fov,
                 int masked_r, float* r, uint8_t* J_lin, uint8_t* It_fov, float* t2f, float* f2t,
                 float* to2from, float* from2to, uint8_t* JJ)
{
    tipl::reg::mm_reg<null_out> m;
    m.It[0] = tipl::image<3, uint8_t>(It, shp(tdim));
    m.I[0] = tipl::image<3, uint8_t>(I, shp(sdim));
    m.Its = shp(tdim);
    m.Is = shp(sdim);
    m.Itvs = tipl::vector<3>(tvs);
    m.Ivs = tipl::vector<3>(svs);
    m.arg = tipl::affine_param<float, 3>(arg);
    m.skip_linear = true;
    m.skip_reverse = !reverse;
    m.match_fov = match_fov != 0;
    m.masked_r = masked_r != 0;
    m.use_cuda = false;
    bool terminated = false;
    m.linear_reg(terminated);
    r[0] = m.r[0];
    std::memcpy(J_lin, m.J[0].data(), m.J[0].size());
    std::memcpy(It_fov, m.It[0].data(), m.It[0].size());
    m.nonlinear_reg(terminated);
    r[1] = m.r[0];
    from_field(m.t2f_dis, t2f);
    if (reverse) {
        from_field(m.f2t_dis, f2t);
        from_field(m.from2to, from2to);
    }
    from_field(m.to2from, to2from);
    std::memcpy(JJ, m.J[0].data(), m.J[0].size());
}

}  // extern "C"

// tipl::io::gz_nifti reading (`>>`: LPS reorientation, type conversion,
// scl) of volume `volume` into a float or uint8 image, with the header
// values the registration uses.  Returns 0 on failure.
template <class T>
static int nifti_read_as(const char* path, int volume, T* out, size_t cap, int* dim, float* vs,
                         float* srow, int* is_mni)
{
    tipl::io::gz_nifti in;
    if (!in.open(path, std::ios::in)) return 0;
    tipl::image<3, T> I;
    for (int v = 0; v <= volume; ++v)
        if (!(in >> I)) return 0;
    if (I.size() > cap) return 0;
    std::copy(I.begin(), I.end(), out);
    tipl::shape<3> s;
    tipl::vector<3> pv;
    tipl::matrix<4, 4> R;
    in >> s >> pv >> R;
    for (int i = 0; i < 3; ++i) dim[i] = int(s[i]);
    for (int i = 0; i < 3; ++i) vs[i] = pv[i];
    std::copy_n(R.begin(), 12, srow);
    *is_mni = in.is_mni() ? 1 : 0;
    return 1;
}

extern "C" {

int tref_nifti_read_f32(const char* path, int volume, float* out, size_t cap, int* dim, float* vs,
                        float* srow, int* is_mni)
{
    return nifti_read_as(path, volume, out, cap, dim, vs, srow, is_mni);
}

int tref_nifti_read_u8(const char* path, int volume, uint8_t* out, size_t cap, int* dim, float* vs,
                       float* srow, int* is_mni)
{
    return nifti_read_as(path, volume, out, cap, dim, vs, srow, is_mni);
}

// `<< vs << R << image` with gz_nifti (gzip when the path ends in .gz).
int tref_nifti_write_f32(const char* path, const float* data, const int* dim, const float* vs,
                         const float* srow)
{
    tipl::io::gz_nifti out;
    if (!out.open(path, std::ios::out)) return 0;
    tipl::image<3> I(tipl::shape<3>(dim[0], dim[1], dim[2]));
    std::copy_n(data, I.size(), I.begin());
    tipl::matrix<4, 4> R;
    R.identity();
    std::copy_n(srow, 12, R.begin());
    out << tipl::vector<3>(vs) << R << I;
    return bool(out) ? 1 : 0;
}

}  // extern "C"



