// tipl_hip_plug.cpp -- the reference's GPU plugs, defined over libtiplhip's C ABI.
//
// The reference selects its GPU registration path through two templates that
// a CUDA translation unit instantiates:
//   linear_reg_imp<out, mr, unsigned char, 3>(std::true_type, ...)
//       declared reg/linear.hpp:606-609, instantiated reg/linear_cu.hpp:7-16;
//   cdm_cuda<out, unsigned char, dim>
//       declared reg/cdm.hpp:525-531, instantiated reg/cdm_cu.hpp:6-17.
// This file replaces that CUDA translation unit.  It is plain C++17, compiled
// by g++ against the reference headers with TIPL_USE_CUDA defined (so
// tipl::use_cuda and tipl::has_gpu are true and the reference's host code
// takes its GPU branches) and linked with libtiplhip.so; no nvcc.  It defines
// the same symbols as explicit specialisations:
//   * linear_reg_imp for mr = true and mr = false: the reference's own
//     linear_reg_runner::optimize with a cost functor that evaluates each
//     candidate transform through tiplhip_cost_batch (one candidate per call,
//     as the reference's optimisers call their functors);
//   * cdm_cuda<out, unsigned char, 3>: the whole cdm() on the device through
//     tiplhip_cdm;
//   * cdm_cuda<out, unsigned char, 2>: the whole 2-D cdm() on the device
//     through tiplhip_cdm2d.
//
// Build (as __graft_entry__.build() does, via oracle/Makefile):
//   g++ -std=c++17 -DTIPL_USE_CUDA -I<reference root> -I<repo>/include -fPIC \
//       -c tipl_hip_plug.cpp   ...   -L<repo>/tipl_amd -ltiplhip
#include <algorithm>  // before the reference headers (mt.hpp uses std::clamp)
#include <cstdint>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "prog.hpp"
#include "reg/cdm.hpp"
#include "reg/linear.hpp"

#include "tiplhip.h"

// Optional observer of every cost the plug returns (logging, tests):
// the 12-float transform (sr[9], shift[3]), the cost kind and the cost.
extern "C" {
void (*tipl_hip_plug_cost_observer)(const float* T, int cost_kind, double cost) = nullptr;
}

namespace {

void hip_check(int rc)
{
    if (rc != TIPLHIP_OK) throw std::runtime_error(std::string("tiplhip: ") + tiplhip_last_error());
}

// The cost functor handed to the reference's linear_reg_runner::optimize in
// place of mutual_information_cuda / correlation_cuda (reg/linear.hpp:41-78,
// 208-252).  One functor per modality and run_optimize call, default
// constructed; the reference calls it from several threads (line search and
// L-BFGS race, reg/linear.hpp:459-499), each call one transform.  The device
// volumes are uploaded on the first call (uint8 as given, the reference's
// rule for uint8 inputs) and live as long as the functor.
template <int kind>
struct hip_cost {
    tiplhip_ctx* ctx = nullptr;
    tiplhip_volume* from = nullptr;
    tiplhip_volume* to = nullptr;
    std::mutex init;

    hip_cost() = default;
    hip_cost(const hip_cost&) = delete;
    hip_cost& operator=(const hip_cost&) = delete;
    ~hip_cost()
    {
        if (from) tiplhip_volume_destroy(from);
        if (to) tiplhip_volume_destroy(to);
        if (ctx) tiplhip_ctx_destroy(ctx);
    }

    template <typename T, typename U, typename V>
    double operator()(const T& from_, const U& to_, const V& transform)
    {
        static_assert(T::dimension == 3 && U::dimension == 3, "libtiplhip registers 3-D images");
        {
            std::scoped_lock<std::mutex> lk(init);
            if (!ctx) {
                hip_check(tiplhip_ctx_create(0, &ctx));
                const int fd[3] = {int(from_.width()), int(from_.height()), int(from_.depth())};
                const int td[3] = {int(to_.width()), int(to_.height()), int(to_.depth())};
                hip_check(tiplhip_volume_create(ctx, &from_[0], TIPLHIP_U8, fd, 255.99f, &from));
                hip_check(tiplhip_volume_create(ctx, &to_[0], TIPLHIP_U8, td, 255.99f, &to));
            }
        }
        // the context serialises concurrent calls; correlation's role swap
        // for a larger `from` happens inside the library (reg/linear.hpp:28-33)
        double c = 0.0;
        hip_check(tiplhip_cost_batch(ctx, from, to, transform.data(), 1, kind, 256, &c));
        if (tipl_hip_plug_cost_observer) tipl_hip_plug_cost_observer(transform.data(), kind, c);
        return c;
    }
};

template <bool mr>
float linear_plug(std::shared_ptr<tipl::reg::linear_reg_runner<3, unsigned char, tipl::out>> reg,
                  tipl::reg::cost_type cost, bool& terminated)
{
    return cost == tipl::reg::mutual_info
               ? reg->template optimize<hip_cost<TIPLHIP_COST_MI>, mr>(terminated)
               : reg->template optimize<hip_cost<TIPLHIP_COST_CORR>, mr>(terminated);
}

}  // namespace

namespace tipl::reg {

template <>
float linear_reg_imp<tipl::out, true, unsigned char, 3>(
    std::true_type, std::shared_ptr<tipl::reg::linear_reg_runner<3, unsigned char, tipl::out>> reg,
    tipl::reg::cost_type cost, bool& terminated)
{
    return linear_plug<true>(reg, cost, terminated);
}

template <>
float linear_reg_imp<tipl::out, false, unsigned char, 3>(
    std::true_type, std::shared_ptr<tipl::reg::linear_reg_runner<3, unsigned char, tipl::out>> reg,
    tipl::reg::cost_type cost, bool& terminated)
{
    return linear_plug<false>(reg, cost, terminated);
}

// cdm() on the device (reg/cdm.hpp:373-474 through tiplhip_cdm).  As the
// reference's cdm_cuda: false when cdm() returns false (zero cost) or on a
// device error; a run the caller terminates keeps its best field.
template <>
bool cdm_cuda<tipl::out, unsigned char, 3>(
    const std::vector<tipl::const_pointer_image<3, unsigned char>>& It,
    const std::vector<tipl::const_pointer_image<3, unsigned char>>& Is,
    tipl::image<3, tipl::vector<3>>& d, bool& terminated, const cdm_param& param)
{
    tiplhip_ctx* ctx = nullptr;
    std::vector<tiplhip_volume*> vt, vs;
    auto release = [&] {
        for (auto* v : vt) tiplhip_volume_destroy(v);
        for (auto* v : vs) tiplhip_volume_destroy(v);
        if (ctx) tiplhip_ctx_destroy(ctx);
        vt.clear();
        vs.clear();
        ctx = nullptr;
    };
    try {
        const size_t n = std::min(It.size(), Is.size());
        if (n == 0 || n > 8) throw std::runtime_error("cdm_cuda: 1 to 8 modalities");
        hip_check(tiplhip_ctx_create(0, &ctx));
        const int dim[3] = {int(It[0].width()), int(It[0].height()), int(It[0].depth())};
        for (size_t i = 0; i < n; ++i) {
            vt.push_back(nullptr);
            hip_check(tiplhip_volume_create(ctx, &It[i][0], TIPLHIP_U8, dim, 255.99f, &vt.back()));
            vs.push_back(nullptr);
            hip_check(tiplhip_volume_create(ctx, &Is[i][0], TIPLHIP_U8, dim, 255.99f, &vs.back()));
        }
        tiplhip_cdm_param p;
        tiplhip_cdm_param_default(&p);
        p.resolution = param.resolution;
        p.speed = param.speed;
        p.smoothing = param.smoothing;
        p.min_dimension = int(param.min_dimension);
        for (size_t i = 0; i < 8; ++i)
            p.gradient_type[i] = i < param.gradient_type.size() ? param.gradient_type[i] : 'r';
        d.resize(It[0].shape());
        int iterations = 0;
        const int rc = tiplhip_cdm(ctx, vt.data(), vs.data(), int(n), &p, &d[0][0],
                                   reinterpret_cast<const volatile uint8_t*>(&terminated), &iterations);
        release();
        if (rc == TIPLHIP_ERR_FAILED) return false;
        if (rc != TIPLHIP_OK && rc != TIPLHIP_ERR_TERMINATED)
            throw std::runtime_error(std::string("tiplhip: ") + tiplhip_last_error());
        return true;
    } catch (std::runtime_error& er) {
        release();
        tipl::out() << "❌️" << er.what();
        return false;
    }
}

// The 2-D cdm() on the device (reg/cdm.hpp:373-474 at dimension 2, through
// tiplhip_cdm2d), with the 3-D plug's return rules.
template <>
bool cdm_cuda<tipl::out, unsigned char, 2>(
    const std::vector<tipl::const_pointer_image<2, unsigned char>>& It,
    const std::vector<tipl::const_pointer_image<2, unsigned char>>& Is,
    tipl::image<2, tipl::vector<2>>& d, bool& terminated, const cdm_param& param)
{
    tiplhip_ctx* ctx = nullptr;
    try {
        const size_t n = std::min(It.size(), Is.size());
        if (n == 0 || n > 8) throw std::runtime_error("cdm_cuda: 1 to 8 modalities");
        hip_check(tiplhip_ctx_create(0, &ctx));
        const int dim[2] = {int(It[0].width()), int(It[0].height())};
        std::vector<const uint8_t*> pt, ps;
        for (size_t i = 0; i < n; ++i) {
            pt.push_back(&It[i][0]);
            ps.push_back(&Is[i][0]);
        }
        tiplhip_cdm_param p;
        tiplhip_cdm_param_default(&p);
        p.resolution = param.resolution;
        p.speed = param.speed;
        p.smoothing = param.smoothing;
        p.min_dimension = int(param.min_dimension);
        for (size_t i = 0; i < 8; ++i)
            p.gradient_type[i] = i < param.gradient_type.size() ? param.gradient_type[i] : 'r';
        d.resize(It[0].shape());
        int iterations = 0;
        const int rc = tiplhip_cdm2d(ctx, pt.data(), ps.data(), int(n), dim, &p, &d[0][0],
                                     reinterpret_cast<const volatile uint8_t*>(&terminated),
                                     &iterations);
        tiplhip_ctx_destroy(ctx);
        ctx = nullptr;
        if (rc == TIPLHIP_ERR_FAILED) return false;
        if (rc != TIPLHIP_OK && rc != TIPLHIP_ERR_TERMINATED)
            throw std::runtime_error(std::string("tiplhip: ") + tiplhip_last_error());
        return true;
    } catch (std::runtime_error& er) {
        if (ctx) tiplhip_ctx_destroy(ctx);
        tipl::out() << "❌️" << er.what();
        return false;
    }
}

}  // namespace tipl::reg
