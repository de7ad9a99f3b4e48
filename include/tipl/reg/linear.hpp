// reg/linear.hpp -- tipl::reg::linear on the MI355X path.
//
// The same entry points, parameters and optimisation schedule as the
// reference's reg/linear.hpp:255-714: linear_reg_param, the bound tables,
// the reg_type stage list (translation -> +scaling -> +rotation -> affine),
// the resolution pyramid above 96^3, line search + L-BFGS per stage.
// Differences, by design:
//   * every optimiser candidate set is evaluated in one batched launch
//     (tiplhip_cost_batch_params) instead of one thread per candidate;
//   * the reference races line_search (on a thread) against repeated L-BFGS
//     and lets them exchange their best point; here line search runs first,
//     then L-BFGS from the better of the two -- deterministic, with the same
//     candidate sets and update rules;
//   * there is no CPU fallback: without a GPU the call throws.
#ifndef TIPL_HIP_REG_LINEAR_HPP
#define TIPL_HIP_REG_LINEAR_HPP

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <functional>
#include <memory>
#include <stdexcept>
#include <vector>

#include "../def.hpp"
#include "../hip.hpp"
#include "../image.hpp"
#include "../optimization.hpp"
#include "../transformation.hpp"

namespace tipl {
namespace reg {

enum reg_type { none = 0, translocation = 1, rotation = 2, rigid_body = 3, scaling = 4,
                translocation_scaling = 5, rigid_scaling = 7, tilt = 8, affine = 15 };
enum cost_type { corr = 0, mutual_info = 1 };

// Bound tables (reg/linear.hpp:258-266): per axis {t+, t-, r+, r-, s+, s-, a+, a-}.
constexpr float narrow_bound[3][8] = {{0.2f, -0.2f, 0.2f, -0.2f, 1.2f, 0.8f, 0.05f, -0.05f},
                                      {0.2f, -0.2f, 0.1f, -0.1f, 1.2f, 0.8f, 0.05f, -0.05f},
                                      {0.2f, -0.2f, 0.1f, -0.1f, 1.2f, 0.8f, 0.05f, -0.05f}};
constexpr float reg_bound[3][8] = {{0.75f, -0.75f, 0.4f, -0.4f, 1.5f, 0.7f, 0.1f, -0.1f},
                                   {0.75f, -0.75f, 0.25f, -0.25f, 1.5f, 0.7f, 0.1f, -0.1f},
                                   {0.75f, -0.75f, 0.25f, -0.25f, 1.5f, 0.7f, 0.1f, -0.1f}};
constexpr float large_bound[3][8] = {{1.0f, -1.0f, 0.5f, -0.5f, 2.0f, 0.5f, 0.25f, -0.25f},
                                     {1.0f, -1.0f, 0.4f, -0.4f, 2.0f, 0.5f, 0.25f, -0.25f},
                                     {1.0f, -1.0f, 0.4f, -0.4f, 2.0f, 0.5f, 0.25f, -0.25f}};

struct linear_reg_param {
    tipl::reg::reg_type reg_type = affine;
    tipl::reg::cost_type cost_type = mutual_info;
    const float (*bound)[8] = reg_bound;
    unsigned int search_count = 16;
    bool cuda = tipl::has_gpu();
    bool absolute_bound = true;
    int mi_bins = 256;  // 256 is the reference; 64 bins (value >> 2) is an extension
};

namespace detail {

// Evaluates K candidate affine parameter vectors (12 floats each) of one
// modality pair: costs[k] = cost of params[k].  The default launches them on
// the caller's context (tiplhip_cost_batch_params); a serving loop installs
// one that merges the candidate sets of concurrent registrations into
// multi-pair launches.  Either way each cost is the same value.
using cost_evaluator =
    std::function<void(const tiplhip_volume* from, const float* from_vs, const tiplhip_volume* to,
                       const float* to_vs, const float* params, int K, int cost, int bins,
                       double* costs)>;

inline cost_evaluator direct_evaluator(hip::context& ctx)
{
    tiplhip_ctx* c = ctx.get();
    return [c](const tiplhip_volume* from, const float* from_vs, const tiplhip_volume* to,
               const float* to_vs, const float* params, int K, int cost, int bins, double* costs) {
        hip::check(tiplhip_cost_batch_params(c, from, from_vs, to, to_vs, params, K, cost, bins, costs));
    };
}

// The device state of one registration: volumes per modality, voxel sizes
// and the evaluator that prices candidate sets on them.  A level provides
// what the runner needs: the first modality's shapes, the number of
// modalities, the costs of K candidates of modality i, and the half-
// resolution level.  (Any type with these members can drive the runner,
// e.g. host images priced by another evaluator.)
struct linear_level {
    std::vector<std::shared_ptr<hip::volume>> from, to;
    vector<3> from_vs, to_vs;
    cost_evaluator eval;
    const shape<3>& from_shape() const { return from[0]->shape(); }
    const shape<3>& to_shape() const { return to[0]->shape(); }
    size_t modalities() const { return std::min(from.size(), to.size()); }
    void costs(size_t i, const float* params, int K, int cost, int bins, double* out) const
    {
        eval(from[i]->get(), from_vs.begin(), to[i]->get(), to_vs.begin(), params, K, cost, bins,
             out);
    }
    linear_level down() const
    {
        linear_level d;
        for (auto& v : from) d.from.push_back(std::make_shared<hip::volume>(v->downsample()));
        for (auto& v : to) d.to.push_back(std::make_shared<hip::volume>(v->downsample()));
        d.from_vs = from_vs;
        d.to_vs = to_vs;
        d.from_vs *= 2.0f;
        d.to_vs *= 2.0f;
        d.eval = eval;
        return d;
    }
};

// linear_reg_runner (reg/linear.hpp:290-561).
template <class Level>
class basic_runner {
public:
    affine_param<float, 3> arg_upper, arg_lower;
    affine_param<float, 3>& arg_min;
    unsigned int search_count = 16;
    double precision = 0.001;
    size_t max_iterations = 128;
    std::vector<reg_type> reg_list = {translocation, translocation_scaling, rigid_scaling, affine};
    cost_type cost = mutual_info;
    int bins = 256;

    basic_runner(Level lv, affine_param<float, 3>& arg) : arg_min(arg), lv_(std::move(lv)) {}

    std::pair<affine_param<float, 3>, affine_param<float, 3>> current_bound(reg_type type) const
    {
        affine_param<float, 3> upper(arg_upper), lower(arg_lower);
        for (int i = 0; i < 3; ++i) {
            if (!(type & translocation)) upper.translocation[i] = lower.translocation[i] = arg_min.translocation[i];
            if (!(type & scaling)) upper.scaling[i] = lower.scaling[i] = arg_min.scaling[i];
            if (!(type & rotation)) upper.rotation[i] = lower.rotation[i] = arg_min.rotation[i];
            if (!(type & tilt)) upper.affine[i] = lower.affine[i] = arg_min.affine[i];
        }
        return {upper, lower};
    }

    void set_bound(reg_type type, const float bound[3][8], bool absolute)
    {
        if (type == translocation) reg_list = {translocation};
        if (type == rotation) reg_list = {rotation};
        if (type == rigid_body) reg_list = {translocation, rigid_body};
        if (type == scaling) reg_list = {scaling};
        if (type == rigid_scaling) reg_list = {translocation, translocation_scaling, rigid_scaling};
        if (type == affine) reg_list = {translocation, translocation_scaling, rigid_scaling, affine};
        if (absolute) {
            arg_upper.clear();
            arg_lower.clear();
        } else {
            arg_upper = arg_lower = arg_min;
        }
        const shape<3>& fs = lv_.from_shape();
        const shape<3>& ts = lv_.to_shape();
        for (int i = 0; i < 3; ++i) {
            if (type & translocation) {
                const float range = std::max<float>(fs[i] * lv_.from_vs[i], ts[i] * lv_.to_vs[i]) * 0.5f;
                arg_upper.translocation[i] += range * bound[i][0];
                arg_lower.translocation[i] += range * bound[i][1];
            }
            if (type & scaling) {
                arg_upper.scaling[i] *= bound[i][4];
                arg_lower.scaling[i] *= bound[i][5];
            }
            if (type & rotation) {
                arg_upper.rotation[i] += 1.57079632679f * bound[i][2];
                arg_lower.rotation[i] += 1.57079632679f * bound[i][3];
            }
            if (type & tilt) {
                arg_upper.affine[i] += bound[i][6];
                arg_lower.affine[i] += bound[i][7];
            }
        }
    }

    // fun over a candidate set: per modality one batched launch, then the
    // reference's mean over modalities divided by 2 (reg/linear.hpp:427-444).
    std::vector<double> evaluate(const std::vector<optimization::param_vec>& xs)
    {
        std::vector<double> total(xs.size(), 0.0);
        if (xs.empty()) return total;
        std::vector<float> P(12 * xs.size());
        for (size_t k = 0; k < xs.size(); ++k) std::copy_n(xs[k].data(), 12, &P[12 * k]);
        std::vector<double> c(xs.size());
        const size_t nmod = lv_.modalities();
        const int kind = cost == mutual_info ? TIPLHIP_COST_MI : TIPLHIP_COST_CORR;
        for (size_t i = 0; i < nmod; ++i) {
            lv_.costs(i, P.data(), int(xs.size()), kind, bins, c.data());
            for (size_t k = 0; k < xs.size(); ++k) total[k] += c[k];
        }
        for (double& v : total) v /= double(float(nmod) * 2.0f);
        evaluations += xs.size();
        return total;
    }

    template <class Term>
    float run_optimize(unsigned int cur_search_count, Term&& is_terminated)
    {
        auto fun = [&](const std::vector<optimization::param_vec>& xs) { return evaluate(xs); };
        optimization::param_vec x(arg_min.begin(), arg_min.end());
        double optimal_value = evaluate({x})[0];
        if (cur_search_count) {
            for (size_t it = 0; it < reg_list.size() && !is_terminated(); ++it) {
                const auto b = current_bound(reg_list[it]);
                optimization::param_vec x2(x);
                double v2 = optimal_value;
                optimization::line_search(x2, b.first.begin(), b.second.begin(),
                                          {0, 0, 0, 1, 1, 1, 1, 1, 1, 1, 1, 1}, fun, v2,
                                          int(cur_search_count), is_terminated);
                if (v2 < optimal_value) {
                    optimal_value = v2;
                    x = x2;
                }
                optimization::lbfgs(x, b.first.begin(), b.second.begin(), fun, optimal_value,
                                    is_terminated, precision, int(max_iterations));
                std::copy_n(x.data(), 12, arg_min.begin());
            }
        } else {
            const auto b = current_bound(reg_list.back());
            optimization::lbfgs(x, b.first.begin(), b.second.begin(), fun, optimal_value,
                                is_terminated, precision * 0.2f, int(max_iterations));
            std::copy_n(x.data(), 12, arg_min.begin());
        }
        return float(optimal_value);
    }

    // run_optimize_mr (reg/linear.hpp:527-550): above 96^3 optimise a
    // half-resolution copy first (recursively); above 256^3 stop there.
    template <class Term>
    float run_optimize_mr(Term&& is_terminated)
    {
        const size_t sz = lv_.from_shape().size();
        const bool is_large = double(sz) > std::pow(96.0, 3);
        if (is_large) {
            basic_runner low(lv_.down(), arg_min);
            low.arg_upper = arg_upper;
            low.arg_lower = arg_lower;
            low.search_count = search_count;
            low.precision = precision;
            low.max_iterations = max_iterations;
            low.reg_list = reg_list;
            low.cost = cost;
            low.bins = bins;
            const float v = low.run_optimize_mr(is_terminated);
            evaluations += low.evaluations;
            if (double(sz) > std::pow(256.0, 3)) return v;
        }
        return run_optimize(is_large ? 0 : search_count, is_terminated);
    }

    size_t evaluations = 0;

private:
    Level lv_;
};

using runner = basic_runner<linear_level>;

// has_mask (numerical/basic_op.hpp:811-830) with bounding_box (:697-731).
template <class I>
bool has_mask(const I& img)
{
    const size_t sz = img.size(), threshold = sz / 5;
    size_t zeros = 0;
    for (size_t i = 0; i < sz && zeros <= threshold; ++i)
        if (img[i] == 0) ++zeros;
    if (zeros > threshold) return true;
    const shape<3>& s = img.shape();
    int lo[3] = {int(s[0]) - 1, int(s[1]) - 1, int(s[2]) - 1}, hi[3] = {0, 0, 0};
    size_t i = 0;
    for (unsigned z = 0; z < s[2]; ++z)
        for (unsigned y = 0; y < s[1]; ++y)
            for (unsigned x = 0; x < s[0]; ++x, ++i) {
                if (img[i] <= 0) continue;
                const int p[3] = {int(x), int(y), int(z)};
                for (int d = 0; d < 3; ++d) {
                    lo[d] = std::min(lo[d], p[d]);
                    hi[d] = std::max(hi[d], p[d]);
                }
            }
    for (int d = 0; d < 3; ++d) {
        if (hi[d] == 0) hi[d] = lo[d] = 0;
        ++hi[d];
        if (lo[d] > 0 && hi[d] < int(s[d])) return true;
    }
    return false;
}

// eigen_decomposition_sym for a symmetric 3x3 matrix
// (numerical/matrix.hpp:1829-1935): the closed-form trigonometric solution
// of the characteristic cubic (O. K. Smith, "Eigenvalues of a symmetric
// 3x3 matrix", CACM 4(4):168, 1961), eigenvectors from the cross product of
// two rows of A - d I, and the reference's shortcuts for matrices with
// zero off-diagonal entries.  Eigenvalues in the reference's order (not
// sorted), eigenvector i in V[3i..3i+2] with the reference's signs, every
// operation in the reference's order, so the moment-based starting affine
// below matches the reference's bit for bit.
inline void eigen_sym3(const double* A, double* V, double* d)
{
    auto identity = [&] {
        std::fill(V, V + 9, 0.0);
        V[0] = V[4] = V[8] = 1.0;
    };
    // one off-diagonal pair (p,q) non-zero, k the decoupled axis
    auto plane = [&](int k, double app, double aqq, double apq, int vk, int v_pq0, int v_pq1,
                     int v_a, int v_b) {
        const double apq2 = apq * apq;
        d[0] = A[k];
        d[1] = d[2] = (app + aqq) * 0.5;
        const double half = (app - aqq) * 0.5;
        const double l = std::sqrt(half * half + apq2);
        d[1] += l;
        d[2] -= l;
        std::fill(V, V + 9, 0.0);
        V[vk] = 1.0;
        const double a_d1 = app - d[1];
        const double length = std::sqrt(apq2 + a_d1 * a_d1);
        V[v_pq0] = V[v_pq1] = -apq / length;
        V[v_a] = a_d1 / length;
        V[v_b] = -V[v_a];
    };
    if (A[1] == 0.0) {
        if (A[2] == 0.0) {
            if (A[5] == 0.0) {
                d[0] = A[0];
                d[1] = A[4];
                d[2] = A[8];
                identity();
                return;
            }
            plane(0, A[4], A[8], A[5], 0, 4, 8, 5, 7);
            return;
        }
        if (A[5] == 0.0) {
            plane(4, A[0], A[8], A[2], 1, 3, 8, 5, 6);
            return;
        }
    } else if (A[2] == 0.0 && A[5] == 0.0) {
        plane(8, A[0], A[4], A[1], 2, 3, 7, 4, 6);
        return;
    }
    const double A2A1 = A[2] * A[1], A1A5 = A[1] * A[5], A2A5 = A[2] * A[5];
    const double A0A4 = A[0] * A[4], A1A1 = A[1] * A[1], A2A2 = A[2] * A[2], A5A5 = A[5] * A[5];
    const double A0_A4 = A[0] + A[4];
    {
        const double I1 = A0_A4 + A[8];
        const double I2 = A0A4 + A0_A4 * A[8] - A1A1 - A2A2 - A5A5;
        const double I3 = A0A4 * A[8] + 2.0 * A2A1 * A[5] - (A[8] * A1A1 + A[4] * A2A2 + A[0] * A5A5);
        const double I1_3 = I1 / 3.0;
        const double I1_3sq = I1_3 * I1_3;
        const double v = I1_3sq - I2 / 3.0;
        const double s = I1_3sq * I1_3 - I1 * I2 / 6.0 + I3 / 2.0;
        if (v == 0.0) {
            std::fill(d, d + 3, 0.0);
            identity();
            return;
        }
        const double sqrt_v = std::sqrt(v);
        const double angle = std::acos(std::max<double>(-1.0, std::min<double>(1.0, s / v / sqrt_v))) / 3.0;
        d[0] = I1_3 + 2.0 * sqrt_v * std::cos(angle);
        d[1] = I1_3 - 2.0 * sqrt_v * std::cos(3.14159265358979323846 / 3.0 + angle);
        d[2] = I1 - d[0] - d[1];
    }
    for (int i = 0; i < 3; ++i) {
        double* e = V + 3 * i;
        const double Ai = A[0] - d[i], Bi = A[4] - d[i], Ci = A[8] - d[i];
        const double q1 = A2A1 - Ai * A[5], q2 = A1A5 - Bi * A[2], q3 = A2A5 - Ci * A[1];
        e[0] = q2 * q3;
        e[1] = q3 * q1;
        e[2] = q2 * q1;
        const double length = std::sqrt(e[0] * e[0] + e[1] * e[1] + e[2] * e[2]);
        e[0] /= length;
        e[1] /= length;
        e[2] /= length;
    }
}

// estimate_affine_param (numerical/transformation.hpp:933-1021): initial
// affine from the centroid and second moments of the non-zero voxels, in the
// reference's double arithmetic and order.
template <class I>
void estimate_affine_param(const I& src, const vector<3>& svs, const I& tgt, const vector<3>& tvs,
                           affine_param<float, 3>& arg)
{
    arg.clear();
    auto moments = [](const I& img, const vector<3>& vs, double* c, double* cov) {
        const shape<3>& s = img.shape();
        size_t n = 0, i = 0;
        c[0] = c[1] = c[2] = 0.0;
        for (unsigned z = 0; z < s[2]; ++z)
            for (unsigned y = 0; y < s[1]; ++y)
                for (unsigned x = 0; x < s[0]; ++x, ++i)
                    if (img[i] > 0) {
                        c[0] += x * vs[0];
                        c[1] += y * vs[1];
                        c[2] += z * vs[2];
                        ++n;
                    }
        if (!n) return n;
        for (int k = 0; k < 3; ++k) c[k] /= double(n);
        std::fill(cov, cov + 9, 0.0);
        i = 0;
        for (unsigned z = 0; z < s[2]; ++z)
            for (unsigned y = 0; y < s[1]; ++y)
                for (unsigned x = 0; x < s[0]; ++x, ++i)
                    if (img[i] > 0) {
                        const double d[3] = {x * vs[0] - c[0], y * vs[1] - c[1], z * vs[2] - c[2]};
                        cov[0] += d[0] * d[0];
                        cov[1] += d[0] * d[1];
                        cov[2] += d[0] * d[2];
                        cov[4] += d[1] * d[1];
                        cov[5] += d[1] * d[2];
                        cov[8] += d[2] * d[2];
                    }
        cov[3] = cov[1];
        cov[6] = cov[2];
        cov[7] = cov[5];
        for (int k = 0; k < 9; ++k) cov[k] /= double(n);
        return n;
    };
    double cs[3], ct[3], Cs[9], Ct[9];
    if (!moments(src, svs, cs, Cs) || !moments(tgt, tvs, ct, Ct)) return;
    double Vs[9], Vt[9], Ls[3], Lt[3], R[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    eigen_sym3(Cs, Vs, Ls);
    eigen_sym3(Ct, Vt, Lt);
    for (int i = 0; i < 3; ++i)
        arg.scaling[i] = float(std::sqrt(std::max(Lt[i], 0.0) / std::max(Ls[i], 1e-6)));
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            for (int k = 0; k < 3; ++k) R[i * 3 + j] += Vt[i + k * 3] * Vs[j + k * 3];
    const double det = R[0] * (R[4] * R[8] - R[5] * R[7]) - R[1] * (R[3] * R[8] - R[5] * R[6]) +
                       R[2] * (R[3] * R[7] - R[4] * R[6]);
    if (det < 0.0)
        for (int i = 0; i < 3; ++i) R[i * 3 + 2] = -R[i * 3 + 2];
    arg.rotation[0] = float(std::atan2(R[7], R[8]));
    arg.rotation[1] = float(std::atan2(-R[6], std::sqrt(R[7] * R[7] + R[8] * R[8])));
    arg.rotation[2] = float(std::atan2(R[3], R[0]));
    const float pi_half = 1.570796327f;
    for (int i = 0; i < 3; ++i) arg.rotation[i] = std::remainder(arg.rotation[i], pi_half);
    double dt[3];
    for (int i = 0; i < 3; ++i)
        dt[i] = (ct[i] - tgt.shape()[i] * tvs[i] * 0.5) / std::max<double>(arg.scaling[i], 1e-6);
    for (int i = 0; i < 3; ++i) {
        const double r = R[i] * dt[0] + R[3 + i] * dt[1] + R[6 + i] * dt[2];
        arg.translocation[i] = float(r - (cs[i] - src.shape()[i] * svs[i] * 0.5));
    }
}

// The body of tipl::reg::linear (reg/linear.hpp:622-700) over any level
// type (see linear_level).  from / to are the host uint8 images (has_mask and
// the moment estimate read them); `base` prices candidates on the same
// images, its voxel sizes are set here.
template <typename out_type = void, class Level>
float linear_search(const std::vector<const_pointer_image<3, unsigned char>>& from, vector<3> from_vs,
                    const std::vector<const_pointer_image<3, unsigned char>>& to, vector<3> to_vs,
                    const Level& base, affine_param<float, 3>& arg, linear_reg_param param,
                    bool& terminated, size_t* evaluations = nullptr)
{
    auto term = [&] { return bool(terminated); };
    auto run = [&](const vector<3>& f_vs, const vector<3>& t_vs, affine_param<float, 3>& a, bool mr) {
        Level lv = base;
        lv.from_vs = f_vs;
        lv.to_vs = t_vs;
        basic_runner<Level> r(std::move(lv), a);
        r.set_bound(param.reg_type, param.bound, param.absolute_bound);
        r.search_count = param.search_count;
        r.cost = param.cost_type;
        r.bins = param.mi_bins;
        if constexpr (!std::is_void_v<out_type>)
            out_type() << (param.reg_type == affine ? "affine" : "rigid body") << " registration by "
                       << (param.cost_type == mutual_info ? "mutual info" : "correlation") << " using gpu";
        const float v = mr ? r.run_optimize_mr(term) : r.run_optimize(0, term);
        if (evaluations) *evaluations += r.evaluations;
        return v;
    };

    if (param.bound == narrow_bound) {
        param.absolute_bound = false;
        param.search_count = 0;
        return run(from_vs, to_vs, arg, false);
    }
    vector<3> new_to_vs = to_vs;
    affine_param<float, 3> surrogate = arg;
    bool adjust_vs = false;
    if (param.reg_type == affine && has_mask(from[0]) && has_mask(to[0])) {
        estimate_affine_param(from[0], from_vs, to[0], to_vs, surrogate);
        arg = surrogate;
        for (int i = 0; i < 3; ++i) {
            new_to_vs[i] = to_vs[i] / std::max<float>(surrogate.scaling[i], 0.01f);
            surrogate.scaling[i] = 1.0f;
        }
        adjust_vs = true;
    }
    param.absolute_bound = true;
    const float result = run(from_vs, new_to_vs, adjust_vs ? surrogate : arg, true);
    if (adjust_vs)
        arg = transformation_matrix<float, 3>(surrogate, from[0].shape(), from_vs, to[0].shape(), new_to_vs)
                  .to_affine_param(from[0].shape(), from_vs, to[0].shape(), to_vs);
    return result;
}

// linear_search on volumes already on the device, priced by `eval`.
template <typename out_type = void>
float linear_impl(const std::vector<const_pointer_image<3, unsigned char>>& from, vector<3> from_vs,
                  const std::vector<const_pointer_image<3, unsigned char>>& to, vector<3> to_vs,
                  const std::vector<std::shared_ptr<hip::volume>>& dev_from,
                  const std::vector<std::shared_ptr<hip::volume>>& dev_to,
                  affine_param<float, 3>& arg, linear_reg_param param, bool& terminated,
                  const cost_evaluator& eval, size_t* evaluations = nullptr)
{
    linear_level base;
    base.from = dev_from;
    base.to = dev_to;
    base.eval = eval;
    return linear_search<out_type>(from, from_vs, to, to_vs, base, arg, param, terminated,
                                   evaluations);
}

}  // namespace detail

// make_list (reg/linear.hpp:566-592): image lists as const views; a vector
// of images stops at its first empty entry.
template <typename image_type>
auto make_list(const image_type& I, const image_type& I2)
{
    auto pI = tipl::make_shared(I);
    if (I2.empty()) return std::vector<decltype(pI)>({pI});
    return std::vector<decltype(pI)>({pI, tipl::make_shared(I2)});
}

template <typename image_type>
auto make_list(const image_type& I)
{
    auto pI = tipl::make_shared(I);
    return std::vector<decltype(pI)>({pI});
}

template <typename T>
auto make_list(const std::vector<T>& data)
{
    std::vector<tipl::const_pointer_image<T::dimension, typename T::value_type>> out;
    for (const auto& each : data) {
        if (each.empty()) break;
        out.push_back(tipl::make_shared(each));
    }
    return out;
}

// tipl::reg::linear (reg/linear.hpp:622-700): registers `from` onto `to`
// (lists of modalities, uint8), arg in/out; returns the final cost.
template <typename out_type = void, int dim>
float linear(std::vector<const_pointer_image<dim, unsigned char>> from, vector<dim> from_vs,
             std::vector<const_pointer_image<dim, unsigned char>> to, vector<dim> to_vs,
             affine_param<float, dim>& arg, linear_reg_param param = linear_reg_param(),
             bool& terminated = tipl::prog_aborted)
{
    static_assert(dim == 3, "the MI355X path registers 3-D volumes");
    if (from.empty() || to.empty()) throw std::invalid_argument("linear: empty image list");
    if (!param.cuda) throw std::runtime_error("linear: only the GPU path is provided (param.cuda)");
    hip::context& ctx = hip::thread_context();
    std::vector<std::shared_ptr<hip::volume>> df, dt;
    for (auto& f : from) df.push_back(std::make_shared<hip::volume>(ctx, f.data(), f.shape()));
    for (auto& t : to) dt.push_back(std::make_shared<hip::volume>(ctx, t.data(), t.shape()));
    return detail::linear_impl<out_type>(from, from_vs, to, to_vs, df, dt, arg, param, terminated,
                                         detail::direct_evaluator(ctx));
}

// The overload returning the transformation (reg/linear.hpp:703-714).
template <typename out_type = void, int dim>
transformation_matrix<float, dim> linear(std::vector<const_pointer_image<dim, unsigned char>> from,
                                         vector<dim> from_vs,
                                         std::vector<const_pointer_image<dim, unsigned char>> to,
                                         vector<dim> to_vs,
                                         const linear_reg_param& param = linear_reg_param(),
                                         bool& terminated = tipl::prog_aborted)
{
    affine_param<float, dim> arg;
    linear<out_type>(from, from_vs, to, to_vs, arg, param, terminated);
    return transformation_matrix<float, dim>(arg, from[0].shape(), from_vs, to[0].shape(), to_vs);
}

}  // namespace reg
}  // namespace tipl

#endif
