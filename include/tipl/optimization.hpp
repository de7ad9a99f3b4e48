// optimization.hpp -- the reference's optimisers with batched cost calls.
//
// Restates numerical/optimization.hpp:16-512 (estimate_change, gradient,
// armijo_line_search, calculate_resolution, lbfgs, line_search).  Where the
// reference evaluates a candidate set with par_for -- one cost call per
// thread -- these evaluate the whole set with one call of `fun`, which takes
// a list of parameter vectors and returns their costs; on the GPU path that
// is one batched kernel launch.  Candidate sets, update rules and the order
// of every floating-point operation on parameters follow the reference;
// where the reference updates its best point from concurrent threads in
// completion order, these scan the candidates in index order.
#ifndef TIPL_HIP_OPTIMIZATION_HPP
#define TIPL_HIP_OPTIMIZATION_HPP

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <vector>

namespace tipl {
namespace optimization {

using param_vec = std::vector<float>;

// fun_x_ei[i] = fun(x + tol_i e_i) for tol_i != 0 (:16-35).
template <class BatchFun>
void estimate_change(const param_vec& x, const param_vec& tol, std::vector<double>& fun_ei,
                     BatchFun&& fun)
{
    std::vector<param_vec> cand;
    std::vector<size_t> idx;
    for (size_t i = 0; i < x.size(); ++i) {
        if (tol[i] == 0) continue;
        param_vec y(x);
        y[i] += tol[i];
        cand.push_back(std::move(y));
        idx.push_back(i);
    }
    if (cand.empty()) return;
    const std::vector<double> c = fun(cand);
    for (size_t k = 0; k < idx.size(); ++k) fun_ei[idx[k]] = c[k];
}

// (fun(x + tol e_i) - fun(x)) / tol_i, 0 where tol_i == 0 (:37-57).
inline void gradient(const param_vec& tol, double fun_x, const std::vector<double>& fun_ei,
                     param_vec& g)
{
    for (size_t i = 0; i < g.size(); ++i) {
        float v = float(fun_ei[i]);
        v = float(double(v) - fun_x);
        g[i] = (tol[i] == 0) ? 0.0f : v / tol[i];
    }
}

// Ten steps 2^-6 .. 2^3 along -g, clamped to the bounds; keeps the best if
// it improves on fun_x (:59-96).
template <class BatchFun>
bool armijo_line_search(param_vec& x, const float* upper, const float* lower, const param_vec& g,
                        double& fun_x, BatchFun&& fun)
{
    std::vector<param_vec> cand;
    for (double step = 0.015625; step <= 10.0; step *= 2) {
        param_vec nx(x);
        for (size_t j = 0; j < nx.size(); ++j) nx[j] += g[j] * -step;  // aypx(g, -step, nx)
        for (size_t j = 0; j < nx.size(); ++j)
            nx[j] = float(std::min<double>(std::max<double>(nx[j], lower[j]), upper[j]));
        cand.push_back(std::move(nx));
    }
    const std::vector<double> c = fun(cand);
    size_t best = 0;
    double best_v = fun_x;
    for (size_t k = 0; k < c.size(); ++k)
        if (c[k] < best_v) {
            best_v = c[k];
            best = k + 1;
        }
    if (best == 0) return false;
    fun_x = best_v;
    x = cand[best - 1];
    return true;
}

inline float norm2(const param_vec& v)
{
    float r = 0.0f;
    for (float a : v) r += a * a;
    return std::sqrt(r);
}

// tols = (upper - lower) * precision; returns |tols| (:98-104).
inline double calculate_resolution(param_vec& tols, const float* upper, const float* lower,
                                   double precision)
{
    for (size_t i = 0; i < tols.size(); ++i) tols[i] = float((upper[i] - lower[i]) * precision);
    return norm2(tols);
}

// L-BFGS with range-scaled preconditioning (:137-270).
template <class BatchFun, class Term>
void lbfgs(param_vec& x, const float* upper, const float* lower, BatchFun&& fun, double& fun_x,
           Term&& terminated, double precision = 0.001, int max_iteration = 100, size_t m = 10)
{
    const size_t size = x.size();
    param_vec tols(size);
    const double tol_length = calculate_resolution(tols, upper, lower, precision);
    std::vector<param_vec> S, Y;
    std::vector<double> RHO;
    param_vec g(size), g_old(size), x_old(size);
    std::vector<double> fun_ei(size);
    estimate_change(x, tols, fun_ei, fun);
    gradient(tols, fun_x, fun_ei, g);
    for (int iter = 0; iter < max_iteration && !terminated(); ++iter) {
        param_vec q = g;
        std::vector<double> alpha(S.size());
        for (int i = int(S.size()) - 1; i >= 0; --i) {
            double sq = 0.0;
            for (size_t j = 0; j < size; ++j) sq += S[size_t(i)][j] * q[j];
            alpha[size_t(i)] = RHO[size_t(i)] * sq;
            for (size_t j = 0; j < size; ++j) q[j] -= alpha[size_t(i)] * Y[size_t(i)][j];
        }
        if (!S.empty()) {
            double yty = 0.0;
            for (size_t j = 0; j < size; ++j) yty += Y.back()[j] * Y.back()[j];
            if (yty > 1e-14) {
                const double gamma = (1.0 / RHO.back()) / yty;
                for (size_t j = 0; j < size; ++j) {
                    const double range = upper[j] - lower[j];
                    q[j] *= gamma * range;
                }
            }
        } else {
            for (size_t j = 0; j < size; ++j) {
                const double range = upper[j] - lower[j];
                q[j] *= range;
            }
            const float length = norm2(q);
            if (length > 0.0f)
                for (float& v : q) v *= tol_length / length;
        }
        for (size_t i = 0; i < S.size(); ++i) {
            double yq = 0.0;
            for (size_t j = 0; j < size; ++j) yq += Y[i][j] * q[j];
            const double beta = RHO[i] * yq;
            for (size_t j = 0; j < size; ++j) q[j] += S[i][j] * (alpha[i] - beta);
        }
        x_old = x;
        g_old = g;
        if (!armijo_line_search(x, upper, lower, q, fun_x, fun)) {
            if (!S.empty()) {
                S.clear();
                Y.clear();
                RHO.clear();
                continue;
            }
            break;
        }
        estimate_change(x, tols, fun_ei, fun);
        gradient(tols, fun_x, fun_ei, g);
        param_vec s_k(size), y_k(size);
        double sty = 0.0;
        for (size_t i = 0; i < size; ++i) {
            s_k[i] = x[i] - x_old[i];
            y_k[i] = g[i] - g_old[i];
            sty += s_k[i] * y_k[i];
        }
        if (sty > 1e-14) {
            if (S.size() == m) {
                S.erase(S.begin());
                Y.erase(Y.begin());
                RHO.erase(RHO.begin());
            }
            S.push_back(std::move(s_k));
            Y.push_back(std::move(y_k));
            RHO.push_back(1.0 / sty);
        }
    }
}

// Coordinate grid search with a resultant-direction pass (:272-512).
// search_strategy[i]: 0 = uniform grid, 1 = cubic (denser near 0).
template <class BatchFun, class Term>
void line_search(param_vec& x, const float* upper, const float* lower,
                 const std::vector<int>& search_strategy, BatchFun&& fun, double& optimal_value,
                 int grid_samples, Term&& is_terminated)
{
    const size_t size = x.size();
    const double ftol = 0.05;
    std::vector<param_vec> xi(size, param_vec(size, 0.0f));
    for (size_t i = 0; i < size; ++i) xi[i][i] = 1.0f;
    param_vec p(x);
    for (int iter = 0; iter < 2; ++iter) {
        if (is_terminated()) return;
        const double fp = optimal_value;
        std::vector<double> dim_steps(size, 0.0), dim_upper(size, 0.0), dim_lower(size, 0.0);
        std::vector<int> active;
        for (size_t i = 0; i < size; ++i) {
            double bound_pos = 1e10, bound_neg = -1e10;
            bool movable = false;
            for (size_t j = 0; j < size; ++j)
                if (std::abs(xi[i][j]) > 1e-8 && upper[j] != lower[j]) {
                    movable = true;
                    double up = (upper[j] - p[j]) / xi[i][j];
                    double lo = (lower[j] - p[j]) / xi[i][j];
                    if (up < lo) std::swap(up, lo);
                    bound_pos = std::min(bound_pos, up);
                    bound_neg = std::max(bound_neg, lo);
                }
            if (!movable) continue;
            if (bound_pos < 0.0) bound_pos = 0.0;
            if (bound_neg > 0.0) bound_neg = 0.0;
            if (bound_pos > 1e9) bound_pos = 1.0;
            if (bound_neg < -1e9) bound_neg = -1.0;
            const double scale = 0.5 / (iter + 1.0);
            const double cu = bound_pos * scale, cl = bound_neg * scale;
            if (cu - cl > 1e-8) {
                active.push_back(int(i));
                dim_upper[i] = cu;
                dim_lower[i] = cl;
            }
        }
        if (is_terminated()) return;
        // stage 2: all grid points of all active dimensions in one batch
        std::vector<double> costs(size * size_t(grid_samples), fp), steps(size * size_t(grid_samples), 0.0);
        std::vector<param_vec> cand;
        std::vector<size_t> where;
        for (int i : active)
            for (int s = 0; s < grid_samples; ++s) {
                const int strategy = size_t(i) < search_strategy.size() ? search_strategy[size_t(i)] : 0;
                const double cu = dim_upper[size_t(i)], cl = dim_lower[size_t(i)];
                double trial;
                if (strategy == 0) {
                    trial = cl + s * ((cu - cl) / (grid_samples - 1));
                } else {
                    const double u = double(s - (grid_samples - 1) / 2.0) / ((grid_samples - 1) / 2.0);
                    const double v = u * u * u;
                    trial = (v > 0.0) ? (v * cu) : (v * -cl);
                }
                const size_t li = size_t(i) * size_t(grid_samples) + size_t(s);
                steps[li] = trial;
                if (std::abs(trial) < 1e-8) continue;
                param_vec t(p);
                for (size_t j = 0; j < size; ++j)
                    t[j] = std::min<float>(upper[j], std::max<float>(lower[j], float(p[j] + trial * xi[size_t(i)][j])));
                cand.push_back(std::move(t));
                where.push_back(li);
            }
        if (!cand.empty()) {
            const std::vector<double> c = fun(cand);
            for (size_t k = 0; k < c.size(); ++k) {
                costs[where[k]] = c[k];
                if (c[k] < optimal_value) {
                    optimal_value = c[k];
                    x = cand[k];
                }
            }
        }
        if (is_terminated()) return;
        for (int i : active) {
            double best_f = fp;
            int s_min = (grid_samples - 1) / 2;
            for (int s = 0; s < grid_samples; ++s) {
                const size_t li = size_t(i) * size_t(grid_samples) + size_t(s);
                if (costs[li] < best_f) {
                    best_f = costs[li];
                    s_min = s;
                }
            }
            dim_steps[size_t(i)] = steps[size_t(i) * size_t(grid_samples) + size_t(s_min)];
        }
        if (is_terminated()) return;
        // stage 4: the resultant direction
        param_vec dir(size, 0.0f);
        bool moving = false;
        for (size_t i = 0; i < size; ++i) {
            if (std::abs(dim_steps[i]) > 1e-8) moving = true;
            for (size_t j = 0; j < size; ++j) dir[j] += dim_steps[i] * xi[i][j];
        }
        if (moving) {
            double bound_pos = 1e10, bound_neg = -1e10;
            bool movable = false;
            for (size_t j = 0; j < size; ++j)
                if (std::abs(dir[j]) > 1e-8 && upper[j] != lower[j]) {
                    movable = true;
                    double up = (upper[j] - p[j]) / dir[j];
                    double lo = (lower[j] - p[j]) / dir[j];
                    if (up < lo) std::swap(up, lo);
                    bound_pos = std::min(bound_pos, up);
                    bound_neg = std::max(bound_neg, lo);
                }
            if (movable) {
                if (bound_pos < 0.0) bound_pos = 0.0;
                if (bound_neg > 0.0) bound_neg = 0.0;
                if (bound_pos > 1e9) bound_pos = 1.0;
                if (bound_neg < -1e9) bound_neg = -1.0;
                const double step = (bound_pos - bound_neg) / (grid_samples - 1);
                if (step > 1e-8) {
                    std::vector<param_vec> c2;
                    for (int s = 0; s < grid_samples; ++s) {
                        const double trial = bound_neg + s * step;
                        if (std::abs(trial) <= 1e-8) continue;
                        param_vec t(p);
                        for (size_t j = 0; j < size; ++j)
                            t[j] = std::min<float>(upper[j], std::max<float>(lower[j], float(p[j] + trial * dir[j])));
                        c2.push_back(std::move(t));
                    }
                    if (!c2.empty()) {
                        const std::vector<double> c = fun(c2);
                        for (size_t k = 0; k < c.size(); ++k)
                            if (c[k] < optimal_value) {
                                optimal_value = c[k];
                                x = c2[k];
                            }
                    }
                }
            }
        }
        if (is_terminated()) return;
        p = x;
        if (2.0 * (fp - optimal_value) <= ftol * (std::abs(fp) + std::abs(optimal_value)) + 1e-10) break;
    }
}

}  // namespace optimization
}  // namespace tipl

#endif
