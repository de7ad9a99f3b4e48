// mirror_shim.cpp -- C entry points into the host logic of the header-only
// C++ mirror (include/tipl/optimization.hpp, include/tipl/reg/linear.hpp),
// for CPU tests.  TEST INFRASTRUCTURE: the optimisers and the linear()
// runner are driven by cost callbacks that the tests supply (the oracle's
// costs), instead of the device, so their candidate sets, evaluation counts
// and results can be compared with the reference's own routines
// (oracle/ref_shim.cpp) and with the GPU path.  Built by build() into
// tests/cpp/libmirror.so, linked against libtiplhip.so only for the host
// geometry helper tiplhip_affine_matrix (no GPU call is made).
#include <cstdint>
#include <memory>
#include <vector>

#include "tipl/reg/linear.hpp"

extern "C" {
typedef void (*mir_cost_cb)(const uint8_t* from, const int* fdim, const uint8_t* to,
                            const int* tdim, const float* T, int K, int cost, int bins,
                            double* costs, void* user);
typedef void (*mir_down_cb)(const uint8_t* in, const int* dim, uint8_t* out, int* odim,
                            void* user);
typedef void (*mir_param_cb)(const float* x, int n, int K, double* costs, void* user);
}

namespace {

using himage = tipl::image<3, uint8_t>;

// A level of host images: candidates are turned into transforms by the
// library's host geometry (the arithmetic tiplhip_cost_batch_params uses)
// and priced by the callback; down() halves the images with the callback.
struct host_level {
    std::vector<std::shared_ptr<const himage>> from, to;
    tipl::vector<3> from_vs, to_vs;
    mir_cost_cb cb = nullptr;
    mir_down_cb down_cb = nullptr;
    void* user = nullptr;

    const tipl::shape<3>& from_shape() const { return from[0]->shape(); }
    const tipl::shape<3>& to_shape() const { return to[0]->shape(); }
    size_t modalities() const { return std::min(from.size(), to.size()); }
    static void dims(const himage& I, int* d)
    {
        for (int i = 0; i < 3; ++i) d[i] = int(I.shape()[i]);
    }
    void costs(size_t i, const float* params, int K, int cost, int bins, double* out) const
    {
        int fd[3], td[3];
        dims(*from[i], fd);
        dims(*to[i], td);
        std::vector<float> T(12 * size_t(K));
        for (int k = 0; k < K; ++k)
            tipl::hip::check(tiplhip_affine_matrix(params + 12 * k, fd, from_vs.begin(), td,
                                                   to_vs.begin(), &T[12 * size_t(k)]));
        cb(from[i]->data(), fd, to[i]->data(), td, T.data(), K, cost, bins, out, user);
    }
    std::shared_ptr<const himage> half(const himage& I) const
    {
        int d[3], od[3];
        dims(I, d);
        auto out = std::make_shared<himage>(
            tipl::shape<3>(unsigned(d[0] + 1) / 2, unsigned(d[1] + 1) / 2, unsigned(d[2] + 1) / 2));
        down_cb(I.data(), d, out->data(), od, user);
        return out;
    }
    host_level down() const
    {
        host_level r = *this;
        for (auto& v : r.from) v = half(*v);
        for (auto& v : r.to) v = half(*v);
        r.from_vs *= 2.0f;
        r.to_vs *= 2.0f;
        return r;
    }
};

}  // namespace

extern "C" {

// tipl::reg::linear's body (detail::linear_search) on one host image pair;
// arg in/out, *evaluations receives the number of cost evaluations.
float mir_linear(const uint8_t* from, const int* fdim, const float* fvs, const uint8_t* to,
                 const int* tdim, const float* tvs, int cost_type, int reg_type, int bins,
                 float* arg, mir_cost_cb cb, mir_down_cb down_cb, void* user, uint64_t* evaluations)
{
    auto shp = [](const int* d) { return tipl::shape<3>(unsigned(d[0]), unsigned(d[1]), unsigned(d[2])); };
    auto F = std::make_shared<himage>(shp(fdim));
    auto T = std::make_shared<himage>(shp(tdim));
    std::copy_n(from, F->size(), F->data());
    std::copy_n(to, T->size(), T->data());
    host_level base;
    base.from = {F};
    base.to = {T};
    base.cb = cb;
    base.down_cb = down_cb;
    base.user = user;
    std::vector<tipl::const_pointer_image<3, unsigned char>> f{tipl::make_shared(*F)};
    std::vector<tipl::const_pointer_image<3, unsigned char>> t{tipl::make_shared(*T)};
    tipl::affine_param<float, 3> a;
    std::copy_n(arg, 12, a.begin());
    tipl::reg::linear_reg_param param;
    param.cost_type = cost_type ? tipl::reg::mutual_info : tipl::reg::corr;
    param.reg_type = tipl::reg::reg_type(reg_type);
    param.mi_bins = bins;
    bool terminated = false;
    size_t n = 0;
    const float r = tipl::reg::detail::linear_search(f, tipl::vector<3>(fvs), t, tipl::vector<3>(tvs),
                                                     base, a, param, terminated, &n);
    std::copy_n(a.begin(), 12, arg);
    if (evaluations) *evaluations = n;
    return r;
}

int mir_has_mask(const uint8_t* I, const int* dim)
{
    tipl::const_pointer_image<3, uint8_t> img(I, tipl::shape<3>(unsigned(dim[0]), unsigned(dim[1]),
                                                                 unsigned(dim[2])));
    return tipl::reg::detail::has_mask(img) ? 1 : 0;
}

void mir_estimate_affine_param(const uint8_t* src, const int* sdim, const float* svs,
                               const uint8_t* tgt, const int* tdim, const float* tvs, float* arg)
{
    auto shp = [](const int* d) { return tipl::shape<3>(unsigned(d[0]), unsigned(d[1]), unsigned(d[2])); };
    tipl::const_pointer_image<3, uint8_t> S(src, shp(sdim)), T(tgt, shp(tdim));
    tipl::affine_param<float, 3> a;
    tipl::reg::detail::estimate_affine_param(S, tipl::vector<3>(svs), T, tipl::vector<3>(tvs), a);
    std::copy_n(a.begin(), 12, arg);
}

// The batched optimisers on n parameters: each candidate set goes to the
// callback in one call (K candidates).
double mir_line_search(float* x, int n, const float* upper, const float* lower,
                       const int* strategy, int grid, double optimal, mir_param_cb cb, void* user)
{
    tipl::optimization::param_vec X(x, x + n);
    std::vector<int> S(strategy, strategy + n);
    auto fun = [&](const std::vector<tipl::optimization::param_vec>& xs) {
        std::vector<float> P;
        for (auto& v : xs) P.insert(P.end(), v.begin(), v.end());
        std::vector<double> c(xs.size());
        if (!xs.empty()) cb(P.data(), n, int(xs.size()), c.data(), user);
        return c;
    };
    tipl::optimization::line_search(X, upper, lower, S, fun, optimal, grid, [] { return false; });
    std::copy(X.begin(), X.end(), x);
    return optimal;
}

double mir_lbfgs(float* x, int n, const float* upper, const float* lower, double fun_x,
                 double precision, int max_iteration, mir_param_cb cb, void* user)
{
    tipl::optimization::param_vec X(x, x + n);
    auto fun = [&](const std::vector<tipl::optimization::param_vec>& xs) {
        std::vector<float> P;
        for (auto& v : xs) P.insert(P.end(), v.begin(), v.end());
        std::vector<double> c(xs.size());
        if (!xs.empty()) cb(P.data(), n, int(xs.size()), c.data(), user);
        return c;
    };
    tipl::optimization::lbfgs(X, upper, lower, fun, fun_x, [] { return false; }, precision,
                              max_iteration);
    std::copy(X.begin(), X.end(), x);
    return fun_x;
}

}  // extern "C"
