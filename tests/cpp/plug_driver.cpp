// plug_driver.cpp -- TEST INFRASTRUCTURE: C entry points that call the
// reference's own registration entry points, tipl::reg::linear<tipl::out>
// (reg/linear.hpp:622-690) and tipl::reg::cdm_common<tipl::out>
// (reg/cdm.hpp:534-558), compiled against the reference headers with
// TIPL_USE_CUDA defined.  Linked with tipl_amd/plug/tipl_hip_plug.cpp into
// oracle/_ref/libtiplplug.so, so those calls reach the GPU through the plug,
// exactly as the reference's own host code would once the plug replaces its
// CUDA translation unit.
#include <algorithm>  // before the reference headers (mt.hpp uses std::clamp)
#include <cstdint>
#include <cstring>
#include <mutex>
#include <vector>

#include "prog.hpp"
#include "reg/cdm.hpp"
#include "reg/linear.hpp"

extern "C" void (*tipl_hip_plug_cost_observer)(const float* T, int cost_kind, double cost);

namespace {

std::mutex g_log_mu;
std::vector<float> g_log_T;
std::vector<double> g_log_cost;
size_t g_log_cap = 0;
int64_t g_calls = 0;

void observe(const float* T, int, double c)
{
    std::lock_guard<std::mutex> lk(g_log_mu);
    ++g_calls;
    if (g_log_cost.size() < g_log_cap) {
        g_log_T.insert(g_log_T.end(), T, T + 12);
        g_log_cost.push_back(c);
    }
}

template <int dim>
tipl::shape<dim> shp(const int* d)
{
    if constexpr (dim == 3)
        return tipl::shape<3>(d[0], d[1], d[2]);
    else
        return tipl::shape<2>(d[0], d[1]);
}

}  // namespace

extern "C" {

// tipl::reg::linear<tipl::out> on one uint8 pair (unit voxel sizes) with the
// GPU selected (param.cuda).  narrow = 1 uses narrow_bound (the mr = false
// plug), 0 the default bounds (mr = true).  Logs up to log_cap (transform,
// cost) pairs the plug returned; returns the number of plug calls.
int64_t plug_linear(const uint8_t* from, const int* fdim, const uint8_t* to, const int* tdim,
                    int mutual_info, int narrow, float* arg_io, float* cost_out, int log_cap,
                    float* log_T, double* log_cost, int* n_log)
{
    {
        std::lock_guard<std::mutex> lk(g_log_mu);
        g_log_T.clear();
        g_log_cost.clear();
        g_log_cap = size_t(log_cap);
        g_calls = 0;
    }
    tipl_hip_plug_cost_observer = observe;
    tipl::const_pointer_image<3, unsigned char> f(from, shp<3>(fdim)), t(to, shp<3>(tdim));
    tipl::reg::linear_reg_param param;
    param.cost_type = mutual_info ? tipl::reg::mutual_info : tipl::reg::corr;
    param.cuda = true;
    if (narrow) param.bound = tipl::reg::narrow_bound;
    tipl::affine_param<float, 3> arg;
    std::memcpy(arg.data(), arg_io, sizeof(float) * 12);
    bool terminated = false;
    *cost_out = tipl::reg::linear<tipl::out>({f}, tipl::vector<3>(1, 1, 1), {t},
                                             tipl::vector<3>(1, 1, 1), arg, param, terminated);
    std::memcpy(arg_io, arg.data(), sizeof(float) * 12);
    tipl_hip_plug_cost_observer = nullptr;
    std::lock_guard<std::mutex> lk(g_log_mu);
    std::memcpy(log_T, g_log_T.data(), sizeof(float) * g_log_T.size());
    std::memcpy(log_cost, g_log_cost.data(), sizeof(double) * g_log_cost.size());
    *n_log = int(g_log_cost.size());
    return g_calls;
}

// tipl::reg::cdm_common<tipl::out> (cuda = true) on one uint8 pair of 2 or 3
// dimensions; field receives vox x dim floats.  Returns cdm_common's result.
int plug_cdm(const uint8_t* It, const uint8_t* Is, const int* dim, int ndim, float resolution,
             float speed, float smoothing, int min_dimension, float* field)
{
    tipl::reg::cdm_param param;
    param.resolution = resolution;
    param.speed = speed;
    param.smoothing = smoothing;
    param.min_dimension = unsigned(min_dimension);
    bool terminated = false;
    if (ndim == 3) {
        tipl::image<3, tipl::vector<3>> d;
        using list = std::vector<tipl::const_pointer_image<3, unsigned char>>;
        const bool ok = tipl::reg::cdm_common<tipl::out>(
            list{tipl::const_pointer_image<3, unsigned char>(It, shp<3>(dim))},
            list{tipl::const_pointer_image<3, unsigned char>(Is, shp<3>(dim))}, d, terminated, param,
            true);
        if (ok) std::memcpy(field, &d[0][0], sizeof(float) * 3 * d.size());
        return ok ? 1 : 0;
    }
    tipl::image<2, tipl::vector<2>> d;
    using list = std::vector<tipl::const_pointer_image<2, unsigned char>>;
    const bool ok = tipl::reg::cdm_common<tipl::out>(
        list{tipl::const_pointer_image<2, unsigned char>(It, shp<2>(dim))},
        list{tipl::const_pointer_image<2, unsigned char>(Is, shp<2>(dim))}, d, terminated, param,
        true);
    if (ok) std::memcpy(field, &d[0][0], sizeof(float) * 2 * d.size());
    return ok ? 1 : 0;
}

// The reference's CPU cdm() on a 2-D pair, for comparison with plug_cdm.
int ref_cdm2(const uint8_t* It, const uint8_t* Is, const int* dim, float* field)
{
    tipl::reg::cdm_param param;
    bool terminated = false;
    tipl::image<2, tipl::vector<2>> d;
    const bool ok = tipl::reg::cdm<void>(
        std::vector<tipl::const_pointer_image<2, unsigned char>>{
            tipl::const_pointer_image<2, unsigned char>(It, shp<2>(dim))},
        std::vector<tipl::const_pointer_image<2, unsigned char>>{
            tipl::const_pointer_image<2, unsigned char>(Is, shp<2>(dim))},
        d, terminated, param);
    if (ok) std::memcpy(field, &d[0][0], sizeof(float) * 2 * d.size());
    return ok ? 1 : 0;
}

}  // extern "C"
