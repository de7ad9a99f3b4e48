// reg/mm_reg.hpp -- tipl::reg::mm_reg's registration flow on the MI355X path
// (reg/mm_reg.hpp of the reference).
//
// subject_image_pre: filter::gaussian (bandwidth 1), then
// segmentation::normalize_otsu_median(I, 255.99f), then uint8 -- on the
// device (tiplhip_volume_create_subject), bit-exact.
// template_image_pre: normalize_upper_lower2(I, out, 255.999f) on the device.
// mm_reg: the members and steps the registration uses -- linear_reg,
// calculate_linear_r / It_match_fov, nonlinear_reg (the reverse cdm on its
// own thread, hence its own context and stream), the mappings,
// calculate_nonlinear_r and apply_warping -- with the reference's names,
// defaults and order.  Image loading, space conversion and export are not
// part of this path.
#ifndef TIPL_HIP_REG_MM_REG_HPP
#define TIPL_HIP_REG_MM_REG_HPP

#include <cmath>
#include <memory>
#include <string>
#include <thread>
#include <vector>

#include "../hip.hpp"
#include "../image.hpp"
#include "../numerical/dif.hpp"
#include "../transformation.hpp"
#include "cdm.hpp"
#include "linear.hpp"

namespace tipl {
namespace reg {

inline image<3, unsigned char> subject_image_pre(const image<3, float>& I)
{
    hip::context& ctx = hip::thread_context();
    const int d[3] = {int(I.width()), int(I.height()), int(I.depth())};
    tiplhip_volume* v = nullptr;
    hip::check(tiplhip_volume_create_subject(ctx.get(), I.data(), TIPLHIP_F32, d, &v));
    return hip::volume(ctx, v).read();
}

inline image<3, unsigned char> template_image_pre(const image<3, float>& I)
{
    hip::context& ctx = hip::thread_context();
    return hip::volume(ctx, I.data(), I.shape(), 255.999f).read();
}

namespace detail {

// tipl::correlation / correlation_ygz of two uint8 images of one shape on
// the device (exact sums, FP64 tail).
inline double device_correlation(const image<3, unsigned char>& x, const image<3, unsigned char>& y,
                                 bool masked)
{
    hip::context& ctx = hip::thread_context();
    hip::volume vx(ctx, x.data(), x.shape()), vy(ctx, y.data(), y.shape());
    double r = 0.0;
    hip::check(tiplhip_correlation(ctx.get(), vx.get(), vy.get(), masked ? 1 : 0, &r));
    return r;
}

// tipl::resample(I, shape, T) of a uint8 image (with the integer-shift crop
// of the CPU path) on the device.
inline image<3, unsigned char> device_resample(const image<3, unsigned char>& I, const shape<3>& s,
                                               const transformation_matrix<float, 3>& T)
{
    hip::context& ctx = hip::thread_context();
    hip::volume src(ctx, I.data(), I.shape());
    const int d[3] = {int(s[0]), int(s[1]), int(s[2])};
    tiplhip_volume* out = nullptr;
    hip::check(tiplhip_volume_resample(ctx.get(), src.get(), d, T.data(), 1, &out));
    return hip::volume(ctx, out).read();
}

}  // namespace detail

template <typename out_type>
struct mm_reg {
    static constexpr int dimension = 3;
    static constexpr int max_modality = 8;  // the reference allows 12; the device cdm takes 8
    using image_type = image<dimension, unsigned char>;
    using mapping_type = image<dimension, vector<dimension>>;

    affine_param<float, dimension> arg;
    linear_reg_param linear_param = {affine, corr};
    cdm_param param;

    bool use_cuda = true;
    bool mm_linear = false;
    bool skip_linear = false;
    bool skip_nonlinear = false;
    bool match_fov = true;
    bool masked_r = true;
    bool skip_reverse = false;
    std::vector<image_type> I, J, It;
    std::vector<float> r;
    mapping_type t2f_dis, to2from, f2t_dis, from2to;
    vector<dimension> Itvs, Ivs;
    shape<dimension> Its, Is;
    mapping_type previous_t2f, previous_f2t;
    std::vector<image_type> previous_It;

    mm_reg() : I(max_modality), J(max_modality), It(max_modality), r(max_modality) {}

    void clear_reg()
    {
        J.assign(max_modality, image_type());
        r.assign(max_modality, 0.0f);
        t2f_dis.clear();
        to2from.clear();
        f2t_dis.clear();
        from2to.clear();
        arg = affine_param<float, dimension>();
    }

    transformation_matrix<float, dimension> T() const
    {
        return transformation_matrix<float, dimension>(arg, Its, Itvs, Is, Ivs);
    }
    transformation_matrix<float, dimension> invT() const
    {
        auto t = T();
        t.inverse();
        return t;
    }
    bool data_ready() const { return !I[0].empty() && !It[0].empty(); }

    // reg/mm_reg.hpp:271-284
    void compute_mapping_from_displacement()
    {
        auto trans = T();
        if (!f2t_dis.empty()) {
            from2to.resize(Is);
            inv_displacement_to_mapping(f2t_dis, from2to, trans);
        }
        if (!t2f_dis.empty()) {
            to2from.resize(Its);
            displacement_to_mapping(t2f_dis, to2from, trans);
        }
    }

    // reg/mm_reg.hpp:285-311
    void calculate_linear_r()
    {
        auto trans = T();
        std::fill(r.begin(), r.end(), 0.0f);
        J.resize(I.size());
        for (size_t i = 0; i < I.size(); ++i)
            if (!I[i].empty()) J[i] = detail::device_resample(I[i], Its, trans);
        for (size_t i = 0; i < I.size(); ++i) {
            if (J[i].empty() && It[i].empty()) continue;
            const auto& x = J[i].empty() ? J[0] : J[i];
            const auto& y = It[i].empty() ? It[0] : It[i];
            r[i] = float(detail::device_correlation(x, y, masked_r));
        }
        if (match_fov) It_match_fov();
    }

    // reg/mm_reg.hpp:313-328
    void It_match_fov()
    {
        auto trans = T();
        hip::context& ctx = hip::thread_context();
        const int d[3] = {int(Is[0]), int(Is[1]), int(Is[2])};
        for (size_t i = 0; i < It.size() && !It[i].empty(); ++i) {
            hip::volume v(ctx, It[i].data(), It[i].shape());
            hip::check(tiplhip_match_fov(ctx.get(), v.get(), trans.data(), d));
            hip::check(tiplhip_volume_read(ctx.get(), v.get(), It[i].data()));
        }
    }

    // reg/mm_reg.hpp:330-346
    void calculate_nonlinear_r()
    {
        std::fill(r.begin(), r.end(), 0.0f);
        if (to2from.empty()) return;
        for (size_t i = 0; i < I.size(); ++i) {
            if (I[i].empty() || It[i].empty()) continue;
            J[i] = compose_mapping(I[i], to2from);
            const auto& y = previous_It.empty() ? It[i] : previous_It[i];
            r[i] = float(detail::device_correlation(J[i], y, masked_r));
        }
    }

    // reg/mm_reg.hpp:348-368
    float linear_reg(bool& terminated)
    {
        if (!data_ready()) return 0.0f;
        float cost = 0.0f;
        linear_param.cuda = use_cuda;
        if (!skip_linear) {
            if (mm_linear)
                cost = linear<out_type>(make_list(It), Itvs, make_list(I), Ivs, arg, linear_param, terminated);
            else
                cost = linear<out_type>(make_list(It[0]), Itvs, make_list(I[0]), Ivs, arg, linear_param,
                                        terminated);
        }
        calculate_linear_r();
        return cost;
    }

    // reg/mm_reg.hpp:370-410: the reverse direction on its own thread (its
    // own thread context, hence its own stream) while the forward one runs
    void nonlinear_reg(bool& terminated)
    {
        if (!data_ready()) return;
        f2t_dis.clear();
        t2f_dis.clear();
        if (skip_nonlinear) {
            f2t_dis.resize(Its);
            t2f_dis.resize(Its);
            return;
        }
        auto param0 = param;
        auto param1 = param;
        std::unique_ptr<std::thread> t;
        if (!skip_reverse)
            t.reset(new std::thread([&] {
                cdm_common<out_type>(make_list(J), make_list(It), f2t_dis, terminated, param1, use_cuda);
            }));
        cdm_common<out_type>(make_list(It), make_list(J), t2f_dis, terminated, param0, use_cuda);
        if (t) t->join();
        if (!previous_f2t.empty() && !previous_t2f.empty()) {
            accumulate_displacement(previous_f2t, f2t_dis);
            accumulate_displacement(previous_t2f, t2f_dis);
        }
        compute_mapping_from_displacement();
        calculate_nonlinear_r();
    }

    // reg/mm_reg.hpp:442-450: direction true warps subject -> template.
    template <bool direction, interpolation itype = linear, typename T_>
    image<3, T_> apply_warping(const image<3, T_>& input) const
    {
        const auto& mapping = direction ? to2from : from2to;
        if (!mapping.empty()) return compose_mapping<itype>(input, mapping);
        const shape<3> out_shape = direction ? Its : Is;
        const auto trans = direction ? T() : invT();
        if constexpr (std::is_same_v<T_, unsigned char>) {
            return detail::device_resample(input, out_shape, trans);
        } else {
            // resample's integer-shift shortcut (numerical/resampling.hpp:
            // 1239-1262): an exact crop with zero fill
            const float* t = trans.data();
            int pos[3];
            bool crop = true;
            for (int i = 0; i < 3 && crop; ++i) {
                const long s = std::lround(t[9 + i]);
                crop = std::fabs(t[9 + i] - float(s)) <= 1.0e-4;
                pos[i] = int(s);
                for (int j = 0; j < 3; ++j)
                    crop = crop && std::fabs(t[3 * i + j] - (i == j ? 1.0f : 0.0f)) <= 1.0e-4;
            }
            if (crop) {
                image<3, T_> out(out_shape);
                const shape<3> is = input.shape();
                for (unsigned z = 0; z < out_shape[2]; ++z)
                    for (unsigned y = 0; y < out_shape[1]; ++y)
                        for (unsigned x = 0; x < out_shape[0]; ++x) {
                            const long sx = long(x) + pos[0], sy = long(y) + pos[1], sz = long(z) + pos[2];
                            if (sx < 0 || sy < 0 || sz < 0 || sx >= long(is[0]) || sy >= long(is[1]) ||
                                sz >= long(is[2]))
                                continue;
                            out[(size_t(z) * out_shape[1] + y) * out_shape[0] + x] =
                                input[(size_t(sz) * is[1] + size_t(sy)) * is[0] + size_t(sx)];
                        }
                return out;
            }
            // otherwise the identity displacement through T, then compose
            mapping_type m(out_shape);
            displacement_to_mapping(m, m, trans);
            return compose_mapping<itype>(input, m);
        }
    }
};

}  // namespace reg
}  // namespace tipl

#endif
