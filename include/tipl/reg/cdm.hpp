// reg/cdm.hpp -- tipl::reg::cdm / cdm_common on the MI355X path.
//
// Same entry points and parameters as the reference's reg/cdm.hpp:36-44,
// 373-558; the whole multi-resolution loop runs in the library
// (tiplhip_cdm) with volumes and fields resident on the device.  As in the
// reference, a terminated run still returns its best field and reports
// success; a zero cost returns false.
#ifndef TIPL_HIP_REG_CDM_HPP
#define TIPL_HIP_REG_CDM_HPP

#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../def.hpp"
#include "../hip.hpp"
#include "../image.hpp"

namespace tipl {
namespace reg {

struct cdm_param {
    float resolution = 2.0f;
    float speed = 0.3f;
    float smoothing = 0.05f;
    unsigned int min_dimension = 16;
    enum { corr = 0, mi = 1 } cost_type = corr;  // unused by the reference loop as well
    std::string gradient_type;                   // per modality 'r' (default) or 'd'
    std::vector<std::pair<vector<3>, vector<3>>> anchor;  // not used by the loop
    int max_iterations = 200;                    // cdm() hard-codes 200 (reg/cdm.hpp:418)
    // fixed schedule, an extension (off by default): exactly `levels` pyramid
    // levels of `level_iterations` iterations each, the field smoothed by
    // filter::gaussian(field_bandwidth) when that is > 0
    int levels = 0;
    int level_iterations = 0;
    int field_bandwidth = 0;
};

namespace detail {

inline tiplhip_cdm_param to_c(const cdm_param& p)
{
    tiplhip_cdm_param c;
    tiplhip_cdm_param_default(&c);
    c.resolution = p.resolution;
    c.speed = p.speed;
    c.smoothing = p.smoothing;
    c.min_dimension = int(p.min_dimension);
    c.max_iterations = p.max_iterations;
    for (size_t i = 0; i < 8; ++i) c.gradient_type[i] = i < p.gradient_type.size() ? p.gradient_type[i] : 'r';
    c.levels = p.levels;
    c.level_iterations = p.level_iterations;
    c.field_bandwidth = p.field_bandwidth;
    return c;
}

}  // namespace detail

// tipl::reg::cdm (reg/cdm.hpp:373-474): It / Is are lists of modalities of
// one shape; dis receives the displacement field (x-fastest, packed
// vector<dim>).  Volumes go through tiplhip_cdm; 2-D uint8 images (the
// reference's cdm_cuda<out,uint8,2>, reg/cdm_cu.hpp:12-17) through
// tiplhip_cdm2d.
template <typename out_type = void, typename value_type, int dim>
bool cdm(const std::vector<const_pointer_image<dim, value_type>>& It,
         const std::vector<const_pointer_image<dim, value_type>>& Is,
         image<dim, vector<dim>>& dis, bool& terminated, cdm_param param = cdm_param())
{
    static_assert(dim == 3 || dim == 2, "cdm registers 2-D or 3-D images");
    const size_t n = std::min(It.size(), Is.size());
    if (n == 0) throw std::invalid_argument("cdm: empty image list");
    hip::context& ctx = hip::thread_context();
    if constexpr (dim == 2) {
        // the reference's 2-D GPU instantiation, cdm_cuda<out,uint8,2>
        // (reg/cdm_cu.hpp:12-17), takes uint8 images
        static_assert(std::is_same_v<value_type, unsigned char>, "2-D cdm takes uint8 images");
        std::vector<const uint8_t*> pt, ps;
        for (size_t i = 0; i < n; ++i) {
            pt.push_back(It[i].data());
            ps.push_back(Is[i].data());
        }
        const int d2[2] = {int(It[0].width()), int(It[0].height())};
        const tiplhip_cdm_param cp = detail::to_c(param);
        dis.resize(It[0].shape());
        static_assert(sizeof(bool) == 1, "the terminate flag is read as one byte");
        int iterations = 0;
        const int rc = tiplhip_cdm2d(ctx.get(), pt.data(), ps.data(), int(n), d2, &cp, &dis[0][0],
                                     reinterpret_cast<const volatile uint8_t*>(&terminated),
                                     &iterations);
        if constexpr (!std::is_void_v<out_type>)
            out_type() << "nonlinear registration using gpu: " << iterations << " iterations";
        if (rc == TIPLHIP_ERR_FAILED) return false;
        if (rc == TIPLHIP_ERR_TERMINATED) return true;
        hip::check(rc);
        return true;
    } else {
        std::vector<std::unique_ptr<hip::volume>> vt, vs;
        std::vector<const tiplhip_volume*> pt, ps;
        for (size_t i = 0; i < n; ++i) {
            if constexpr (std::is_same_v<value_type, unsigned char>) {
                vt.emplace_back(new hip::volume(ctx, It[i].data(), It[i].shape()));
                vs.emplace_back(new hip::volume(ctx, Is[i].data(), Is[i].shape()));
            } else {
                vt.emplace_back(new hip::volume(ctx, It[i].data(), It[i].shape(), 255.99f));
                vs.emplace_back(new hip::volume(ctx, Is[i].data(), Is[i].shape(), 255.99f));
            }
            pt.push_back(vt.back()->get());
            ps.push_back(vs.back()->get());
        }
        const tiplhip_cdm_param cp = detail::to_c(param);
        dis.resize(It[0].shape());
        // the flag is polled in place once per iteration (a bool is one byte)
        static_assert(sizeof(bool) == 1, "the terminate flag is read as one byte");
        const volatile uint8_t* stop = reinterpret_cast<const volatile uint8_t*>(&terminated);
        int iterations = 0;
        const int rc = tiplhip_cdm(ctx.get(), pt.data(), ps.data(), int(n), &cp, &dis[0][0], stop,
                                   &iterations);
        if constexpr (!std::is_void_v<out_type>)
            out_type() << "nonlinear registration using gpu: " << iterations << " iterations";
        if (rc == TIPLHIP_ERR_FAILED) return false;
        if (rc == TIPLHIP_ERR_TERMINATED) return true;
        hip::check(rc);
        return true;
    }
}

// tipl::reg::cdm_common (reg/cdm.hpp:534-558).
template <typename out_type = void, typename value_type, int dim>
bool cdm_common(std::vector<const_pointer_image<dim, value_type>> It,
                std::vector<const_pointer_image<dim, value_type>> Is, image<dim, vector<dim>>& dis,
                bool& terminated, cdm_param param = cdm_param(), bool cuda = true)
{
    if (It.size() < Is.size()) Is.resize(It.size());
    if (Is.size() < It.size()) It.resize(Is.size());
    if (!(cuda && tipl::has_gpu()))
        throw std::runtime_error("cdm_common: only the GPU path is provided");
    return cdm<out_type>(It, Is, dis, terminated, param);
}

}  // namespace reg
}  // namespace tipl

#endif
