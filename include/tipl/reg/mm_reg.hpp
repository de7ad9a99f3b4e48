// reg/mm_reg.hpp -- the image preprocessing of tipl::reg::mm_reg on the
// MI355X path (reg/mm_reg.hpp:14-40 of the reference).
//
// subject_image_pre: filter::gaussian (bandwidth 1), then
// segmentation::normalize_otsu_median(I, 255.99f), then uint8 -- on the
// device (tiplhip_volume_create_subject), bit-exact.
// template_image_pre: normalize_upper_lower2(I, out, 255.999f) on the device.
#ifndef TIPL_HIP_REG_MM_REG_HPP
#define TIPL_HIP_REG_MM_REG_HPP

#include "../hip.hpp"
#include "../image.hpp"

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

}  // namespace reg
}  // namespace tipl

#endif
