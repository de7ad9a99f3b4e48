// filter/gaussian.hpp -- tipl::filter::gaussian on the MI355X path.
//
// The reference's separable binomial filter (filter/gaussian.hpp:18-105):
// `bandwidth` rounds of [1 2 1] along x, y, z with replicate edges, scaled by
// 2^(-6*bandwidth).  Float images and vector<3> fields run on the device
// (tiplhip_gaussian) with the reference's rounding, bit for bit.
#ifndef TIPL_HIP_FILTER_GAUSSIAN_HPP
#define TIPL_HIP_FILTER_GAUSSIAN_HPP

#include <type_traits>

#include "../hip.hpp"
#include "../image.hpp"

namespace tipl {
namespace filter {

template <typename image_type>
image_type& gaussian(image_type& src, unsigned int bandwidth = 1)
{
    using T = typename image_type::value_type;
    static_assert(image_type::dimension == 3, "the MI355X path filters 3-D images");
    static_assert(std::is_same_v<T, float> || std::is_same_v<T, tipl::vector<3, float>>,
                  "float images and vector<3> fields run on the device");
    if (src.empty() || !bandwidth) return src;
    const int d[3] = {int(src.width()), int(src.height()), int(src.depth())};
    const int ch = std::is_same_v<T, float> ? 1 : 3;
    hip::check(tiplhip_gaussian(hip::thread_context().get(), reinterpret_cast<float*>(src.data()), d,
                                ch, bandwidth));
    return src;
}

}  // namespace filter
}  // namespace tipl

#endif
