// numerical/dif.hpp -- displacement/mapping utilities on the MI355X path.
//
// Same names and argument meaning as the reference's numerical/dif.hpp
// (:9-86 mapping conversion and compose_mapping, :166-186 accumulate, :356-465
// Jacobian determinants), computed by libtiplhip.so (bit-exact, FP32 in the
// reference's operation order).  Fields and mappings are
// image<3, vector<3>> (packed float3, x fastest).  Only linear interpolation
// is provided, the reference's default.
#ifndef TIPL_HIP_NUMERICAL_DIF_HPP
#define TIPL_HIP_NUMERICAL_DIF_HPP

#include <type_traits>

#include "../hip.hpp"
#include "../image.hpp"
#include "../transformation.hpp"

namespace tipl {

enum interpolation { nearest, majority, linear, cubic, check };

using field3 = image<3, vector<3>>;

namespace detail {

inline void dims_of(const shape<3>& s, int* d)
{
    d[0] = int(s[0]);
    d[1] = int(s[1]);
    d[2] = int(s[2]);
}

// vector<3> is three packed floats, so a field is a float array
inline float* fdata(field3& f) { return f.empty() ? nullptr : reinterpret_cast<float*>(f.data()); }
inline const float* fdata(const field3& f)
{
    return f.empty() ? nullptr : reinterpret_cast<const float*>(f.data());
}

}  // namespace detail

// displacement_to_mapping(dis): in place, mapping(x) = x + dis(x).
inline void displacement_to_mapping(field3& dis)
{
    int d[3];
    detail::dims_of(dis.shape(), d);
    hip::check(tiplhip_displacement_to_mapping(hip::thread_context().get(), detail::fdata(dis), d,
                                               nullptr, detail::fdata(dis)));
}

// displacement_to_mapping(dis, mapping).
inline void displacement_to_mapping(const field3& dis, field3& mapping)
{
    mapping.resize(dis.shape());
    int d[3];
    detail::dims_of(dis.shape(), d);
    hip::check(tiplhip_displacement_to_mapping(hip::thread_context().get(), detail::fdata(dis), d,
                                               nullptr, detail::fdata(mapping)));
}

// displacement_to_mapping(dis, mapping, T): mapping(x) = T(x + dis(x)).
inline void displacement_to_mapping(const field3& dis, field3& mapping,
                                    const transformation_matrix<float, 3>& T)
{
    mapping.resize(dis.shape());
    int d[3];
    detail::dims_of(dis.shape(), d);
    hip::check(tiplhip_displacement_to_mapping(hip::thread_context().get(), detail::fdata(dis), d,
                                               T.data(), detail::fdata(mapping)));
}

// mapping_to_displacement(mapping): in place, dis(x) = mapping(x) - x.
inline void mapping_to_displacement(field3& mapping)
{
    int d[3];
    detail::dims_of(mapping.shape(), d);
    hip::check(tiplhip_mapping_to_displacement(hip::thread_context().get(), detail::fdata(mapping), d,
                                               detail::fdata(mapping)));
}

// inv_displacement_to_mapping(inv_dis, inv_mapping, T): over inv_mapping's
// shape (set by the caller, as in the reference), p = T^-1(x) plus inv_dis(p).
template <interpolation itype = linear>
void inv_displacement_to_mapping(const field3& inv_dis, field3& inv_mapping,
                                 const transformation_matrix<float, 3>& T)
{
    static_assert(itype == linear, "only linear interpolation is provided");
    int dd[3], od[3];
    detail::dims_of(inv_dis.shape(), dd);
    detail::dims_of(inv_mapping.shape(), od);
    hip::check(tiplhip_inv_displacement_to_mapping(hip::thread_context().get(), detail::fdata(inv_dis),
                                                   dd, T.data(), od, detail::fdata(inv_mapping)));
}

// compose_mapping(from, mapping, to): to(x) = from(mapping(x)), 0 outside.
template <interpolation itype = linear, typename T>
void compose_mapping(const image<3, T>& from, const field3& mapping, image<3, T>& to)
{
    static_assert(itype == linear, "only linear interpolation is provided");
    static_assert(std::is_same_v<T, unsigned char> || std::is_same_v<T, float>,
                  "compose_mapping takes uint8 or float images");
    to.resize(mapping.shape());
    int fd[3], md[3];
    detail::dims_of(from.shape(), fd);
    detail::dims_of(mapping.shape(), md);
    hip::check(tiplhip_compose_mapping(hip::thread_context().get(), from.data(),
                                       std::is_same_v<T, float> ? TIPLHIP_F32 : TIPLHIP_U8, fd,
                                       detail::fdata(mapping), md, to.data()));
}

template <interpolation itype = linear, typename T>
image<3, T> compose_mapping(const image<3, T>& from, const field3& mapping)
{
    image<3, T> to;
    compose_mapping<itype>(from, mapping, to);
    return to;
}

// accumulate_displacement(dis, new_dis) (numerical/dif.hpp:166-186):
// new(x) = m(x + new(x)) - x with m(y) = dis(y) + y; dis(x) where new(x) is 0
// or the sample is outside.
inline void accumulate_displacement(const field3& dis, field3& new_dis)
{
    int d[3];
    detail::dims_of(dis.shape(), d);
    hip::check(tiplhip_accumulate_displacement(hip::thread_context().get(), detail::fdata(dis),
                                               detail::fdata(new_dis), d, 1.0f));
}

// jacobian_determinant(src, dest) of a mapping; edge voxels are 0.
inline void jacobian_determinant(const field3& src, image<3, float>& dest)
{
    dest.resize(src.shape());
    int d[3];
    detail::dims_of(src.shape(), d);
    hip::check(tiplhip_jacobian_determinant(hip::thread_context().get(), detail::fdata(src), d,
                                            dest.data()));
}

inline image<3, float> jacobian_determinant(const field3& src)
{
    image<3, float> J;
    jacobian_determinant(src, J);
    return J;
}

// jacobian_determinant_dis(displacement, dest); edge voxels are 1.
inline void jacobian_determinant_dis(const field3& displacement, image<3, float>& dest)
{
    dest.resize(displacement.shape());
    int d[3];
    detail::dims_of(displacement.shape(), d);
    hip::check(tiplhip_jacobian_determinant_dis(hip::thread_context().get(), detail::fdata(displacement),
                                                d, dest.data()));
}

}  // namespace tipl

#endif
