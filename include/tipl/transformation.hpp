// transformation.hpp -- affine_param and transformation_matrix.
//
// Same layouts as the reference (numerical/transformation.hpp:421-437,
// 589-604): affine_param = translocation[3], rotation[3], scaling[3],
// affine[3]; transformation_matrix = sr[9] (row-major) then shift[3], mapping
// a voxel of the `from` grid to a position in the `to` volume.  The matrix
// built from parameters and its inverse come from the library
// (tiplhip_affine_matrix / tiplhip_matrix_inverse), which computes them in the
// reference's float operation order with FMA contraction off, so host and
// device agree with the reference bit for bit whatever flags this header is
// compiled with.
#ifndef TIPL_HIP_TRANSFORMATION_HPP
#define TIPL_HIP_TRANSFORMATION_HPP

#include <algorithm>
#include <cmath>
#include <initializer_list>
#include <ostream>
#include <stdexcept>
#include <vector>

#include "../tiplhip.h"
#include "image.hpp"

namespace tipl {

template <typename value_type_ = float, int dim = 3>
class affine_param;

template <>
class affine_param<float, 3> {
public:
    using value_type = float;
    static constexpr int dimension = 3;
    static constexpr int total_size = 12;
    union {
        struct {
            float translocation[3];
            float rotation[3];
            float scaling[3];
            float affine[3];
        };
        float data_[12];
    };
    affine_param() { clear(); }
    affine_param(std::initializer_list<float> v)
    {
        clear();
        size_t i = 0;
        for (float f : v)
            if (i < 12) data_[i++] = f;
    }
    template <class T>
    explicit affine_param(const T* p)
    {
        std::copy_n(p, 12, data_);
    }
    affine_param(const std::vector<float>& v) { std::copy_n(v.data(), 12, data_); }
    affine_param(const affine_param& r) { std::copy_n(r.data_, 12, data_); }
    affine_param& operator=(const affine_param& r)
    {
        std::copy_n(r.data_, 12, data_);
        return *this;
    }
    void clear()
    {
        std::fill_n(data_, 12, 0.0f);
        scaling[0] = scaling[1] = scaling[2] = 1.0f;
    }
    float operator[](unsigned i) const { return data_[i]; }
    float& operator[](unsigned i) { return data_[i]; }
    float* begin() { return data_; }
    float* end() { return data_ + 12; }
    const float* begin() const { return data_; }
    const float* end() const { return data_ + 12; }
    float* data() { return data_; }
    const float* data() const { return data_; }
    unsigned size() const { return 12; }
    bool operator==(const affine_param& r) const { return std::equal(data_, data_ + 12, r.data_); }
    bool operator!=(const affine_param& r) const { return !(*this == r); }
    friend std::ostream& operator<<(std::ostream& out, const affine_param& T)
    {
        out << "t:" << T.translocation[0] << " " << T.translocation[1] << " " << T.translocation[2]
            << " r:" << T.rotation[0] << " " << T.rotation[1] << " " << T.rotation[2]
            << " s:" << T.scaling[0] << " " << T.scaling[1] << " " << T.scaling[2]
            << " a:" << T.affine[0] << " " << T.affine[1] << " " << T.affine[2];
        return out;
    }
};

// tipl::matrix<4,4> (numerical/matrix.hpp): the 4x4 row-major image
// transformation of an image space (srow_x, srow_y, srow_z, 0 0 0 1).
template <int rows, int cols, typename T = float>
struct matrix {
    T data_[rows * cols] = {};
    void identity()
    {
        std::fill_n(data_, rows * cols, T(0));
        for (int i = 0; i < std::min(rows, cols); ++i) data_[i * cols + i] = T(1);
    }
    T* begin() { return data_; }
    T* end() { return data_ + rows * cols; }
    const T* begin() const { return data_; }
    const T* end() const { return data_ + rows * cols; }
    T operator[](unsigned i) const { return data_[i]; }
    T& operator[](unsigned i) { return data_[i]; }
    bool operator==(const matrix& r) const { return std::equal(begin(), end(), r.begin()); }
    bool operator!=(const matrix& r) const { return !(*this == r); }
};

template <typename value_type_ = float, int dim = 3>
struct transformation_matrix;

template <>
struct transformation_matrix<float, 3> {
    using value_type = float;
    static constexpr int dimension = 3;
    union {
        struct {
            float sr[9];
            float shift[3];
        };
        float data_[12];
    };
    transformation_matrix() { identity(); }
    transformation_matrix(const transformation_matrix& r) { std::copy_n(r.data_, 12, data_); }
    transformation_matrix& operator=(const transformation_matrix& r)
    {
        std::copy_n(r.data_, 12, data_);
        return *this;
    }
    template <class T>
    explicit transformation_matrix(const T* p)
    {
        std::copy_n(p, 12, data_);
    }
    // transformation_matrix(affine_param, from, from_vs, to, to_vs)
    // (numerical/transformation.hpp:612-666).
    transformation_matrix(const affine_param<float, 3>& arg, const shape<3>& from,
                          const vector<3>& from_vs, const shape<3>& to, const vector<3>& to_vs)
    {
        const int fd[3] = {int(from[0]), int(from[1]), int(from[2])};
        const int td[3] = {int(to[0]), int(to[1]), int(to[2])};
        if (tiplhip_affine_matrix(arg.data(), fd, from_vs.begin(), td, to_vs.begin(), data_) != 0)
            throw std::runtime_error(tiplhip_last_error());
    }
    void identity()
    {
        std::fill_n(data_, 12, 0.0f);
        sr[0] = sr[4] = sr[8] = 1.0f;
    }
    float* begin() { return data_; }
    float* end() { return data_ + 12; }
    const float* begin() const { return data_; }
    const float* end() const { return data_ + 12; }
    float* data() { return data_; }
    const float* data() const { return data_; }
    size_t size() const { return 12; }
    float operator[](unsigned i) const { return data_[i]; }
    float& operator[](unsigned i) { return data_[i]; }

    // transformation_matrix::inverse (numerical/transformation.hpp:820-832):
    // false and unchanged when singular.
    bool inverse()
    {
        float t[12];
        std::copy_n(data_, 12, t);
        if (tiplhip_matrix_inverse(t) != 0) return false;
        std::copy_n(t, 12, data_);
        return true;
    }

    // p' = sr p + shift with the reference's summation order.
    vector<3> operator()(const vector<3>& p) const
    {
        return vector<3>(p[0] * sr[0] + p[1] * sr[1] + p[2] * sr[2] + shift[0],
                         p[0] * sr[3] + p[1] * sr[4] + p[2] * sr[5] + shift[1],
                         p[0] * sr[6] + p[1] * sr[7] + p[2] * sr[8] + shift[2]);
    }

    // to_affine_param (numerical/transformation.hpp:670-715): undo the
    // voxel-size and centring, then split the 3x3 part as
    // (unit upper shear) x (scaling) x (rotation) by an RQ factorisation in
    // double and read the Euler angles back (rotation_matrix_to_angles,
    // :226-238).  Agrees with the reference to float rounding.
    affine_param<float, 3> to_affine_param(const shape<3>& from, const vector<3>& from_vs,
                                           const shape<3>& to, const vector<3>& to_vs) const
    {
        double R[9], t[3];
        for (int i = 0; i < 9; ++i) R[i] = sr[i];
        for (int i = 0; i < 3; ++i) t[i] = double(shift[i]) - double(to[i]) * 0.5;
        for (int r = 0; r < 3; ++r) {
            for (int c = 0; c < 3; ++c) R[r * 3 + c] = R[r * 3 + c] * double(to_vs[r]) / double(from_vs[c]);
            t[r] *= double(to_vs[r]);
        }
        double iR[9];
        if (!inv3(R, iR)) return affine_param<float, 3>();
        affine_param<float, 3> rb;
        for (int r = 0; r < 3; ++r)
            rb.translocation[r] = float(iR[r * 3] * t[0] + iR[r * 3 + 1] * t[1] + iR[r * 3 + 2] * t[2] +
                                        double(from[r]) * 0.5 * double(from_vs[r]));
        // R = U Q with U upper triangular (positive diagonal), Q a rotation:
        // Gram-Schmidt on the rows of R from the last row upwards.
        double Q[9], U[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        for (int r = 2; r >= 0; --r) {
            double v[3] = {R[r * 3], R[r * 3 + 1], R[r * 3 + 2]};
            for (int k = r + 1; k < 3; ++k) {
                const double d = v[0] * Q[k * 3] + v[1] * Q[k * 3 + 1] + v[2] * Q[k * 3 + 2];
                U[r * 3 + k] = d;
                for (int c = 0; c < 3; ++c) v[c] -= d * Q[k * 3 + c];
            }
            const double n = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
            U[r * 3 + r] = n;
            for (int c = 0; c < 3; ++c) Q[r * 3 + c] = n > 0 ? v[c] / n : 0.0;
        }
        rb.scaling[0] = float(U[0]);
        rb.scaling[1] = float(U[4]);
        rb.scaling[2] = float(U[8]);
        rb.affine[0] = float(U[1] / U[4]);
        rb.affine[1] = float(U[2] / U[8]);
        rb.affine[2] = float(U[5] / U[8]);
        const double sy = std::sqrt(Q[0] * Q[0] + Q[1] * Q[1]);
        if (sy > 1.0e-6) {
            rb.rotation[0] = float(std::atan2(Q[5], Q[8]));
            rb.rotation[2] = float(std::atan2(Q[1], Q[0]));
        } else {
            rb.rotation[0] = float(std::atan2(-Q[7], Q[4]));
            rb.rotation[2] = 0.0f;
        }
        rb.rotation[1] = float(std::atan2(-Q[2], sy));
        return rb;
    }

private:
    static bool inv3(const double* a, double* o)
    {
        const double det = a[0] * (a[4] * a[8] - a[5] * a[7]) + a[1] * (a[5] * a[6] - a[3] * a[8]) +
                           a[2] * (a[3] * a[7] - a[4] * a[6]);
        if (det == 0.0) return false;
        o[0] = (a[4] * a[8] - a[5] * a[7]) / det;
        o[1] = (a[2] * a[7] - a[1] * a[8]) / det;
        o[2] = (a[1] * a[5] - a[2] * a[4]) / det;
        o[3] = (a[5] * a[6] - a[3] * a[8]) / det;
        o[4] = (a[0] * a[8] - a[2] * a[6]) / det;
        o[5] = (a[2] * a[3] - a[0] * a[5]) / det;
        o[6] = (a[3] * a[7] - a[4] * a[6]) / det;
        o[7] = (a[1] * a[6] - a[0] * a[7]) / det;
        o[8] = (a[0] * a[4] - a[1] * a[3]) / det;
        return true;
    }
};

// tipl::inverse(arg, from, from_vs, to, to_vs) (numerical/transformation.hpp:898-907).
inline void inverse(affine_param<float, 3>& arg, const shape<3>& from, const vector<3>& from_vs,
                    const shape<3>& to, const vector<3>& to_vs)
{
    transformation_matrix<float, 3> T(arg, from, from_vs, to, to_vs);
    T.inverse();
    arg = T.to_affine_param(to, to_vs, from, from_vs);
}

}  // namespace tipl

#endif
