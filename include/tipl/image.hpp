// image.hpp -- the container types of TIPL's registration entry points.
//
// Mirrors the parts of the reference's containers that tipl::reg::linear and
// tipl::reg::cdm take and return (utility/shape.hpp, utility/pixel_index.hpp:
// 259-640, utility/basic_image.hpp:136-317, 616-693, 930-947): shape<2|3>,
// vector<2|3,T>, image<dim,T> (x fastest, std::vector storage),
// pointer_image / const_pointer_image (non-owning views) and make_shared.
#ifndef TIPL_HIP_IMAGE_HPP
#define TIPL_HIP_IMAGE_HPP

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <vector>

namespace tipl {

template <int dim>
class shape;

template <>
class shape<3> {
public:
    static constexpr int dimension = 3;
    unsigned int dim[3] = {0, 0, 0};
    shape() = default;
    shape(unsigned int w, unsigned int h, unsigned int d) : dim{w, h, d} {}
    template <class T>
    explicit shape(const T* p) : dim{unsigned(p[0]), unsigned(p[1]), unsigned(p[2])} {}
    unsigned int operator[](int i) const { return dim[i]; }
    unsigned int& operator[](int i) { return dim[i]; }
    unsigned int width() const { return dim[0]; }
    unsigned int height() const { return dim[1]; }
    unsigned int depth() const { return dim[2]; }
    size_t plane_size() const { return size_t(dim[0]) * dim[1]; }
    size_t size() const { return size_t(dim[0]) * dim[1] * dim[2]; }
    const unsigned int* begin() const { return dim; }
    const unsigned int* end() const { return dim + 3; }
    bool operator==(const shape& r) const { return std::equal(dim, dim + 3, r.dim); }
    bool operator!=(const shape& r) const { return !(*this == r); }
};

template <>
class shape<2> {
public:
    static constexpr int dimension = 2;
    unsigned int dim[2] = {0, 0};
    shape() = default;
    shape(unsigned int w, unsigned int h) : dim{w, h} {}
    template <class T>
    explicit shape(const T* p) : dim{unsigned(p[0]), unsigned(p[1])} {}
    unsigned int operator[](int i) const { return dim[i]; }
    unsigned int& operator[](int i) { return dim[i]; }
    unsigned int width() const { return dim[0]; }
    unsigned int height() const { return dim[1]; }
    size_t plane_size() const { return size_t(dim[0]) * dim[1]; }
    size_t size() const { return size_t(dim[0]) * dim[1]; }
    const unsigned int* begin() const { return dim; }
    const unsigned int* end() const { return dim + 2; }
    bool operator==(const shape& r) const { return std::equal(dim, dim + 2, r.dim); }
    bool operator!=(const shape& r) const { return !(*this == r); }
};

template <int dim, typename T = float>
class vector;

// tipl::vector<2,T>: two packed components (the 2-D displacement field pixel).
template <typename T>
class vector<2, T> {
public:
    static constexpr int dimension = 2;
    using value_type = T;
    T v[2] = {T(0), T(0)};
    vector() = default;
    vector(T x, T y) : v{x, y} {}
    template <class U>
    explicit vector(const U* p) : v{T(p[0]), T(p[1])} {}
    T operator[](int i) const { return v[i]; }
    T& operator[](int i) { return v[i]; }
    T* begin() { return v; }
    T* end() { return v + 2; }
    const T* begin() const { return v; }
    const T* end() const { return v + 2; }
    vector& operator*=(T s)
    {
        for (T& c : v) c *= s;
        return *this;
    }
    vector& operator+=(const vector& r)
    {
        for (int i = 0; i < 2; ++i) v[i] += r.v[i];
        return *this;
    }
    double length() const { return std::sqrt(v[0] * v[0] + v[1] * v[1]); }
};

// tipl::vector<3,T>: three packed components (the displacement field voxel).
template <typename T>
class vector<3, T> {
public:
    static constexpr int dimension = 3;
    using value_type = T;
    T v[3] = {T(0), T(0), T(0)};
    vector() = default;
    vector(T x, T y, T z) : v{x, y, z} {}
    template <class U>
    explicit vector(const U* p) : v{T(p[0]), T(p[1]), T(p[2])} {}
    T operator[](int i) const { return v[i]; }
    T& operator[](int i) { return v[i]; }
    T* begin() { return v; }
    T* end() { return v + 3; }
    const T* begin() const { return v; }
    const T* end() const { return v + 3; }
    vector& operator*=(T s)
    {
        for (T& c : v) c *= s;
        return *this;
    }
    vector& operator+=(const vector& r)
    {
        for (int i = 0; i < 3; ++i) v[i] += r.v[i];
        return *this;
    }
    double length() const { return std::sqrt(double(v[0] * v[0] + v[1] * v[1] + v[2] * v[2])); }
};

// a displacement field is handed to the library as a packed float array
static_assert(sizeof(vector<3, float>) == 3 * sizeof(float), "vector<3> must be three packed floats");
static_assert(sizeof(vector<2, float>) == 2 * sizeof(float), "vector<2> must be two packed floats");

// Owning x-fastest image (depth() exists for dim 3 only).
template <int dim, typename T>
class image {
public:
    static constexpr int dimension = dim;
    using value_type = T;
    using buffer_type = image<dim, T>;
    image() = default;
    explicit image(const tipl::shape<dim>& s) : sp_(s), data_(s.size()) {}
    image(const T* p, const tipl::shape<dim>& s) : sp_(s), data_(p, p + s.size()) {}
    const tipl::shape<dim>& shape() const { return sp_; }
    unsigned int width() const { return sp_.width(); }
    unsigned int height() const { return sp_.height(); }
    unsigned int depth() const { return sp_.depth(); }
    size_t plane_size() const { return sp_.plane_size(); }
    size_t size() const { return data_.size(); }
    bool empty() const { return data_.empty(); }
    void resize(const tipl::shape<dim>& s)
    {
        sp_ = s;
        data_.resize(s.size());
    }
    void clear()
    {
        data_.clear();
        sp_ = tipl::shape<dim>();
    }
    T* data() { return data_.data(); }
    const T* data() const { return data_.data(); }
    T& operator[](size_t i) { return data_[i]; }
    const T& operator[](size_t i) const { return data_[i]; }
    auto begin() { return data_.begin(); }
    auto end() { return data_.end(); }
    auto begin() const { return data_.begin(); }
    auto end() const { return data_.end(); }
    void swap(image& r)
    {
        std::swap(sp_, r.sp_);
        data_.swap(r.data_);
    }

private:
    tipl::shape<dim> sp_;
    std::vector<T> data_;
};

// Non-owning read-only view (const_pointer_image, utility/basic_image.hpp:666-693).
template <int dim, typename T>
class const_pointer_image {
public:
    static constexpr int dimension = dim;
    using value_type = T;
    using buffer_type = image<dim, T>;
    const_pointer_image() = default;
    const_pointer_image(const T* p, const tipl::shape<dim>& s) : p_(p), sp_(s) {}
    template <class I, class = decltype(std::declval<const I&>().data())>
    const_pointer_image(const I& img) : p_(img.data()), sp_(img.shape()) {}
    const tipl::shape<dim>& shape() const { return sp_; }
    unsigned int width() const { return sp_.width(); }
    unsigned int height() const { return sp_.height(); }
    unsigned int depth() const { return sp_.depth(); }
    size_t size() const { return sp_.size(); }
    bool empty() const { return sp_.size() == 0; }
    const T* data() const { return p_; }
    const T& operator[](size_t i) const { return p_[i]; }
    const T* begin() const { return p_; }
    const T* end() const { return p_ + sp_.size(); }

private:
    const T* p_ = nullptr;
    tipl::shape<dim> sp_;
};

// Non-owning writable view (pointer_image, utility/basic_image.hpp:616-656).
template <int dim, typename T>
class pointer_image {
public:
    static constexpr int dimension = dim;
    using value_type = T;
    using buffer_type = image<dim, T>;
    pointer_image() = default;
    pointer_image(T* p, const tipl::shape<dim>& s) : p_(p), sp_(s) {}
    template <class I, class = decltype(std::declval<I&>().data())>
    pointer_image(I& img) : p_(img.data()), sp_(img.shape()) {}
    const tipl::shape<dim>& shape() const { return sp_; }
    size_t size() const { return sp_.size(); }
    T* data() const { return p_; }
    T& operator[](size_t i) const { return p_[i]; }

private:
    T* p_ = nullptr;
    tipl::shape<dim> sp_;
};

template <int dim, typename T>
const_pointer_image<dim, T> make_shared(const image<dim, T>& I)
{
    return const_pointer_image<dim, T>(I.data(), I.shape());
}

template <int dim, typename T>
pointer_image<dim, T> make_shared(image<dim, T>& I)
{
    return pointer_image<dim, T>(I.data(), I.shape());
}

}  // namespace tipl

#endif
