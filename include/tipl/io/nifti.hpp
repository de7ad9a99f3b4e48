// io/nifti.hpp -- tipl::io::nifti / gz_nifti on the MI355X path.
//
// The reading and writing interface of the reference's nifti_base
// (io/nifti.hpp:385-1350) that registration callers use -- open, `>>` of an
// image (the next volume, reoriented to LPS, converted, scaled), of the
// voxel size, shape and image transformation, is_mni, dim, select_volume,
// and `<<` of voxel size, transformation, MNI flag and image -- backed by
// the library's reader and writer (tiplhip_nifti_*), which read NIfTI-1,
// NIfTI-2 and Analyze files, either byte order, gzip-compressed or not.
#ifndef TIPL_HIP_IO_NIFTI_HPP
#define TIPL_HIP_IO_NIFTI_HPP

#include <algorithm>
#include <filesystem>
#include <ios>
#include <string>
#include <type_traits>
#include <vector>

#include "../../tiplhip.h"
#include "../image.hpp"
#include "../transformation.hpp"

namespace tipl {
namespace io {

class nifti {
public:
    mutable std::string error_msg;

    nifti() = default;
    template <class Mode>
    nifti(const std::filesystem::path& file_name, Mode mode)
    {
        open(file_name, mode);
    }

    template <class Mode>
    bool open(const std::filesystem::path& file_name, Mode mode)
    {
        path_ = file_name.string();
        volume_ = 0;
        writing_ = (mode & std::ios::out) != 0;
        error_msg.clear();
        if (writing_) return true;
        if (tiplhip_nifti_read_info(path_.c_str(), &info_) != TIPLHIP_OK)
            return error_msg = tiplhip_last_error(), false;
        return true;
    }
    explicit operator bool() const { return error_msg.empty(); }
    bool operator!() const { return !error_msg.empty(); }

    // header queries (after the LPS reorientation, as the reference reports them)
    unsigned short dim(unsigned int i) const { return (unsigned short)info_.dim[i]; }
    unsigned short width() const { return dim(1); }
    unsigned short height() const { return dim(2); }
    unsigned short depth() const { return dim(3); }
    bool is_mni() const { return info_.is_mni != 0; }
    bool is_integer() const { return info_.datatype != 16 && info_.datatype != 64; }
    bool is_int8() const { return info_.datatype == 2; }
    const tiplhip_nifti_info& info() const { return info_; }
    bool select_volume(size_t i)
    {
        volume_ = i;
        return true;
    }
    void get_image_transformation(matrix<4, 4, float>& R) const
    {
        R.identity();
        std::copy_n(info_.srow, 12, R.begin());
    }

    // `>>`: an image reads the next volume (uint8, int16 or float voxels)
    template <class T>
    nifti& operator>>(image<3, T>& I)
    {
        static_assert(std::is_same_v<T, unsigned char> || std::is_same_v<T, short> ||
                          std::is_same_v<T, float>,
                      "nifti reads uint8, int16 or float images");
        if (!error_msg.empty()) return *this;
        I.resize(shape<3>(unsigned(info_.dim[1]), unsigned(info_.dim[2]), unsigned(info_.dim[3])));
        const int dt = std::is_same_v<T, float> ? TIPLHIP_F32 : std::is_same_v<T, short> ? TIPLHIP_I16 : TIPLHIP_U8;
        if (tiplhip_nifti_read(path_.c_str(), int(volume_), dt, I.data(), &info_) != TIPLHIP_OK)
            error_msg = tiplhip_last_error();
        ++volume_;
        return *this;
    }
    nifti& operator>>(vector<3>& vs)
    {
        vs = vector<3>(info_.pixdim[1], info_.pixdim[2], info_.pixdim[3]);
        return *this;
    }
    nifti& operator>>(shape<3>& s)
    {
        s = shape<3>(unsigned(info_.dim[1]), unsigned(info_.dim[2]), unsigned(info_.dim[3]));
        return *this;
    }
    nifti& operator>>(matrix<4, 4, float>& R)
    {
        get_image_transformation(R);
        return *this;
    }
    template <class T>
    T toImage()
    {
        T I;
        *this >> I;
        return I;
    }

    // `<<`: voxel size, transformation and MNI flag first, then the image writes the file
    nifti& operator<<(const vector<3>& vs)
    {
        std::copy(vs.begin(), vs.end(), vs_);
        return *this;
    }
    nifti& operator<<(const matrix<4, 4, float>& R)
    {
        std::copy_n(R.begin(), 12, srow_);
        has_R_ = true;
        return *this;
    }
    nifti& operator<<(bool mni)
    {
        mni_ = mni;
        return *this;
    }
    template <class T>
    nifti& operator<<(const image<3, T>& I)
    {
        static_assert(std::is_same_v<T, unsigned char> || std::is_same_v<T, short> ||
                          std::is_same_v<T, float>,
                      "nifti writes uint8, int16 or float images");
        if (!error_msg.empty()) return *this;
        if (!writing_) return error_msg = "file not opened for writing", *this;
        const int d[3] = {int(I.width()), int(I.height()), int(I.depth())};
        float srow[12] = {vs_[0], 0, 0, 0, 0, vs_[1], 0, 0, 0, 0, vs_[2], 0};
        if (has_R_) std::copy_n(srow_, 12, srow);
        const int dt = std::is_same_v<T, float> ? TIPLHIP_F32 : std::is_same_v<T, short> ? TIPLHIP_I16 : TIPLHIP_U8;
        if (tiplhip_nifti_write(path_.c_str(), I.data(), dt, d, vs_, srow, mni_ ? 1 : 0) != TIPLHIP_OK)
            error_msg = tiplhip_last_error();
        return *this;
    }

private:
    std::string path_;
    tiplhip_nifti_info info_{};
    size_t volume_ = 0;
    bool writing_ = false;
    float vs_[3] = {1.0f, 1.0f, 1.0f};
    float srow_[12] = {};
    bool has_R_ = false;
    bool mni_ = false;
};

// the reference's gzip-capable alias; the library handles .gz either way
using gz_nifti = nifti;

}  // namespace io
}  // namespace tipl

#endif
