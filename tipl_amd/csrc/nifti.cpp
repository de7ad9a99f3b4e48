// nifti.cpp -- NIfTI-1 / NIfTI-2 / Analyze reading and writing on the host,
// so that real image files can be handed to the registration path.
//
// Follows tipl::io::nifti_base (io/nifti.hpp:385-1350) for what a caller sees:
//   * the header: NIfTI-2 fields narrowed into a NIfTI-1 header
//     (:619-687), big-endian files byte-swapped (:472-593), `n+1` / `n+2`
//     single files read from vox_offset, `ni1` and Analyze pairs from the
//     `.img` file (:690-720);
//   * handle_qform (:1219-1255): with qform_code > 0 and sform_code == 0 the
//     srows are built from the quaternion in float;
//   * toLPS (:132-160): the swap / flip sequence that turns the stored axes
//     into LPS, applied to dim, pixdim and srow here and to the voxels by
//     tipl::flip (numerical/basic_op.hpp:1001-1204);
//   * save_to_buffer (:1037-1122): the file's voxel type converted to the
//     output type by value conversion (raw copy when the types are
//     compatible), then scl_slope / scl_inter (:1136-1142) in the output type;
//   * writing (:937-985): the default header of init_header (:835-857),
//     set_voxel_size, set_image_transformation, 4 zero bytes of extension,
//     then the voxels; `.gz` paths are gzip-compressed (io/gz_stream.hpp:535-590).
// Compressed files are recognised by the `.gz` extension or the gzip magic,
// as gz_istream::open does (io/gz_stream.hpp:321-347).
#include <zlib.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <fstream>
#include <string>
#include <type_traits>
#include <vector>

#include "internal.hpp"

namespace tiplhip {
namespace {

// NIfTI-1 header (348 bytes, natural alignment gives the standard offsets).
struct Nifti1 {
    int32_t sizeof_hdr;
    char data_type[10];
    char db_name[18];
    int32_t extents;
    int16_t session_error;
    char regular;
    char dim_info;
    int16_t dim[8];
    float intent_p1, intent_p2, intent_p3;
    int16_t intent_code, datatype, bitpix, slice_start;
    float pixdim[8];
    float vox_offset, scl_slope, scl_inter;
    int16_t slice_end;
    char slice_code, xyzt_units;
    float cal_max, cal_min, slice_duration, toffset;
    int32_t glmax, glmin;
    char descrip[80];
    char aux_file[24];
    int16_t qform_code, sform_code;
    float quatern_b, quatern_c, quatern_d, qoffset_x, qoffset_y, qoffset_z;
    float srow_x[4], srow_y[4], srow_z[4];
    char intent_name[16];
    char magic[4];
};
static_assert(sizeof(Nifti1) == 348, "NIfTI-1 header layout");

template <class T>
void swap_bytes(T& v)
{
    char* p = reinterpret_cast<char*>(&v);
    std::reverse(p, p + sizeof(T));
}
template <class T, size_t N>
void swap_bytes(T (&a)[N])
{
    for (auto& v : a) swap_bytes(v);
}

void swap_header(Nifti1& h)
{
    swap_bytes(h.sizeof_hdr);
    swap_bytes(h.extents);
    swap_bytes(h.session_error);
    swap_bytes(h.dim);
    swap_bytes(h.intent_p1);
    swap_bytes(h.intent_p2);
    swap_bytes(h.intent_p3);
    swap_bytes(h.intent_code);
    swap_bytes(h.datatype);
    swap_bytes(h.bitpix);
    swap_bytes(h.slice_start);
    swap_bytes(h.pixdim);
    swap_bytes(h.vox_offset);
    swap_bytes(h.scl_slope);
    swap_bytes(h.scl_inter);
    swap_bytes(h.slice_end);
    swap_bytes(h.cal_max);
    swap_bytes(h.cal_min);
    swap_bytes(h.slice_duration);
    swap_bytes(h.toffset);
    swap_bytes(h.glmax);
    swap_bytes(h.glmin);
    swap_bytes(h.qform_code);
    swap_bytes(h.sform_code);
    swap_bytes(h.quatern_b);
    swap_bytes(h.quatern_c);
    swap_bytes(h.quatern_d);
    swap_bytes(h.qoffset_x);
    swap_bytes(h.qoffset_y);
    swap_bytes(h.qoffset_z);
    swap_bytes(h.srow_x);
    swap_bytes(h.srow_y);
    swap_bytes(h.srow_z);
}

bool ends_with(const std::string& s, const char* suffix)
{
    const size_t n = std::strlen(suffix);
    return s.size() >= n && s.compare(s.size() - n, n, suffix) == 0;
}

// The whole (decompressed) content of a file.
std::vector<char> read_file(const std::string& path)
{
    std::ifstream in(path, std::ios::binary);
    TIPLHIP_REQUIRE(bool(in), "cannot open file: " + path);
    unsigned char magic[2] = {0, 0};
    in.read(reinterpret_cast<char*>(magic), 2);
    const bool gz = ends_with(path, ".gz") || (in.gcount() == 2 && magic[0] == 0x1F && magic[1] == 0x8B);
    std::vector<char> out;
    if (!gz) {
        in.seekg(0, std::ios::end);
        const std::streamoff n = in.tellg();
        in.seekg(0, std::ios::beg);
        out.resize(size_t(n));
        if (n > 0) in.read(out.data(), n);
        TIPLHIP_REQUIRE(bool(in), "I/O error reading " + path);
        return out;
    }
    in.close();
    gzFile f = gzopen(path.c_str(), "rb");
    TIPLHIP_REQUIRE(f != nullptr, "cannot open file: " + path);
    std::vector<char> buf(size_t(1) << 24);
    for (;;) {
        const int got = gzread(f, buf.data(), unsigned(buf.size()));
        if (got < 0) {
            gzclose(f);
            throw Error(TIPLHIP_ERR_INVALID, "gzip error reading " + path);
        }
        if (got == 0) break;
        out.insert(out.end(), buf.data(), buf.data() + got);
    }
    gzclose(f);
    return out;
}

struct Parsed {
    Nifti1 h;
    bool big_endian = false;
    int version = 1;          // 1, 2, or 0 for Analyze / ni1 pairs
    std::vector<char> file;   // file holding the voxels
    size_t data_begin = 0;    // where the stream stands after open
    short untouched[8];       // dims before the LPS sequence
    std::vector<char> seq;    // the LPS swap/flip sequence
};

// handle_qform (io/nifti.hpp:1219-1255), in float.
void handle_qform(Nifti1& h)
{
    if (!(h.qform_code > 0 && h.sform_code == 0)) return;
    const float b = h.quatern_b, c = h.quatern_c, d = h.quatern_d;
    const float b2 = b * b, c2 = c * c, d2 = d * d;
    const float sum = b2 + c2 + d2;
    const float a2 = (sum > 1.0f) ? 0.0f : 1.0f - sum;
    const float a = std::sqrt(a2);
    const float ab2 = 2.0f * a * b, ac2 = 2.0f * a * c, ad2 = 2.0f * a * d;
    const float bc2 = 2.0f * b * c, bd2 = 2.0f * b * d, cd2 = 2.0f * c * d;
    float qfac = h.pixdim[0];
    if (qfac == 0.0f) qfac = 1.0f;
    h.srow_x[0] = (a2 + b2 - c2 - d2) * h.pixdim[1];
    h.srow_x[1] = (bc2 - ad2) * h.pixdim[2];
    h.srow_x[2] = (bd2 + ac2) * h.pixdim[3] * qfac;
    h.srow_x[3] = h.qoffset_x;
    h.srow_y[0] = (bc2 + ad2) * h.pixdim[1];
    h.srow_y[1] = (a2 + c2 - b2 - d2) * h.pixdim[2];
    h.srow_y[2] = (cd2 - ab2) * h.pixdim[3] * qfac;
    h.srow_y[3] = h.qoffset_y;
    h.srow_z[0] = (bd2 - ac2) * h.pixdim[1];
    h.srow_z[1] = (cd2 + ab2) * h.pixdim[2];
    h.srow_z[2] = (a2 + d2 - c2 - b2) * h.pixdim[3] * qfac;
    h.srow_z[3] = h.qoffset_z;
    h.sform_code = h.qform_code;
}

// toLPS (io/nifti.hpp:132-160) on dim[1..3], pixdim[1..3] and the 12
// srow floats: 3 swaps axes 0,1; 5 swaps 0,2; 4 swaps 1,2; 0,1,2 flip.
std::vector<char> to_lps(int16_t* dim, float* pixdim, float* R)
{
    std::vector<char> seq;
    auto swp = [&](int a, int b, char op) {
        seq.push_back(op);
        for (int r : {0, 4, 8}) std::swap(R[r + a], R[r + b]);
        std::swap(pixdim[a], pixdim[b]);
        std::swap(dim[a], dim[b]);
    };
    auto flp = [&](int a, char op) {
        seq.push_back(op);
        for (int r : {0, 4, 8}) {
            R[r + 3] += R[r + a] * float(dim[a] - 1);
            R[r + a] = -R[r + a];
        }
    };
    if (std::fabs(R[5]) < std::fabs(R[1]) && std::fabs(R[9]) < std::fabs(R[1])) swp(0, 1, 3);
    if (std::fabs(R[6]) < std::fabs(R[2]) && std::fabs(R[10]) < std::fabs(R[2])) swp(0, 2, 5);
    if (std::fabs(R[2]) < std::fabs(R[6]) && std::fabs(R[10]) < std::fabs(R[6])) swp(1, 2, 4);
    if (R[0] > 0.0f) flp(0, 0);
    if (R[5] > 0.0f) flp(1, 1);
    if (R[10] < 0.0f) flp(2, 2);
    return seq;
}

template <class T>
T get(const std::vector<char>& b, size_t off)
{
    T v;
    std::memcpy(&v, b.data() + off, sizeof(T));
    return v;
}

std::string sibling(const std::string& path, const char* ext)
{
    std::string base = path;
    if (ends_with(base, ".gz")) base.resize(base.size() - 3);
    const size_t dot = base.find_last_of('.');
    const size_t slash = base.find_last_of('/');
    if (dot != std::string::npos && (slash == std::string::npos || dot > slash)) base.resize(dot);
    return base + ext;
}

Parsed parse(const std::string& path)
{
    Parsed p;
    std::vector<char> f = read_file(path);
    TIPLHIP_REQUIRE(f.size() >= 348, "invalid NIFTI format: file shorter than a header");
    int32_t sz = get<int32_t>(f, 0);
    if (sz != 540 && sz != 348) {
        swap_bytes(sz);
        TIPLHIP_REQUIRE(sz == 540 || sz == 348, "invalid NIFTI format: size of header is not 540 or 348");
        p.big_endian = true;
    }
    Nifti1& h = p.h;
    std::memset(&h, 0, sizeof(h));
    if (sz == 540) {
        TIPLHIP_REQUIRE(f.size() >= 540, "invalid NIFTI format: file shorter than a NIfTI-2 header");
        auto g16 = [&](size_t o) { int16_t v = get<int16_t>(f, o); if (p.big_endian) swap_bytes(v); return v; };
        auto g32 = [&](size_t o) { int32_t v = get<int32_t>(f, o); if (p.big_endian) swap_bytes(v); return v; };
        auto g64 = [&](size_t o) { int64_t v = get<int64_t>(f, o); if (p.big_endian) swap_bytes(v); return v; };
        auto gd = [&](size_t o) { double v = get<double>(f, o); if (p.big_endian) swap_bytes(v); return v; };
        TIPLHIP_REQUIRE(f[4] == 'n' && f[5] == '+' && f[6] == '2', "invalid NIFTI format: no NIFTI tag found.");
        // the reference's field-by-field narrowing into the NIfTI-1 header (io/nifti.hpp:629-681)
        h.sizeof_hdr = 348;
        h.datatype = g16(12);
        h.bitpix = g16(14);
        for (int i = 0; i < 8; ++i) h.dim[i] = int16_t(g64(16 + 8 * size_t(i)));
        h.intent_p1 = float(gd(80));
        h.intent_p2 = float(gd(88));
        h.intent_p3 = float(gd(96));
        for (int i = 0; i < 8; ++i) h.pixdim[i] = float(gd(104 + 8 * size_t(i)));
        const int64_t vox_offset = g64(168);
        h.vox_offset = float(vox_offset);
        h.scl_slope = float(gd(176));
        h.scl_inter = float(gd(184));
        h.cal_max = float(gd(192));
        h.cal_min = float(gd(200));
        h.slice_duration = float(gd(208));
        h.toffset = float(gd(216));
        h.slice_start = int16_t(g64(224));
        h.slice_end = int16_t(g64(232));
        std::memcpy(h.descrip, f.data() + 240, 80);
        std::memcpy(h.aux_file, f.data() + 320, 24);
        h.qform_code = int16_t(g32(344));
        h.sform_code = int16_t(g32(348));
        h.quatern_b = float(gd(352));
        h.quatern_c = float(gd(360));
        h.quatern_d = float(gd(368));
        h.qoffset_x = float(gd(376));
        h.qoffset_y = float(gd(384));
        h.qoffset_z = float(gd(392));
        for (int i = 0; i < 4; ++i) {
            h.srow_x[i] = float(gd(400 + 8 * size_t(i)));
            h.srow_y[i] = float(gd(432 + 8 * size_t(i)));
            h.srow_z[i] = float(gd(464 + 8 * size_t(i)));
        }
        h.slice_code = char(g32(496));
        h.xyzt_units = char(g32(500));
        h.intent_code = int16_t(g32(504));
        std::memcpy(h.intent_name, f.data() + 508, 16);
        h.dim_info = f[524];
        p.version = 2;
        TIPLHIP_REQUIRE(vox_offset >= 0, "invalid NIFTI format: negative vox_offset");
        p.data_begin = size_t(vox_offset);
        p.file.swap(f);
    } else {
        std::memcpy(&h, f.data(), 348);
        if (p.big_endian) swap_header(h);
        const bool is_nii = h.magic[0] == 'n' && (h.magic[2] == '1' || h.magic[2] == '2') &&
                            (h.magic[1] == 'i' || h.magic[1] == '+');
        if (is_nii && h.magic[1] == '+') {
            p.version = 1;
            TIPLHIP_REQUIRE(h.vox_offset >= 0.0f && h.vox_offset < 4.0e18f,
                            "invalid NIFTI format: vox_offset out of range");
            p.data_begin = size_t(h.vox_offset);
            p.file.swap(f);
        } else {
            // ni1 / Analyze: the voxels are in the .img file, read from its start
            p.version = 0;
            p.file = read_file(sibling(path, ".img"));
            p.data_begin = 0;
        }
    }
    handle_qform(h);
    std::copy(h.dim, h.dim + 8, p.untouched);
    p.seq = to_lps(h.dim + 1, h.pixdim + 1, h.srow_x);
    return p;
}

size_t nvox(const int16_t* d) { return size_t(uint16_t(d[1])) * size_t(uint16_t(d[2])) * size_t(uint16_t(d[3])); }

bool compatible(int a, int b)
{
    if (a == b) return true;
    auto pair = [&](int x, int y) { return (a == x && b == y) || (a == y && b == x); };
    return pair(2, 256) || pair(4, 512) || pair(8, 768) || pair(1024, 1280);
}

template <class Out, class In>
void convert(const char* src, size_t n, bool swap, Out* out)
{
    for (size_t i = 0; i < n; ++i) {
        In v;
        std::memcpy(&v, src + i * sizeof(In), sizeof(In));
        if (swap) swap_bytes(v);
        out[i] = Out(v);
    }
}

template <class Out>
void read_voxels(const Parsed& p, size_t volume, Out* out)
{
    const Nifti1& h = p.h;
    const size_t bpp = size_t(h.bitpix) / 8;
    TIPLHIP_REQUIRE(bpp != 0, "invalid bitpix value: " + std::to_string(h.bitpix));
    const size_t n = nvox(p.untouched);
    TIPLHIP_REQUIRE(n != 0, "invalid dim value");
    // select_volume (io/nifti.hpp:767-775) seeks from vox_offset for every
    // volume; the first volume is read where open() left the stream
    size_t bytes = 0, skip = 0, begin = p.data_begin;
    bool over = __builtin_mul_overflow(bpp, n, &bytes);
    if (volume != 0) {
        TIPLHIP_REQUIRE(h.vox_offset >= 0.0f && h.vox_offset < 4.0e18f, "vox_offset out of range");
        over = over || __builtin_mul_overflow(bytes, volume, &skip) ||
               __builtin_add_overflow(size_t(h.vox_offset), skip, &begin);
    }
    size_t end = 0;
    over = over || __builtin_add_overflow(begin, bytes, &end);
    TIPLHIP_REQUIRE(!over && end <= p.file.size(), "end of file reached. The file is smaller than expected.");
    const char* src = p.file.data() + begin;
    constexpr int kOut = std::is_same<Out, uint8_t>::value ? 2 : std::is_same<Out, int16_t>::value ? 4 : 16;
    if (compatible(kOut, h.datatype) && bpp == sizeof(Out)) {
        std::memcpy(out, src, n * sizeof(Out));
        if (p.big_endian && sizeof(Out) > 1)
            for (size_t i = 0; i < n; ++i) swap_bytes(out[i]);
        return;
    }
    const bool sw = p.big_endian;
    // every conversion below reads n elements of the datatype's size: that
    // size must be the checked bytes per voxel (a malformed header could
    // otherwise make it read past the file)
    const int dt = h.datatype;
    const size_t elem = dt == 2 || dt == 256 ? 1 : dt == 4 || dt == 512 ? 2 : dt == 8 || dt == 16 || dt == 768 ? 4
                        : dt == 64 || dt == 1024 || dt == 1280 ? 8 : 0;
    TIPLHIP_REQUIRE(elem == 0 || elem == bpp, "invalid NIFTI format: bitpix " + std::to_string(h.bitpix) +
                                                  " does not match datatype " + std::to_string(dt));
    switch (h.datatype) {
    case 2: convert<Out, uint8_t>(src, n, sw, out); break;
    case 4: convert<Out, int16_t>(src, n, sw, out); break;
    case 8: convert<Out, int32_t>(src, n, sw, out); break;
    case 16: convert<Out, float>(src, n, sw, out); break;
    case 64: convert<Out, double>(src, n, sw, out); break;
    case 256: convert<Out, int8_t>(src, n, sw, out); break;
    case 512: convert<Out, uint16_t>(src, n, sw, out); break;
    case 768: convert<Out, uint32_t>(src, n, sw, out); break;
    case 1024: convert<Out, int64_t>(src, n, sw, out); break;
    case 1280: convert<Out, uint64_t>(src, n, sw, out); break;
    default: throw Error(TIPLHIP_ERR_INVALID, "unsupported NIfTI datatype " + std::to_string(h.datatype));
    }
}

// tipl::flip / swap_xy / swap_yz / swap_xz on an x-fastest volume of dims d.
template <class T>
void apply_op(std::vector<T>& v, int* d, char op)
{
    const size_t w = size_t(d[0]), hh = size_t(d[1]), dd = size_t(d[2]);
    std::vector<T> o(v.size());
    auto at = [&](size_t x, size_t y, size_t z) { return (z * hh + y) * w + x; };
    switch (op) {
    case 0:
        for (size_t z = 0; z < dd; ++z)
            for (size_t y = 0; y < hh; ++y)
                for (size_t x = 0; x < w; ++x) o[at(x, y, z)] = v[at(w - 1 - x, y, z)];
        break;
    case 1:
        for (size_t z = 0; z < dd; ++z)
            for (size_t y = 0; y < hh; ++y)
                for (size_t x = 0; x < w; ++x) o[at(x, y, z)] = v[at(x, hh - 1 - y, z)];
        break;
    case 2:
        for (size_t z = 0; z < dd; ++z)
            for (size_t y = 0; y < hh; ++y)
                for (size_t x = 0; x < w; ++x) o[at(x, y, z)] = v[at(x, y, dd - 1 - z)];
        break;
    case 3:  // new (y, x, z): new width h
        for (size_t z = 0; z < dd; ++z)
            for (size_t y = 0; y < hh; ++y)
                for (size_t x = 0; x < w; ++x) o[(z * w + x) * hh + y] = v[at(x, y, z)];
        std::swap(d[0], d[1]);
        break;
    case 4:  // new (x, z, y)
        for (size_t z = 0; z < dd; ++z)
            for (size_t y = 0; y < hh; ++y)
                for (size_t x = 0; x < w; ++x) o[(y * dd + z) * w + x] = v[at(x, y, z)];
        std::swap(d[1], d[2]);
        break;
    case 5:  // new (z, y, x)
        for (size_t z = 0; z < dd; ++z)
            for (size_t y = 0; y < hh; ++y)
                for (size_t x = 0; x < w; ++x) o[(x * hh + y) * dd + z] = v[at(x, y, z)];
        std::swap(d[0], d[2]);
        break;
    }
    v.swap(o);
}

void fill_info(const Parsed& p, tiplhip_nifti_info* info)
{
    if (!info) return;
    std::memset(info, 0, sizeof(*info));
    const Nifti1& h = p.h;
    for (int i = 0; i < 8; ++i) {
        info->dim[i] = h.dim[i];
        info->pixdim[i] = h.pixdim[i];
    }
    std::copy(h.srow_x, h.srow_x + 4, info->srow);
    std::copy(h.srow_y, h.srow_y + 4, info->srow + 4);
    std::copy(h.srow_z, h.srow_z + 4, info->srow + 8);
    info->datatype = h.datatype;
    info->bitpix = h.bitpix;
    info->scl_slope = h.scl_slope;
    info->scl_inter = h.scl_inter;
    info->sform_code = h.sform_code;
    info->qform_code = h.qform_code;
    // is_mni (io/nifti.hpp:807-814)
    info->is_mni = h.sform_code >= 4 && h.srow_x[3] != 0.0f && h.srow_y[3] != 0.0f &&
                   h.srow_z[3] != 0.0f && h.srow_x[1] == 0.0f && h.srow_x[2] == 0.0f &&
                   h.srow_y[0] == 0.0f && h.srow_y[2] == 0.0f && h.srow_z[0] == 0.0f &&
                   h.srow_z[1] == 0.0f;
    info->version = p.version;
    info->big_endian = p.big_endian ? 1 : 0;
    info->flip_swap_n = int(p.seq.size());
    for (size_t i = 0; i < p.seq.size() && i < 6; ++i) info->flip_swap[i] = p.seq[i];
    std::memcpy(info->descrip, h.descrip, 80);
    info->descrip[79] = 0;
}

template <class Out>
void read_volume(const Parsed& p, size_t volume, Out* out)
{
    const size_t n = nvox(p.untouched);
    std::vector<Out> v(n);
    read_voxels(p, volume, v.data());
    const Nifti1& h = p.h;
    if (h.scl_slope != 0.0f) {
        if (h.scl_slope != 1.0f)
            for (auto& x : v) x *= h.scl_slope;
        if (h.scl_inter != 0.0f)
            for (auto& x : v) x += h.scl_inter;
    }
    int d[3] = {uint16_t(p.untouched[1]), uint16_t(p.untouched[2]), uint16_t(p.untouched[3])};
    for (char op : p.seq) apply_op(v, d, op);
    std::copy(v.begin(), v.end(), out);
}

void write_all(const std::string& path, const char* data, size_t bytes)
{
    if (ends_with(path, ".gz")) {
        gzFile f = gzopen(path.c_str(), "wb");
        TIPLHIP_REQUIRE(f != nullptr, "cannot write to " + path);
        size_t pos = 0;
        while (pos < bytes) {
            const unsigned chunk = unsigned(std::min<size_t>(bytes - pos, size_t(1) << 26));
            if (gzwrite(f, data + pos, chunk) <= 0) {
                gzclose(f);
                throw Error(TIPLHIP_ERR_INVALID, "insufficient disk space writing " + path);
            }
            pos += chunk;
        }
        TIPLHIP_REQUIRE(gzclose(f) == Z_OK, "cannot finish " + path);
        return;
    }
    std::ofstream out(path, std::ios::binary);
    TIPLHIP_REQUIRE(bool(out), "cannot write to " + path);
    out.write(data, std::streamsize(bytes));
    TIPLHIP_REQUIRE(bool(out), "insufficient disk space writing " + path);
}

}  // namespace
}  // namespace tiplhip

using namespace tiplhip;

extern "C" {

int tiplhip_nifti_read_info(const char* path, tiplhip_nifti_info* info)
{
    return guarded([&] {
        TIPLHIP_REQUIRE(path && info, "null argument");
        fill_info(parse(path), info);
    });
}

int tiplhip_nifti_read(const char* path, int volume, int dtype, void* out, tiplhip_nifti_info* info)
{
    return guarded([&] {
        TIPLHIP_REQUIRE(path && out, "null argument");
        TIPLHIP_REQUIRE(volume >= 0, "negative volume index");
        const Parsed p = parse(path);
        const int volumes = std::max<int>(1, p.untouched[4]);
        TIPLHIP_REQUIRE(p.untouched[0] < 4 || volume < volumes, "volume index beyond the file's volumes");
        if (dtype == TIPLHIP_U8)
            read_volume(p, size_t(volume), static_cast<uint8_t*>(out));
        else if (dtype == TIPLHIP_I16)
            read_volume(p, size_t(volume), static_cast<int16_t*>(out));
        else if (dtype == TIPLHIP_F32)
            read_volume(p, size_t(volume), static_cast<float*>(out));
        else
            throw Error(TIPLHIP_ERR_INVALID, "dtype must be U8, I16 or F32");
        fill_info(p, info);
    });
}

int tiplhip_nifti_write(const char* path, const void* data, int dtype, const int dim[3],
                        const float vs[3], const float srow[12], int is_mni)
{
    return guarded([&] {
        TIPLHIP_REQUIRE(path && data && dim && vs && srow, "null argument");
        TIPLHIP_REQUIRE(dim[0] > 0 && dim[1] > 0 && dim[2] > 0 && dim[0] < 32768 && dim[1] < 32768 &&
                            dim[2] < 32768,
                        "dims must be in [1, 32767]");
        // init_header (io/nifti.hpp:835-857)
        Nifti1 h;
        std::memset(&h, 0, sizeof(h));
        h.sizeof_hdr = 348;
        h.vox_offset = 352;
        h.scl_slope = 1.0f;
        h.sform_code = 1;
        h.quatern_c = 1;
        h.srow_x[0] = 1.0f;
        h.srow_y[1] = 1.0f;
        h.srow_z[2] = 1.0f;
        h.magic[0] = 'n';
        h.magic[1] = '+';
        h.magic[2] = '1';
        // set_voxel_size (:785-797) with (1,1,1) then with vs: pixdim[0] = 3
        h.pixdim[0] = 3.0f;
        for (int i = 1; i < 8; ++i) h.pixdim[i] = i <= 3 ? vs[i - 1] : 1.0f;
        // set_image_transformation (:799-806)
        h.sform_code = is_mni ? 4 : 1;
        h.qform_code = 0;
        std::copy(srow, srow + 4, h.srow_x);
        std::copy(srow + 4, srow + 8, h.srow_y);
        std::copy(srow + 8, srow + 12, h.srow_z);
        // write(image) (:947-985): type, set_dim, header, 4 zero bytes, voxels
        size_t es = 0;
        if (dtype == TIPLHIP_U8) {
            h.datatype = 2;
            h.bitpix = 8;
            es = 1;
        } else if (dtype == TIPLHIP_I16) {
            h.datatype = 4;
            h.bitpix = 16;
            es = 2;
        } else if (dtype == TIPLHIP_F32) {
            h.datatype = 16;
            h.bitpix = 32;
            es = 4;
        } else {
            throw Error(TIPLHIP_ERR_INVALID, "dtype must be U8, I16 or F32");
        }
        std::fill(h.dim, h.dim + 8, int16_t(1));
        h.dim[0] = 3;
        for (int i = 0; i < 3; ++i) h.dim[i + 1] = int16_t(dim[i]);
        const size_t bytes = es * vox(dim);
        std::vector<char> buf(352 + bytes, 0);
        std::memcpy(buf.data(), &h, 348);
        std::memcpy(buf.data() + 352, data, bytes);
        write_all(path, buf.data(), buf.size());
    });
}

}  // extern "C"
