// reg_cli.cpp -- drives the header-only C++ entry points (include/tipl) the
// way a TIPL user would, for the GPU tests in tests/test_gpu_cpp_api.py.
//
//   reg_cli linear <from.u8> <to.u8> <w> <h> <d> <mi|corr> <out.txt>
//       tipl::reg::linear on from -> to (voxel size 1), writes the 12 affine
//       parameters and the final cost.
//   reg_cli cdm <It.u8> <Is.u8> <w> <h> <d> <out.f32>
//       tipl::reg::cdm_common, writes the displacement field (w*h*d*3 floats).
//   reg_cli cdm2d <It.u8> <Is.u8> <w> <h> <out.f32>
//       tipl::reg::cdm_common on 2-D images, writes the field (w*h*2 floats).
//   reg_cli gaussian <in.f32> <w> <h> <d> <channels> <bandwidth> <out.f32>
//       tipl::filter::gaussian on a float image (1) or vector<3> field (3).
//   reg_cli subject <in.f32> <w> <h> <d> <out.u8>
//       tipl::reg::subject_image_pre.
//   reg_cli mmreg <It.u8> <tw> <th> <td> <I.u8> <sw> <sh> <sd> <param.f32> <out>
//       tipl::reg::mm_reg with skip_linear: param holds arg[12], Itvs[3],
//       Ivs[3], then levels, level_iterations, smoothing, field_bandwidth of
//       the fixed-schedule cdm.  Writes <out>.r (r after the linear and the
//       nonlinear step), .t2f, .f2t, .to2from, .from2to (float3 fields),
//       .JJ (the warped subject) and .jac (jacobian_determinant_dis of t2f).
//   reg_cli nifti <in.nii[.gz]> <out>
//       tipl::io::gz_nifti: reads the first volume as float (`>>` image, voxel
//       size, shape, transformation), writes <out>.f32 and <out>.txt (w h d,
//       vs, 12 srow values, is_mni), then writes the image back with `<<` as
//       <out>.nii.gz.
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "tipl/tipl.hpp"

static std::vector<unsigned char> read_u8(const char* path, size_t n)
{
    std::vector<unsigned char> v(n);
    std::ifstream f(path, std::ios::binary);
    if (!f.read(reinterpret_cast<char*>(v.data()), std::streamsize(n)))
        throw std::runtime_error(std::string("cannot read ") + path);
    return v;
}

template <class T>
static std::vector<T> read_raw(const char* path, size_t n)
{
    std::vector<T> v(n);
    std::ifstream f(path, std::ios::binary);
    if (!f.read(reinterpret_cast<char*>(v.data()), std::streamsize(n * sizeof(T))))
        throw std::runtime_error(std::string("cannot read ") + path);
    return v;
}

template <class T>
static void write_raw(const char* path, const T* p, size_t n)
{
    std::ofstream f(path, std::ios::binary);
    f.write(reinterpret_cast<const char*>(p), std::streamsize(n * sizeof(T)));
}

int main(int argc, char** argv)
{
    try {
        if (argc >= 9 && std::string(argv[1]) == "gaussian") {
            const tipl::shape<3> s(unsigned(std::atoi(argv[3])), unsigned(std::atoi(argv[4])),
                                   unsigned(std::atoi(argv[5])));
            const int ch = std::atoi(argv[6]);
            const unsigned bw = unsigned(std::atoi(argv[7]));
            auto v = read_raw<float>(argv[2], s.size() * size_t(ch));
            if (ch == 3) {
                tipl::image<3, tipl::vector<3>> F(reinterpret_cast<const tipl::vector<3>*>(v.data()), s);
                tipl::filter::gaussian(F, bw);
                write_raw(argv[8], reinterpret_cast<const float*>(F.data()), 3 * s.size());
            } else {
                tipl::image<3, float> I(v.data(), s);
                tipl::filter::gaussian(I, bw);
                write_raw(argv[8], I.data(), s.size());
            }
            return 0;
        }
        if (argc >= 7 && std::string(argv[1]) == "subject") {
            const tipl::shape<3> s(unsigned(std::atoi(argv[3])), unsigned(std::atoi(argv[4])),
                                   unsigned(std::atoi(argv[5])));
            auto v = read_raw<float>(argv[2], s.size());
            tipl::image<3, float> I(v.data(), s);
            auto J = tipl::reg::subject_image_pre(I);
            write_raw(argv[6], J.data(), J.size());
            return 0;
        }
        if (argc >= 12 && std::string(argv[1]) == "mmreg") {
            const tipl::shape<3> ts(unsigned(std::atoi(argv[3])), unsigned(std::atoi(argv[4])),
                                    unsigned(std::atoi(argv[5])));
            const tipl::shape<3> ss(unsigned(std::atoi(argv[7])), unsigned(std::atoi(argv[8])),
                                    unsigned(std::atoi(argv[9])));
            auto It = read_u8(argv[2], ts.size());
            auto I = read_u8(argv[6], ss.size());
            auto p = read_raw<float>(argv[10], 22);
            tipl::reg::mm_reg<void> m;
            m.It[0] = tipl::image<3, unsigned char>(It.data(), ts);
            m.I[0] = tipl::image<3, unsigned char>(I.data(), ss);
            m.Its = ts;
            m.Is = ss;
            m.arg = tipl::affine_param<float, 3>(p.data());
            m.Itvs = tipl::vector<3>(p.data() + 12);
            m.Ivs = tipl::vector<3>(p.data() + 15);
            m.param.levels = int(p[18]);
            m.param.level_iterations = int(p[19]);
            m.param.smoothing = p[20];
            m.param.field_bandwidth = int(p[21]);
            m.skip_linear = true;
            bool terminated = false;
            m.linear_reg(terminated);
            float r[2] = {m.r[0], 0.0f};
            m.nonlinear_reg(terminated);
            r[1] = m.r[0];
            const std::string out = argv[11];
            write_raw((out + ".r").c_str(), r, 2);
            auto field = [&](const char* ext, const tipl::image<3, tipl::vector<3>>& f) {
                write_raw((out + ext).c_str(), reinterpret_cast<const float*>(f.data()), 3 * f.size());
            };
            field(".t2f", m.t2f_dis);
            field(".f2t", m.f2t_dis);
            field(".to2from", m.to2from);
            field(".from2to", m.from2to);
            write_raw((out + ".JJ").c_str(), m.J[0].data(), m.J[0].size());
            tipl::image<3, float> jac;
            tipl::jacobian_determinant_dis(m.t2f_dis, jac);
            write_raw((out + ".jac").c_str(), jac.data(), jac.size());
            return 0;
        }
        if (argc >= 4 && std::string(argv[1]) == "nifti") {
            tipl::io::gz_nifti in;
            if (!in.open(argv[2], std::ios::in)) throw std::runtime_error(in.error_msg);
            tipl::image<3, float> I;
            tipl::vector<3> vs;
            tipl::shape<3> s;
            tipl::matrix<4, 4> R;
            in >> I >> vs >> s >> R;
            if (!in) throw std::runtime_error(in.error_msg);
            const std::string out = argv[3];
            write_raw((out + ".f32").c_str(), I.data(), I.size());
            std::ofstream txt(out + ".txt");
            txt.precision(9);
            txt << s[0] << " " << s[1] << " " << s[2] << "\n" << vs[0] << " " << vs[1] << " " << vs[2] << "\n";
            for (int i = 0; i < 12; ++i) txt << R[unsigned(i)] << " ";
            txt << "\n" << (in.is_mni() ? 1 : 0) << "\n";
            tipl::io::gz_nifti o;
            o.open(out + ".nii.gz", std::ios::out);
            o << vs << R << I;
            if (!o) throw std::runtime_error(o.error_msg);
            return 0;
        }
        if (argc >= 7 && std::string(argv[1]) == "cdm2d") {
            const tipl::shape<2> s2(unsigned(std::atoi(argv[4])), unsigned(std::atoi(argv[5])));
            auto a = read_u8(argv[2], s2.size());
            auto b = read_u8(argv[3], s2.size());
            tipl::image<2, tipl::vector<2>> dis;
            bool terminated = false;
            const bool ok = tipl::reg::cdm_common<void>(
                std::vector<tipl::const_pointer_image<2, unsigned char>>{{a.data(), s2}},
                std::vector<tipl::const_pointer_image<2, unsigned char>>{{b.data(), s2}}, dis,
                terminated);
            std::ofstream out(argv[6], std::ios::binary);
            out.write(reinterpret_cast<const char*>(&dis[0][0]),
                      std::streamsize(sizeof(float) * 2 * dis.size()));
            return ok ? 0 : 1;
        }
        if (argc < 8) {
            std::cerr << "usage: reg_cli linear|cdm ...\n";
            return 2;
        }
        const std::string mode = argv[1];
        const tipl::shape<3> s(unsigned(std::atoi(argv[4])), unsigned(std::atoi(argv[5])),
                               unsigned(std::atoi(argv[6])));
        auto a = read_u8(argv[2], s.size());
        auto b = read_u8(argv[3], s.size());
        tipl::const_pointer_image<3, unsigned char> A(a.data(), s), B(b.data(), s);
        if (mode == "linear" && argc >= 9) {
            tipl::reg::linear_reg_param p;
            p.cost_type = std::string(argv[7]) == "corr" ? tipl::reg::corr : tipl::reg::mutual_info;
            tipl::affine_param<float, 3> arg;
            bool terminated = false;
            const float cost = tipl::reg::linear<void>(
                std::vector<tipl::const_pointer_image<3, unsigned char>>{A}, tipl::vector<3>(1, 1, 1),
                std::vector<tipl::const_pointer_image<3, unsigned char>>{B}, tipl::vector<3>(1, 1, 1),
                arg, p, terminated);
            std::ofstream out(argv[8]);
            out.precision(9);
            for (float v : arg) out << v << " ";
            out << "\n" << cost << "\n";
            return 0;
        }
        if (mode == "cdm") {
            tipl::image<3, tipl::vector<3>> dis;
            bool terminated = false;
            const bool ok = tipl::reg::cdm_common<void>(
                std::vector<tipl::const_pointer_image<3, unsigned char>>{A},
                std::vector<tipl::const_pointer_image<3, unsigned char>>{B}, dis, terminated);
            std::ofstream out(argv[7], std::ios::binary);
            out.write(reinterpret_cast<const char*>(&dis[0][0]),
                      std::streamsize(sizeof(float) * 3 * dis.size()));
            return ok ? 0 : 1;
        }
        std::cerr << "unknown mode\n";
        return 2;
    } catch (const std::exception& e) {
        std::cerr << "error: " << e.what() << "\n";
        return 3;
    }
}
