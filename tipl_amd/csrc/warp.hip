// warp.hip -- post-registration warping and mm_reg's r on gfx950.
//
// Restates, MI355X-first:
//   displacement_to_mapping        numerical/dif.hpp:9-28, 41-50
//   mapping_to_displacement        numerical/dif.hpp:30-38
//   inv_displacement_to_mapping    numerical/dif.hpp:52-64
//   compose_mapping<linear>        numerical/dif.hpp:66-86
//   jacobian_determinant           numerical/dif.hpp:356-397
//   jacobian_determinant_dis<3>    numerical/dif.hpp:399-465
//   mm_reg::It_match_fov           reg/mm_reg.hpp:313-328
//   correlation / correlation_ygz  numerical/statistics.hpp:569-619
//
// Every kernel is one pass over the output grid with one thread per voxel on
// the 64 x 4 field tiles; all are HBM-bound gathers or stencils.  FP32
// operations are in the reference's order with contraction off, so results
// are bit-exact (the Jacobian of a displacement field keeps the reference's
// mixed float/double expression types).
#include "device.hpp"
#include "field.hpp"
#include "internal.hpp"
#include "warp.hpp"

#include <algorithm>
#include <cmath>
#include <vector>

namespace tiplhip {

namespace {

// transformation_matrix::operator() (numerical/transformation.hpp:834-850)
// with vector_transformation<3> (:58-73): ((x*r0 + y*r1) + z*r2) + s0.
__device__ inline float3 apply(const Affine& T, float x, float y, float z)
{
    return make_float3(x * T.m[0] + y * T.m[1] + z * T.m[2] + T.m[9],
                       x * T.m[3] + y * T.m[4] + z * T.m[5] + T.m[10],
                       x * T.m[6] + y * T.m[7] + z * T.m[8] + T.m[11]);
}

// mapping(x) = T(x + dis(x)); identity when has_T is 0.  out may alias dis.
__global__ __launch_bounds__(256) void mapping_kernel(const float* dis, FieldGeo g, Affine T,
                                                      int has_T, float* out)
{
    const Vox v = this_voxel(g);
    if (!v.ok) return;
    const float3 d = ld3(dis, v.p);
    float3 m = make_float3(d.x + float(v.x), d.y + float(v.y), d.z + float(v.z));
    if (has_T) m = apply(T, m.x, m.y, m.z);
    st3(out, v.p, m);
}

// dis(x) = mapping(x) - x.  out may alias mapping.
__global__ __launch_bounds__(256) void unmapping_kernel(const float* mapping, FieldGeo g,
                                                        float* out)
{
    const Vox v = this_voxel(g);
    if (!v.ok) return;
    const float3 m = ld3(mapping, v.p);
    st3(out, v.p, make_float3(m.x - float(v.x), m.y - float(v.y), m.z - float(v.z)));
}

// Trilinear sample of a float3 field (weighted_sum over vector<3>:
// r = f0*w0, then r += fk*wk in corner order); zero vector outside.
__device__ inline float3 sample_field(const float* f, const FieldGeo& g, float x, float y, float z)
{
    Tri t;
    float3 r = make_float3(0.0f, 0.0f, 0.0f);
    if (!locate(g, x, y, z, t)) return r;
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        const float3 a = ld3(f, corner(g, t.i0, c));
        if (c == 0) {
            r.x = a.x * t.w[0];
            r.y = a.y * t.w[0];
            r.z = a.z * t.w[0];
        } else {
            r.x = r.x + a.x * t.w[c];
            r.y = r.y + a.y * t.w[c];
            r.z = r.z + a.z * t.w[c];
        }
    }
    return r;
}

// inv_mapping(x) = p + inv_dis(p), p = iT(x), over the output grid go.
__global__ __launch_bounds__(256) void inv_mapping_kernel(const float* __restrict__ inv_dis,
                                                          FieldGeo gd, FieldGeo go, Affine iT,
                                                          float* __restrict__ out)
{
    const Vox v = this_voxel(go);
    if (!v.ok) return;
    const float3 p = apply(iT, float(v.x), float(v.y), float(v.z));
    const float3 d = sample_field(inv_dis, gd, p.x, p.y, p.z);
    st3(out, v.p, make_float3(p.x + d.x, p.y + d.y, p.z + d.z));
}

// to(x) = from(mapping(x)) for a uint8 image (interp_type<uint8> truncation).
__global__ __launch_bounds__(256) void compose_mapping_u8_kernel(const uint8_t* __restrict__ from,
                                                                 VolGeo gf, uint32_t from_n,
                                                                 const float* __restrict__ mapping,
                                                                 FieldGeo g,
                                                                 uint8_t* __restrict__ out)
{
    const Vox v = this_voxel(g);
    if (!v.ok) return;
    const float3 m = ld3(mapping, v.p);
    const __amdgpu_buffer_rsrc_t r = make_rsrc(from, from_n);
    out[v.p] = uint8_t(sample_u8(r, gf, m.x, m.y, m.z));
}

// The same for a float image (no truncation).
__global__ __launch_bounds__(256) void compose_mapping_f32_kernel(const float* __restrict__ from,
                                                                  FieldGeo gf,
                                                                  const float* __restrict__ mapping,
                                                                  FieldGeo g,
                                                                  float* __restrict__ out)
{
    const Vox v = this_voxel(g);
    if (!v.ok) return;
    const float3 m = ld3(mapping, v.p);
    Tri t;
    float r = 0.0f;
    if (locate(gf, m.x, m.y, m.z, t)) {
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            const float a = from[corner(gf, t.i0, c)];
            r = c == 0 ? a * t.w[0] : r + a * t.w[c];
        }
    }
    out[v.p] = r;
}

__device__ inline bool is_edge(const FieldGeo& g, const Vox& v)
{
    return v.x == 0 || v.y == 0 || v.z == 0 || v.x + 1 == g.w || v.y + 1 == g.h || v.z + 1 == g.d;
}

// jacobian_determinant_dis_at<3>: central differences (not halved), +1 on
// the diagonal.  Float differences; the diagonal terms are double, so every
// product with one of them is double while the others stay float.
__global__ __launch_bounds__(256) void jacobian_dis_kernel(const float* __restrict__ d, FieldGeo g,
                                                           float* __restrict__ out)
{
    const Vox v = this_voxel(g);
    if (!v.ok) return;
    if (is_edge(g, v)) {
        out[v.p] = 1.0f;
        return;
    }
    const uint32_t w = uint32_t(g.w);
    const float3 a0 = ld3(d, v.p + 1), a1 = ld3(d, v.p - 1);
    const float3 b0 = ld3(d, v.p + w), b1 = ld3(d, v.p - w);
    const float3 c0 = ld3(d, v.p + g.wh), c1 = ld3(d, v.p - g.wh);
    const float d2_0 = b0.x - b1.x;
    const double d2_1 = double(b0.y - b1.y) + 1.0;
    const float d2_2 = b0.z - b1.z;
    const float d3_0 = c0.x - c1.x;
    const float d3_1 = c0.y - c1.y;
    const double d3_2 = double(c0.z - c1.z) + 1.0;
    const double ax = double(a0.x - a1.x) + 1.0;
    const float ay = a0.y - a1.y;
    const float az = a0.z - a1.z;
    const float p22_31 = d2_2 * d3_1;
    const float p22_30 = d2_2 * d3_0;
    const float p20_31 = d2_0 * d3_1;
    const double r = ax * (d2_1 * d3_2 - double(p22_31)) +
                     double(ay) * (double(p22_30) - double(d2_0) * d3_2) +
                     double(az) * (double(p20_31) - d2_1 * double(d3_0));
    out[v.p] = float(r);
}

// jacobian_determinant of a mapping: float central differences; the
// reference's 0.125f applies to the third product only.  Edges are 0.
__global__ __launch_bounds__(256) void jacobian_kernel(const float* __restrict__ m, FieldGeo g,
                                                       float* __restrict__ out)
{
    const Vox v = this_voxel(g);
    if (!v.ok) return;
    if (is_edge(g, v)) {
        out[v.p] = 0.0f;
        return;
    }
    const uint32_t w = uint32_t(g.w);
    const float3 a0 = ld3(m, v.p + 1), a1 = ld3(m, v.p - 1);
    const float3 b0 = ld3(m, v.p + w), b1 = ld3(m, v.p - w);
    const float3 c0 = ld3(m, v.p + g.wh), c1 = ld3(m, v.p - g.wh);
    const float3 dx = make_float3(a0.x - a1.x, a0.y - a1.y, a0.z - a1.z);
    const float3 dy = make_float3(b0.x - b1.x, b0.y - b1.y, b0.z - b1.z);
    const float3 dz = make_float3(c0.x - c1.x, c0.y - c1.y, c0.z - c1.z);
    const float t0 = dx.x * (dy.y * dz.z - dy.z * dz.y);
    const float t1 = dx.y * (dy.z * dz.x - dy.x * dz.z);
    const float t2 = dx.z * (dy.x * dz.y - dy.y * dz.x) * 0.125f;
    out[v.p] = t0 + t1 + t2;
}

// It_match_fov: zero template voxels whose T(x) is outside the subject grid
// (shape::is_valid: 0 <= p < dim as float comparisons).
__global__ __launch_bounds__(256) void match_fov_kernel(uint8_t* __restrict__ It, FieldGeo g,
                                                        Affine T, float sw, float sh, float sd)
{
    const Vox v = this_voxel(g);
    if (!v.ok) return;
    const float3 p = apply(T, float(v.x), float(v.y), float(v.z));
    const bool valid = p.x >= 0.0f && p.y >= 0.0f && p.z >= 0.0f && p.x < sw && p.y < sh && p.z < sd;
    if (!valid) It[v.p] = 0;
}

// Exact correlation sums of two uint8 arrays: per workgroup {n, Sx, Sxx, Sy,
// Syy, Sxy} over the voxels with y > 0 (masked) or all voxels.  VEC reads 16
// voxels per thread per step (both arrays 16-byte aligned), with 32-bit sums
// per step and 64-bit sums per thread; the n % 16 tail and the unaligned
// case go voxel by voxel.
constexpr int kCorrThreads = 256;

__device__ inline void corr_add(uint64_t* s, uint32_t a, uint32_t b, int masked)
{
    if (masked && b == 0u) return;
    s[0] += 1;
    s[1] += a;
    s[2] += a * a;
    s[3] += b;
    s[4] += b * b;
    s[5] += a * b;
}

template <bool VEC>
__global__ __launch_bounds__(kCorrThreads) void corr_partial_kernel(
    const uint8_t* __restrict__ x, const uint8_t* __restrict__ y, uint32_t n, int masked,
    uint64_t* __restrict__ partials)
{
    uint64_t s[6] = {0, 0, 0, 0, 0, 0};
    const uint32_t stride = gridDim.x * kCorrThreads;
    const uint32_t first = blockIdx.x * kCorrThreads + threadIdx.x;
    uint32_t done = 0;
    if constexpr (VEC) {
        const uint32_t chunks = n / 16u;
        for (uint32_t c = first; c < chunks; c += stride) {
            const uint4 xv = reinterpret_cast<const uint4*>(x)[c];
            const uint4 yv = reinterpret_cast<const uint4*>(y)[c];
            const uint32_t xw[4] = {xv.x, xv.y, xv.z, xv.w};
            const uint32_t yw[4] = {yv.x, yv.y, yv.z, yv.w};
            uint32_t cnt = 0, sx = 0, sxx = 0, sy = 0, syy = 0, sxy = 0;
#pragma unroll
            for (int k = 0; k < 16; ++k) {
                const uint32_t a = (xw[k >> 2] >> (8 * (k & 3))) & 0xFFu;
                const uint32_t b = (yw[k >> 2] >> (8 * (k & 3))) & 0xFFu;
                const uint32_t on = (!masked || b != 0u) ? 1u : 0u;
                const uint32_t am = a * on, bm = b * on;
                cnt += on;
                sx += am;
                sxx += am * a;
                sy += bm;
                syy += bm * b;
                sxy += am * b;
            }
            s[0] += cnt;
            s[1] += sx;
            s[2] += sxx;
            s[3] += sy;
            s[4] += syy;
            s[5] += sxy;
        }
        done = chunks * 16u;
    }
    for (uint32_t i = done + first; i < n; i += stride) corr_add(s, x[i], y[i], masked);
    __shared__ uint64_t red[kCorrThreads / kWave][6];
    const int lane = int(threadIdx.x) & (kWave - 1), wv = int(threadIdx.x) / kWave;
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        const uint64_t t = wave_sum_u64(s[k]);
        if (lane == 0) red[wv][k] = t;
    }
    __syncthreads();
    if (threadIdx.x < 6) {
        uint64_t t = 0;
        for (int w = 0; w < kCorrThreads / kWave; ++w) t += red[w][threadIdx.x];
        partials[6 * size_t(blockIdx.x) + threadIdx.x] = t;
    }
}

}  // namespace

Affine to_affine(const float* T)
{
    Affine a;
    std::copy_n(T, 12, a.m);
    return a;
}

void launch_mapping(tiplhip_ctx* ctx, const float* dis, const int* dim, const float* T, float* out)
{
    const FieldGeo g = field_geo(dim);
    Affine a{};
    if (T) a = to_affine(T);
    hipLaunchKernelGGL(mapping_kernel, tile_grid(g), dim3(kTX, kTY), 0, ctx->stream, dis, g, a,
                       T ? 1 : 0, out);
    TIPLHIP_CHECK(hipGetLastError());
}

void launch_inv_mapping(tiplhip_ctx* ctx, const float* inv_dis, const int* dis_dim, const float* T,
                        const int* out_dim, float* out)
{
    float iT[12];
    std::copy_n(T, 12, iT);
    (void)matrix_inverse(iT);  // transformation_matrix::inverse's result is not checked
    const FieldGeo gd = field_geo(dis_dim), go = field_geo(out_dim);
    hipLaunchKernelGGL(inv_mapping_kernel, tile_grid(go), dim3(kTX, kTY), 0, ctx->stream, inv_dis,
                       gd, go, to_affine(iT), out);
    TIPLHIP_CHECK(hipGetLastError());
}

void launch_compose_mapping_u8(tiplhip_ctx* ctx, const uint8_t* from, const int* from_dim,
                               const float* mapping, const int* dim, uint8_t* out)
{
    const FieldGeo gf = field_geo(from_dim), g = field_geo(dim);
    const VolGeo vg{gf.w, gf.h, gf.d, gf.wh, gf.wm1, gf.hm1, gf.dm1};
    hipLaunchKernelGGL(compose_mapping_u8_kernel, tile_grid(g), dim3(kTX, kTY), 0, ctx->stream,
                       from, vg, gf.n, mapping, g, out);
    TIPLHIP_CHECK(hipGetLastError());
}

void launch_match_fov(tiplhip_ctx* ctx, uint8_t* It, const int* dim, const float* T,
                      const int* moving_dim)
{
    const FieldGeo g = field_geo(dim);
    hipLaunchKernelGGL(match_fov_kernel, tile_grid(g), dim3(kTX, kTY), 0, ctx->stream, It, g,
                       to_affine(T), float(unsigned(moving_dim[0])), float(unsigned(moving_dim[1])),
                       float(unsigned(moving_dim[2])));
    TIPLHIP_CHECK(hipGetLastError());
}

// tipl::correlation from the exact sums s = {n, Sx, Sxx, Sy, Syy, Sxy}
// (numerical/statistics.hpp:569-597, 405-443, 536-545); 0 for n = 0.
double corr_r_from_sums(const uint64_t* s)
{
    if (s[0] == 0) return 0.0;
    const double dn = double(s[0]);
    const double mean_x = double(int64_t(s[1])) / dn;
    const double mean_y = double(int64_t(s[3])) / dn;
    const double var_x = double(int64_t(s[2])) / dn - mean_x * mean_x;
    const double var_y = double(int64_t(s[4])) / dn - mean_y * mean_y;
    const double sd1 = var_x > 0.0 ? std::sqrt(var_x) : 0.0;
    const double sd2 = var_y > 0.0 ? std::sqrt(var_y) : 0.0;
    if (sd1 == 0.0 || sd2 == 0.0) return 0.0;
    const double cov = double(int64_t(s[5])) / dn - mean_x * mean_y;
    return cov / sd1 / sd2;
}

double device_correlation(tiplhip_ctx* ctx, const uint8_t* x, const uint8_t* y, size_t n,
                          bool masked)
{
    TIPLHIP_REQUIRE(n < (size_t(1) << 31), "volume exceeds 2^31 voxels");
    const uint32_t chunks = uint32_t(n / 16);
    const uint32_t blocks =
        std::max(1u, std::min(uint32_t(ctx->cu_count) * 4u, (chunks + kCorrThreads - 1) / kCorrThreads));
    DevMem part(ctx, sizeof(uint64_t) * 6 * blocks);
    std::vector<uint64_t> h(6 * size_t(blocks));
    const bool vec = ((reinterpret_cast<uintptr_t>(x) | reinterpret_cast<uintptr_t>(y)) & 15u) == 0;
    if (vec)
        hipLaunchKernelGGL(corr_partial_kernel<true>, dim3(blocks), dim3(kCorrThreads), 0,
                           ctx->stream, x, y, uint32_t(n), masked ? 1 : 0, part.as<uint64_t>());
    else
        hipLaunchKernelGGL(corr_partial_kernel<false>, dim3(blocks), dim3(kCorrThreads), 0,
                           ctx->stream, x, y, uint32_t(n), masked ? 1 : 0, part.as<uint64_t>());
    TIPLHIP_CHECK(hipGetLastError());
    read_back(ctx, part.p, h.data(), sizeof(uint64_t) * h.size());
    uint64_t s[6] = {0, 0, 0, 0, 0, 0};
    for (uint32_t b = 0; b < blocks; ++b)
        for (int k = 0; k < 6; ++k) s[k] += h[6 * size_t(b) + size_t(k)];
    return corr_r_from_sums(s);
}

}  // namespace tiplhip

using namespace tiplhip;

namespace {

size_t field_bytes(const int* dim) { return sizeof(float) * 3 * vox(dim); }

void check_T(const float* T)
{
    for (int i = 0; i < 12; ++i) TIPLHIP_REQUIRE(std::isfinite(T[i]), "transform must be finite");
}

}  // namespace

extern "C" {

int tiplhip_displacement_to_mapping(tiplhip_ctx* ctx, const float* dis, const int dim[3],
                                    const float* T, float* mapping)
{
    return guarded([&] {
        TIPLHIP_REQUIRE(ctx && dis && mapping, "null argument");
        check_dim(dim);
        if (T) check_T(T);
        std::lock_guard<std::recursive_mutex> lk(ctx->mu);
        DeviceGuard dg(ctx->device);
        DevMem d(ctx, field_bytes(dim));
        TIPLHIP_CHECK(hipMemcpyAsync(d.p, dis, field_bytes(dim), hipMemcpyHostToDevice, ctx->stream));
        launch_mapping(ctx, d.as<float>(), dim, T, d.as<float>());
        TIPLHIP_CHECK(hipMemcpyAsync(mapping, d.p, field_bytes(dim), hipMemcpyDeviceToHost, ctx->stream));
        TIPLHIP_CHECK(hipStreamSynchronize(ctx->stream));
    });
}

int tiplhip_mapping_to_displacement(tiplhip_ctx* ctx, const float* mapping, const int dim[3],
                                    float* dis)
{
    return guarded([&] {
        TIPLHIP_REQUIRE(ctx && dis && mapping, "null argument");
        check_dim(dim);
        std::lock_guard<std::recursive_mutex> lk(ctx->mu);
        DeviceGuard dg(ctx->device);
        const FieldGeo g = field_geo(dim);
        DevMem d(ctx, field_bytes(dim));
        TIPLHIP_CHECK(hipMemcpyAsync(d.p, mapping, field_bytes(dim), hipMemcpyHostToDevice, ctx->stream));
        hipLaunchKernelGGL(unmapping_kernel, tile_grid(g), dim3(kTX, kTY), 0, ctx->stream,
                           d.as<float>(), g, d.as<float>());
        TIPLHIP_CHECK(hipGetLastError());
        TIPLHIP_CHECK(hipMemcpyAsync(dis, d.p, field_bytes(dim), hipMemcpyDeviceToHost, ctx->stream));
        TIPLHIP_CHECK(hipStreamSynchronize(ctx->stream));
    });
}

int tiplhip_inv_displacement_to_mapping(tiplhip_ctx* ctx, const float* inv_dis,
                                        const int dis_dim[3], const float T[12],
                                        const int out_dim[3], float* mapping)
{
    return guarded([&] {
        TIPLHIP_REQUIRE(ctx && inv_dis && T && mapping, "null argument");
        check_dim(dis_dim);
        check_dim(out_dim);
        check_T(T);
        std::lock_guard<std::recursive_mutex> lk(ctx->mu);
        DeviceGuard dg(ctx->device);
        DevMem d(ctx, field_bytes(dis_dim)), o(ctx, field_bytes(out_dim));
        TIPLHIP_CHECK(hipMemcpyAsync(d.p, inv_dis, field_bytes(dis_dim), hipMemcpyHostToDevice,
                                     ctx->stream));
        launch_inv_mapping(ctx, d.as<float>(), dis_dim, T, out_dim, o.as<float>());
        TIPLHIP_CHECK(hipMemcpyAsync(mapping, o.p, field_bytes(out_dim), hipMemcpyDeviceToHost,
                                     ctx->stream));
        TIPLHIP_CHECK(hipStreamSynchronize(ctx->stream));
    });
}

int tiplhip_compose_mapping(tiplhip_ctx* ctx, const void* from, int dtype, const int from_dim[3],
                            const float* mapping, const int dim[3], void* out)
{
    return guarded([&] {
        TIPLHIP_REQUIRE(ctx && from && mapping && out, "null argument");
        TIPLHIP_REQUIRE(dtype == TIPLHIP_U8 || dtype == TIPLHIP_F32, "dtype must be U8 or F32");
        check_dim(from_dim);
        check_dim(dim);
        std::lock_guard<std::recursive_mutex> lk(ctx->mu);
        DeviceGuard dg(ctx->device);
        const size_t es = dtype == TIPLHIP_U8 ? 1 : sizeof(float);
        DevMem f(ctx, es * vox(from_dim)), m(ctx, field_bytes(dim)), o(ctx, es * vox(dim));
        TIPLHIP_CHECK(hipMemcpyAsync(f.p, from, es * vox(from_dim), hipMemcpyHostToDevice, ctx->stream));
        TIPLHIP_CHECK(hipMemcpyAsync(m.p, mapping, field_bytes(dim), hipMemcpyHostToDevice, ctx->stream));
        if (dtype == TIPLHIP_U8) {
            launch_compose_mapping_u8(ctx, f.as<uint8_t>(), from_dim, m.as<float>(), dim,
                                      o.as<uint8_t>());
        } else {
            const FieldGeo gf = field_geo(from_dim), g = field_geo(dim);
            hipLaunchKernelGGL(compose_mapping_f32_kernel, tile_grid(g), dim3(kTX, kTY), 0,
                               ctx->stream, f.as<float>(), gf, m.as<float>(), g, o.as<float>());
            TIPLHIP_CHECK(hipGetLastError());
        }
        TIPLHIP_CHECK(hipMemcpyAsync(out, o.p, es * vox(dim), hipMemcpyDeviceToHost, ctx->stream));
        TIPLHIP_CHECK(hipStreamSynchronize(ctx->stream));
    });
}

int tiplhip_jacobian_determinant_dis(tiplhip_ctx* ctx, const float* dis, const int dim[3], float* out)
{
    return guarded([&] {
        TIPLHIP_REQUIRE(ctx && dis && out, "null argument");
        check_dim(dim);
        std::lock_guard<std::recursive_mutex> lk(ctx->mu);
        DeviceGuard dg(ctx->device);
        const FieldGeo g = field_geo(dim);
        DevMem d(ctx, field_bytes(dim)), o(ctx, sizeof(float) * g.n);
        TIPLHIP_CHECK(hipMemcpyAsync(d.p, dis, field_bytes(dim), hipMemcpyHostToDevice, ctx->stream));
        hipLaunchKernelGGL(jacobian_dis_kernel, tile_grid(g), dim3(kTX, kTY), 0, ctx->stream,
                           d.as<float>(), g, o.as<float>());
        TIPLHIP_CHECK(hipGetLastError());
        TIPLHIP_CHECK(hipMemcpyAsync(out, o.p, sizeof(float) * g.n, hipMemcpyDeviceToHost, ctx->stream));
        TIPLHIP_CHECK(hipStreamSynchronize(ctx->stream));
    });
}

int tiplhip_jacobian_determinant(tiplhip_ctx* ctx, const float* mapping, const int dim[3], float* out)
{
    return guarded([&] {
        TIPLHIP_REQUIRE(ctx && mapping && out, "null argument");
        check_dim(dim);
        std::lock_guard<std::recursive_mutex> lk(ctx->mu);
        DeviceGuard dg(ctx->device);
        const FieldGeo g = field_geo(dim);
        DevMem d(ctx, field_bytes(dim)), o(ctx, sizeof(float) * g.n);
        TIPLHIP_CHECK(hipMemcpyAsync(d.p, mapping, field_bytes(dim), hipMemcpyHostToDevice, ctx->stream));
        hipLaunchKernelGGL(jacobian_kernel, tile_grid(g), dim3(kTX, kTY), 0, ctx->stream,
                           d.as<float>(), g, o.as<float>());
        TIPLHIP_CHECK(hipGetLastError());
        TIPLHIP_CHECK(hipMemcpyAsync(out, o.p, sizeof(float) * g.n, hipMemcpyDeviceToHost, ctx->stream));
        TIPLHIP_CHECK(hipStreamSynchronize(ctx->stream));
    });
}

int tiplhip_match_fov(tiplhip_ctx* ctx, tiplhip_volume* It, const float T[12], const int moving_dim[3])
{
    return guarded([&] {
        TIPLHIP_REQUIRE(ctx && It && T, "null argument");
        check_dim(moving_dim);
        check_T(T);
        TIPLHIP_REQUIRE(It->ctx && It->ctx->device == ctx->device, "volume is on another device");
        std::lock_guard<std::recursive_mutex> lk(ctx->mu);
        DeviceGuard dg(ctx->device);
        launch_match_fov(ctx, It->d, It->dim, T, moving_dim);
        volume_invalidate(ctx, It);
        TIPLHIP_CHECK(hipStreamSynchronize(ctx->stream));
    });
}

int tiplhip_correlation(tiplhip_ctx* ctx, const tiplhip_volume* x, const tiplhip_volume* y,
                        int masked, double* r)
{
    return guarded([&] {
        TIPLHIP_REQUIRE(ctx && x && y && r, "null argument");
        TIPLHIP_REQUIRE(x->n == y->n && x->dim[0] == y->dim[0] && x->dim[1] == y->dim[1],
                        "volumes must have the same dimensions");
        TIPLHIP_REQUIRE(x->ctx->device == ctx->device && y->ctx->device == ctx->device,
                        "volume is on another device");
        std::lock_guard<std::recursive_mutex> lk(ctx->mu);
        DeviceGuard dg(ctx->device);
        *r = device_correlation(ctx, x->d, y->d, x->n, masked != 0);
    });
}

}  // extern "C"
