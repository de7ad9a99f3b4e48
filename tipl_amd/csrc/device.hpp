// device.hpp -- device-side building blocks shared by the kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#pragma clang fp contract(off)

namespace tiplhip {

constexpr int kWave = 64;

// Raw buffer resource over `bytes` bytes: out-of-range loads return 0 and
// never fault.  Word 3 = 0x00020000 is the gfx9-family (incl. gfx950) raw
// buffer format.
__device__ inline __amdgpu_buffer_rsrc_t make_rsrc(const void* p, uint32_t bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), 0, int(bytes), 0x00020000);
}

__device__ inline uint32_t load_u8(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff)
{
    return __builtin_amdgcn_raw_buffer_load_b8(r, int(voff), int(soff), 0);
}

// Geometry of a uint8 volume sampled by trilinear interpolation.
struct VolGeo {
    int w, h, d;
    uint32_t wh;       // w*h
    float wm1, hm1, dm1;  // float(w-1) etc.
};

// interpolator::linear<3> + weighted_sum + interp_type<uint8>::assign
// (numerical/interpolation.hpp:25-73, 337-401) for a uint8 volume held in a
// raw buffer.  The sample is inside iff 0 <= x < w-1 (and the same for y, z):
// for x >= 0 that is exactly the reference's `size_t(x)+1 < w` test, and a
// NaN coordinate is outside as on the x86 reference.  All 8 taps share one
// vector offset; the +1 / +w / +wh neighbours come from scalar and immediate
// offsets.  Every float operation is in the reference's order and contract
// is off, so the truncated result is bit-exact.
__device__ inline uint32_t sample_u8(__amdgpu_buffer_rsrc_t r, const VolGeo& g, float x, float y,
                                     float z)
{
    const bool in = (fminf(fminf(x, y), z) >= 0.0f) & (x < g.wm1) & (y < g.hm1) & (z < g.dm1);
    if (!in) return 0u;
    uint32_t ix = uint32_t(x), iy = uint32_t(y), iz = uint32_t(z);
    float p0 = x - float(ix);
    float p1 = y - float(iy);
    float p2 = z - float(iz);
    float n0 = 1.0f - p0;
    float n1 = 1.0f - p1;
    float n2 = 1.0f - p2;
    uint32_t base = iz * g.wh + iy * uint32_t(g.w) + ix;
    uint32_t w = uint32_t(g.w);
    uint32_t v0 = load_u8(r, base, 0);
    uint32_t v1 = load_u8(r, base + 1, 0);
    uint32_t v2 = load_u8(r, base, w);
    uint32_t v3 = load_u8(r, base + 1, w);
    uint32_t v4 = load_u8(r, base, g.wh);
    uint32_t v5 = load_u8(r, base + 1, g.wh);
    uint32_t v6 = load_u8(r, base, g.wh + w);
    uint32_t v7 = load_u8(r, base + 1, g.wh + w);
    float n0n1 = n0 * n1;
    float p0n1 = p0 * n1;
    float n0p1 = n0 * p1;
    float p0p1 = p0 * p1;
    float acc = float(v0) * (n0n1 * n2);
    acc = acc + float(v1) * (p0n1 * n2);
    acc = acc + float(v2) * (n0p1 * n2);
    acc = acc + float(v3) * (p0p1 * n2);
    acc = acc + float(v4) * (n0n1 * p2);
    acc = acc + float(v5) * (p0n1 * p2);
    acc = acc + float(v6) * (n0p1 * p2);
    acc = acc + float(v7) * (p0p1 * p2);
    // v <= 0 -> 0, else truncate; the conversion already maps (-inf,1) to 0.
    return acc <= 0.0f ? 0u : uint32_t(acc);
}

// The same interpolation with the 8 taps fetched as one 8-byte word from the
// packed copy of the volume (byte k of word v = tap k of the cell at v, in
// the reference's dindex order).  Same arithmetic, one gather per sample,
// and fewer instructions for the same results:
//   * inside iff 0 <= x < w-1 (and y, z) is one unsigned compare per axis:
//     non-negative floats order like their bit patterns, and negative values
//     and NaNs have patterns above any non-negative finite float.  This
//     needs x != -0.0 (which the reference counts as inside): callers build
//     x = ((x*r0 + y*r1) + z*r2) + s0 with every zero shift component s0
//     stored as +0.0, and a round-to-nearest sum with a +0.0 term, or of
//     two non-zero terms, is never -0.0.  (The sign of a zero x does not
//     change the sampled value.)
//   * x - float(uint(x)) is exactly x - floor(x) for 0 <= x < 2^24, which
//     v_fract_f32 returns in one instruction.
//   * taps and weights are non-negative, so the reference's `v <= 0 ? 0 :
//     T(v)` is plain truncation.
__device__ inline uint32_t sample_u8_packed(__amdgpu_buffer_rsrc_t r8, const VolGeo& g, float x,
                                            float y, float z)
{
    const bool in = (__float_as_uint(x) < __float_as_uint(g.wm1)) &
                    (__float_as_uint(y) < __float_as_uint(g.hm1)) &
                    (__float_as_uint(z) < __float_as_uint(g.dm1));
    if (!in) return 0u;
    const uint32_t ix = uint32_t(x), iy = uint32_t(y), iz = uint32_t(z);
    const float p0 = __builtin_amdgcn_fractf(x);
    const float p1 = __builtin_amdgcn_fractf(y);
    const float p2 = __builtin_amdgcn_fractf(z);
    const float n0 = 1.0f - p0;
    const float n1 = 1.0f - p1;
    const float n2 = 1.0f - p2;
    // volumes are < 2^24 voxels per plane and per row index: 24-bit products
    const uint32_t base = __umul24(iz, g.wh) + __umul24(iy, uint32_t(g.w)) + ix;
    const auto q = __builtin_amdgcn_raw_buffer_load_b64(r8, int(base * 8u), 0, 0);
    const uint32_t lo = q[0], hi = q[1];
    const float n0n1 = n0 * n1;
    const float p0n1 = p0 * n1;
    const float n0p1 = n0 * p1;
    const float p0p1 = p0 * p1;
    float acc = float(lo & 0xffu) * (n0n1 * n2);
    acc = acc + float((lo >> 8) & 0xffu) * (p0n1 * n2);
    acc = acc + float((lo >> 16) & 0xffu) * (n0p1 * n2);
    acc = acc + float(lo >> 24) * (p0p1 * n2);
    acc = acc + float(hi & 0xffu) * (n0n1 * p2);
    acc = acc + float((hi >> 8) & 0xffu) * (p0n1 * p2);
    acc = acc + float((hi >> 16) & 0xffu) * (n0p1 * p2);
    acc = acc + float(hi >> 24) * (p0p1 * p2);
    return uint32_t(acc);
}

// Two samples of sample_u8_packed with both cells' loads issued before
// either is used (two independent chains per thread hide more of the
// gather latency).  A sample outside the volume still loads: its offset is
// whatever the truncated coordinates give, and the buffer's range check
// makes any offset safe (out of range reads return 0).  Its value is then
// forced to 0, so no lane diverges; samples inside compute exactly as
// sample_u8_packed.
// v_cvt_u32_f32: truncation for coordinates inside the volume; any other
// input (negative, too large, NaN) gives a defined, saturated value rather
// than the undefined result of a C++ conversion.
__device__ inline uint32_t cvt_u32_sat(float x)
{
    uint32_t r;
    asm("v_cvt_u32_f32 %0, %1" : "=v"(r) : "v"(x));
    return r;
}

__device__ inline void sample2_u8_packed(__amdgpu_buffer_rsrc_t r8, const VolGeo& g, float xa,
                                         float ya, float za, float xb, float yb, float zb,
                                         uint32_t& ta, uint32_t& tb)
{
    const bool ina = (__float_as_uint(xa) < __float_as_uint(g.wm1)) &
                     (__float_as_uint(ya) < __float_as_uint(g.hm1)) &
                     (__float_as_uint(za) < __float_as_uint(g.dm1));
    const bool inb = (__float_as_uint(xb) < __float_as_uint(g.wm1)) &
                     (__float_as_uint(yb) < __float_as_uint(g.hm1)) &
                     (__float_as_uint(zb) < __float_as_uint(g.dm1));
    const uint32_t oa = (__umul24(cvt_u32_sat(za), g.wh) + __umul24(cvt_u32_sat(ya), uint32_t(g.w)) +
                         cvt_u32_sat(xa)) * 8u;
    const uint32_t ob = (__umul24(cvt_u32_sat(zb), g.wh) + __umul24(cvt_u32_sat(yb), uint32_t(g.w)) +
                         cvt_u32_sat(xb)) * 8u;
    const auto qa = __builtin_amdgcn_raw_buffer_load_b64(r8, int(oa), 0, 0);
    const auto qb = __builtin_amdgcn_raw_buffer_load_b64(r8, int(ob), 0, 0);
    auto blend = [](uint32_t lo, uint32_t hi, float x, float y, float z) {
        const float p0 = __builtin_amdgcn_fractf(x);
        const float p1 = __builtin_amdgcn_fractf(y);
        const float p2 = __builtin_amdgcn_fractf(z);
        const float n0 = 1.0f - p0;
        const float n1 = 1.0f - p1;
        const float n2 = 1.0f - p2;
        const float n0n1 = n0 * n1;
        const float p0n1 = p0 * n1;
        const float n0p1 = n0 * p1;
        const float p0p1 = p0 * p1;
        float acc = float(lo & 0xffu) * (n0n1 * n2);
        acc = acc + float((lo >> 8) & 0xffu) * (p0n1 * n2);
        acc = acc + float((lo >> 16) & 0xffu) * (n0p1 * n2);
        acc = acc + float(lo >> 24) * (p0p1 * n2);
        acc = acc + float(hi & 0xffu) * (n0n1 * p2);
        acc = acc + float((hi >> 8) & 0xffu) * (p0n1 * p2);
        acc = acc + float((hi >> 16) & 0xffu) * (n0p1 * p2);
        acc = acc + float(hi >> 24) * (p0p1 * p2);
        return uint32_t(acc);
    };
    const uint32_t va = blend(qa[0], qa[1], xa, ya, za);
    const uint32_t vb = blend(qb[0], qb[1], xb, yb, zb);
    ta = ina ? va : 0u;
    tb = inb ? vb : 0u;
}

// The sampler's bounds test alone: 0 <= x < w-1, and y, z alike (the same
// unsigned compares as sample_u8_packed, with the same -0.0 proviso).
__device__ inline bool inside_u8(const VolGeo& g, float x, float y, float z)
{
    return (__float_as_uint(x) < __float_as_uint(g.wm1)) &
           (__float_as_uint(y) < __float_as_uint(g.hm1)) &
           (__float_as_uint(z) < __float_as_uint(g.dm1));
}

// sample2_u8_packed for two samples the caller knows are inside the volume:
// no bounds test and no select, and the cell index formed exactly in FP32.
// For 0 <= x < 2^23, v_fract gives x - floor(x) exactly, so x - fract(x) is
// floor(x); with w*h*d <= 2^24 every product and sum of
// (floor(z)*wh + floor(y)*w) + floor(x) is an integer below 2^24, so the two
// fused multiply-adds are exact and one conversion replaces three
// conversions and two 24-bit multiplies.  fw, fwh: w and w*h as floats.
// The interpolation is sample_u8_packed's, operation for operation.
__device__ inline void sample2_u8_packed_in(__amdgpu_buffer_rsrc_t r8, float fw, float fwh,
                                            float xa, float ya, float za, float xb, float yb,
                                            float zb, uint32_t& ta, uint32_t& tb)
{
    const float pxa = __builtin_amdgcn_fractf(xa), pya = __builtin_amdgcn_fractf(ya),
                pza = __builtin_amdgcn_fractf(za);
    const float pxb = __builtin_amdgcn_fractf(xb), pyb = __builtin_amdgcn_fractf(yb),
                pzb = __builtin_amdgcn_fractf(zb);
    const float ia = __builtin_fmaf(za - pza, fwh, __builtin_fmaf(ya - pya, fw, xa - pxa));
    const float ib = __builtin_fmaf(zb - pzb, fwh, __builtin_fmaf(yb - pyb, fw, xb - pxb));
    const auto qa = __builtin_amdgcn_raw_buffer_load_b64(r8, int(cvt_u32_sat(ia) * 8u), 0, 0);
    const auto qb = __builtin_amdgcn_raw_buffer_load_b64(r8, int(cvt_u32_sat(ib) * 8u), 0, 0);
    auto blend = [](uint32_t lo, uint32_t hi, float p0, float p1, float p2) {
        const float n0 = 1.0f - p0;
        const float n1 = 1.0f - p1;
        const float n2 = 1.0f - p2;
        const float n0n1 = n0 * n1;
        const float p0n1 = p0 * n1;
        const float n0p1 = n0 * p1;
        const float p0p1 = p0 * p1;
        float acc = float(lo & 0xffu) * (n0n1 * n2);
        acc = acc + float((lo >> 8) & 0xffu) * (p0n1 * n2);
        acc = acc + float((lo >> 16) & 0xffu) * (n0p1 * n2);
        acc = acc + float(lo >> 24) * (p0p1 * n2);
        acc = acc + float(hi & 0xffu) * (n0n1 * p2);
        acc = acc + float((hi >> 8) & 0xffu) * (p0n1 * p2);
        acc = acc + float((hi >> 16) & 0xffu) * (n0p1 * p2);
        acc = acc + float(hi >> 24) * (p0p1 * p2);
        return uint32_t(acc);
    };
    ta = blend(qa[0], qa[1], pxa, pya, pza);
    tb = blend(qb[0], qb[1], pxb, pyb, pzb);
}

// sample2_u8_packed with the cell index formed in FP32 as in
// sample2_u8_packed_in (floor as x - fract(x), two exact fused multiply-adds,
// one conversion), for a `to` volume of at most 2^24 voxels.  An outside
// sample gets an arbitrary index (the buffer's range check makes any offset
// safe) and is then forced to 0, as in sample2_u8_packed.  Measured at
// 182x218x182 with 64 candidates: the 256-bin histogram kernel 1-3% faster,
// the 64-bin one 4.5% slower (which keeps sample2_u8_packed).
__device__ inline void sample2_u8_packed_fi(__amdgpu_buffer_rsrc_t r8, const VolGeo& g, float fw,
                                            float fwh, float xa, float ya, float za, float xb,
                                            float yb, float zb, uint32_t& ta, uint32_t& tb)
{
    const bool ina = (__float_as_uint(xa) < __float_as_uint(g.wm1)) &
                     (__float_as_uint(ya) < __float_as_uint(g.hm1)) &
                     (__float_as_uint(za) < __float_as_uint(g.dm1));
    const bool inb = (__float_as_uint(xb) < __float_as_uint(g.wm1)) &
                     (__float_as_uint(yb) < __float_as_uint(g.hm1)) &
                     (__float_as_uint(zb) < __float_as_uint(g.dm1));
    uint32_t va, vb;
    sample2_u8_packed_in(r8, fw, fwh, xa, ya, za, xb, yb, zb, va, vb);
    ta = ina ? va : 0u;
    tb = inb ? vb : 0u;
}

// Hardware dispatches workgroup b of a 1-D grid to XCD b % 8.  This remap
// gives each XCD a contiguous range of logical workgroups so that the
// volume region a range touches stays in that XCD's L2.  Bijective on
// [0, total).
__device__ inline uint32_t xcd_remap(uint32_t b, uint32_t total)
{
    uint32_t q = total >> 3, r = total & 7u;
    uint32_t xcd = b & 7u, slot = b >> 3;
    return xcd * q + (xcd < r ? xcd : r) + slot;
}

__device__ inline uint64_t wave_sum_u64(uint64_t v)
{
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
    return v;
}

__device__ inline double wave_sum_f64(double v)
{
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
    return v;
}

__device__ inline float wave_max_f32(float v)
{
    for (int off = 32; off > 0; off >>= 1) v = fmaxf(v, __shfl_xor(v, off, kWave));
    return v;
}

}  // namespace tiplhip
