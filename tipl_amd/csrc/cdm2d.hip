// cdm2d.hip -- the 2-D CDM loop on gfx950: tipl::reg::cdm for image<2,uint8>,
// which the reference also instantiates for its GPU as
// cdm_cuda<out,uint8,2> (reg/cdm_cu.hpp:12-17).
//
// Restates, for 2-D images (the reference's templates at dimension 2):
//   compose_displacement        numerical/dif.hpp:90-107
//   interpolator::linear<2>     numerical/interpolation.hpp:283-306, weighted_sum :61-73
//   cdm_get_gradient (r, d)     reg/cdm.hpp:46-106 with the 2-D
//                               get_window_at_width<2> (numerical/window.hpp:73-84),
//                               linear_regression (numerical/statistics.hpp:694-713)
//                               and the 2-D gradient_at (numerical/numerical.hpp:112-117)
//   cdm_solve_poisson           reg/cdm.hpp:158-254 (5-point stencil, -1/4)
//   cdm_max_displacement_length reg/cdm.hpp:260-296, vector<2>::length
//                               (utility/pixel_index.hpp:340)
//   accumulate_displacement     numerical/dif.hpp:166-186
//   cdm_smooth                  reg/cdm.hpp:302-367 (weights s/4 and 1-s)
//   downsample_with_padding, upsample_with_padding
//                               numerical/resampling.hpp:575-673, 849-891
//   cdm / cdm_cuda              reg/cdm.hpp:373-532
//
// A slice is 10^4..10^6 pixels, a fraction of the chip, so the loop is bound
// by launches: each step is one kernel with one thread per pixel on 64 x 4
// tiles, except the Jacobi solve, which runs four (or two) sweeps per launch
// over flat-index tiles held in LDS.  The loop control (cost history,
// cdm_improved, smoothing schedule) runs on the host from the modality costs,
// which the cost reduction stores into coherent host memory while the Jacobi
// sweeps are already queued, as in the 3-D loop.  Fields are packed float2
// per pixel, x-fastest (the reference's image<2,vector<2>>).  Every FP32
// operation is a separately rounded operation in the reference's order.
#include "cdm_control.hpp"
#include "device.hpp"
#include "field.hpp"
#include "host_wait.hpp"
#include "internal.hpp"

#include <algorithm>
#include <chrono>
#include <cstring>
#include <limits>
#include <memory>
#include <vector>

namespace tiplhip {
namespace {

struct Geo2 {
    int w, h;
    uint32_t n;
    float wm1, hm1;
};

Geo2 geo2(int w, int h)
{
    Geo2 g;
    g.w = w;
    g.h = h;
    g.n = uint32_t(w) * uint32_t(h);
    g.wm1 = float(w - 1);
    g.hm1 = float(h - 1);
    return g;
}

constexpr int kBX = 64, kBY = 4;
constexpr uint32_t kThreads2 = kBX * kBY;

// A quiet-NaN float pattern the cost reduction never writes, and how long
// the host spins for an iteration's costs before it synchronises the stream.
constexpr uint32_t kCostPending2 = 0x7FC0C057u;
constexpr int kCostSpinUs2 = 2000;

dim3 grid2(const Geo2& g)
{
    return dim3(uint32_t((g.w + kBX - 1) / kBX), uint32_t((g.h + kBY - 1) / kBY));
}

uint32_t blocks2(const Geo2& g)
{
    const dim3 t = grid2(g);
    return t.x * t.y;
}

struct Pix {
    int x, y;
    uint32_t p;
    bool ok;
};

__device__ inline Pix this_pixel(const Geo2& g)
{
    Pix q;
    q.x = int(blockIdx.x) * kBX + int(threadIdx.x);
    q.y = int(blockIdx.y) * kBY + int(threadIdx.y);
    q.ok = q.x < g.w && q.y < g.h;
    q.p = uint32_t(q.y) * uint32_t(g.w) + uint32_t(q.x);
    return q;
}

__device__ inline float2 ld2(const float* f, uint32_t p)
{
    return make_float2(f[2 * size_t(p)], f[2 * size_t(p) + 1]);
}

__device__ inline void st2(float* f, uint32_t p, float2 v)
{
    f[2 * size_t(p)] = v.x;
    f[2 * size_t(p) + 1] = v.y;
}

// interpolator::linear<2>::get_location: inside when 0 <= x < w-1 and
// 0 <= y < h-1 (size_t(x) + 1 < w); a NaN coordinate fails every test.
struct Bil {
    uint32_t i0;
    int ix, iy;
    float w[4];
};

__device__ inline bool locate2(const Geo2& g, float x, float y, Bil& b)
{
    const bool in = (x >= 0.0f) & (y >= 0.0f) & (x < g.wm1) & (y < g.hm1);
    if (!in) return false;
    const uint32_t ix = uint32_t(x), iy = uint32_t(y);
    const float p0 = x - float(ix), p1 = y - float(iy);
    const float n0 = 1.0f - p0, n1 = 1.0f - p1;
    b.i0 = iy * uint32_t(g.w) + ix;
    b.ix = int(ix);
    b.iy = int(iy);
    b.w[0] = n0 * n1;
    b.w[1] = p0 * n1;
    b.w[2] = n0 * p1;
    b.w[3] = p0 * p1;
    return true;
}

// taps in the reference's order: i0, i0+1, i0+w, i0+w+1
__device__ inline uint32_t tap2(const Geo2& g, uint32_t i0, int k)
{
    return i0 + uint32_t(k & 1) + ((k & 2) ? uint32_t(g.w) : 0u);
}

// weighted_sum of uint8 taps, then the uint8 assignment (v <= 0 -> 0, else
// truncation; interpolation.hpp:25-35).
__device__ inline uint8_t bil_u8(const uint8_t* I, const Geo2& g, const Bil& b)
{
    float r = float(I[b.i0]) * b.w[0];
#pragma unroll
    for (int k = 1; k < 4; ++k) {
        float v = float(I[tap2(g, b.i0, k)]);
        v *= b.w[k];
        r += v;
    }
    return r <= 0.0f ? uint8_t(0) : uint8_t(r);
}

// ---- pyramid ---------------------------------------------------------------

// downsample_with_padding of a uint8 image: the uint16 sum of each 2 x 2
// block, a missing far-edge tap replaced by the tap without the missing
// axes, shifted right by 2.
__global__ __launch_bounds__(kThreads2) void down2_u8_kernel(const uint8_t* __restrict__ in,
                                                             Geo2 gi, uint8_t* __restrict__ out,
                                                             Geo2 go)
{
    const Pix q = this_pixel(go);
    if (!q.ok) return;
    const int x2 = 2 * q.x, y2 = 2 * q.y;
    const uint32_t base = uint32_t(y2) * uint32_t(gi.w) + uint32_t(x2);
    const uint32_t shift[4] = {0u, 1u, uint32_t(gi.w), 1u + uint32_t(gi.w)};
    const int has = (x2 + 1 < gi.w ? 1 : 0) + (y2 + 1 < gi.h ? 2 : 0);
    uint8_t buf[4];
    uint32_t sum = buf[0] = in[base];
#pragma unroll
    for (int i = 1; i < 4; ++i) {
        const int hh = has & i;
        buf[i] = (hh == i) ? in[base + shift[i]] : buf[hh];
        sum += buf[i];
    }
    out[q.p] = uint8_t((sum & 0xffffu) >> 2);
}

// upsample_with_padding<linear> of a float2 field whose taps are first
// multiplied by `scale` (cdm() doubles the coarse field before it upsamples:
// reg/cdm.hpp:403-404); out(p) = in(p * 0.5), 0 outside.
__global__ __launch_bounds__(kThreads2) void up2_kernel(const float* __restrict__ in, Geo2 gi,
                                                        float scale, float* __restrict__ out,
                                                        Geo2 go)
{
    const Pix q = this_pixel(go);
    if (!q.ok) return;
    float px = float(q.x), py = float(q.y);
    px *= 0.5f;
    py *= 0.5f;
    Bil b;
    float2 r = make_float2(0.0f, 0.0f);
    if (locate2(gi, px, py, b)) {
        float2 t = ld2(in, b.i0);
        t.x *= scale;
        t.y *= scale;
        r.x = t.x * b.w[0];
        r.y = t.y * b.w[0];
#pragma unroll
        for (int k = 1; k < 4; ++k) {
            float2 v = ld2(in, tap2(gi, b.i0, k));
            v.x *= scale;
            v.y *= scale;
            v.x *= b.w[k];
            v.y *= b.w[k];
            r.x += v.x;
            r.y += v.y;
        }
    }
    st2(out, q.p, r);
}

// ---- the loop body ----------------------------------------------------------

// compose_displacement: Js(x) = Is(x + d(x)); an exact copy where d(x) == 0,
// 0 where the sample is outside.
__global__ __launch_bounds__(kThreads2) void compose2_kernel(const uint8_t* __restrict__ Is,
                                                             Geo2 g, const float* __restrict__ d,
                                                             uint8_t* __restrict__ Js)
{
    const Pix q = this_pixel(g);
    if (!q.ok) return;
    const float2 v = ld2(d, q.p);
    if (v.x == 0.0f && v.y == 0.0f) {
        Js[q.p] = Is[q.p];
        return;
    }
    float px = float(q.x), py = float(q.y);
    px += v.x;
    py += v.y;
    Bil b;
    Js[q.p] = locate2(g, px, py, b) ? bil_u8(Is, g, b) : uint8_t(0);
}

// gradient_at (2-D): central differences, 0 on the first and last column/row.
__device__ inline float2 gradient_at2(const uint8_t* I, const Geo2& g, const Pix& q)
{
    float2 r;
    r.x = (q.x + 1 < g.w && q.x) ? float(int(I[q.p + 1]) - int(I[q.p - 1])) : 0.0f;
    r.y = (q.y + 1 < g.h && q.y) ? float(int(I[q.p + uint32_t(g.w)]) - int(I[q.p - uint32_t(g.w)]))
                                 : 0.0f;
    return r;
}

// Sum of a double over the block in a fixed tree order (deterministic).
__device__ inline double block_sum_f64(double v, double* sh)
{
    const uint32_t t = threadIdx.y * kBX + threadIdx.x;
    sh[t] = v;
    __syncthreads();
    for (uint32_t s = kThreads2 / 2; s > 0; s >>= 1) {
        if (t < s) sh[t] += sh[t + s];
        __syncthreads();
    }
    return sh[0];
}

// cdm_get_gradient: new_d(x) for 'r' (windowed regression of It on Js) or
// 'd' (SSD); `accumulate` adds it to nd (the modality sum new_d += dd, with
// dd zero where a pixel contributes nothing).  Each block's FP64 cost sum
// goes to part[block].
__global__ __launch_bounds__(kThreads2) void gradient2_kernel(const uint8_t* __restrict__ Js,
                                                              const uint8_t* __restrict__ It,
                                                              Geo2 g, char type, int accumulate,
                                                              float* __restrict__ nd,
                                                              double* __restrict__ part)
{
    __shared__ double sh[kThreads2];
    const Pix q = this_pixel(g);
    float cost = 0.0f;
    if (q.ok) {
        float2 val = make_float2(0.0f, 0.0f);
        if (type == 'd') {
            const float2 gr = gradient_at2(Js, g, q);
            const int dif = int(Js[q.p]) - int(It[q.p]);
            val.x = 0.0f + gr.x * float(dif);
            val.y = 0.0f + gr.y * float(dif);
            cost = float(dif * dif);
        } else if (Js[q.p] != 0 && It[q.p] != 0) {
            const int x0 = max(q.x - 2, 0), x1 = min(q.x + 2, g.w - 1);
            const int y0 = max(q.y - 2, 0), y1 = min(q.y + 2, g.h - 1);
            // at most 25 * 255^2 per sum: exact in int32
            int sx = 0, sxx = 0, sy = 0, syy = 0, sxy = 0;
            for (int yy = y0; yy <= y1; ++yy) {
                const uint32_t row = uint32_t(yy) * uint32_t(g.w);
                for (int xx = x0; xx <= x1; ++xx) {
                    const int a = Js[row + uint32_t(xx)], b = It[row + uint32_t(xx)];
                    sx += a;
                    sxx += a * a;
                    sy += b;
                    syy += b * b;
                    sxy += a * b;
                }
            }
            const double dn = double((x1 - x0 + 1) * (y1 - y0 + 1));
            const double mean_x = double(sx) / dn;
            const double mean_y = double(sy) / dn;
            const double x_var = double(sxx) / dn - mean_x * mean_x;
            const double y_var = double(syy) / dn - mean_y * mean_y;
            float a, b, r2;
            if (x_var == 0.0 || y_var == 0.0) {
                a = 0.0f;
                b = float(mean_y);
                r2 = 0.0f;
            } else {
                const double cov = double(sxy) / dn - mean_x * mean_y;
                a = float(cov / x_var);
                b = float(mean_y - double(a) * mean_x);
                r2 = float(cov * cov / x_var / y_var);
            }
            if (a > 0.0f) {
                const float2 gr = gradient_at2(Js, g, q);
                float s = float(Js[q.p]) * a;
                s = s + b;
                s = s - float(It[q.p]);
                s = r2 * s;
                val.x = 0.0f + gr.x * s;
                val.y = 0.0f + gr.y * s;
                cost = -r2;
            }
        }
        if (accumulate) {
            const float2 o = ld2(nd, q.p);
            val.x = o.x + val.x;
            val.y = o.y + val.y;
        }
        st2(nd, q.p, val);
    }
    const double s = block_sum_f64(double(cost), sh);
    if (threadIdx.x == 0 && threadIdx.y == 0) part[blockIdx.y * gridDim.x + blockIdx.x] = s;
}

// float(mean(cost_map)): the block sums added in block order, then / n;
// also stored into host-visible memory when host_out is set.
__global__ __launch_bounds__(kThreads2) void cost2_final_kernel(const double* __restrict__ part,
                                                                uint32_t nparts, uint32_t n,
                                                                float* __restrict__ out,
                                                                float* __restrict__ host_out)
{
    __shared__ double sh[kThreads2];
    const uint32_t t = threadIdx.y * kBX + threadIdx.x;
    double s = 0.0;
    for (uint32_t i = t; i < nparts; i += kThreads2) s += part[i];
    const double tot = block_sum_f64(s, sh);
    if (t == 0) {
        *out = float(tot / double(n));
        if (host_out) *host_out = float(tot / double(n));
    }
}

// One Jacobi sweep of cdm_solve_poisson in 2-D: v = new_d - (the flat-index
// neighbours -1, +1, -w, +w of the previous solve, dropped past the array
// ends), times -1/4.  The first sweep reads the initial solve new_d * -1/4
// as it goes (the same products the reference stores first).
__global__ __launch_bounds__(kThreads2) void poisson2d_kernel(const float* __restrict__ nd,
                                                              const float* __restrict__ src,
                                                              Geo2 g, int first,
                                                              float* __restrict__ dst)
{
    const Pix q = this_pixel(g);
    if (!q.ok) return;
    constexpr float inv_d2 = -0.5f / 2.0f;
    const uint32_t p = q.p, w = uint32_t(g.w), n = g.n;
    auto at = [&](uint32_t i) {
        if (first) {
            float2 a = ld2(nd, i);
            a.x *= inv_d2;
            a.y *= inv_d2;
            return a;
        }
        return ld2(src, i);
    };
    float2 v = ld2(nd, p);
    if (p >= 1) {
        const float2 a = at(p - 1);
        v.x -= a.x;
        v.y -= a.y;
    }
    if (p + 1 < n) {
        const float2 a = at(p + 1);
        v.x -= a.x;
        v.y -= a.y;
    }
    if (p >= w) {
        const float2 a = at(p - w);
        v.x -= a.x;
        v.y -= a.y;
    }
    if (p + w < n) {
        const float2 a = at(p + w);
        v.x -= a.x;
        v.y -= a.y;
    }
    v.x *= inv_d2;
    v.y *= inv_d2;
    st2(dst, p, v);
}

// K Jacobi sweeps in one launch over a flat-index segment [a, a + L) of the
// field.  The stencil reaches p +- 1 and p +- w, so sweep s (1..K) is
// evaluated on [a - (K-s)w, a + L + (K-s)w) from sweep s-1's values one w
// wider on each side; the input solve (sweep 0) covers [a - Kw, a + L + Kw).
// Two LDS buffers of L + 2Kw values alternate.  Positions outside the array
// hold +0.0: subtracting +0.0 leaves a value unchanged bit for bit
// (x - (+0) == x, -0 included), so every neighbour is subtracted, in the
// reference's order, and the dropped ones change nothing.  `first`: the
// input solve is nd * -1/4 (the initial solve).  Dynamic LDS:
// 2 (L + 2Kw) float2.
template <int K>
__global__ __launch_bounds__(kThreads2) void poisson2d_xk_kernel(const float* __restrict__ nd,
                                                                 const float* __restrict__ src,
                                                                 uint32_t n, uint32_t w, uint32_t L,
                                                                 int first, float* __restrict__ dst)
{
    extern __shared__ float2 sm2[];
    constexpr float inv_d2 = -0.5f / 2.0f;
    const int64_t a = int64_t(blockIdx.x) * int64_t(L);
    const int64_t W = int64_t(w), span = int64_t(L) + 2 * K * W;
    float2* buf[2] = {sm2, sm2 + span};
    const uint32_t t = threadIdx.y * kBX + threadIdx.x;
    {
        const int64_t lo = a - K * W;
        for (int64_t i = t; i < span; i += kThreads2) {
            const int64_t q = lo + i;
            float2 v = make_float2(0.0f, 0.0f);
            if (q >= 0 && q < int64_t(n)) {
                if (first) {
                    v = ld2(nd, uint32_t(q));
                    v.x *= inv_d2;
                    v.y *= inv_d2;
                } else {
                    v = ld2(src, uint32_t(q));
                }
            }
            buf[0][i] = v;
        }
    }
#pragma unroll
    for (int s = 1; s <= K; ++s) {
        __syncthreads();
        const float2* in = buf[(s - 1) & 1];
        float2* out = buf[s & 1];
        const int64_t lo = a - (K - s) * W, len = int64_t(L) + 2 * (K - s) * W;
        for (int64_t i = t; i < len; i += kThreads2) {
            const int64_t q = lo + i;
            if (s == K && q >= int64_t(n)) break;
            float2 v = make_float2(0.0f, 0.0f);
            if (q >= 0 && q < int64_t(n)) {
                const int64_t k = i + W;  // in's index of q
                v = ld2(nd, uint32_t(q));
                const float2 a1 = in[k - 1], a2 = in[k + 1], a3 = in[k - W], a4 = in[k + W];
                v.x -= a1.x;
                v.y -= a1.y;
                v.x -= a2.x;
                v.y -= a2.y;
                v.x -= a3.x;
                v.y -= a3.y;
                v.x -= a4.x;
                v.y -= a4.y;
                v.x *= inv_d2;
                v.y *= inv_d2;
            }
            if (s == K)
                st2(dst, uint32_t(q), v);
            else
                out[i] = v;
        }
    }
}

// cdm_max_displacement_length: the largest float(sqrt(x*x + y*y)) as float
// bits in *out (lengths are >= 0, so their bit patterns order as uints).
__global__ __launch_bounds__(kThreads2) void maxlen2_kernel(const float* __restrict__ d, Geo2 g,
                                                            unsigned int* __restrict__ out)
{
    const Pix q = this_pixel(g);
    float len = 0.0f;
    if (q.ok) {
        const float2 v = ld2(d, q.p);
        float s2 = v.x * v.x;
        s2 = s2 + v.y * v.y;
        // std::sqrt of the float sum; sqrt through double then to float is
        // the same correctly rounded value
        len = float(sqrt(double(s2)));
    }
    len = wave_max_f32(len);
    if ((threadIdx.x & 63u) == 0) atomicMax(out, __float_as_uint(len));
}

// new_d *= k, then accumulate_displacement(cur, new_d): with the mapping
// m(x) = cur(x) + x, new(x) = m(x + new(x)) - x by bilinear interpolation;
// where the scaled new(x) is 0 or the sample is outside, new(x) = cur(x).
// Out of place: the smoothing after it reads the neighbours of new.
__global__ __launch_bounds__(kThreads2) void acc2_kernel(const float* __restrict__ cur,
                                                         const float* __restrict__ nd, float k,
                                                         Geo2 g, float* __restrict__ out)
{
    const Pix q = this_pixel(g);
    if (!q.ok) return;
    float2 d = ld2(nd, q.p);
    d.x *= k;
    d.y *= k;
    if (d.x != 0.0f || d.y != 0.0f) {
        const float px = d.x + float(q.x), py = d.y + float(q.y);
        Bil b;
        if (locate2(g, px, py, b)) {
            float2 r;
#pragma unroll
            for (int kk = 0; kk < 4; ++kk) {
                const uint32_t t = tap2(g, b.i0, kk);
                float2 m = ld2(cur, t);
                m.x = m.x + float(b.ix + (kk & 1));
                m.y = m.y + float(b.iy + ((kk & 2) ? 1 : 0));
                if (kk == 0) {
                    r.x = m.x * b.w[0];
                    r.y = m.y * b.w[0];
                } else {
                    m.x *= b.w[kk];
                    m.y *= b.w[kk];
                    r.x += m.x;
                    r.y += m.y;
                }
            }
            r.x = r.x - float(q.x);
            r.y = r.y - float(q.y);
            st2(out, q.p, r);
            return;
        }
    }
    st2(out, q.p, ld2(cur, q.p));
}

// cdm_smooth: out = a*(1-s) + (s/4)*(a[+1] + a[-1] + a[+w] + a[-w]) over
// flat-index neighbours, added in that order (dropped past the array ends).
__global__ __launch_bounds__(kThreads2) void smooth2_kernel(const float* __restrict__ a, Geo2 g,
                                                            float w_4, float w_1,
                                                            float* __restrict__ out)
{
    const Pix q = this_pixel(g);
    if (!q.ok) return;
    const uint32_t p = q.p, w = uint32_t(g.w), n = g.n;
    float2 v = make_float2(0.0f, 0.0f);
    if (p + 1 < n) {
        const float2 t = ld2(a, p + 1);
        v.x += t.x;
        v.y += t.y;
    }
    if (p >= 1) {
        const float2 t = ld2(a, p - 1);
        v.x += t.x;
        v.y += t.y;
    }
    if (p + w < n) {
        const float2 t = ld2(a, p + w);
        v.x += t.x;
        v.y += t.y;
    }
    if (p >= w) {
        const float2 t = ld2(a, p - w);
        v.x += t.x;
        v.y += t.y;
    }
    v.x *= w_4;
    v.y *= w_4;
    const float2 c = ld2(a, p);
    float2 r;
    r.x = c.x * w_1 + v.x;
    r.y = c.y * w_1 + v.y;
    st2(out, p, r);
}

// ---- host ---------------------------------------------------------------------

void check2(const int* dim)
{
    TIPLHIP_REQUIRE(dim && dim[0] > 0 && dim[1] > 0, "dims must be positive");
    TIPLHIP_REQUIRE(size_t(dim[0]) * size_t(dim[1]) < (size_t(1) << 30), "image exceeds 2^30 pixels");
    // rows go to grid y in fours (at most 65535 of them)
    TIPLHIP_REQUIRE(dim[1] <= 4 * 65535, "image height exceeds 262140");
}

void launch_compose2(tiplhip_ctx* ctx, const uint8_t* Is, const float* d, const Geo2& g, uint8_t* Js)
{
    hipLaunchKernelGGL(compose2_kernel, grid2(g), dim3(kBX, kBY), 0, ctx->stream, Is, g, d, Js);
    TIPLHIP_CHECK(hipGetLastError());
}

// The gradient of one modality into nd and its float cost into *cost
// (device) and, when host_cost is set, into host-visible memory.
void launch_gradient2(tiplhip_ctx* ctx, const uint8_t* Js, const uint8_t* It, const Geo2& g,
                      char type, bool accumulate, float* nd, double* part, float* cost,
                      float* host_cost = nullptr)
{
    hipLaunchKernelGGL(gradient2_kernel, grid2(g), dim3(kBX, kBY), 0, ctx->stream, Js, It, g, type,
                       accumulate ? 1 : 0, nd, part);
    TIPLHIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(cost2_final_kernel, dim3(1), dim3(kBX, kBY), 0, ctx->stream, part, blocks2(g),
                       g.n, cost, host_cost);
    TIPLHIP_CHECK(hipGetLastError());
}

// Flat outputs per multi-sweep launch; a launch runs K = 4 sweeps while the
// two LDS buffers (2 (L + 8w) float2) fit in 64 KB, K = 2 up to twice that
// width, and one sweep per launch beyond.
constexpr uint32_t kSweepTile = 1024;
constexpr uint32_t kMaxX4Width = 384;
constexpr uint32_t kMaxX2Width = 768;

// `sweeps` Jacobi sweeps ping-ponging a/b (several per launch where the
// tile fits); returns the buffer holding the result.
float* launch_poisson2(tiplhip_ctx* ctx, const float* nd, float* a, float* b, const Geo2& g, int sweeps)
{
    float* src = a;
    float* dst = a;
    const uint32_t w = uint32_t(g.w);
    const dim3 tiles((g.n + kSweepTile - 1) / kSweepTile);
    int done = 0;
    for (int i = 0; done < sweeps; ++i) {
        dst = (i & 1) ? b : a;
        if (w <= kMaxX4Width && sweeps - done >= 4) {
            const size_t lds = sizeof(float2) * 2 * (size_t(kSweepTile) + 8 * size_t(w));
            hipLaunchKernelGGL(poisson2d_xk_kernel<4>, tiles, dim3(kBX, kBY), lds, ctx->stream, nd, src,
                               g.n, w, kSweepTile, done == 0 ? 1 : 0, dst);
            done += 4;
        } else if (w <= kMaxX2Width && sweeps - done >= 2) {
            const size_t lds = sizeof(float2) * 2 * (size_t(kSweepTile) + 4 * size_t(w));
            hipLaunchKernelGGL(poisson2d_xk_kernel<2>, tiles, dim3(kBX, kBY), lds, ctx->stream, nd, src,
                               g.n, w, kSweepTile, done == 0 ? 1 : 0, dst);
            done += 2;
        } else {
            hipLaunchKernelGGL(poisson2d_kernel, grid2(g), dim3(kBX, kBY), 0, ctx->stream, nd, src, g,
                               done == 0 ? 1 : 0, dst);
            done += 1;
        }
        TIPLHIP_CHECK(hipGetLastError());
        src = dst;
    }
    return dst;
}

float max_length2(tiplhip_ctx* ctx, const float* d, const Geo2& g, unsigned int* dev_word)
{
    TIPLHIP_CHECK(hipMemsetAsync(dev_word, 0, sizeof(unsigned int), ctx->stream));
    hipLaunchKernelGGL(maxlen2_kernel, grid2(g), dim3(kBX, kBY), 0, ctx->stream, d, g, dev_word);
    TIPLHIP_CHECK(hipGetLastError());
    unsigned int bits = 0;
    read_back(ctx, dev_word, &bits, sizeof(bits));
    float v;
    std::memcpy(&v, &bits, sizeof(v));
    return v;
}

void launch_acc2(tiplhip_ctx* ctx, const float* cur, const float* nd, float k, const Geo2& g, float* out)
{
    hipLaunchKernelGGL(acc2_kernel, grid2(g), dim3(kBX, kBY), 0, ctx->stream, cur, nd, k, g, out);
    TIPLHIP_CHECK(hipGetLastError());
}

void launch_smooth2(tiplhip_ctx* ctx, const float* a, float smoothing, const Geo2& g, float* out)
{
    const float w_4 = smoothing / (2.0f * 2.0f);
    const float w_1 = 1.0f - smoothing;
    hipLaunchKernelGGL(smooth2_kernel, grid2(g), dim3(kBX, kBY), 0, ctx->stream, a, g, w_4, w_1, out);
    TIPLHIP_CHECK(hipGetLastError());
}

void launch_down2(tiplhip_ctx* ctx, const uint8_t* in, const Geo2& gi, uint8_t* out, const Geo2& go)
{
    hipLaunchKernelGGL(down2_u8_kernel, grid2(go), dim3(kBX, kBY), 0, ctx->stream, in, gi, out, go);
    TIPLHIP_CHECK(hipGetLastError());
}

void launch_up2(tiplhip_ctx* ctx, const float* in, const Geo2& gi, float scale, float* out, const Geo2& go)
{
    hipLaunchKernelGGL(up2_kernel, grid2(go), dim3(kBX, kBY), 0, ctx->stream, in, gi, scale, out, go);
    TIPLHIP_CHECK(hipGetLastError());
}

void copy2(tiplhip_ctx* ctx, void* dst, const void* src, size_t bytes)
{
    TIPLHIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, ctx->stream));
}

struct Cdm2Run {
    tiplhip_ctx* ctx;
    tiplhip_cdm_param param;
    int nmod;
    const volatile uint8_t* terminated;
    int iterations = 0;
    bool stopped = false;
};

// One call of cdm() (reg/cdm.hpp:373-474) at shape g: the coarser level
// first while min(w, h) > min_dimension, then this level's loop unless
// resolution > 1.  best is this level's field (zero on entry).
bool cdm2_level(Cdm2Run& run, const std::vector<const uint8_t*>& It,
                const std::vector<const uint8_t*>& Is, const Geo2& g, float* best, float resolution)
{
    tiplhip_ctx* ctx = run.ctx;
    const int nmod = run.nmod;
    if (std::min(g.w, g.h) > run.param.min_dimension) {
        const Geo2 rg = geo2((g.w + 1) / 2, (g.h + 1) / 2);
        DevMem rimg(ctx, size_t(2) * size_t(nmod) * rg.n);
        std::vector<const uint8_t*> rIt(static_cast<size_t>(nmod)), rIs(static_cast<size_t>(nmod));
        for (int i = 0; i < nmod; ++i) {
            uint8_t* t = rimg.as<uint8_t>() + size_t(2 * i) * rg.n;
            uint8_t* s = t + rg.n;
            launch_down2(ctx, It[size_t(i)], g, t, rg);
            launch_down2(ctx, Is[size_t(i)], g, s, rg);
            rIt[size_t(i)] = t;
            rIs[size_t(i)] = s;
        }
        DevMem rbest(ctx, sizeof(float) * 2 * rg.n);
        TIPLHIP_CHECK(hipMemsetAsync(rbest.p, 0, sizeof(float) * 2 * rg.n, ctx->stream));
        if (!cdm2_level(run, rIt, rIs, rg, rbest.as<float>(), resolution / 2.0f)) return false;
        launch_up2(ctx, rbest.as<float>(), rg, 2.0f, best, g);
        if (resolution > 1.0f) return true;
    }

    const size_t fb = sizeof(float) * 2 * g.n;
    DevMem cur(ctx, fb), nd(ctx, fb), acc(ctx, fb), pa(ctx, fb), pb(ctx, fb), Js(ctx, g.n);
    DevMem part(ctx, sizeof(double) * blocks2(g)), costs(ctx, sizeof(float) * size_t(nmod)),
        word(ctx, sizeof(unsigned int));
    // the cost reductions also store into ctx->h_cdm: if a step throws while
    // one is queued, drain the stream before the slots (and the buffers
    // above) can be reused
    auto drain = [ctx] { (void)hipStreamSynchronize(ctx->stream); };
    DrainOnUnwind<decltype(drain)> unwind(drain);
    ctx->h_cdm.reserve(sizeof(float) * 16);
    copy2(ctx, cur.p, best, fb);
    float theta = 0.0f;
    std::vector<float> cost;
    float best_cost = std::numeric_limits<float>::max();
    float cur_smoothing = 0.8f;
    std::vector<float> c(static_cast<size_t>(nmod));
    for (int index = 0; index < run.param.max_iterations; ++index) {
        if (run.terminated && *run.terminated) {
            run.stopped = true;
            break;
        }
        ++run.iterations;
        // Each modality's cost also goes to coherent host memory, whose slots
        // start as a NaN pattern the reduction never writes; the host spins
        // on them (a stream synchronisation after kCostSpinUs2).  The Jacobi
        // sweeps do not depend on the host's decision (the reference solves
        // the same new_d whatever cdm_improved says), so they are queued
        // before the wait and run during the round trip.
        uint32_t* slots = ctx->h_cdm.as<uint32_t>();
        std::fill(slots, slots + nmod, kCostPending2);
        float* h_cost = static_cast<float*>(ctx->h_cdm.dev);
        for (int i = 0; i < nmod; ++i) {
            launch_compose2(ctx, Is[size_t(i)], cur.as<float>(), g, Js.as<uint8_t>());
            const char type = run.param.gradient_type[i] == 'd' ? 'd' : 'r';
            launch_gradient2(ctx, Js.as<uint8_t>(), It[size_t(i)], g, type, i > 0, nd.as<float>(),
                             part.as<double>(), costs.as<float>() + i, h_cost + i);
        }
        float* solved = launch_poisson2(ctx, nd.as<float>(), pa.as<float>(), pb.as<float>(), g, 12);
        wait_for_slots(static_cast<const volatile uint32_t*>(slots), size_t(nmod), kCostPending2,
                       std::chrono::microseconds(kCostSpinUs2),
                       [ctx] { TIPLHIP_CHECK(hipStreamSynchronize(ctx->stream)); });
        for (int i = 0; i < nmod; ++i) std::memcpy(&c[size_t(i)], &slots[i], sizeof(float));
        float sum_cost = 0.0f;
        for (int i = 0; i < nmod; ++i) {
            sum_cost += c[size_t(i)];
            if (sum_cost == 0.0f) return false;
        }
        cost.push_back(sum_cost / float(nmod));
        if (cost.back() < best_cost) {
            best_cost = cost.back();
            if (index) copy2(ctx, best, cur.p, fb);
        }
        if (!cdm_improved(cost)) {
            if (cur_smoothing <= run.param.smoothing) break;
            copy2(ctx, cur.p, best, fb);
            cur_smoothing *= 0.5f;
            cost.clear();
        }
        if (theta == 0.0f) theta = max_length2(ctx, solved, g, word.as<unsigned int>());
        if (theta == 0.0f) break;
        const float k = run.param.speed / theta;
        launch_acc2(ctx, cur.as<float>(), solved, k, g, acc.as<float>());
        // cdm_smooth(new_d, cur_d, 0) leaves cur_d as it was
        if (cur_smoothing != 0.0f) launch_smooth2(ctx, acc.as<float>(), cur_smoothing, g, cur.as<float>());
    }
    return true;
}

// Device buffers of host images: one upload each.
struct Upload2 {
    DevMem mem;
    std::vector<const uint8_t*> It, Is;
    Upload2(tiplhip_ctx* ctx, const uint8_t* const* hIt, const uint8_t* const* hIs, int n, const Geo2& g)
        : mem(ctx, size_t(2) * size_t(n) * g.n)
    {
        for (int i = 0; i < n; ++i) {
            TIPLHIP_REQUIRE(hIt[i] && hIs[i], "null image");
            uint8_t* t = mem.as<uint8_t>() + size_t(2 * i) * g.n;
            uint8_t* s = t + g.n;
            TIPLHIP_CHECK(hipMemcpyAsync(t, hIt[i], g.n, hipMemcpyHostToDevice, ctx->stream));
            TIPLHIP_CHECK(hipMemcpyAsync(s, hIs[i], g.n, hipMemcpyHostToDevice, ctx->stream));
            It.push_back(t);
            Is.push_back(s);
        }
    }
};

}  // namespace
}  // namespace tiplhip

using namespace tiplhip;

extern "C" {

int tiplhip_cdm2d(tiplhip_ctx* ctx, const uint8_t* const* It, const uint8_t* const* Is, int n,
                  const int dim[2], const tiplhip_cdm_param* param, float* field,
                  const volatile uint8_t* terminated, int* iterations)
{
    return guarded([&] {
        TIPLHIP_REQUIRE(ctx && It && Is && field, "null argument");
        TIPLHIP_REQUIRE(n >= 1 && n <= 8, "between 1 and 8 modality pairs");
        check2(dim);
        std::lock_guard<std::recursive_mutex> lk(ctx->mu);
        DeviceGuard dg(ctx->device);
        Cdm2Run run{ctx, {}, n, terminated};
        if (param)
            run.param = *param;
        else
            tiplhip_cdm_param_default(&run.param);
        TIPLHIP_REQUIRE(run.param.max_iterations >= 0, "negative iteration cap");
        TIPLHIP_REQUIRE(run.param.levels == 0, "the fixed schedule is 3-D only");
        const Geo2 g = geo2(dim[0], dim[1]);
        Upload2 up(ctx, It, Is, n, g);
        const size_t fb = sizeof(float) * 2 * g.n;
        DevMem best(ctx, fb);
        TIPLHIP_CHECK(hipMemsetAsync(best.p, 0, fb, ctx->stream));
        const bool ok = cdm2_level(run, up.It, up.Is, g, best.as<float>(), run.param.resolution);
        if (iterations) *iterations = run.iterations;
        if (!ok) {
            TIPLHIP_CHECK(hipStreamSynchronize(ctx->stream));
            throw Error(TIPLHIP_ERR_FAILED, "cdm returned false (zero cost)");
        }
        TIPLHIP_CHECK(hipMemcpyAsync(field, best.p, fb, hipMemcpyDeviceToHost, ctx->stream));
        TIPLHIP_CHECK(hipStreamSynchronize(ctx->stream));
        // As in the reference, a terminated run still returns its best field.
        if (run.stopped) throw Error(TIPLHIP_ERR_TERMINATED, "terminated by the caller");
    });
}

// ---- single steps on host buffers (2-D) --------------------------------------

int tiplhip_cdm2d_compose(tiplhip_ctx* ctx, const uint8_t* Is, const int dim[2], const float* d,
                          uint8_t* Js)
{
    return guarded([&] {
        TIPLHIP_REQUIRE(ctx && Is && d && Js, "null argument");
        check2(dim);
        std::lock_guard<std::recursive_mutex> lk(ctx->mu);
        DeviceGuard dg(ctx->device);
        const Geo2 g = geo2(dim[0], dim[1]);
        DevMem dI(ctx, g.n), dD(ctx, sizeof(float) * 2 * g.n), dJ(ctx, g.n);
        TIPLHIP_CHECK(hipMemcpyAsync(dI.p, Is, g.n, hipMemcpyHostToDevice, ctx->stream));
        TIPLHIP_CHECK(hipMemcpyAsync(dD.p, d, sizeof(float) * 2 * g.n, hipMemcpyHostToDevice, ctx->stream));
        launch_compose2(ctx, dI.as<uint8_t>(), dD.as<float>(), g, dJ.as<uint8_t>());
        TIPLHIP_CHECK(hipMemcpyAsync(Js, dJ.p, g.n, hipMemcpyDeviceToHost, ctx->stream));
        TIPLHIP_CHECK(hipStreamSynchronize(ctx->stream));
    });
}

int tiplhip_cdm2d_gradient(tiplhip_ctx* ctx, const uint8_t* Js, const uint8_t* It, const int dim[2],
                           char type, float* new_d, float* cost)
{
    return guarded([&] {
        TIPLHIP_REQUIRE(ctx && Js && It && new_d, "null argument");
        TIPLHIP_REQUIRE(type == 'r' || type == 'd', "gradient type must be 'r' or 'd'");
        check2(dim);
        std::lock_guard<std::recursive_mutex> lk(ctx->mu);
        DeviceGuard dg(ctx->device);
        const Geo2 g = geo2(dim[0], dim[1]);
        DevMem dJ(ctx, g.n), dI(ctx, g.n), dN(ctx, sizeof(float) * 2 * g.n);
        DevMem part(ctx, sizeof(double) * blocks2(g)), dc(ctx, sizeof(float));
        TIPLHIP_CHECK(hipMemcpyAsync(dJ.p, Js, g.n, hipMemcpyHostToDevice, ctx->stream));
        TIPLHIP_CHECK(hipMemcpyAsync(dI.p, It, g.n, hipMemcpyHostToDevice, ctx->stream));
        launch_gradient2(ctx, dJ.as<uint8_t>(), dI.as<uint8_t>(), g, type, false, dN.as<float>(),
                         part.as<double>(), dc.as<float>());
        TIPLHIP_CHECK(hipMemcpyAsync(new_d, dN.p, sizeof(float) * 2 * g.n, hipMemcpyDeviceToHost, ctx->stream));
        float c = 0.0f;
        TIPLHIP_CHECK(hipMemcpyAsync(&c, dc.p, sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
        TIPLHIP_CHECK(hipStreamSynchronize(ctx->stream));
        if (cost) *cost = c;
    });
}

int tiplhip_cdm2d_poisson(tiplhip_ctx* ctx, float* new_d, const int dim[2], int sweeps,
                          float* max_length)
{
    return guarded([&] {
        TIPLHIP_REQUIRE(ctx && new_d, "null argument");
        TIPLHIP_REQUIRE(sweeps >= 1 && sweeps <= 1000, "sweeps must be in [1, 1000]");
        check2(dim);
        std::lock_guard<std::recursive_mutex> lk(ctx->mu);
        DeviceGuard dg(ctx->device);
        const Geo2 g = geo2(dim[0], dim[1]);
        const size_t fb = sizeof(float) * 2 * g.n;
        DevMem dN(ctx, fb), a(ctx, fb), b(ctx, fb), w(ctx, sizeof(unsigned int));
        TIPLHIP_CHECK(hipMemcpyAsync(dN.p, new_d, fb, hipMemcpyHostToDevice, ctx->stream));
        float* out = launch_poisson2(ctx, dN.as<float>(), a.as<float>(), b.as<float>(), g, sweeps);
        const float ml = max_length2(ctx, out, g, w.as<unsigned int>());
        TIPLHIP_CHECK(hipMemcpyAsync(new_d, out, fb, hipMemcpyDeviceToHost, ctx->stream));
        TIPLHIP_CHECK(hipStreamSynchronize(ctx->stream));
        if (max_length) *max_length = ml;
    });
}

int tiplhip_cdm2d_accumulate(tiplhip_ctx* ctx, const float* cur, float* new_d, const int dim[2],
                             float scale)
{
    return guarded([&] {
        TIPLHIP_REQUIRE(ctx && cur && new_d, "null argument");
        check2(dim);
        std::lock_guard<std::recursive_mutex> lk(ctx->mu);
        DeviceGuard dg(ctx->device);
        const Geo2 g = geo2(dim[0], dim[1]);
        const size_t fb = sizeof(float) * 2 * g.n;
        DevMem dC(ctx, fb), dN(ctx, fb), dO(ctx, fb);
        TIPLHIP_CHECK(hipMemcpyAsync(dC.p, cur, fb, hipMemcpyHostToDevice, ctx->stream));
        TIPLHIP_CHECK(hipMemcpyAsync(dN.p, new_d, fb, hipMemcpyHostToDevice, ctx->stream));
        launch_acc2(ctx, dC.as<float>(), dN.as<float>(), scale, g, dO.as<float>());
        TIPLHIP_CHECK(hipMemcpyAsync(new_d, dO.p, fb, hipMemcpyDeviceToHost, ctx->stream));
        TIPLHIP_CHECK(hipStreamSynchronize(ctx->stream));
    });
}

int tiplhip_cdm2d_smooth(tiplhip_ctx* ctx, const float* d, float* dd, const int dim[2], float smoothing)
{
    return guarded([&] {
        TIPLHIP_REQUIRE(ctx && d && dd, "null argument");
        check2(dim);
        if (smoothing == 0.0f) return;  // cdm_smooth leaves dd untouched
        std::lock_guard<std::recursive_mutex> lk(ctx->mu);
        DeviceGuard dg(ctx->device);
        const Geo2 g = geo2(dim[0], dim[1]);
        const size_t fb = sizeof(float) * 2 * g.n;
        DevMem dD(ctx, fb), dO(ctx, fb);
        TIPLHIP_CHECK(hipMemcpyAsync(dD.p, d, fb, hipMemcpyHostToDevice, ctx->stream));
        launch_smooth2(ctx, dD.as<float>(), smoothing, g, dO.as<float>());
        TIPLHIP_CHECK(hipMemcpyAsync(dd, dO.p, fb, hipMemcpyDeviceToHost, ctx->stream));
        TIPLHIP_CHECK(hipStreamSynchronize(ctx->stream));
    });
}

int tiplhip_cdm2d_downsample(tiplhip_ctx* ctx, const uint8_t* in, const int dim[2], uint8_t* out)
{
    return guarded([&] {
        TIPLHIP_REQUIRE(ctx && in && out, "null argument");
        check2(dim);
        std::lock_guard<std::recursive_mutex> lk(ctx->mu);
        DeviceGuard dg(ctx->device);
        const Geo2 gi = geo2(dim[0], dim[1]), go = geo2((dim[0] + 1) / 2, (dim[1] + 1) / 2);
        DevMem dI(ctx, gi.n), dO(ctx, go.n);
        TIPLHIP_CHECK(hipMemcpyAsync(dI.p, in, gi.n, hipMemcpyHostToDevice, ctx->stream));
        launch_down2(ctx, dI.as<uint8_t>(), gi, dO.as<uint8_t>(), go);
        TIPLHIP_CHECK(hipMemcpyAsync(out, dO.p, go.n, hipMemcpyDeviceToHost, ctx->stream));
        TIPLHIP_CHECK(hipStreamSynchronize(ctx->stream));
    });
}

int tiplhip_cdm2d_upsample(tiplhip_ctx* ctx, const float* in, const int in_dim[2], float* out,
                           const int out_dim[2])
{
    return guarded([&] {
        TIPLHIP_REQUIRE(ctx && in && out, "null argument");
        check2(in_dim);
        check2(out_dim);
        std::lock_guard<std::recursive_mutex> lk(ctx->mu);
        DeviceGuard dg(ctx->device);
        const Geo2 gi = geo2(in_dim[0], in_dim[1]), go = geo2(out_dim[0], out_dim[1]);
        DevMem dI(ctx, sizeof(float) * 2 * gi.n), dO(ctx, sizeof(float) * 2 * go.n);
        TIPLHIP_CHECK(hipMemcpyAsync(dI.p, in, sizeof(float) * 2 * gi.n, hipMemcpyHostToDevice, ctx->stream));
        launch_up2(ctx, dI.as<float>(), gi, 1.0f, dO.as<float>(), go);
        TIPLHIP_CHECK(hipMemcpyAsync(out, dO.p, sizeof(float) * 2 * go.n, hipMemcpyDeviceToHost, ctx->stream));
        TIPLHIP_CHECK(hipStreamSynchronize(ctx->stream));
    });
}

}  // extern "C"
