// cdm.hip -- the CDM nonlinear registration loop on gfx950.
//
// Restates, MI355X-first:
//   compose_displacement        numerical/dif.hpp:90-115, 260-285
//   cdm_get_gradient (r, d)     reg/cdm.hpp:46-152 with get_window_at_width<2>
//                               (numerical/window.hpp:92-110), linear_regression
//                               (numerical/statistics.hpp:694-713), gradient_at
//                               (numerical/numerical.hpp:102-109)
//   cdm_solve_poisson           reg/cdm.hpp:158-254
//   cdm_max_displacement_length reg/cdm.hpp:260-296
//   accumulate_displacement     numerical/dif.hpp:166-186, 311-329
//   cdm_smooth                  reg/cdm.hpp:302-367
//   upsample_with_padding       numerical/resampling.hpp:849-891
//   cdm / cdm_cuda              reg/cdm.hpp:373-532
//
// Fields are packed float3 per voxel (the reference's image<3,vector<3>>).
// Kernels are launched on a 3-D grid of 64 x 4 thread tiles (x = one wave),
// so no index division is needed.  The per-iteration control (cost history,
// cdm_improved, smoothing schedule) runs on the host from one float cost
// read back per iteration; everything else stays on the device.
#include "cdm_control.hpp"
#include "device.hpp"
#include "field.hpp"
#include "host_wait.hpp"
#include "internal.hpp"
#include "step_gate.hpp"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <map>
#include <memory>
#include <mutex>
#include <vector>

namespace tiplhip {
namespace {

// ---------------------------------------------------------------------------
// compose_displacement: Js(x) = Is(x + d(x)); exact copy where d(x) == 0.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void compose_body(const uint8_t* __restrict__ Is,
                                             const float* __restrict__ d, const FieldGeo& g,
                                             uint8_t* __restrict__ Js, const Block& blk)
{
    const Vox v = voxel_of(g, blk);
    if (!v.ok) return;
    const float3 dv = ld3(d, v.p);
    if (dv.x == 0.0f && dv.y == 0.0f && dv.z == 0.0f) {
        Js[v.p] = Is[v.p];
        return;
    }
    const __amdgpu_buffer_rsrc_t r = make_rsrc(Is, g.n);
    VolGeo vg{g.w, g.h, g.d, g.wh, g.wm1, g.hm1, g.dm1};
    Js[v.p] = uint8_t(sample_u8(r, vg, float(v.x) + dv.x, float(v.y) + dv.y, float(v.z) + dv.z));
}

__global__ __launch_bounds__(256) void compose_kernel(const uint8_t* __restrict__ Is,
                                                      const float* __restrict__ d, FieldGeo g,
                                                      uint8_t* __restrict__ Js)
{
    compose_body(Is, d, g, Js, logical_block());
}

// ---------------------------------------------------------------------------
// cdm_get_gradient.  'r': local linear regression of Js onto It over the
// border-clipped 5x5x5 window (exact integer sums, then FP64 as in
// linear_regression), gradient scaled by r2*(a*Js+b-It) where a > 0.
// 'd': SSD gradient.  accumulate = 0 writes nd = 0 + g (a fresh field);
// accumulate = 1 adds a further modality: nd = nd + (0 + g).
// Per-block FP64 partial sums of the cost map go to `partials`.
// ---------------------------------------------------------------------------
template <char TYPE>
__global__ __launch_bounds__(256) void gradient_kernel(const uint8_t* __restrict__ Js,
                                                       const uint8_t* __restrict__ It, FieldGeo g,
                                                       int accumulate, float* __restrict__ nd,
                                                       double* __restrict__ partials)
{
    const Vox v = this_voxel(g);
    float3 gr = make_float3(0.0f, 0.0f, 0.0f);
    float cost = 0.0f;
    if (v.ok) {
        const uint32_t js = Js[v.p], it = It[v.p];
        bool active;
        float s = 0.0f;
        if (TYPE == 'd') {
            const int dif = int(js) - int(it);
            s = float(dif);
            cost = float(dif * dif);
            active = true;
        } else {
            active = false;
            if (js != 0u && it != 0u) {
                const int x0 = max(v.x - 2, 0), x1 = min(v.x + 2, g.w - 1);
                const int y0 = max(v.y - 2, 0), y1 = min(v.y + 2, g.h - 1);
                const int z0 = max(v.z - 2, 0), z1 = min(v.z + 2, g.d - 1);
                const __amdgpu_buffer_rsrc_t rj = make_rsrc(Js, g.n);
                const __amdgpu_buffer_rsrc_t ri = make_rsrc(It, g.n);
                uint32_t sx = 0, sxx = 0, sy = 0, syy = 0, sxy = 0;
                for (int zz = z0; zz <= z1; ++zz)
                    for (int yy = y0; yy <= y1; ++yy) {
                        const uint32_t row = uint32_t(zz) * g.wh + uint32_t(yy) * uint32_t(g.w);
                        for (int xx = x0; xx <= x1; ++xx) {
                            const uint32_t a = load_u8(rj, row + uint32_t(xx), 0);
                            const uint32_t b = load_u8(ri, row + uint32_t(xx), 0);
                            sx += a;
                            sxx += a * a;
                            sy += b;
                            syy += b * b;
                            sxy += a * b;
                        }
                    }
                const double dn = double((x1 - x0 + 1) * (y1 - y0 + 1) * (z1 - z0 + 1));
                const double mean_x = double(sx) / dn;
                const double mean_y = double(sy) / dn;
                const double x_var = double(sxx) / dn - mean_x * mean_x;
                const double y_var = double(syy) / dn - mean_y * mean_y;
                if (!(x_var == 0.0 || y_var == 0.0)) {
                    const double cov = double(sxy) / dn - mean_x * mean_y;
                    const float a = float(cov / x_var);
                    if (a > 0.0f) {
                        const float b = float(mean_y - double(a) * mean_x);
                        const float r2 = float(cov * cov / x_var / y_var);
                        s = float(js) * a;
                        s = s + b;
                        s = s - float(it);
                        s = r2 * s;
                        cost = -r2;
                        active = true;
                    }
                }
            }
        }
        if (active) {
            // gradient_at: central differences, 0 on the first/last slab.
            const uint32_t w = uint32_t(g.w);
            const float gx = (v.x + 1 < g.w && v.x) ? float(int(Js[v.p + 1]) - int(Js[v.p - 1])) : 0.0f;
            const float gy = (v.y + 1 < g.h && v.y) ? float(int(Js[v.p + w]) - int(Js[v.p - w])) : 0.0f;
            const float gz = (v.z + 1 < g.d && v.z) ? float(int(Js[v.p + g.wh]) - int(Js[v.p - g.wh])) : 0.0f;
            gr.x = 0.0f + gx * s;
            gr.y = 0.0f + gy * s;
            gr.z = 0.0f + gz * s;
        }
        if (accumulate) {
            float3 o = ld3(nd, v.p);
            o.x = o.x + gr.x;
            o.y = o.y + gr.y;
            o.z = o.z + gr.z;
            st3(nd, v.p, o);
        } else {
            st3(nd, v.p, gr);
        }
    }
    // deterministic per-block partial sum of the cost map
    __shared__ double part[4];
    double c = wave_sum_f64(double(cost));
    const uint32_t tid = threadIdx.y * blockDim.x + threadIdx.x;
    if ((tid & 63u) == 0) part[tid >> 6] = c;
    __syncthreads();
    if (tid == 0) partials[logical_block().lin] = (part[0] + part[1]) + (part[2] + part[3]);
}

// ---------------------------------------------------------------------------
// cdm_get_gradient 'r' with the window sums done as box filters.  A block
// owns a 64 x 4 column of outputs over kZC planes and marches z: for every
// plane it stages Js and It (with a 2-voxel halo, zero outside the volume)
// in LDS, forms 5-tap x sums, then 5-tap y sums, and keeps the last five
// plane sums in registers for the z sum.  Zero padding leaves the five
// integer sums of the border-clipped window unchanged; the window size is
// the product of the clipped extents, as get_window_at_width counts it.
// ---------------------------------------------------------------------------
constexpr int kZC = 16;

// A quiet-NaN float pattern the cost reduction never writes, and how long
// the host spins for an iteration's cost before it synchronises the stream.
constexpr uint32_t kCostPending = 0x7FC0C057u;
constexpr int kCostSpinUs = 2000;
constexpr int kPX = kTX + 4, kPY = kTY + 4;

struct Sums5 {
    uint32_t j, i, jj, ii, ji;
};

__device__ inline void regress_and_store(const uint8_t* __restrict__ Js, const FieldGeo& g,
                                         const Vox& v, uint32_t js, uint32_t it, const Sums5& s,
                                         int accumulate, float* __restrict__ nd, float& cost)
{
    float3 gr = make_float3(0.0f, 0.0f, 0.0f);
    if (js != 0u && it != 0u) {
        const int nx = min(v.x + 2, g.w - 1) - max(v.x - 2, 0) + 1;
        const int ny = min(v.y + 2, g.h - 1) - max(v.y - 2, 0) + 1;
        const int nz = min(v.z + 2, g.d - 1) - max(v.z - 2, 0) + 1;
        const double dn = double(nx * ny * nz);
        const double mean_x = double(s.j) / dn;
        const double mean_y = double(s.i) / dn;
        const double x_var = double(s.jj) / dn - mean_x * mean_x;
        const double y_var = double(s.ii) / dn - mean_y * mean_y;
        if (!(x_var == 0.0 || y_var == 0.0)) {
            const double cov = double(s.ji) / dn - mean_x * mean_y;
            const float a = float(cov / x_var);
            if (a > 0.0f) {
                const float b = float(mean_y - double(a) * mean_x);
                const float r2 = float(cov * cov / x_var / y_var);
                float sc = float(js) * a;
                sc = sc + b;
                sc = sc - float(it);
                sc = r2 * sc;
                cost = -r2;
                const uint32_t w = uint32_t(g.w);
                const float gx =
                    (v.x + 1 < g.w && v.x) ? float(int(Js[v.p + 1]) - int(Js[v.p - 1])) : 0.0f;
                const float gy =
                    (v.y + 1 < g.h && v.y) ? float(int(Js[v.p + w]) - int(Js[v.p - w])) : 0.0f;
                const float gz =
                    (v.z + 1 < g.d && v.z) ? float(int(Js[v.p + g.wh]) - int(Js[v.p - g.wh])) : 0.0f;
                gr.x = 0.0f + gx * sc;
                gr.y = 0.0f + gy * sc;
                gr.z = 0.0f + gz * sc;
            }
        }
    }
    if (accumulate) {
        float3 o = ld3(nd, v.p);
        o.x = o.x + gr.x;
        o.y = o.y + gr.y;
        o.z = o.z + gr.z;
        st3(nd, v.p, o);
    } else {
        st3(nd, v.p, gr);
    }
}

constexpr int kStage = (kPX * kPY + kTX * kTY - 1) / (kTX * kTY);  // staged bytes per thread

__device__ __forceinline__ void gradient_r_box_body(const uint8_t* __restrict__ Js,
                                                    const uint8_t* __restrict__ It,
                                                    const FieldGeo& g, int accumulate,
                                                    float* __restrict__ nd,
                                                    double* __restrict__ partials, int zc,
                                                    const Block& blk)
{
    __shared__ uint8_t pj[kPY][kPX], pi[kPY][kPX];
    __shared__ Sums5 xs[kPY][kTX];
    const int tx = int(threadIdx.x), ty = int(threadIdx.y);
    const int tid = ty * kTX + tx;
    const int x0 = blk.x * kTX, y0 = blk.y * kTY;
    const int z0 = blk.z * zc, z1 = min(z0 + zc, g.d);
    Sums5 ring[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) ring[k] = Sums5{0u, 0u, 0u, 0u, 0u};
    float cost_acc = 0.0f;
    double cost_sum = 0.0;
    // This thread's staging slots (plane-independent): LDS cell c = tid + 256 k
    // of the 2-voxel-haloed tile, its in-plane offset, and whether it lies
    // inside the volume (zero outside).  The next plane's bytes are loaded
    // into registers while the current plane is summed.
    uint32_t soff[kStage];
    bool sin[kStage];
#pragma unroll
    for (int k = 0; k < kStage; ++k) {
        const int c = tid + k * kTX * kTY;
        const int cy = c / kPX, cx = c - cy * kPX;
        const int gx = x0 - 2 + cx, gy = y0 - 2 + cy;
        sin[k] = c < kPX * kPY && gx >= 0 && gx < g.w && gy >= 0 && gy < g.h;
        soff[k] = sin[k] ? uint32_t(gy) * uint32_t(g.w) + uint32_t(gx) : 0u;
    }
    uint8_t na[kStage], nb[kStage];
    {
        const int pz = z0 - 2;
#pragma unroll
        for (int k = 0; k < kStage; ++k) {
            const bool in = sin[k] && pz >= 0 && pz < g.d;
            const uint32_t q = uint32_t(in ? pz : 0) * g.wh + soff[k];
            na[k] = in ? Js[q] : uint8_t(0);
            nb[k] = in ? It[q] : uint8_t(0);
        }
    }
    for (int pz = z0 - 2; pz < z1 + 2; ++pz) {
        Sums5 P{0u, 0u, 0u, 0u, 0u};
        const bool plane_in = pz >= 0 && pz < g.d;
        if (plane_in) {
            // stage the plane with a 2-voxel halo (zero outside the volume)
#pragma unroll
            for (int k = 0; k < kStage; ++k) {
                const int c = tid + k * kTX * kTY;
                if (c < kPX * kPY) {
                    const int cy = c / kPX, cx = c - cy * kPX;
                    pj[cy][cx] = na[k];
                    pi[cy][cx] = nb[k];
                }
            }
        }
        {
            const int nz = pz + 1;
            if (nz < z1 + 2) {
#pragma unroll
                for (int k = 0; k < kStage; ++k) {
                    const bool in = sin[k] && nz >= 0 && nz < g.d;
                    const uint32_t q = uint32_t(in ? nz : 0) * g.wh + soff[k];
                    na[k] = in ? Js[q] : uint8_t(0);
                    nb[k] = in ? It[q] : uint8_t(0);
                }
            }
        }
        if (plane_in) {
            __syncthreads();
            // x sums for the 8 staged rows (2 per thread)
            for (int r = ty; r < kPY; r += kTY) {
                Sums5 s{0u, 0u, 0u, 0u, 0u};
#pragma unroll
                for (int k = 0; k < 5; ++k) {
                    const uint32_t a = pj[r][tx + k], b = pi[r][tx + k];
                    s.j += a;
                    s.i += b;
                    s.jj += a * a;
                    s.ii += b * b;
                    s.ji += a * b;
                }
                xs[r][tx] = s;
            }
            __syncthreads();
#pragma unroll
            for (int k = 0; k < 5; ++k) {
                const Sums5 s = xs[ty + k][tx];
                P.j += s.j;
                P.i += s.i;
                P.jj += s.jj;
                P.ii += s.ii;
                P.ji += s.ji;
            }
            __syncthreads();  // pj/pi/xs are rewritten for the next plane
        }
#pragma unroll
        for (int k = 0; k < 4; ++k) ring[k] = ring[k + 1];
        ring[4] = P;
        const int z = pz - 2;
        if (z >= z0) {
            Sums5 S = ring[0];
#pragma unroll
            for (int k = 1; k < 5; ++k) {
                S.j += ring[k].j;
                S.i += ring[k].i;
                S.jj += ring[k].jj;
                S.ii += ring[k].ii;
                S.ji += ring[k].ji;
            }
            Vox v;
            v.x = x0 + tx;
            v.y = y0 + ty;
            v.z = z;
            v.ok = v.x < g.w && v.y < g.h;
            v.p = uint32_t(v.z) * g.wh + uint32_t(v.y) * uint32_t(g.w) + uint32_t(v.x);
            if (v.ok) {
                float cost = 0.0f;
                regress_and_store(Js, g, v, Js[v.p], It[v.p], S, accumulate, nd, cost);
                cost_sum += double(cost);
            }
        }
    }
    (void)cost_acc;
    __shared__ double part[4];
    double c = wave_sum_f64(cost_sum);
    if ((tid & 63) == 0) part[tid >> 6] = c;
    __syncthreads();
    if (tid == 0) partials[blk.lin] = (part[0] + part[1]) + (part[2] + part[3]);
}

__global__ __launch_bounds__(256) void gradient_r_box_kernel(const uint8_t* __restrict__ Js,
                                                             const uint8_t* __restrict__ It,
                                                             FieldGeo g, int accumulate,
                                                             float* __restrict__ nd,
                                                             double* __restrict__ partials, int zc)
{
    gradient_r_box_body(Js, It, g, accumulate, nd, partials, zc, logical_block());
}

// Planes per block of the box-sum gradient: kZC for large volumes; on
// small pyramid levels fewer (down to 4), so about 1024 blocks share the
// work instead of a few blocks marching long z chains.
int box_zc(const FieldGeo& g)
{
    const long tiles = long((g.w + kTX - 1) / kTX) * long((g.h + kTY - 1) / kTY);
    return int(std::min<long>(kZC, std::max<long>(4, long(g.d) * tiles / 1024)));
}

dim3 box_grid(const FieldGeo& g)
{
    const int zc = box_zc(g);
    return dim3(uint32_t((g.w + kTX - 1) / kTX), uint32_t((g.h + kTY - 1) / kTY),
                uint32_t((g.d + zc - 1) / zc));
}

// Sums the block partials in a fixed order and stores float(sum / n).
__device__ __forceinline__ void cost_reduce_body(const double* __restrict__ partials,
                                                 uint32_t count, double n, float* __restrict__ out,
                                                 float* __restrict__ host_out)
{
    double acc = 0.0;
    for (uint32_t i = threadIdx.x; i < count; i += blockDim.x) acc += partials[i];
    __shared__ double part[16];
    acc = wave_sum_f64(acc);
    if ((threadIdx.x & 63u) == 0) part[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (uint32_t w = 0; w < (blockDim.x >> 6); ++w) s += part[w];
        out[0] = float(s / n);
        if (host_out) host_out[0] = float(s / n);
    }
}

__global__ __launch_bounds__(1024) void cost_reduce_kernel(const double* __restrict__ partials,
                                                           uint32_t count, double n,
                                                           float* __restrict__ out,
                                                           float* __restrict__ host_out)
{
    cost_reduce_body(partials, count, n, out, host_out);
}

// Max of the per-block maxima (non-negative floats), as float bits.
__device__ __forceinline__ void max_reduce_body(const float* __restrict__ parts, uint32_t count,
                                                unsigned int* __restrict__ out,
                                                unsigned int* __restrict__ host_out)
{
    float m = 0.0f;
    for (uint32_t i = threadIdx.x; i < count; i += blockDim.x) m = fmaxf(m, parts[i]);
    __shared__ float part[16];
    m = wave_max_f32(m);
    if ((threadIdx.x & 63u) == 0) part[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        float r = 0.0f;
        for (uint32_t w = 0; w < (blockDim.x >> 6); ++w) r = fmaxf(r, part[w]);
        out[0] = __float_as_uint(r);
        if (host_out) host_out[0] = __float_as_uint(r);
    }
}

__global__ __launch_bounds__(1024) void max_reduce_kernel(const float* __restrict__ parts,
                                                          uint32_t count,
                                                          unsigned int* __restrict__ out)
{
    max_reduce_body(parts, count, out, nullptr);
}

// ---------------------------------------------------------------------------
// cdm_solve_poisson: one Jacobi sweep, 7-point stencil with neighbours taken
// by flat index and dropped at the array ends.  FIRST: the previous iterate
// is nd * (-1/6), formed on the fly.  MAXLEN: also the max of
// float(sqrt(double(|v|^2))) (cdm_max_displacement_length) as float bits.
// ---------------------------------------------------------------------------
constexpr float kInvD2 = -0.5f / 3.0f;

template <bool FIRST, bool MAXLEN>
__global__ __launch_bounds__(256) void poisson_kernel(const float* __restrict__ nd,
                                                      const float* __restrict__ a, FieldGeo g,
                                                      float* __restrict__ b,
                                                      float* __restrict__ maxparts)
{
    const Vox v = this_voxel(g);
    float len = 0.0f;
    if (v.ok) {
        auto A = [&](uint32_t q) -> float3 {
            if (FIRST) {
                float3 t = ld3(nd, q);
                return make_float3(t.x * kInvD2, t.y * kInvD2, t.z * kInvD2);
            }
            return ld3(a, q);
        };
        const uint32_t p = v.p, n = g.n, w = uint32_t(g.w), wh = g.wh;
        float3 r = ld3(nd, p);
        auto sub = [&](float3 t) {
            r.x = r.x - t.x;
            r.y = r.y - t.y;
            r.z = r.z - t.z;
        };
        if (p >= 1u) sub(A(p - 1));
        if (p < n - 1u) sub(A(p + 1));
        if (p >= w) sub(A(p - w));
        if (p < n - w) sub(A(p + w));
        if (p >= wh) sub(A(p - wh));
        if (p < n - wh) sub(A(p + wh));
        r.x = r.x * kInvD2;
        r.y = r.y * kInvD2;
        r.z = r.z * kInvD2;
        st3(b, p, r);
        if (MAXLEN) {
            float s2 = r.x * r.x + r.y * r.y;
            s2 = s2 + r.z * r.z;
            len = float(sqrt(double(s2)));
        }
    }
    if (MAXLEN) {
        __shared__ float part[4];
        const uint32_t tid = threadIdx.y * blockDim.x + threadIdx.x;
        float m = wave_max_f32(len);
        if ((tid & 63u) == 0) part[tid >> 6] = m;
        __syncthreads();
        if (tid == 0) {
            maxparts[logical_block().lin] = fmaxf(fmaxf(part[0], part[1]), fmaxf(part[2], part[3]));
        }
    }
}

// ---------------------------------------------------------------------------
// The same Jacobi sweep, marching z.  A block owns a 64 x 8 column of voxels
// over kPZ planes; each thread keeps the previous iterate of its own column
// at z-1, z, z+1 in registers (the +-wh neighbours), takes x-1 / x+1 from
// the neighbouring lanes of its wave, y-1 / y+1 from the block's LDS copy of
// the plane, and loads the neighbours that fall outside the tile by flat
// index -- which is exactly the reference's flat-index neighbourhood, row and
// plane wrap included.  Every value is combined in the reference's order.
// ---------------------------------------------------------------------------
constexpr int kPTY = 8, kPZ = 16;

dim3 poisson_grid(const FieldGeo& g)
{
    return dim3(uint32_t((g.w + kTX - 1) / kTX), uint32_t((g.h + kPTY - 1) / kPTY),
                uint32_t((g.d + kPZ - 1) / kPZ));
}

template <bool FIRST, bool MAXLEN>
__global__ __launch_bounds__(kTX * kPTY) void poisson_zm_kernel(const float* __restrict__ nd,
                                                                const float* __restrict__ a,
                                                                FieldGeo g, float* __restrict__ b,
                                                                float* __restrict__ maxparts)
{
    __shared__ float3 plane[kPTY][kTX];
    const int tx = int(threadIdx.x), ty = int(threadIdx.y);
    const Block blk = logical_block();
    const int x = blk.x * kTX + tx, y = blk.y * kPTY + ty;
    const int z0 = blk.z * kPZ, z1 = min(z0 + kPZ, g.d);
    const bool ok = x < g.w && y < g.h;
    const uint32_t n = g.n, w = uint32_t(g.w), wh = g.wh;
    auto A = [&](uint32_t q) -> float3 {
        if (FIRST) {
            const float3 t = ld3(nd, q);
            return make_float3(t.x * kInvD2, t.y * kInvD2, t.z * kInvD2);
        }
        return ld3(a, q);
    };
    const float3 zero3 = make_float3(0.0f, 0.0f, 0.0f);
    uint32_t p = uint32_t(z0) * wh + uint32_t(y) * w + uint32_t(x);
    float3 am = (ok && z0 >= 1) ? A(p - wh) : zero3;
    float3 ac = ok ? A(p) : zero3;
    float len_max = 0.0f;
    for (int z = z0; z < z1; ++z) {
        const float3 ap = (ok && z + 1 < g.d) ? A(p + wh) : zero3;
        plane[ty][tx] = ac;
        // x neighbours from the adjacent lanes of this row's wave
        const float lx0 = __shfl_up(ac.x, 1, kTX), lx1 = __shfl_up(ac.y, 1, kTX),
                    lx2 = __shfl_up(ac.z, 1, kTX);
        const float rx0 = __shfl_down(ac.x, 1, kTX), rx1 = __shfl_down(ac.y, 1, kTX),
                    rx2 = __shfl_down(ac.z, 1, kTX);
        __syncthreads();
        if (ok) {
            float3 r = ld3(nd, p);
            auto sub = [&](float3 t) {
                r.x = r.x - t.x;
                r.y = r.y - t.y;
                r.z = r.z - t.z;
            };
            if (p >= 1u) sub(tx > 0 ? make_float3(lx0, lx1, lx2) : A(p - 1u));
            if (p < n - 1u) sub((tx < kTX - 1 && x + 1 < g.w) ? make_float3(rx0, rx1, rx2) : A(p + 1u));
            if (p >= w) sub(ty > 0 ? plane[ty - 1][tx] : A(p - w));
            if (p < n - w) sub((ty < kPTY - 1 && y + 1 < g.h) ? plane[ty + 1][tx] : A(p + w));
            if (p >= wh) sub(am);
            if (p < n - wh) sub(ap);
            r.x = r.x * kInvD2;
            r.y = r.y * kInvD2;
            r.z = r.z * kInvD2;
            st3(b, p, r);
            if (MAXLEN) {
                float s2 = r.x * r.x + r.y * r.y;
                s2 = s2 + r.z * r.z;
                len_max = fmaxf(len_max, float(sqrt(double(s2))));
            }
        }
        __syncthreads();  // the plane buffer is rewritten next step
        am = ac;
        ac = ap;
        p += wh;
    }
    if (MAXLEN) {
        __shared__ float part[kPTY];
        const int tid = ty * kTX + tx;
        const float m = wave_max_f32(len_max);
        if ((tid & 63) == 0) part[tid >> 6] = m;
        __syncthreads();
        if (tid == 0) {
            float r = part[0];
            for (int i = 1; i < kPTY; ++i) r = fmaxf(r, part[i]);
            maxparts[blk.lin] = r;
        }
    }
}

// ---------------------------------------------------------------------------
// Two Jacobi sweeps per launch (s_t -> s_t+1 -> s_t+2), marching z.
//
// Tiles are laid out in flat-index space: the voxel at offset (i, j) of a
// tile of plane m is flat index P0(m) + j*w + i, P0(m) the flat index of the
// tile's first voxel.  The reference's neighbours p-1, p+1, p-w, p+w are then
// offsets (i-1, j), (i+1, j), (i, j-1), (i, j+1) -- row and plane wrap
// included -- and p-+wh is the same offset in planes m-+1.  A neighbour
// outside [0, n) is dropped by the reference; here it holds +0.0, and
// r - (+0.0) == r for every float r, so the results are bit-identical.
//
// Per plane m the block holds s_t on its tile grown by 2 (S ring, 3 planes),
// computes s_t+1 on the tile grown by 1 (T ring, 3 planes), then s_t+2 of
// plane m-1 on the tile itself.  The loop is software-pipelined: while it
// works on plane m from LDS, each thread's global loads for plane m+2 of s_t
// and for the new_d values of plane m+1 are in flight in registers.  Thread
// (tx, ty) computes s_t+1 at its own voxel (tx, ty) and keeps that new_d
// value for its s_t+2 output one plane later; threads 0..147 also compute
// the 148 positions of the ring around the tile.  HBM traffic per two
// sweeps: new_d and s_t read once, s_t+2 written once (36 B per voxel
// instead of 72).  A workgroup covers zq planes; small volumes use a small
// zq so that they still get enough workgroups.
// ---------------------------------------------------------------------------
constexpr int kQX = 64, kQY = 8, kQZMax = 16;
constexpr int kS2X = kQX + 4, kS2Y = kQY + 4;  // tile grown by 2
constexpr int kS1X = kQX + 2, kS1Y = kQY + 2;  // tile grown by 1
constexpr int kRing = kS1X * kS1Y - kQX * kQY;  // 148 positions around the tile
constexpr int kS2N = kS2X * kS2Y;               // 816

// planes per workgroup: about 512 workgroups or more, 2..16 planes each
int poisson2_zq(const FieldGeo& g)
{
    const long long tiles = (long long)((g.w + kQX - 1) / kQX) * ((g.h + kQY - 1) / kQY);
    const long long want = std::max<long long>(1, (tiles * g.d + 511) / 512);
    return int(std::min<long long>(kQZMax, std::max<long long>(2, want)));
}

dim3 poisson2_grid(const FieldGeo& g, int zq)
{
    return dim3(uint32_t((g.w + kQX - 1) / kQX), uint32_t((g.h + kQY - 1) / kQY),
                uint32_t((g.d + zq - 1) / zq));
}

// offset (i, j) of ring position r (-1 <= i <= kQX, -1 <= j <= kQY, not inside)
__device__ inline void ring_offset(int r, int& i, int& j)
{
    if (r < kS1X) { i = r - 1; j = -1; }
    else if (r < 2 * kS1X) { i = r - kS1X - 1; j = kQY; }
    else if (r < 2 * kS1X + kQY) { i = -1; j = r - 2 * kS1X; }
    else { i = kQX; j = r - 2 * kS1X - kQY; }
}

template <bool FIRST, bool MAXLEN>
__device__ __forceinline__ void poisson2_body(const float* __restrict__ nd, const float* __restrict__ a,
                                              const FieldGeo& g, int zq, float* __restrict__ b,
                                              float* __restrict__ maxparts, const Block& blk)
{
    __shared__ float3 S[3][kS2Y][kS2X];
    __shared__ float3 T[3][kS1Y][kS1X];
    const int tx = int(threadIdx.x), ty = int(threadIdx.y), tid = ty * kQX + tx;
    const int x0 = blk.x * kQX, y0 = blk.y * kQY;
    const int z0 = blk.z * zq, z1 = min(z0 + zq, g.d);
    const long long n = g.n, w = g.w, wh = g.wh;
    const long long base = (long long)y0 * w + x0;
    const float3 zero = make_float3(0.0f, 0.0f, 0.0f);

    // this thread's S positions (two at most) and its ring position
    const int e0 = tid, e1 = tid + kQX * kQY;
    const long long so0 = (long long)(e0 / kS2X - 2) * w + (e0 % kS2X - 2);
    const long long so1 = (long long)(e1 / kS2X - 2) * w + (e1 % kS2X - 2);
    const bool has1 = e1 < kS2N;
    int ri = 0, rj = 0;
    const bool has_ring = tid < kRing;
    if (has_ring) ring_offset(tid, ri, rj);
    const long long ro = (long long)rj * w + ri;
    const long long io = (long long)ty * w + tx;  // this thread's own voxel

    auto load_s = [&](long long q) -> float3 {
        if (q < 0 || q >= n) return make_float3(0.0f, 0.0f, 0.0f);
        const float3 t = ld3(FIRST ? nd : a, uint32_t(q));
        if (FIRST) return make_float3(t.x * kInvD2, t.y * kInvD2, t.z * kInvD2);
        return t;
    };
    auto load_r = [&](long long q) -> float3 {
        if (q < 0 || q >= n) return make_float3(0.0f, 0.0f, 0.0f);
        return ld3(nd, uint32_t(q));
    };
    auto valid = [&](long long q) { return q >= 0 && q < n; };
    // one Jacobi update in the reference's order from the six neighbours
    auto update = [](float3 r, float3 xm, float3 xp, float3 ym, float3 yp, float3 zm, float3 zp) {
        r.x = r.x - xm.x; r.y = r.y - xm.y; r.z = r.z - xm.z;
        r.x = r.x - xp.x; r.y = r.y - xp.y; r.z = r.z - xp.z;
        r.x = r.x - ym.x; r.y = r.y - ym.y; r.z = r.z - ym.z;
        r.x = r.x - yp.x; r.y = r.y - yp.y; r.z = r.z - yp.z;
        r.x = r.x - zm.x; r.y = r.y - zm.y; r.z = r.z - zm.z;
        r.x = r.x - zp.x; r.y = r.y - zp.y; r.z = r.z - zp.z;
        return make_float3(r.x * kInvD2, r.y * kInvD2, r.z * kInvD2);
    };
    auto store_s = [&](int m, float3 v0, float3 v1) {
        float3(*dst)[kS2X] = S[(m + 3) % 3];
        dst[e0 / kS2X][e0 % kS2X] = v0;
        if (has1) dst[e1 / kS2X][e1 % kS2X] = v1;
    };

    // prologue: S planes z0-2, z0-1 in LDS; S plane z0 and new_d of plane z0-1 in registers
    {
        const long long p = (long long)(z0 - 2) * wh + base;
        store_s(z0 - 2, load_s(p + so0), has1 ? load_s(p + so1) : zero);
        const long long p1 = p + wh;
        store_s(z0 - 1, load_s(p1 + so0), has1 ? load_s(p1 + so1) : zero);
    }
    long long pz = (long long)z0 * wh + base;  // plane m+1 during iteration m
    float3 s0 = load_s(pz + so0), s1 = has1 ? load_s(pz + so1) : zero;
    const long long pm = pz - wh;              // plane z0-1
    float3 r_in = load_r(pm + io), r_ring = has_ring ? load_r(pm + ro) : zero;
    bool v_in = valid(pm + io), v_ring = has_ring && valid(pm + ro);
    float3 r_prev = zero;  // new_d at this thread's voxel of plane m-1
    float len_max = 0.0f;

    for (int m = z0 - 1; m <= z1; ++m) {
        // S(m+1) from registers into the ring slot of plane m-2, then the next prefetch
        store_s(m + 1, s0, s1);
        const float3 c_in = r_in, c_ring = r_ring;
        const bool cv_in = v_in, cv_ring = v_ring;
        pz += wh;  // plane m+2
        s0 = load_s(pz + so0);
        s1 = has1 ? load_s(pz + so1) : zero;
        const long long pn = pz - wh;  // plane m+1
        r_in = load_r(pn + io);
        r_ring = has_ring ? load_r(pn + ro) : zero;
        v_in = valid(pn + io);
        v_ring = has_ring && valid(pn + ro);
        __syncthreads();
        // s_t+1 of plane m: this thread's voxel and its ring position
        {
            float3(*sm)[kS2X] = S[(m + 2) % 3];
            float3(*sc)[kS2X] = S[(m + 3) % 3];
            float3(*sp)[kS2X] = S[(m + 4) % 3];
            float3(*dst)[kS1X] = T[(m + 3) % 3];
            {
                const int ci = tx + 2, cj = ty + 2;
                dst[ty + 1][tx + 1] = cv_in ? update(c_in, sc[cj][ci - 1], sc[cj][ci + 1], sc[cj - 1][ci],
                                                     sc[cj + 1][ci], sm[cj][ci], sp[cj][ci])
                                            : zero;
            }
            if (has_ring) {
                const int ci = ri + 2, cj = rj + 2;
                dst[rj + 1][ri + 1] = cv_ring ? update(c_ring, sc[cj][ci - 1], sc[cj][ci + 1],
                                                       sc[cj - 1][ci], sc[cj + 1][ci], sm[cj][ci],
                                                       sp[cj][ci])
                                              : zero;
            }
        }
        __syncthreads();
        // s_t+2 of plane m-1 at this thread's voxel
        const int z = m - 1;
        if (z >= z0 && z < z1 && x0 + tx < g.w && y0 + ty < g.h) {
            float3(*tm)[kS1X] = T[(z + 2) % 3];
            float3(*tc)[kS1X] = T[(z + 3) % 3];
            float3(*tp)[kS1X] = T[(z + 4) % 3];
            const int cj = ty + 1, ci = tx + 1;
            const float3 r = update(r_prev, tc[cj][ci - 1], tc[cj][ci + 1], tc[cj - 1][ci],
                                    tc[cj + 1][ci], tm[cj][ci], tp[cj][ci]);
            st3(b, uint32_t((long long)z * wh + base + io), r);
            if (MAXLEN) {
                float s2 = r.x * r.x + r.y * r.y;
                s2 = s2 + r.z * r.z;
                len_max = fmaxf(len_max, float(sqrt(double(s2))));
            }
        }
        r_prev = c_in;
    }
    if (MAXLEN) {
        __shared__ float part[kQY];
        const float mx = wave_max_f32(len_max);
        if ((tid & 63) == 0) part[tid >> 6] = mx;
        __syncthreads();
        if (tid == 0) {
            float r = part[0];
            for (int i = 1; i < kQY; ++i) r = fmaxf(r, part[i]);
            maxparts[blk.lin] = r;
        }
    }
}

template <bool FIRST, bool MAXLEN>
__global__ __launch_bounds__(kQX * kQY) void poisson2_kernel(const float* __restrict__ nd,
                                                             const float* __restrict__ a,
                                                             FieldGeo g, int zq,
                                                             float* __restrict__ b,
                                                             float* __restrict__ maxparts)
{
    poisson2_body<FIRST, MAXLEN>(nd, a, g, zq, b, maxparts, logical_block());
}

// ---------------------------------------------------------------------------
// new_d *= k, then accumulate_displacement(cur, new_d) in place:
// new(x) = m(x + new(x)) - x with m(y) = cur(y) + y by trilinear
// interpolation; cur(x) where new(x) == 0 or the sample is outside.
// ---------------------------------------------------------------------------
// The accumulated displacement at voxel p = (x, y, z).
__device__ inline float3 accumulated(const float* __restrict__ cur, const float* __restrict__ nd,
                                     float k, const FieldGeo& g, uint32_t p, int x, int y, int z)
{
    float3 d = ld3(nd, p);
    d.x = d.x * k;
    d.y = d.y * k;
    d.z = d.z * k;
    if (d.x != 0.0f || d.y != 0.0f || d.z != 0.0f) {
        Tri t;
        if (locate(g, d.x + float(x), d.y + float(y), d.z + float(z), t)) {
            const int ix = t.ix, iy = t.iy, iz = t.iz;
            float3 acc = make_float3(0.0f, 0.0f, 0.0f);
#pragma unroll
            for (int c = 0; c < 8; ++c) {
                const float3 m = ld3(cur, corner(g, t.i0, c));
                const float mx = m.x + float(ix + (c & 1));
                const float my = m.y + float(iy + ((c >> 1) & 1));
                const float mz = m.z + float(iz + ((c >> 2) & 1));
                if (c == 0) {
                    acc.x = mx * t.w[0];
                    acc.y = my * t.w[0];
                    acc.z = mz * t.w[0];
                } else {
                    acc.x = acc.x + mx * t.w[c];
                    acc.y = acc.y + my * t.w[c];
                    acc.z = acc.z + mz * t.w[c];
                }
            }
            return make_float3(acc.x - float(x), acc.y - float(y), acc.z - float(z));
        }
    }
    return ld3(cur, p);
}

__global__ __launch_bounds__(256) void accumulate_kernel(const float* __restrict__ cur,
                                                         float* __restrict__ nd, float k,
                                                         FieldGeo g)
{
    const Vox v = this_voxel(g);
    if (!v.ok) return;
    st3(nd, v.p, accumulated(cur, nd, k, g, v.p, v.x, v.y, v.z));
}

// One packed float3 voxel through a raw buffer resource: a load past the
// end returns 0 and never faults.
__device__ inline float3 ld3_rsrc(__amdgpu_buffer_rsrc_t r, uint32_t p)
{
    const auto v = __builtin_amdgcn_raw_buffer_load_b96(r, int(p * 12u), 0, 0);
    return make_float3(__uint_as_float(v[0]), __uint_as_float(v[1]), __uint_as_float(v[2]));
}

// accumulated() without branches, so that several evaluations per thread
// keep their loads in flight together: the eight corners are always loaded
// (at the voxel itself when the sample is not taken, so corner 0 is then
// cur(p); corners past the field's end read 0 and are unused) and the
// result is selected.  Where the sample is taken every operation is
// accumulated()'s, in its order.  Needs 12 n < 2^32.
__device__ inline float3 accumulated_bf(__amdgpu_buffer_rsrc_t rcur, __amdgpu_buffer_rsrc_t rnd,
                                        float k, const FieldGeo& g, uint32_t p, int x, int y, int z)
{
    float3 d = ld3_rsrc(rnd, p);
    d.x = d.x * k;
    d.y = d.y * k;
    d.z = d.z * k;
    const bool moved = d.x != 0.0f || d.y != 0.0f || d.z != 0.0f;
    const float sx = d.x + float(x), sy = d.y + float(y), sz = d.z + float(z);
    const bool in = (fminf(fminf(sx, sy), sz) >= 0.0f) & (sx < g.wm1) & (sy < g.hm1) & (sz < g.dm1);
    const bool use = moved && in;
    const uint32_t ix = use ? cvt_u32_sat(sx) : 0u, iy = use ? cvt_u32_sat(sy) : 0u,
                   iz = use ? cvt_u32_sat(sz) : 0u;
    const uint32_t i0 = use ? iz * g.wh + iy * uint32_t(g.w) + ix : p;
    float3 m[8];
#pragma unroll
    for (int c = 0; c < 8; ++c) m[c] = ld3_rsrc(rcur, corner(g, i0, c));
    const float p0 = sx - float(ix), p1 = sy - float(iy), p2 = sz - float(iz);
    const float n0 = 1.0f - p0, n1 = 1.0f - p1, n2 = 1.0f - p2;
    const float n0n1 = n0 * n1, p0n1 = p0 * n1, n0p1 = n0 * p1, p0p1 = p0 * p1;
    const float w[8] = {n0n1 * n2, p0n1 * n2, n0p1 * n2, p0p1 * n2,
                        n0n1 * p2, p0n1 * p2, n0p1 * p2, p0p1 * p2};
    float3 acc;
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        const float mx = m[c].x + float(int(ix) + (c & 1));
        const float my = m[c].y + float(int(iy) + ((c >> 1) & 1));
        const float mz = m[c].z + float(int(iz) + ((c >> 2) & 1));
        if (c == 0) {
            acc.x = mx * w[0];
            acc.y = my * w[0];
            acc.z = mz * w[0];
        } else {
            acc.x = acc.x + mx * w[c];
            acc.y = acc.y + my * w[c];
            acc.z = acc.z + mz * w[c];
        }
    }
    return use ? make_float3(acc.x - float(x), acc.y - float(y), acc.z - float(z)) : m[0];
}

// ---------------------------------------------------------------------------
// accumulate_displacement then cdm_smooth in one pass, marching z:
// out = smooth(acc), acc = the accumulated field of accumulate_kernel.
// Tiles are laid out in flat-index space as in poisson2_kernel, so smooth's
// neighbours p-+1 and p-+w are offsets (i-+1, j), (i, j-+1) of the tile grown
// by 1 (row and plane wrap included) and p-+wh the same offset in planes
// m-+1.  A position outside [0, n) holds +0.0, which adds nothing to the
// neighbour sum (that sum starts at +0.0 and is never -0.0), as the
// reference drops it.  Per plane the block computes acc once on the tile and
// its 148-position ring (two branch-free evaluations per thread, their loads
// in flight together) and keeps the planes in an LDS ring.  HBM traffic per
// voxel: new_d read, cur gathered, out written (36 B, against 36 + 24 B for
// the two passes).  Measured at 182x218x182, whole iterations: 1558
// iterations/s against 1546 with the two passes, 1490 with branching
// evaluations and 1538 with two planes a step (four evaluations, 125
// VGPRs, half the occupancy).  out must not alias cur or nd.  With keep,
// cur is also copied into keep (the best field taking the running one
// before the step, as the reference's loop does), keep aliasing neither.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void accumulate_smooth_body(
    const float* __restrict__ cur, const float* __restrict__ nd, float k, float w_6, float w_1,
    const FieldGeo& g, int zq, float* __restrict__ out, float* __restrict__ keep, const Block& blk)
{
    __shared__ float3 A[4][kS1Y][kS1X];
    const int tx = int(threadIdx.x), ty = int(threadIdx.y), tid = ty * kQX + tx;
    const int x0 = blk.x * kQX, y0 = blk.y * kQY;
    const int z0 = blk.z * zq, z1 = min(z0 + zq, g.d);
    const long long n = g.n, w = g.w, wh = g.wh;
    const long long base = (long long)y0 * w + x0;
    const __amdgpu_buffer_rsrc_t rcur = make_rsrc(cur, g.n * 12u);
    const __amdgpu_buffer_rsrc_t rnd = make_rsrc(nd, g.n * 12u);
    int ri = 0, rj = 0;
    const bool has_ring = tid < kRing;
    if (has_ring) ring_offset(tid, ri, rj);
    const long long ro = (long long)rj * w + ri;
    const long long io = (long long)ty * w + tx;
    const float3 zero = make_float3(0.0f, 0.0f, 0.0f);
    auto acc_at = [&](long long q) -> float3 {
        const bool ok = q >= 0 && q < n;
        const uint32_t p = ok ? uint32_t(q) : 0u;
        const uint32_t z = p / g.wh, r = p - z * g.wh;
        const uint32_t y = r / uint32_t(g.w), x = r - y * uint32_t(g.w);
        const float3 v = accumulated_bf(rcur, rnd, k, g, p, int(x), int(y), int(z));
        return ok ? v : zero;
    };
    auto fill = [&](int m) {
        const long long pm = (long long)m * wh + base;
        const float3 a_in = acc_at(pm + io);
        const float3 a_r = has_ring ? acc_at(pm + ro) : zero;
        float3(*da)[kS1X] = A[(m + 4) & 3];
        da[ty + 1][tx + 1] = a_in;
        if (has_ring) da[rj + 1][ri + 1] = a_r;
    };
    auto smooth = [&](int m) {
        if (m >= z1 || x0 + tx >= g.w || y0 + ty >= g.h) return;
        float3(*am)[kS1X] = A[(m + 3) & 3];
        float3(*ac)[kS1X] = A[(m + 4) & 3];
        float3(*ap)[kS1X] = A[(m + 5) & 3];
        const int ci = tx + 1, cj = ty + 1;
        const float3 t[6] = {ac[cj][ci + 1], ac[cj][ci - 1], ac[cj + 1][ci],
                             ac[cj - 1][ci], ap[cj][ci],     am[cj][ci]};
        float3 v = zero;
#pragma unroll
        for (int e = 0; e < 6; ++e) {
            v.x = v.x + t[e].x;
            v.y = v.y + t[e].y;
            v.z = v.z + t[e].z;
        }
        const float3 c = ac[cj][ci];
        const uint32_t p = uint32_t((long long)m * wh + base + io);
        st3(out, p, make_float3(c.x * w_1 + v.x * w_6, c.y * w_1 + v.y * w_6, c.z * w_1 + v.z * w_6));
        if (keep) st3(keep, p, ld3(cur, p));
    };
    fill(z0 - 1);
    fill(z0);
    for (int m = z0; m < z1; ++m) {
        fill(m + 1);
        __syncthreads();
        smooth(m);
        __syncthreads();
    }
}

__global__ __launch_bounds__(kQX * kQY) void accumulate_smooth_kernel(
    const float* __restrict__ cur, const float* __restrict__ nd, float k, float w_6, float w_1,
    FieldGeo g, int zq, float* __restrict__ out)
{
    accumulate_smooth_body(cur, nd, k, w_6, w_1, g, zq, out, nullptr, logical_block());
}

// ---------------------------------------------------------------------------
// Multi-field launches.  One launch runs one step of the CDM iteration for
// up to kMaxSegs fields -- the concurrent registrations of a server, each
// with its own shape and buffers.  The launch's workgroups are the fields'
// grids one after another (field s from begin[s]); each workgroup finds its
// field and its logical block within the field's grid, and runs the same
// body as the single-field kernel, so every field's results, its block
// partials included, are those of a launch of its own.  The table travels
// as a kernel argument (no upload).
// ---------------------------------------------------------------------------
constexpr int kMaxSegs = 8;

struct CdmStep {
    uint32_t nseg = 0, total = 0;
    uint32_t begin[kMaxSegs];
    uint32_t gx[kMaxSegs], gy[kMaxSegs];
    int zq[kMaxSegs];
    FieldGeo g[kMaxSegs];
    const uint8_t* Is[kMaxSegs];
    const uint8_t* It[kMaxSegs];
    uint8_t* Js[kMaxSegs];
    const float* cur[kMaxSegs];
    float* nd[kMaxSegs];
    const float* src[kMaxSegs];  // Jacobi: the sweeps' input
    float* dst[kMaxSegs];        // Jacobi: output; accumulate+smooth: the new field
    float* keep[kMaxSegs];       // accumulate+smooth: receives cur when set
    double* partials[kMaxSegs];
    float* maxparts[kMaxSegs];
    float* cost[kMaxSegs];
    float* host_cost[kMaxSegs];
    unsigned* maxlen[kMaxSegs];
    unsigned* host_maxlen[kMaxSegs];
    uint32_t count[kMaxSegs];    // reductions: partials of the field
    double n[kMaxSegs];
    float k[kMaxSegs], w_6[kMaxSegs], w_1[kMaxSegs];
};
static_assert(sizeof(CdmStep) <= 4096, "CdmStep travels as a kernel argument (4 KB at most)");

// This workgroup's field s and logical block within the field's grid
// (hardware order remapped so that each XCD takes a contiguous range).
__device__ inline Block seg_block(const CdmStep& st, int& s)
{
    const uint32_t L = xcd_remap(blockIdx.x, st.total);
    int k = 0;
#pragma unroll
    for (int i = 1; i < kMaxSegs; ++i)
        if (uint32_t(i) < st.nseg && st.begin[i] <= L) k = i;
    s = k;
    const uint32_t local = L - st.begin[k];
    const uint32_t gx = st.gx[k], gy = st.gy[k];
    Block b;
    b.x = int(local % gx);
    b.y = int((local / gx) % gy);
    b.z = int(local / (gx * gy));
    b.lin = local;
    return b;
}

__global__ __launch_bounds__(256) void compose_seg_kernel(const CdmStep st)
{
    int s;
    const Block b = seg_block(st, s);
    compose_body(st.Is[s], st.cur[s], st.g[s], st.Js[s], b);
}

__global__ __launch_bounds__(256) void gradient_seg_kernel(const CdmStep st)
{
    int s;
    const Block b = seg_block(st, s);
    gradient_r_box_body(st.Js[s], st.It[s], st.g[s], 0, st.nd[s], st.partials[s], st.zq[s], b);
}

// one workgroup per field
__global__ __launch_bounds__(1024) void cost_reduce_seg_kernel(const CdmStep st)
{
    const int s = int(blockIdx.x);
    cost_reduce_body(st.partials[s], st.count[s], st.n[s], st.cost[s], st.host_cost[s]);
}

template <bool FIRST, bool MAXLEN>
__global__ __launch_bounds__(kQX * kQY) void poisson2_seg_kernel(const CdmStep st)
{
    int s;
    const Block b = seg_block(st, s);
    poisson2_body<FIRST, MAXLEN>(st.nd[s], st.src[s], st.g[s], st.zq[s], st.dst[s], st.maxparts[s], b);
}

// one workgroup per field; fields without a max_len slot are skipped
__global__ __launch_bounds__(1024) void max_reduce_seg_kernel(const CdmStep st)
{
    const int s = int(blockIdx.x);
    if (!st.maxlen[s]) return;
    max_reduce_body(st.maxparts[s], st.count[s], st.maxlen[s], st.host_maxlen[s]);
}

__global__ __launch_bounds__(kQX * kQY) void accumulate_smooth_seg_kernel(const CdmStep st)
{
    int s;
    const Block b = seg_block(st, s);
    accumulate_smooth_body(st.cur[s], st.nd[s], st.k[s], st.w_6[s], st.w_1[s], st.g[s], st.zq[s],
                           st.dst[s], st.keep[s], b);
}

// cdm_smooth: dd = d*(1-s) + (s/6)*(neighbours +1, -1, +w, -w, +wh, -wh by
// flat index, dropped at the array ends).
__global__ __launch_bounds__(256) void smooth_kernel(const float* __restrict__ d,
                                                     float* __restrict__ dd, float w_6, float w_1,
                                                     FieldGeo g)
{
    const Vox v = this_voxel(g);
    if (!v.ok) return;
    const uint32_t p = v.p, n = g.n, w = uint32_t(g.w), wh = g.wh;
    float3 s = make_float3(0.0f, 0.0f, 0.0f);
    auto add = [&](uint32_t q) {
        const float3 t = ld3(d, q);
        s.x = s.x + t.x;
        s.y = s.y + t.y;
        s.z = s.z + t.z;
    };
    if (p + 1u < n) add(p + 1u);
    if (p >= 1u) add(p - 1u);
    if (p + w < n) add(p + w);
    if (p >= w) add(p - w);
    if (p + wh < n) add(p + wh);
    if (p >= wh) add(p - wh);
    const float3 c = ld3(d, p);
    st3(dd, p, make_float3(c.x * w_1 + s.x * w_6, c.y * w_1 + s.y * w_6, c.z * w_1 + s.z * w_6));
}

// best_d *= 2 then upsample_with_padding<linear>: out(p) = in2(p * 0.5),
// 0 outside.
__global__ __launch_bounds__(256) void upsample_kernel(const float* __restrict__ in, FieldGeo gi,
                                                       FieldGeo go, float* __restrict__ out)
{
    const Vox v = this_voxel(go);
    if (!v.ok) return;
    float x = float(v.x), y = float(v.y), z = float(v.z);
    x = x * 0.5f;
    y = y * 0.5f;
    z = z * 0.5f;
    Tri t;
    float3 acc = make_float3(0.0f, 0.0f, 0.0f);
    if (locate(gi, x, y, z, t)) {
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            float3 m = ld3(in, corner(gi, t.i0, c));
            m.x = m.x * 2.0f;
            m.y = m.y * 2.0f;
            m.z = m.z * 2.0f;
            if (c == 0) {
                acc.x = m.x * t.w[0];
                acc.y = m.y * t.w[0];
                acc.z = m.z * t.w[0];
            } else {
                acc.x = acc.x + m.x * t.w[c];
                acc.y = acc.y + m.y * t.w[c];
                acc.z = acc.z + m.z * t.w[c];
            }
        }
    }
    st3(out, v.p, acc);
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------

// Scratch memory of one CDM call, stream-ordered on the context's stream.

void zero(tiplhip_ctx* ctx, void* p, size_t bytes)
{
    TIPLHIP_CHECK(hipMemsetAsync(p, 0, bytes, ctx->stream));
}

void copy_d2d(tiplhip_ctx* ctx, void* dst, const void* src, size_t bytes)
{
    TIPLHIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, ctx->stream));
}

uint32_t block_count(const FieldGeo& g)
{
    dim3 t = tile_grid(g);
    return t.x * t.y * t.z;
}

void launch_compose(tiplhip_ctx* ctx, const uint8_t* Is, const float* d, const FieldGeo& g,
                    uint8_t* Js)
{
    hipLaunchKernelGGL(compose_kernel, tile_grid(g), dim3(kTX, kTY), 0, ctx->stream, Is, d, g, Js);
    TIPLHIP_CHECK(hipGetLastError());
}

void launch_gradient(tiplhip_ctx* ctx, const uint8_t* Js, const uint8_t* It, const FieldGeo& g,
                     char type, int accumulate, float* nd, double* partials, float* cost_out,
                     float* host_cost = nullptr)
{
    uint32_t parts;
    if (type == 'd') {
        hipLaunchKernelGGL(gradient_kernel<'d'>, tile_grid(g), dim3(kTX, kTY), 0, ctx->stream, Js,
                           It, g, accumulate, nd, partials);
        parts = block_count(g);
    } else {
        const dim3 bg = box_grid(g);
        hipLaunchKernelGGL(gradient_r_box_kernel, bg, dim3(kTX, kTY), 0, ctx->stream, Js, It, g,
                           accumulate, nd, partials, box_zc(g));
        parts = bg.x * bg.y * bg.z;
    }
    TIPLHIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(cost_reduce_kernel, dim3(1), dim3(1024), 0, ctx->stream, partials, parts,
                       double(g.n), cost_out, host_cost);
    TIPLHIP_CHECK(hipGetLastError());
}

// `sweeps` Jacobi sweeps, two per launch (poisson2_kernel) and a single
// z-marching sweep for an odd remainder; returns the buffer holding the
// result.  With maxlen, the last launch also forms per-block maxima of the
// displacement length and max_reduce_kernel writes their maximum.
float* launch_poisson(tiplhip_ctx* ctx, const float* nd, float* a, float* b, const FieldGeo& g,
                      int sweeps, unsigned int* maxlen, float* maxparts)
{
    float* src = a;
    float* dst = b;
    uint32_t last_blocks = 0;
    int done = 0;
    while (done < sweeps) {
        const bool first = done == 0;
        const bool pair = sweeps - done >= 2;
        const bool last = done + (pair ? 2 : 1) == sweeps && maxlen;
        if (pair) {
            const int zq = poisson2_zq(g);
            const dim3 grid = poisson2_grid(g, zq), block(kQX, kQY);
            if (first && last)
                hipLaunchKernelGGL((poisson2_kernel<true, true>), grid, block, 0, ctx->stream, nd, src,
                                   g, zq, dst, maxparts);
            else if (first)
                hipLaunchKernelGGL((poisson2_kernel<true, false>), grid, block, 0, ctx->stream, nd,
                                   src, g, zq, dst, maxparts);
            else if (last)
                hipLaunchKernelGGL((poisson2_kernel<false, true>), grid, block, 0, ctx->stream, nd,
                                   src, g, zq, dst, maxparts);
            else
                hipLaunchKernelGGL((poisson2_kernel<false, false>), grid, block, 0, ctx->stream, nd,
                                   src, g, zq, dst, maxparts);
            last_blocks = grid.x * grid.y * grid.z;
            done += 2;
        } else {
            const dim3 grid = poisson_grid(g), block(kTX, kPTY);
            if (first && last)
                hipLaunchKernelGGL((poisson_zm_kernel<true, true>), grid, block, 0, ctx->stream, nd,
                                   src, g, dst, maxparts);
            else if (first)
                hipLaunchKernelGGL((poisson_zm_kernel<true, false>), grid, block, 0, ctx->stream, nd,
                                   src, g, dst, maxparts);
            else if (last)
                hipLaunchKernelGGL((poisson_zm_kernel<false, true>), grid, block, 0, ctx->stream, nd,
                                   src, g, dst, maxparts);
            else
                hipLaunchKernelGGL((poisson_zm_kernel<false, false>), grid, block, 0, ctx->stream, nd,
                                   src, g, dst, maxparts);
            last_blocks = grid.x * grid.y * grid.z;
            done += 1;
        }
        TIPLHIP_CHECK(hipGetLastError());
        std::swap(src, dst);
    }
    if (maxlen) {
        hipLaunchKernelGGL(max_reduce_kernel, dim3(1), dim3(1024), 0, ctx->stream, maxparts,
                           last_blocks, maxlen);
        TIPLHIP_CHECK(hipGetLastError());
    }
    return src;
}

void launch_accumulate(tiplhip_ctx* ctx, const float* cur, float* nd, float k, const FieldGeo& g)
{
    hipLaunchKernelGGL(accumulate_kernel, tile_grid(g), dim3(kTX, kTY), 0, ctx->stream, cur, nd, k, g);
    TIPLHIP_CHECK(hipGetLastError());
}

// accumulate then cdm_smooth with `smoothing` (non-zero) in one pass into out.
void launch_accumulate_smooth(tiplhip_ctx* ctx, const float* cur, const float* nd, float k,
                              float smoothing, const FieldGeo& g, float* out)
{
    const float w_6 = smoothing / float(2.0f * 3.0f);
    const float w_1 = 1.0f - smoothing;
    const int zq = poisson2_zq(g);
    hipLaunchKernelGGL(accumulate_smooth_kernel, poisson2_grid(g, zq), dim3(kQX, kQY), 0, ctx->stream,
                       cur, nd, k, w_6, w_1, g, zq, out);
    TIPLHIP_CHECK(hipGetLastError());
}

void launch_smooth(tiplhip_ctx* ctx, const float* d, float* dd, float smoothing, const FieldGeo& g)
{
    const float w_6 = smoothing / float(2.0f * 3.0f);
    const float w_1 = 1.0f - smoothing;
    hipLaunchKernelGGL(smooth_kernel, tile_grid(g), dim3(kTX, kTY), 0, ctx->stream, d, dd, w_6, w_1, g);
    TIPLHIP_CHECK(hipGetLastError());
}

void launch_upsample(tiplhip_ctx* ctx, const float* in, const FieldGeo& gi, const FieldGeo& go,
                     float* out)
{
    hipLaunchKernelGGL(upsample_kernel, tile_grid(go), dim3(kTX, kTY), 0, ctx->stream, in, gi, go, out);
    TIPLHIP_CHECK(hipGetLastError());
}

struct CdmRun {
    tiplhip_ctx* ctx;
    tiplhip_cdm_param param;
    int nmod;
    const volatile uint8_t* terminated;
    int iterations = 0;
    bool no_early_exit = false;  // timing mode: run max_iterations exactly
    bool stopped = false;        // terminated by the caller
};

bool is_terminated(const CdmRun& r) { return r.terminated && *r.terminated; }

// ---------------------------------------------------------------------------
// Batched CDM iterations (the serving loop).  A registration whose context
// is attached to a CdmBatcher hands the device work of each iteration of its
// cdm() level loop to the batcher as one step; the batcher launches the
// steps of up to kMaxSegs registrations together, one *_seg launch per
// kernel, on its own stream.  Each registration keeps its own host control
// (cost history, cdm_improved, smoothing schedule) and each field's results
// are those of launches of its own, so a batched cdm() equals a plain one
// bit for bit.
// ---------------------------------------------------------------------------

// One registration's level loop as the batcher sees it.  The owner sets
// the step fields before CdmBatcher::run; the launch reads them under the
// gate's lock.
struct CdmJob : GateSlot {
    FieldGeo g{};
    dim3 tile, box, p2;       // grids: compose, gradient, Jacobi and accumulate+smooth
    int zq_box = 0, zq_p2 = 0;
    const uint8_t* It = nullptr;
    const uint8_t* Is = nullptr;
    uint8_t* Js = nullptr;
    float* nd = nullptr;
    float* best = nullptr;
    float* cur = nullptr;     // the running field the iteration body reads
    float* pa = nullptr;      // Jacobi buffers: the solved new_d ends in pa
    float* pb = nullptr;
    double* partials = nullptr;
    float* maxparts = nullptr;
    float* d_cost = nullptr;
    unsigned* d_maxlen = nullptr;
    float* h_cost = nullptr;  // coherent host slots (device addresses)
    unsigned* h_maxlen = nullptr;
    hipEvent_t ready = nullptr;  // on the owner's stream: the level's inputs are written
    hipEvent_t done = nullptr;   // on the batch stream: the job's last step
    bool waited = false;         // the lane's stream has waited for `ready`
    int lane = 0;                // the batcher lane the job joined
    // the step: accumulate+smooth from s_src into s_out (s_keep <- s_src when
    // set) or plain copies, then the iteration body on cur
    bool smooth = false;
    const float* s_src = nullptr;
    const float* s_nd = nullptr;
    float* s_out = nullptr;
    float* s_keep = nullptr;
    float k = 0.0f, w_6 = 0.0f, w_1 = 0.0f;
    bool copy_best = false;   // without smoothing: best <- cur, then
    bool restore = false;     // cur <- best
    bool need_theta = false;  // the last Jacobi launch forms max_len
};

// Fills the workgroup ranges of a multi-field launch from one grid per field.
void set_grids(CdmStep& st, CdmJob* const* J, int n, dim3 CdmJob::*grid)
{
    uint32_t total = 0;
    for (int s = 0; s < n; ++s) {
        const dim3 d = J[s]->*grid;
        st.begin[s] = total;
        st.gx[s] = d.x;
        st.gy[s] = d.y;
        const uint64_t t = uint64_t(total) + uint64_t(d.x) * d.y * d.z;
        TIPLHIP_REQUIRE(t < (uint64_t(1) << 31), "batched CDM launch too large");
        total = uint32_t(t);
    }
    st.total = total;
}

// Queues one step of every job in J (at most kMaxSegs) on `stream`, in the
// order of the reference's loop body: the previous iteration's update of the
// running field (reg/cdm.hpp:447-469), then compose_displacement, the
// gradient and cost, and cdm_solve_poisson's 12 sweeps (reg/cdm.hpp:420-462),
// exactly as cdm_level queues them for one field.
void launch_cdm_steps(hipStream_t stream, CdmJob* const* J, int n)
{
    TIPLHIP_REQUIRE(n >= 1 && n <= kMaxSegs, "1 to 8 fields per batched CDM step");
    for (int s = 0; s < n; ++s) {
        CdmJob& j = *J[s];
        if (!j.waited) {
            TIPLHIP_CHECK(hipStreamWaitEvent(stream, j.ready, 0));
            j.waited = true;
        }
        if (j.smooth) continue;
        const size_t fb = sizeof(float) * 3 * j.g.n;
        if (j.copy_best)
            TIPLHIP_CHECK(hipMemcpyAsync(j.best, j.cur, fb, hipMemcpyDeviceToDevice, stream));
        if (j.restore)
            TIPLHIP_CHECK(hipMemcpyAsync(j.cur, j.best, fb, hipMemcpyDeviceToDevice, stream));
    }
    {
        CdmStep st;
        CdmJob* sm[kMaxSegs];
        int m = 0;
        for (int s = 0; s < n; ++s)
            if (J[s]->smooth) sm[m++] = J[s];
        if (m) {
            st.nseg = uint32_t(m);
            set_grids(st, sm, m, &CdmJob::p2);
            for (int s = 0; s < m; ++s) {
                const CdmJob& j = *sm[s];
                st.g[s] = j.g;
                st.zq[s] = j.zq_p2;
                st.cur[s] = j.s_src;
                st.nd[s] = const_cast<float*>(j.s_nd);
                st.dst[s] = j.s_out;
                st.keep[s] = j.s_keep;
                st.k[s] = j.k;
                st.w_6[s] = j.w_6;
                st.w_1[s] = j.w_1;
            }
            hipLaunchKernelGGL(accumulate_smooth_seg_kernel, dim3(st.total), dim3(kQX, kQY), 0, stream, st);
            TIPLHIP_CHECK(hipGetLastError());
        }
    }
    CdmStep st;
    st.nseg = uint32_t(n);
    bool theta = false;
    for (int s = 0; s < n; ++s) {
        const CdmJob& j = *J[s];
        st.g[s] = j.g;
        st.Is[s] = j.Is;
        st.It[s] = j.It;
        st.Js[s] = j.Js;
        st.cur[s] = j.cur;
        st.nd[s] = j.nd;
        st.partials[s] = j.partials;
        st.maxparts[s] = j.maxparts;
        st.cost[s] = j.d_cost;
        st.host_cost[s] = j.h_cost;
        st.maxlen[s] = j.need_theta ? j.d_maxlen : nullptr;
        st.host_maxlen[s] = j.h_maxlen;
        st.n[s] = double(j.g.n);
        theta = theta || j.need_theta;
    }
    set_grids(st, J, n, &CdmJob::tile);
    hipLaunchKernelGGL(compose_seg_kernel, dim3(st.total), dim3(kTX, kTY), 0, stream, st);
    TIPLHIP_CHECK(hipGetLastError());
    set_grids(st, J, n, &CdmJob::box);
    for (int s = 0; s < n; ++s) {
        st.zq[s] = J[s]->zq_box;
        st.count[s] = J[s]->box.x * J[s]->box.y * J[s]->box.z;
    }
    hipLaunchKernelGGL(gradient_seg_kernel, dim3(st.total), dim3(kTX, kTY), 0, stream, st);
    TIPLHIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(cost_reduce_seg_kernel, dim3(st.nseg), dim3(1024), 0, stream, st);
    TIPLHIP_CHECK(hipGetLastError());
    set_grids(st, J, n, &CdmJob::p2);
    for (int s = 0; s < n; ++s) {
        st.zq[s] = J[s]->zq_p2;
        st.count[s] = J[s]->p2.x * J[s]->p2.y * J[s]->p2.z;
    }
    // 12 sweeps, two per launch: launch i reads pa (i even) or pb and writes
    // the other, so the solution ends in pa
    for (int i = 0; i < 6; ++i) {
        for (int s = 0; s < n; ++s) {
            st.src[s] = (i & 1) ? J[s]->pb : J[s]->pa;
            st.dst[s] = (i & 1) ? J[s]->pa : J[s]->pb;
        }
        const dim3 grid(st.total), block(kQX, kQY);
        if (i == 0)
            hipLaunchKernelGGL((poisson2_seg_kernel<true, false>), grid, block, 0, stream, st);
        else if (i == 5 && theta)
            hipLaunchKernelGGL((poisson2_seg_kernel<false, true>), grid, block, 0, stream, st);
        else
            hipLaunchKernelGGL((poisson2_seg_kernel<false, false>), grid, block, 0, stream, st);
        TIPLHIP_CHECK(hipGetLastError());
    }
    if (theta) {
        hipLaunchKernelGGL(max_reduce_seg_kernel, dim3(st.nseg), dim3(1024), 0, stream, st);
        TIPLHIP_CHECK(hipGetLastError());
    }
}

}  // namespace

// The batch executor of one device: `lanes` streams, each with a StepGate
// (at most kMaxSegs fields a batch, kGatherUs); a job keeps the lane it
// joined (the one with the fewest members), whose stream orders its steps.
// Levels above max_voxels run unbatched on their owner's stream.  Measured
// on 32 uneven affine+CDM pairs served on 16 slots (pairs/s, two runs
// each): unbatched 57.6-61.2; 1 lane (up to 8 fields a step, 5.5 on
// average) 56.1-58.0 -- every registration held to the pace of the
// largest field in its batch; 2 lanes 62.3-64.3; 3 lanes 64.6-65.6;
// 4 lanes 65.6-68.8; 6 lanes 62.3; 8 lanes 61.6-61.9; 16 lanes (one field
// a step) 58.9-61.3.  Leaving levels above 256K or 512K voxels unbatched
// did not help (63.8-67.6).
class CdmBatcher {
public:
    static constexpr int kGatherUs = 300;
    static constexpr int kDefaultLanes = 4;
    static constexpr size_t kDefaultMaxVoxels = size_t(1) << 25;

    CdmBatcher(int device, int lanes, size_t max_voxels) : device_(device), max_voxels_(max_voxels)
    {
        TIPLHIP_REQUIRE(lanes >= 1 && lanes <= 16, "1 to 16 CDM batch lanes");
        DeviceGuard dg(device);
        lanes_.resize(size_t(lanes));
        for (Lane& l : lanes_) {
            const hipError_t e = hipStreamCreateWithFlags(&l.stream, hipStreamNonBlocking);
            if (e != hipSuccess) {
                destroy_streams();  // the destructor does not run for a throwing constructor
                TIPLHIP_CHECK(e);
            }
            hipStream_t st = l.stream;
            l.gate.reset(new StepGate<CdmJob>(kMaxSegs, std::chrono::microseconds(kGatherUs),
                                              [this, st](CdmJob* const* J, int n) {
                                                  DeviceGuard g(device_);
                                                  launch_cdm_steps(st, J, n);
                                              }));
        }
    }
    ~CdmBatcher() { destroy_streams(); }
    CdmBatcher(const CdmBatcher&) = delete;
    CdmBatcher& operator=(const CdmBatcher&) = delete;

    int device() const { return device_; }
    size_t max_voxels() const { return max_voxels_; }
    hipStream_t stream(int lane) const { return lanes_[size_t(lane)].stream; }
    // Makes j a member of the least occupied lane; returns the lane.
    int join(CdmJob& j)
    {
        {
            std::lock_guard<std::mutex> lk(mu_);
            size_t best = 0;
            for (size_t i = 1; i < lanes_.size(); ++i)
                if (lanes_[i].members < lanes_[best].members) best = i;
            ++lanes_[best].members;
            j.lane = int(best);
        }
        lanes_[size_t(j.lane)].gate->join();
        return j.lane;
    }
    // After the job's last step: the owner's stream waits for the lane's
    // stream, so the owner's later work and frees follow every step.
    void leave(CdmJob& j, hipStream_t owner)
    {
        Lane& l = lanes_[size_t(j.lane)];
        l.gate->leave([&] {
            TIPLHIP_CHECK(hipEventRecord(j.done, l.stream));
            TIPLHIP_CHECK(hipStreamWaitEvent(owner, j.done, 0));
        });
        std::lock_guard<std::mutex> lk(mu_);
        --l.members;
    }
    // Queues j's step; returns once it is launched (rethrows its launch error).
    // A failed launch fails every job of its batch: their steps may be partly
    // queued.
    void run(CdmJob& j) { lanes_[size_t(j.lane)].gate->run(j); }
    int64_t launches() const
    {
        int64_t n = 0;
        for (const Lane& l : lanes_) n += l.gate->batches();
        return n;
    }
    int64_t fields() const
    {
        int64_t n = 0;
        for (const Lane& l : lanes_) n += l.gate->steps();
        return n;
    }

private:
    struct Lane {
        hipStream_t stream = nullptr;
        std::unique_ptr<StepGate<CdmJob>> gate;
        int members = 0;
    };
    void destroy_streams()
    {
        int prev = -1;
        (void)hipGetDevice(&prev);
        (void)hipSetDevice(device_);
        for (Lane& l : lanes_) {
            if (!l.stream) continue;
            (void)hipStreamSynchronize(l.stream);
            (void)hipStreamDestroy(l.stream);
            l.stream = nullptr;
        }
        if (prev >= 0) (void)hipSetDevice(prev);
    }
    int device_;
    size_t max_voxels_;
    std::mutex mu_;
    std::vector<Lane> lanes_;
};

void cdm_batcher_counts(const CdmBatcher& b, int64_t* steps, int64_t* fields)
{
    *steps = b.launches();
    *fields = b.fields();
}

bool cdm_batching_enabled()
{
    const char* e = std::getenv("TIPLHIP_SERVER_CDM_BATCH");
    return !(e && e[0] == '0');
}

// TIPLHIP_CDM_LANES and TIPLHIP_CDM_BATCH_VOXELS, read when a device's
// batcher is created, override the lane count and the largest batched level.
std::shared_ptr<CdmBatcher> shared_cdm_batcher(int device)
{
    static std::mutex mu;
    static std::map<int, std::weak_ptr<CdmBatcher>> live;
    std::lock_guard<std::mutex> lk(mu);
    std::shared_ptr<CdmBatcher> b = live[device].lock();
    if (!b) {
        int lanes = CdmBatcher::kDefaultLanes;
        size_t max_voxels = CdmBatcher::kDefaultMaxVoxels;
        if (const char* e = std::getenv("TIPLHIP_CDM_LANES")) lanes = std::atoi(e);
        if (const char* e = std::getenv("TIPLHIP_CDM_BATCH_VOXELS")) max_voxels = size_t(std::atoll(e));
        b = std::make_shared<CdmBatcher>(device, lanes, max_voxels);
        live[device] = b;
    }
    return b;
}

namespace {

// A job's membership of the batcher for one level: events created, `ready`
// recorded on the owner's stream and the job counted active; leave() (or
// the destructor) ends it.
struct CdmMembership {
    CdmBatcher& b;
    CdmJob& j;
    hipStream_t owner;
    bool joined = false;
    CdmMembership(CdmBatcher& batcher, CdmJob& job, hipStream_t s) : b(batcher), j(job), owner(s)
    {
        TIPLHIP_CHECK(hipEventCreateWithFlags(&j.ready, hipEventDisableTiming));
        TIPLHIP_CHECK(hipEventCreateWithFlags(&j.done, hipEventDisableTiming));
        TIPLHIP_CHECK(hipEventRecord(j.ready, owner));
        b.join(j);
        joined = true;
    }
    void leave()
    {
        if (!joined) return;
        joined = false;
        b.leave(j, owner);
    }
    ~CdmMembership()
    {
        try {
            leave();
        } catch (...) {
            // the lane is drained before the owner's buffers are freed
            (void)hipStreamSynchronize(b.stream(j.lane));
        }
        if (j.ready) (void)hipEventDestroy(j.ready);
        if (j.done) (void)hipEventDestroy(j.done);
    }
};

// cdm_level's iteration loop for one modality with the 'r' gradient, its
// device work run as batched steps.  The host control is cdm_level's, line
// for line; a decision's copies and the field update are queued with the
// next iteration's body (fused into accumulate+smooth where there is one).
bool cdm_iterate_batched(CdmRun& run, const uint8_t* It, const uint8_t* Is, const FieldGeo& g,
                         float* best_d)
{
    tiplhip_ctx* ctx = run.ctx;
    CdmBatcher& batch = *ctx->cdm_batch;
    TIPLHIP_REQUIRE(batch.device() == ctx->device, "CDM batcher of another device");
    const size_t fb = sizeof(float) * 3 * g.n;
    DevMem f0(ctx, fb), f1(ctx, fb), f2(ctx, fb), nd(ctx, fb), Js(ctx, g.n);
    const uint32_t nblocks = block_count(g);
    DevMem partials(ctx, sizeof(double) * nblocks), maxparts(ctx, sizeof(float) * nblocks);
    DevMem scal(ctx, sizeof(float) * 16 + sizeof(unsigned int));
    ctx->h_cdm.reserve(sizeof(float) * 16);
    uint32_t* slots = ctx->h_cdm.as<uint32_t>();

    CdmJob job;
    job.g = g;
    job.tile = tile_grid(g);
    job.box = box_grid(g);
    job.zq_box = box_zc(g);
    job.zq_p2 = poisson2_zq(g);
    job.p2 = poisson2_grid(g, job.zq_p2);
    job.It = It;
    job.Is = Is;
    job.Js = Js.as<uint8_t>();
    job.nd = nd.as<float>();
    job.best = best_d;
    job.cur = f0.as<float>();
    job.pa = f1.as<float>();
    job.pb = f2.as<float>();
    job.partials = partials.as<double>();
    job.maxparts = maxparts.as<float>();
    job.d_cost = scal.as<float>();
    job.d_maxlen = reinterpret_cast<unsigned int*>(scal.as<float>() + 16);
    job.h_cost = static_cast<float*>(ctx->h_cdm.dev);
    job.h_maxlen = static_cast<unsigned int*>(ctx->h_cdm.dev) + 1;
    copy_d2d(ctx, job.cur, best_d, fb);

    CdmMembership member(batch, job, ctx->stream);
    // kernels of this job may be queued on either stream when a step throws
    const hipStream_t lane = batch.stream(job.lane);
    auto drain = [ctx, lane] {
        (void)hipStreamSynchronize(ctx->stream);
        (void)hipStreamSynchronize(lane);
    };
    DrainOnUnwind<decltype(drain)> unwind(drain);
    auto sync_batch = [lane] { TIPLHIP_CHECK(hipStreamSynchronize(lane)); };

    float theta = 0.0f;
    std::vector<float> cost;
    float best_cost = 3.402823466e+38f;
    float cur_smoothing = 0.8f;
    // the last decision, carried out with the next step
    bool copy_best = false, restore = false, update = false;
    bool failed = false;
    float k = 0.0f;
    for (int index = 0; index < run.param.max_iterations; ++index) {
        if (is_terminated(run)) {
            run.stopped = true;
            break;
        }
        ++run.iterations;
        job.smooth = update && cur_smoothing != 0.0f;
        job.copy_best = job.restore = false;
        if (job.smooth) {
            // accumulate+smooth of the running field (best when it was
            // reset to it; best <- cur fused in) into the spare buffer,
            // which becomes the running field
            job.s_src = (restore && !copy_best) ? job.best : job.cur;
            job.s_keep = copy_best ? job.best : nullptr;
            job.s_nd = job.pa;
            job.s_out = job.pb;
            job.k = k;
            job.w_6 = cur_smoothing / float(2.0f * 3.0f);
            job.w_1 = 1.0f - cur_smoothing;
            float* spare = job.pb;
            job.pb = job.pa;
            job.pa = job.cur;
            job.cur = spare;
        } else {
            job.copy_best = copy_best;
            job.restore = restore;
        }
        copy_best = restore = update = false;
        job.need_theta = theta == 0.0f;
        slots[0] = kCostPending;
        slots[1] = kCostPending;
        batch.run(job);
        wait_for_slots(static_cast<const volatile uint32_t*>(slots), 1, kCostPending,
                       std::chrono::microseconds(kCostSpinUs), sync_batch);
        float sum_cost = 0.0f;
        std::memcpy(&sum_cost, &slots[0], sizeof(float));
        // zero cost: cdm() fails (its step's later kernels still run, and
        // may store into the slots: the exit below waits for them)
        if (sum_cost == 0.0f) {
            failed = true;
            break;
        }
        cost.push_back(sum_cost);
        if (cost.back() < best_cost) {
            best_cost = cost.back();
            if (index) copy_best = true;
        }
        if (!cdm_improved(cost)) {
            if (cur_smoothing <= run.param.smoothing) break;
            restore = true;
            cur_smoothing *= 0.5f;
            cost.clear();
        }
        if (job.need_theta) {
            wait_for_slots(static_cast<const volatile uint32_t*>(slots + 1), 1, kCostPending,
                           std::chrono::microseconds(kCostSpinUs), sync_batch);
            std::memcpy(&theta, &slots[1], sizeof(theta));
        }
        if (theta == 0.0f) break;
        k = run.param.speed / theta;
        update = true;
    }
    member.leave();
    // a decision not yet carried out: only its best <- cur matters now
    if (copy_best && !failed) copy_d2d(ctx, best_d, job.cur, fb);
    // every step has finished (and stored into the slots) before the
    // context's next call resets them
    TIPLHIP_CHECK(hipStreamSynchronize(ctx->stream));
    return !failed;
}

// One cdm() call at shape dim.  best_d is a zeroed device field of vox(dim).
bool cdm_level(CdmRun& run, const std::vector<const uint8_t*>& It,
               const std::vector<const uint8_t*>& Is, const int* dim, float* best_d,
               float resolution)
{
    tiplhip_ctx* ctx = run.ctx;
    const FieldGeo g = field_geo(dim);
    const size_t n = g.n, fb = sizeof(float) * 3 * n;
    if (!run.no_early_exit && std::min({dim[0], dim[1], dim[2]}) > run.param.min_dimension) {
        const int rd[3] = {(dim[0] + 1) / 2, (dim[1] + 1) / 2, (dim[2] + 1) / 2};
        const size_t rn = vox(rd);
        DevMem rbuf(ctx, 2 * size_t(run.nmod) * rn);
        std::vector<const uint8_t*> rIt, rIs;
        for (int i = 0; i < run.nmod; ++i) {
            uint8_t* t = rbuf.as<uint8_t>() + size_t(2 * i) * rn;
            uint8_t* s = rbuf.as<uint8_t>() + size_t(2 * i + 1) * rn;
            launch_downsample_u8(ctx, It[size_t(i)], dim, t);
            launch_downsample_u8(ctx, Is[size_t(i)], dim, s);
            rIt.push_back(t);
            rIs.push_back(s);
        }
        DevMem rbest(ctx, sizeof(float) * 3 * rn);
        zero(ctx, rbest.p, sizeof(float) * 3 * rn);
        if (!cdm_level(run, rIt, rIs, rd, rbest.as<float>(), resolution / 2.0f)) return false;
        launch_upsample(ctx, rbest.as<float>(), field_geo(rd), g, best_d);
        if (resolution > 1.0f) return true;
    }
    if (run.stopped) return true;
    if (run.ctx->cdm_batch && run.nmod == 1 && run.param.gradient_type[0] != 'd' &&
        !run.no_early_exit && n <= run.ctx->cdm_batch->max_voxels())
        return cdm_iterate_batched(run, It[0], Is[0], g, best_d);

    // three fields rotate: the running field and the two Jacobi buffers;
    // accumulate+smooth writes the new running field into the Jacobi buffer
    // that does not hold the solved new_d
    DevMem f0(ctx, fb), f1(ctx, fb), f2(ctx, fb), nd(ctx, fb), Js(ctx, n);
    float* cur = f0.as<float>();
    float* pa = f1.as<float>();
    float* pb = f2.as<float>();
    const uint32_t nblocks = block_count(g);
    DevMem partials(ctx, sizeof(double) * nblocks), maxparts(ctx, sizeof(float) * nblocks);
    DevMem scal(ctx, sizeof(float) * 16 + sizeof(unsigned int));
    float* d_cost = scal.as<float>();
    unsigned int* d_maxlen = reinterpret_cast<unsigned int*>(scal.as<float>() + 16);
    // the cost reductions store into ctx->h_cdm: if a step throws while one
    // is queued, drain the stream before the slots (and the buffers above)
    // can be reused
    auto drain = [ctx] { (void)hipStreamSynchronize(ctx->stream); };
    DrainOnUnwind<decltype(drain)> unwind(drain);
    copy_d2d(ctx, cur, best_d, fb);

    float theta = 0.0f;
    std::vector<float> cost;
    float best_cost = 3.402823466e+38f;
    float cur_smoothing = 0.8f;
    std::vector<float> mod_cost(size_t(run.nmod));
    TIPLHIP_REQUIRE(run.nmod >= 1 && run.nmod <= 16, "1 to 16 modalities");
    ctx->h_cdm.reserve(sizeof(float) * 16);
    for (int index = 0; index < run.param.max_iterations; ++index) {
        if (is_terminated(run)) {
            run.stopped = true;
            break;
        }
        ++run.iterations;
        // Each modality's cost is also written into coherent host memory,
        // whose slots start as a NaN pattern the reduction never writes; the
        // host spins on them (a stream synchronisation after kCostSpinUs).
        // The Jacobi sweeps do not depend on the host's decision (the
        // reference solves the same new_d whatever cdm_improved says), so
        // they are queued before the host waits and run during the round trip.
        uint32_t* slots = ctx->h_cdm.as<uint32_t>();
        std::fill(slots, slots + run.nmod, kCostPending);
        float* h_cost = static_cast<float*>(ctx->h_cdm.dev);
        for (int i = 0; i < run.nmod; ++i) {
            launch_compose(ctx, Is[size_t(i)], cur, g, Js.as<uint8_t>());
            const char type = run.param.gradient_type[std::min(i, 7)] == 'd' ? 'd' : 'r';
            launch_gradient(ctx, Js.as<uint8_t>(), It[size_t(i)], g, type, i > 0 ? 1 : 0,
                            nd.as<float>(), partials.as<double>(), d_cost + i, h_cost + i);
        }
        const bool need_theta = theta == 0.0f;
        float* solved = launch_poisson(ctx, nd.as<float>(), pa, pb, g, 12,
                                       need_theta ? d_maxlen : nullptr, maxparts.as<float>());
        wait_for_slots(static_cast<const volatile uint32_t*>(slots), size_t(run.nmod), kCostPending,
                       std::chrono::microseconds(kCostSpinUs),
                       [ctx] { TIPLHIP_CHECK(hipStreamSynchronize(ctx->stream)); });
        for (int i = 0; i < run.nmod; ++i) std::memcpy(&mod_cost[size_t(i)], &slots[i], sizeof(float));
        float sum_cost = 0.0f;
        for (int i = 0; i < run.nmod; ++i) {
            sum_cost += mod_cost[size_t(i)];
            if (sum_cost == 0.0f && !run.no_early_exit) return false;
        }
        cost.push_back(sum_cost / float(run.nmod));
        if (cost.back() < best_cost) {
            best_cost = cost.back();
            if (index) copy_d2d(ctx, best_d, cur, fb);
        }
        if (!cdm_improved(cost) && !run.no_early_exit) {
            if (cur_smoothing <= run.param.smoothing) break;
            copy_d2d(ctx, cur, best_d, fb);
            cur_smoothing *= 0.5f;
            cost.clear();
        }
        if (need_theta) {
            unsigned int bits = 0;
            read_back(ctx, d_maxlen, &bits, sizeof(bits));
            std::memcpy(&theta, &bits, sizeof(theta));
        }
        if (theta == 0.0f) break;
        const float k = run.param.speed / theta;
        // cdm_smooth with weight 0 leaves the running field as it was (the
        // accumulated field is then unused), as in the reference
        if (cur_smoothing != 0.0f) {
            float* spare = solved == pa ? pb : pa;
            launch_accumulate_smooth(ctx, cur, solved, k, cur_smoothing, g, spare);
            pa = cur;
            pb = solved;
            cur = spare;
        }
    }
    // timing mode reports the running field after the fixed iteration count
    if (run.no_early_exit) copy_d2d(ctx, best_d, cur, fb);
    TIPLHIP_CHECK(hipStreamSynchronize(ctx->stream));
    return true;
}

// The fixed schedule (an extension; see tiplhip_cdm_param): `levels`
// pyramid levels, all run, coarsest first, each for exactly
// param.level_iterations iterations of the loop body with no cost-driven
// control; the running field is kept and, with field_bandwidth > 0, smoothed
// by filter::gaussian each iteration in place of cdm_smooth.  `field` is a
// device field of vox(dim), overwritten.
void cdm_fixed(CdmRun& run, const std::vector<const uint8_t*>& It,
               const std::vector<const uint8_t*>& Is, const int* dim, float* field, int levels)
{
    tiplhip_ctx* ctx = run.ctx;
    const FieldGeo g = field_geo(dim);
    const size_t n = g.n, fb = sizeof(float) * 3 * n;
    if (levels > 1) {
        const int rd[3] = {(dim[0] + 1) / 2, (dim[1] + 1) / 2, (dim[2] + 1) / 2};
        const size_t rn = vox(rd);
        DevMem rbuf(ctx, 2 * size_t(run.nmod) * rn), rfield(ctx, sizeof(float) * 3 * rn);
        std::vector<const uint8_t*> rIt, rIs;
        for (int i = 0; i < run.nmod; ++i) {
            uint8_t* t = rbuf.as<uint8_t>() + size_t(2 * i) * rn;
            uint8_t* s = rbuf.as<uint8_t>() + size_t(2 * i + 1) * rn;
            launch_downsample_u8(ctx, It[size_t(i)], dim, t);
            launch_downsample_u8(ctx, Is[size_t(i)], dim, s);
            rIt.push_back(t);
            rIs.push_back(s);
        }
        cdm_fixed(run, rIt, rIs, rd, rfield.as<float>(), levels - 1);
        zero(ctx, field, fb);  // far-edge slabs of the upsampled field stay 0
        launch_upsample(ctx, rfield.as<float>(), field_geo(rd), g, field);
    } else {
        zero(ctx, field, fb);
    }
    if (run.stopped) return;
    DevMem nd(ctx, fb), pa(ctx, fb), pb(ctx, fb), Js(ctx, n);
    const uint32_t nblocks = block_count(g);
    DevMem partials(ctx, sizeof(double) * nblocks), maxparts(ctx, sizeof(float) * nblocks);
    DevMem scal(ctx, sizeof(float) * 16 + sizeof(unsigned int));
    float* d_cost = scal.as<float>();
    unsigned int* d_maxlen = reinterpret_cast<unsigned int*>(scal.as<float>() + 16);
    float theta = 0.0f;
    for (int index = 0; index < run.param.level_iterations; ++index) {
        if (is_terminated(run)) {
            run.stopped = true;
            return;
        }
        ++run.iterations;
        for (int i = 0; i < run.nmod; ++i) {
            launch_compose(ctx, Is[size_t(i)], field, g, Js.as<uint8_t>());
            const char type = run.param.gradient_type[std::min(i, 7)] == 'd' ? 'd' : 'r';
            launch_gradient(ctx, Js.as<uint8_t>(), It[size_t(i)], g, type, i > 0 ? 1 : 0,
                            nd.as<float>(), partials.as<double>(), d_cost + i);
        }
        const bool need_theta = theta == 0.0f;
        float* solved = launch_poisson(ctx, nd.as<float>(), pa.as<float>(), pb.as<float>(), g, 12,
                                       need_theta ? d_maxlen : nullptr, maxparts.as<float>());
        if (need_theta) {
            unsigned int bits = 0;
            read_back(ctx, d_maxlen, &bits, sizeof(bits));
            std::memcpy(&theta, &bits, sizeof(theta));
        }
        if (theta == 0.0f) break;
        launch_accumulate(ctx, field, solved, run.param.speed / theta, g);
        if (run.param.field_bandwidth > 0) {
            float* spare = solved == pa.as<float>() ? pb.as<float>() : pa.as<float>();
            launch_gaussian_to(ctx, solved, field, spare, dim, 3, unsigned(run.param.field_bandwidth));
        } else if (run.param.smoothing != 0.0f) {
            // cdm_smooth with weight 0 leaves the field untouched, as in the reference
            launch_smooth(ctx, solved, field, run.param.smoothing, g);
        }
    }
    TIPLHIP_CHECK(hipStreamSynchronize(ctx->stream));
}

std::vector<const uint8_t*> volume_ptrs(tiplhip_ctx* ctx, const tiplhip_volume* const* v, int n,
                                        const int* dim)
{
    std::vector<const uint8_t*> out;
    for (int i = 0; i < n; ++i) {
        TIPLHIP_REQUIRE(v[i] != nullptr, "null volume");
        TIPLHIP_REQUIRE(v[i]->ctx && v[i]->ctx->device == ctx->device, "volume lives on another device");
        TIPLHIP_REQUIRE(std::equal(dim, dim + 3, v[i]->dim), "all CDM volumes must share one shape");
        out.push_back(v[i]->d);
    }
    return out;
}

}  // namespace

int cdm_device(tiplhip_ctx* ctx, const tiplhip_volume* const* It, const tiplhip_volume* const* Is,
               int n, const tiplhip_cdm_param* param, float* d_field,
               const volatile uint8_t* terminated, int* iterations)
{
    TIPLHIP_REQUIRE(n >= 1 && n <= 8, "between 1 and 8 modality pairs");
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    DeviceGuard dg(ctx->device);
    const int* dim = It[0] ? It[0]->dim : nullptr;
    TIPLHIP_REQUIRE(dim != nullptr, "null volume");
    auto vt = volume_ptrs(ctx, It, n, dim);
    auto vs = volume_ptrs(ctx, Is, n, dim);
    CdmRun run{ctx, {}, n, terminated};
    if (param)
        run.param = *param;
    else
        tiplhip_cdm_param_default(&run.param);
    TIPLHIP_REQUIRE(run.param.max_iterations >= 0, "negative iteration cap");
    TIPLHIP_REQUIRE(run.param.levels >= 0 && run.param.levels <= 16, "levels must be in [0, 16]");
    TIPLHIP_REQUIRE(run.param.field_bandwidth >= 0 &&
                        run.param.field_bandwidth <= int(kMaxGaussianBandwidth),
                    "field_bandwidth must be in [0, 16]");
    zero(ctx, d_field, sizeof(float) * 3 * vox(dim));
    bool ok = true;
    if (run.param.levels > 0) {
        TIPLHIP_REQUIRE(run.param.level_iterations >= 0, "negative iteration count");
        cdm_fixed(run, vt, vs, dim, d_field, run.param.levels);
    } else {
        ok = cdm_level(run, vt, vs, dim, d_field, run.param.resolution);
    }
    if (iterations) *iterations = run.iterations;
    if (!ok) return TIPLHIP_ERR_FAILED;
    return run.stopped ? TIPLHIP_ERR_TERMINATED : TIPLHIP_OK;
}

}  // namespace tiplhip

using namespace tiplhip;

extern "C" {

void tiplhip_cdm_param_default(tiplhip_cdm_param* p)
{
    if (!p) return;
    p->resolution = 2.0f;
    p->speed = 0.3f;
    p->smoothing = 0.05f;
    p->min_dimension = 16;
    p->max_iterations = 200;
    for (char& c : p->gradient_type) c = 'r';
    p->levels = 0;
    p->level_iterations = 0;
    p->field_bandwidth = 0;
}

int tiplhip_cdm(tiplhip_ctx* ctx, const tiplhip_volume* const* It, const tiplhip_volume* const* Is,
                int n, const tiplhip_cdm_param* param, float* field, const volatile uint8_t* terminated,
                int* iterations)
{
    return guarded([&] {
        TIPLHIP_REQUIRE(ctx && It && Is && field, "null argument");
        TIPLHIP_REQUIRE(n >= 1 && It[0], "null volume");
        std::lock_guard<std::recursive_mutex> lk(ctx->mu);
        DeviceGuard dg(ctx->device);
        const size_t fb = sizeof(float) * 3 * vox(It[0]->dim);
        DevMem best(ctx, fb);
        const int rc = cdm_device(ctx, It, Is, n, param, best.as<float>(), terminated, iterations);
        if (rc == TIPLHIP_ERR_FAILED) throw Error(TIPLHIP_ERR_FAILED, "cdm returned false (zero cost)");
        TIPLHIP_CHECK(hipMemcpyAsync(field, best.p, fb, hipMemcpyDeviceToHost, ctx->stream));
        TIPLHIP_CHECK(hipStreamSynchronize(ctx->stream));
        // As in the reference, a terminated run still returns its best field.
        if (rc == TIPLHIP_ERR_TERMINATED) throw Error(TIPLHIP_ERR_TERMINATED, "terminated by the caller");
    });
}

int tiplhip_cdm_level_iterations(tiplhip_ctx* ctx, const tiplhip_volume* It,
                                 const tiplhip_volume* Is, const tiplhip_cdm_param* param,
                                 int iterations, float* field)
{
    return guarded([&] {
        TIPLHIP_REQUIRE(ctx && It && Is, "null argument");
        TIPLHIP_REQUIRE(iterations >= 0, "negative iteration count");
        std::lock_guard<std::recursive_mutex> lk(ctx->mu);
        DeviceGuard dg(ctx->device);
        const int* dim = It->dim;
        auto vt = volume_ptrs(ctx, &It, 1, dim);
        auto vs = volume_ptrs(ctx, &Is, 1, dim);
        CdmRun run{ctx, {}, 1, nullptr};
        if (param)
            run.param = *param;
        else
            tiplhip_cdm_param_default(&run.param);
        run.param.max_iterations = iterations;
        run.no_early_exit = true;
        const size_t fb = sizeof(float) * 3 * vox(dim);
        DevMem best(ctx, fb);
        zero(ctx, best.p, fb);
        (void)cdm_level(run, vt, vs, dim, best.as<float>(), 1.0f);
        if (field) {
            TIPLHIP_CHECK(hipMemcpyAsync(field, best.p, fb, hipMemcpyDeviceToHost, ctx->stream));
            TIPLHIP_CHECK(hipStreamSynchronize(ctx->stream));
        }
    });
}

// ---- single steps on host buffers ----------------------------------------

int tiplhip_compose_displacement(tiplhip_ctx* ctx, const uint8_t* Is, const int dim[3],
                                 const float* d, uint8_t* Js)
{
    return guarded([&] {
        TIPLHIP_REQUIRE(ctx && Is && d && Js, "null argument");
        check_dim(dim);
        std::lock_guard<std::recursive_mutex> lk(ctx->mu);
        DeviceGuard dg(ctx->device);
        const FieldGeo g = field_geo(dim);
        DevMem dI(ctx, g.n), dD(ctx, sizeof(float) * 3 * g.n), dJ(ctx, g.n);
        TIPLHIP_CHECK(hipMemcpyAsync(dI.p, Is, g.n, hipMemcpyHostToDevice, ctx->stream));
        TIPLHIP_CHECK(hipMemcpyAsync(dD.p, d, sizeof(float) * 3 * g.n, hipMemcpyHostToDevice, ctx->stream));
        launch_compose(ctx, dI.as<uint8_t>(), dD.as<float>(), g, dJ.as<uint8_t>());
        TIPLHIP_CHECK(hipMemcpyAsync(Js, dJ.p, g.n, hipMemcpyDeviceToHost, ctx->stream));
        TIPLHIP_CHECK(hipStreamSynchronize(ctx->stream));
    });
}

int tiplhip_cdm_gradient(tiplhip_ctx* ctx, const uint8_t* Js, const uint8_t* It, const int dim[3],
                         char type, float* new_d, float* cost)
{
    return guarded([&] {
        TIPLHIP_REQUIRE(ctx && Js && It && new_d, "null argument");
        TIPLHIP_REQUIRE(type == 'r' || type == 'd', "gradient type must be 'r' or 'd'");
        check_dim(dim);
        std::lock_guard<std::recursive_mutex> lk(ctx->mu);
        DeviceGuard dg(ctx->device);
        const FieldGeo g = field_geo(dim);
        DevMem dJ(ctx, g.n), dI(ctx, g.n), dN(ctx, sizeof(float) * 3 * g.n);
        DevMem part(ctx, sizeof(double) * block_count(g)), dc(ctx, sizeof(float));
        TIPLHIP_CHECK(hipMemcpyAsync(dJ.p, Js, g.n, hipMemcpyHostToDevice, ctx->stream));
        TIPLHIP_CHECK(hipMemcpyAsync(dI.p, It, g.n, hipMemcpyHostToDevice, ctx->stream));
        launch_gradient(ctx, dJ.as<uint8_t>(), dI.as<uint8_t>(), g, type, 0, dN.as<float>(),
                        part.as<double>(), dc.as<float>());
        TIPLHIP_CHECK(hipMemcpyAsync(new_d, dN.p, sizeof(float) * 3 * g.n, hipMemcpyDeviceToHost, ctx->stream));
        float c = 0.0f;
        TIPLHIP_CHECK(hipMemcpyAsync(&c, dc.p, sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
        TIPLHIP_CHECK(hipStreamSynchronize(ctx->stream));
        if (cost) *cost = c;
    });
}

int tiplhip_cdm_poisson(tiplhip_ctx* ctx, float* new_d, const int dim[3], int sweeps,
                        float* max_length)
{
    return guarded([&] {
        TIPLHIP_REQUIRE(ctx && new_d, "null argument");
        TIPLHIP_REQUIRE(sweeps >= 1 && sweeps <= 1000, "sweeps must be in [1, 1000]");
        check_dim(dim);
        std::lock_guard<std::recursive_mutex> lk(ctx->mu);
        DeviceGuard dg(ctx->device);
        const FieldGeo g = field_geo(dim);
        const size_t fb = sizeof(float) * 3 * g.n;
        DevMem dN(ctx, fb), a(ctx, fb), b(ctx, fb), ml(ctx, sizeof(unsigned int)), mp(ctx, sizeof(float) * block_count(g));
        TIPLHIP_CHECK(hipMemcpyAsync(dN.p, new_d, fb, hipMemcpyHostToDevice, ctx->stream));
        float* out = launch_poisson(ctx, dN.as<float>(), a.as<float>(), b.as<float>(), g, sweeps,
                                    ml.as<unsigned int>(), mp.as<float>());
        TIPLHIP_CHECK(hipMemcpyAsync(new_d, out, fb, hipMemcpyDeviceToHost, ctx->stream));
        unsigned int bits = 0;
        TIPLHIP_CHECK(hipMemcpyAsync(&bits, ml.p, sizeof(bits), hipMemcpyDeviceToHost, ctx->stream));
        TIPLHIP_CHECK(hipStreamSynchronize(ctx->stream));
        if (max_length) std::memcpy(max_length, &bits, sizeof(float));
    });
}

int tiplhip_accumulate_displacement(tiplhip_ctx* ctx, const float* cur, float* new_d,
                                    const int dim[3], float scale)
{
    return guarded([&] {
        TIPLHIP_REQUIRE(ctx && cur && new_d, "null argument");
        check_dim(dim);
        std::lock_guard<std::recursive_mutex> lk(ctx->mu);
        DeviceGuard dg(ctx->device);
        const FieldGeo g = field_geo(dim);
        const size_t fb = sizeof(float) * 3 * g.n;
        DevMem dC(ctx, fb), dN(ctx, fb);
        TIPLHIP_CHECK(hipMemcpyAsync(dC.p, cur, fb, hipMemcpyHostToDevice, ctx->stream));
        TIPLHIP_CHECK(hipMemcpyAsync(dN.p, new_d, fb, hipMemcpyHostToDevice, ctx->stream));
        launch_accumulate(ctx, dC.as<float>(), dN.as<float>(), scale, g);
        TIPLHIP_CHECK(hipMemcpyAsync(new_d, dN.p, fb, hipMemcpyDeviceToHost, ctx->stream));
        TIPLHIP_CHECK(hipStreamSynchronize(ctx->stream));
    });
}

int tiplhip_cdm_smooth(tiplhip_ctx* ctx, const float* d, float* dd, const int dim[3], float smoothing)
{
    return guarded([&] {
        TIPLHIP_REQUIRE(ctx && d && dd, "null argument");
        check_dim(dim);
        if (smoothing == 0.0f) return;  // cdm_smooth leaves dd untouched
        std::lock_guard<std::recursive_mutex> lk(ctx->mu);
        DeviceGuard dg(ctx->device);
        const FieldGeo g = field_geo(dim);
        const size_t fb = sizeof(float) * 3 * g.n;
        DevMem dD(ctx, fb), dO(ctx, fb);
        TIPLHIP_CHECK(hipMemcpyAsync(dD.p, d, fb, hipMemcpyHostToDevice, ctx->stream));
        launch_smooth(ctx, dD.as<float>(), dO.as<float>(), smoothing, g);
        TIPLHIP_CHECK(hipMemcpyAsync(dd, dO.p, fb, hipMemcpyDeviceToHost, ctx->stream));
        TIPLHIP_CHECK(hipStreamSynchronize(ctx->stream));
    });
}

int tiplhip_upsample_field(tiplhip_ctx* ctx, const float* in, const int in_dim[3], float* out,
                           const int out_dim[3])
{
    return guarded([&] {
        TIPLHIP_REQUIRE(ctx && in && out, "null argument");
        check_dim(in_dim);
        check_dim(out_dim);
        std::lock_guard<std::recursive_mutex> lk(ctx->mu);
        DeviceGuard dg(ctx->device);
        const FieldGeo gi = field_geo(in_dim), go = field_geo(out_dim);
        DevMem dI(ctx, sizeof(float) * 3 * gi.n), dO(ctx, sizeof(float) * 3 * go.n);
        TIPLHIP_CHECK(hipMemcpyAsync(dI.p, in, sizeof(float) * 3 * gi.n, hipMemcpyHostToDevice, ctx->stream));
        launch_upsample(ctx, dI.as<float>(), gi, go, dO.as<float>());
        TIPLHIP_CHECK(hipMemcpyAsync(out, dO.p, sizeof(float) * 3 * go.n, hipMemcpyDeviceToHost, ctx->stream));
        TIPLHIP_CHECK(hipStreamSynchronize(ctx->stream));
    });
}

}  // extern "C"
