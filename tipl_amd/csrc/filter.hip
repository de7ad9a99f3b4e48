// filter.hip -- the binomial Gaussian filter of filter/gaussian.hpp:18-105 on gfx950.
//
// gaussian(I, bandwidth) applies `bandwidth` rounds of the separable [1 2 1]
// stencil, x then y then z, with replicate edges, and scales the result by
// 2^(-2*bandwidth*3) (bandwidth b has variance b/2 per axis).  Each pass of
// the reference writes a float array, so every intermediate value is rounded
// to float; one launch here runs one whole round (three passes) and keeps
// those roundings: the x pass of a tile (with a one-voxel halo) is written to
// LDS, the y pass of it stays in registers, and the z pass combines three
// consecutive y-pass planes while the block marches along z.  HBM traffic per
// round is one read (plus halo) and one write of the image instead of the
// reference's three of each.
//
// Works on float images of CH interleaved channels: CH = 1 for a scalar
// volume, CH = 3 for a vector<3,float> displacement field.
#include "device.hpp"
#include "internal.hpp"

#include <algorithm>
#include <cmath>

namespace tiplhip {
namespace {

constexpr int kGX = 64;        // output columns per block (one wave)
constexpr int kGY = 8;         // output rows per block (2 per thread row)
constexpr int kGRows = kGY / 4;
constexpr int kGThreads = 256;

struct GaussGeo {
    int w, h, d;
    uint32_t wh;
};

// One round over a 64 x 8 column of `zq` output planes.  Each thread owns
// fixed slots of the haloed plane tile A (AR x AC floats): their in-plane
// source offsets (with the replicate clamp) are computed once, and the next
// plane's values are loaded into registers while the current plane is
// filtered, so the global-load latency overlaps the LDS passes.
template <int CH>
__global__ __launch_bounds__(kGThreads) void gaussian_round_kernel(const float* __restrict__ a,
                                                                   float* __restrict__ b, GaussGeo g,
                                                                   float scale, int zq)
{
    constexpr int AR = kGY + 2, AC = (kGX + 2) * CH, XC = kGX * CH;
    constexpr int NA = AR * AC;
    constexpr int NL = (NA + kGThreads - 1) / kGThreads;
    __shared__ float A[NA];
    __shared__ float X[AR][XC];
    const int tx = threadIdx.x, ty = threadIdx.y, tid = ty * kGX + tx;
    // each XCD takes a contiguous range of blocks (x fastest), so the halo
    // rows and planes adjacent blocks share are read through one L2
    const uint32_t gx = gridDim.x, gy = gridDim.y;
    const uint32_t L = xcd_remap((blockIdx.z * gy + blockIdx.y) * gx + blockIdx.x, gx * gy * gridDim.z);
    const int x0 = int(L % gx) * kGX, y0 = int((L / gx) % gy) * kGY, z0 = int(L / (gx * gy)) * zq;
    const int zend = min(z0 + zq, g.d);
    const uint32_t plane = g.wh * uint32_t(CH);

    // slot k holds A[tid + 256 k]: row r = e / AC, column (x0-1+col) and
    // channel comp; a clamped neighbour is the centre itself (the
    // reference's replicate edge)
    uint32_t off[NL];
#pragma unroll
    for (int k = 0; k < NL; ++k) {
        const int e = min(tid + kGThreads * k, NA - 1);
        const int r = e / AC, cc = e - r * AC;
        const int col = cc / CH, comp = cc - col * CH;
        const int gy = min(max(y0 - 1 + r, 0), g.h - 1);
        const int gx = min(max(x0 - 1 + col, 0), g.w - 1);
        off[k] = (uint32_t(gy) * uint32_t(g.w) + uint32_t(gx)) * uint32_t(CH) + uint32_t(comp);
    }
    float v[NL];
    {
        const uint32_t pz = uint32_t(min(max(z0 - 1, 0), g.d - 1));
#pragma unroll
        for (int k = 0; k < NL; ++k) v[k] = a[size_t(pz) * plane + off[k]];
    }

    float Yp[kGRows][CH], Yc[kGRows][CH];
#pragma unroll
    for (int j = 0; j < kGRows; ++j)
#pragma unroll
        for (int c = 0; c < CH; ++c) Yp[j][c] = Yc[j][c] = 0.0f;

    for (int zz = z0 - 1; zz <= zend; ++zz) {
#pragma unroll
        for (int k = 0; k < NL; ++k)
            if (tid + kGThreads * k < NA) A[tid + kGThreads * k] = v[k];
        if (zz < zend) {
            const uint32_t pz = uint32_t(min(zz + 1, g.d - 1));
#pragma unroll
            for (int k = 0; k < NL; ++k) v[k] = a[size_t(pz) * plane + off[k]];
        }
        __syncthreads();
        // x pass: (left + (c + c)) + right
        for (int e = tid; e < AR * kGX; e += kGThreads) {
            const int r = e / kGX, col = e - r * kGX;
#pragma unroll
            for (int c = 0; c < CH; ++c) {
                float center = A[r * AC + (col + 1) * CH + c];
                center += center;
                X[r][col * CH + c] = (A[r * AC + col * CH + c] + center) + A[r * AC + (col + 2) * CH + c];
            }
        }
        __syncthreads();
        // y pass of this thread's rows of plane zz
        float Yn[kGRows][CH];
#pragma unroll
        for (int j = 0; j < kGRows; ++j) {
            const int r = ty + 4 * j;
#pragma unroll
            for (int c = 0; c < CH; ++c) {
                float center = X[r + 1][tx * CH + c];
                center += center;
                Yn[j][c] = (X[r][tx * CH + c] + center) + X[r + 2][tx * CH + c];
            }
        }
        // z pass of plane zz-1 from the y passes of planes zz-2, zz-1, zz
        if (zz > z0) {
            const int x = x0 + tx;
            const uint32_t oz = uint32_t(zz - 1);
#pragma unroll
            for (int j = 0; j < kGRows; ++j) {
                const int y = y0 + ty + 4 * j;
                if (x < g.w && y < g.h) {
                    const size_t o = size_t(oz * g.wh + uint32_t(y) * uint32_t(g.w) + uint32_t(x)) * CH;
#pragma unroll
                    for (int c = 0; c < CH; ++c) {
                        float center = Yc[j][c];
                        center += center;
                        b[o + c] = ((Yp[j][c] + center) + Yn[j][c]) * scale;
                    }
                }
            }
        }
#pragma unroll
        for (int j = 0; j < kGRows; ++j)
#pragma unroll
            for (int c = 0; c < CH; ++c) {
                Yp[j][c] = Yc[j][c];
                Yc[j][c] = Yn[j][c];
            }
    }
}

GaussGeo gauss_geo(const int* dim)
{
    return GaussGeo{dim[0], dim[1], dim[2], uint32_t(dim[0]) * uint32_t(dim[1])};
}

// Planes per block: about 1024 blocks over the volume (each block reads two
// halo planes beyond its zq), at least 4 and at most 16 planes.
int gauss_zq(const GaussGeo& g)
{
    const long tiles = long((g.w + kGX - 1) / kGX) * long((g.h + kGY - 1) / kGY);
    const long zq = long(g.d) * tiles / 1024;
    return int(std::min(16L, std::max(4L, zq)));
}

void launch_round(tiplhip_ctx* ctx, const float* in, float* out, const GaussGeo& g, int channels,
                  float scale)
{
    const int zq = gauss_zq(g);
    const dim3 grid(uint32_t((g.w + kGX - 1) / kGX), uint32_t((g.h + kGY - 1) / kGY),
                    uint32_t((g.d + zq - 1) / zq));
    const dim3 block(kGX, 4);
    if (channels == 1)
        hipLaunchKernelGGL(gaussian_round_kernel<1>, grid, block, 0, ctx->stream, in, out, g, scale, zq);
    else
        hipLaunchKernelGGL(gaussian_round_kernel<3>, grid, block, 0, ctx->stream, in, out, g, scale, zq);
    TIPLHIP_CHECK(hipGetLastError());
}

void check_gaussian(const int* dim, int channels, unsigned bandwidth)
{
    TIPLHIP_REQUIRE(channels == 1 || channels == 3, "gaussian: channels must be 1 or 3");
    TIPLHIP_REQUIRE(bandwidth <= kMaxGaussianBandwidth, "gaussian: bandwidth above 16");
    TIPLHIP_REQUIRE(vox(dim) * size_t(channels) < (size_t(1) << 32),
                    "gaussian: image exceeds 2^32 floats");
}

}  // namespace

void launch_gaussian_to(tiplhip_ctx* ctx, const float* src, float* dst, float* tmp,
                        const int* dim, int channels, unsigned bandwidth)
{
    check_gaussian(dim, channels, bandwidth);
    const size_t bytes = sizeof(float) * size_t(channels) * vox(dim);
    if (bandwidth == 0) {
        TIPLHIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, ctx->stream));
        return;
    }
    const GaussGeo g = gauss_geo(dim);
    const float last_scale = std::ldexp(1.0f, -int(6 * bandwidth));
    // round k writes dst when bandwidth-1-k is even, tmp otherwise, so the
    // last round lands in dst and src is only read
    const float* in = src;
    for (unsigned k = 0; k < bandwidth; ++k) {
        float* out = ((bandwidth - 1 - k) % 2 == 0) ? dst : tmp;
        launch_round(ctx, in, out, g, channels, k + 1 == bandwidth ? last_scale : 1.0f);
        in = out;
    }
}

float* launch_gaussian(tiplhip_ctx* ctx, float* a, float* tmp, const int* dim, int channels,
                       unsigned bandwidth)
{
    check_gaussian(dim, channels, bandwidth);
    if (bandwidth == 0 || vox(dim) == 0) return a;
    const GaussGeo g = gauss_geo(dim);
    // the scale is a power of two with a normal float value for bandwidth <= 16,
    // so scaling in float equals the reference's float(double(v) * scale)
    const float last_scale = std::ldexp(1.0f, -int(6 * bandwidth));
    float *src = a, *dst = tmp;
    for (unsigned k = 0; k < bandwidth; ++k) {
        launch_round(ctx, src, dst, g, channels, (k + 1 == bandwidth) ? last_scale : 1.0f);
        std::swap(src, dst);
    }
    return src;
}

}  // namespace tiplhip

using namespace tiplhip;

extern "C" {

int tiplhip_gaussian(tiplhip_ctx* ctx, float* data, const int dim[3], int channels,
                     unsigned bandwidth)
{
    return guarded([&] {
        TIPLHIP_REQUIRE(ctx && data && dim, "null argument");
        TIPLHIP_REQUIRE(dim[0] > 0 && dim[1] > 0 && dim[2] > 0, "dims must be positive");
        TIPLHIP_REQUIRE(channels == 1 || channels == 3, "gaussian: channels must be 1 or 3");
        TIPLHIP_REQUIRE(bandwidth <= kMaxGaussianBandwidth, "gaussian: bandwidth above 16");
        if (bandwidth == 0) return;  // the reference returns the image unchanged
        std::lock_guard<std::recursive_mutex> lk(ctx->mu);
        DeviceGuard dg(ctx->device);
        const size_t bytes = sizeof(float) * size_t(channels) * vox(dim);
        DeviceBuffer a, t;
        a.reserve(bytes);
        t.reserve(bytes);
        TIPLHIP_CHECK(hipMemcpyAsync(a.p, data, bytes, hipMemcpyHostToDevice, ctx->stream));
        float* r = launch_gaussian(ctx, a.as<float>(), t.as<float>(), dim, channels, bandwidth);
        TIPLHIP_CHECK(hipMemcpyAsync(data, r, bytes, hipMemcpyDeviceToHost, ctx->stream));
        TIPLHIP_CHECK(hipStreamSynchronize(ctx->stream));
    });
}

}  // extern "C"
