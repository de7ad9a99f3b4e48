// field.hpp -- displacement-field geometry and trilinear helpers shared by
// the CDM loop (cdm.hip) and the post-registration warping kernels (warp.hip).
//
// Fields are packed float3 per voxel (the reference's image<3,vector<3>>),
// x-fastest.  Field kernels run on a 3-D grid of 64 x 4 thread tiles
// (x = one wave, grid z = one plane), so no index division is needed.
#pragma once

#include "device.hpp"
#include "internal.hpp"

namespace tiplhip {

struct FieldGeo {
    int w, h, d;
    uint32_t wh;
    uint32_t n;
    float wm1, hm1, dm1;
};

inline FieldGeo field_geo(const int* dim)
{
    FieldGeo g;
    g.w = dim[0];
    g.h = dim[1];
    g.d = dim[2];
    g.wh = uint32_t(dim[0]) * uint32_t(dim[1]);
    g.n = g.wh * uint32_t(dim[2]);
    g.wm1 = float(dim[0] - 1);
    g.hm1 = float(dim[1] - 1);
    g.dm1 = float(dim[2] - 1);
    return g;
}

constexpr int kTX = 64, kTY = 4;

inline dim3 tile_grid(const FieldGeo& g)
{
    return dim3(uint32_t((g.w + kTX - 1) / kTX), uint32_t((g.h + kTY - 1) / kTY), uint32_t(g.d));
}

struct Vox {
    int x, y, z;
    uint32_t p;
    bool ok;
};

// The block's logical position in a 3-D grid and its x-fastest linear
// index.  Hardware hands workgroup b (in x-fastest order) to XCD b % 8;
// through xcd_remap each XCD works through a contiguous range of logical
// blocks instead, so the halo and neighbourhood reads of adjacent tiles and
// planes meet in one XCD's L2.  Per-block partial results are stored at the
// logical index, so reductions over them keep their order.
struct Block {
    int x, y, z;
    uint32_t lin;
};
__device__ inline Block logical_block()
{
    const uint32_t gx = gridDim.x, gy = gridDim.y;
    const uint32_t hw = (blockIdx.z * gy + blockIdx.y) * gx + blockIdx.x;
    const uint32_t L = xcd_remap(hw, gx * gy * gridDim.z);
    Block b;
    b.x = int(L % gx);
    b.y = int((L / gx) % gy);
    b.z = int(L / (gx * gy));
    b.lin = L;
    return b;
}

// This thread's voxel of the 64 x 4 tile of block b.
__device__ inline Vox voxel_of(const FieldGeo& g, const Block& b)
{
    Vox v;
    v.x = b.x * kTX + int(threadIdx.x);
    v.y = b.y * kTY + int(threadIdx.y);
    v.z = b.z;
    v.ok = v.x < g.w && v.y < g.h;
    v.p = uint32_t(v.z) * g.wh + uint32_t(v.y) * uint32_t(g.w) + uint32_t(v.x);
    return v;
}

__device__ inline Vox this_voxel(const FieldGeo& g)
{
    const Block b = logical_block();
    Vox v;
    v.x = b.x * kTX + int(threadIdx.x);
    v.y = b.y * kTY + int(threadIdx.y);
    v.z = b.z;
    v.ok = v.x < g.w && v.y < g.h;
    v.p = uint32_t(v.z) * g.wh + uint32_t(v.y) * uint32_t(g.w) + uint32_t(v.x);
    return v;
}

__device__ inline float3 ld3(const float* f, uint32_t p)
{
    return make_float3(f[3 * size_t(p)], f[3 * size_t(p) + 1], f[3 * size_t(p) + 2]);
}

__device__ inline void st3(float* f, uint32_t p, float3 v)
{
    f[3 * size_t(p)] = v.x;
    f[3 * size_t(p) + 1] = v.y;
    f[3 * size_t(p) + 2] = v.z;
}

// Trilinear location as interpolator::linear<3>::get_location
// (numerical/interpolation.hpp:337-383); false when outside.
struct Tri {
    uint32_t i0;
    int ix, iy, iz;
    float w[8];
};

__device__ inline bool locate(const FieldGeo& g, float x, float y, float z, Tri& t)
{
    // min(x,y,z) >= 0 is x,y,z >= 0 (a NaN coordinate fails its upper test)
    const bool in = (fminf(fminf(x, y), z) >= 0.0f) & (x < g.wm1) & (y < g.hm1) & (z < g.dm1);
    if (!in) return false;
    const uint32_t ix = uint32_t(x), iy = uint32_t(y), iz = uint32_t(z);
    const float p0 = x - float(ix), p1 = y - float(iy), p2 = z - float(iz);
    const float n0 = 1.0f - p0, n1 = 1.0f - p1, n2 = 1.0f - p2;
    t.i0 = iz * g.wh + iy * uint32_t(g.w) + ix;
    t.ix = int(ix);
    t.iy = int(iy);
    t.iz = int(iz);
    const float n0n1 = n0 * n1, p0n1 = p0 * n1, n0p1 = n0 * p1, p0p1 = p0 * p1;
    t.w[0] = n0n1 * n2;
    t.w[1] = p0n1 * n2;
    t.w[2] = n0p1 * n2;
    t.w[3] = p0p1 * n2;
    t.w[4] = n0n1 * p2;
    t.w[5] = p0n1 * p2;
    t.w[6] = n0p1 * p2;
    t.w[7] = p0p1 * p2;
    return true;
}

__device__ inline uint32_t corner(const FieldGeo& g, uint32_t i0, int k)
{
    return i0 + uint32_t(k & 1) + ((k & 2) ? uint32_t(g.w) : 0u) + ((k & 4) ? g.wh : 0u);
}


// Stream-ordered scratch from the context's pool, freed on scope exit.
struct DevMem {
    tiplhip_ctx* ctx = nullptr;
    void* p = nullptr;
    DevMem(tiplhip_ctx* c, size_t bytes) : ctx(c), p(pool_alloc(c, bytes)) {}
    DevMem(const DevMem&) = delete;
    DevMem& operator=(const DevMem&) = delete;
    DevMem(DevMem&& o) noexcept : ctx(o.ctx), p(o.p) { o.p = nullptr; }
    ~DevMem() { pool_free(ctx, p); }
    template <class T>
    T* as() const { return static_cast<T*>(p); }
};

inline void check_dim(const int* dim)
{
    TIPLHIP_REQUIRE(dim && dim[0] > 0 && dim[1] > 0 && dim[2] > 0, "dims must be positive");
    TIPLHIP_REQUIRE(vox(dim) < (size_t(1) << 31), "volume exceeds 2^31 voxels");
}

}  // namespace tiplhip
