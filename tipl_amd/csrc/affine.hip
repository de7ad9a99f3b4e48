// affine.hip -- affine resampling, correlation sums and the MI joint
// histogram for K candidate transforms per launch (gfx950).
//
// Restates, MI355X-first:
//   tipl::resample            numerical/resampling.hpp:1239-1299
//   reg::correlation(_cuda)   reg/linear.hpp:23-78  + tipl::correlation
//                             numerical/statistics.hpp:186-625
//   reg::get_mutual_info      reg/linear.hpp:138-169
//   reg::get_mutual_info_mean reg/linear.hpp:87-134
//
// Work decomposition.  The `from` grid is cut into wave tiles of 16 x 4
// voxels of one plane, each marched through a run of planes (an item).  A
// launch covers one or more image pairs (segments), each with its own K
// candidate transforms; a segment contributes K x C workgroups, each adding
// a chunk of items for one candidate.  Workgroups are remapped so that each
// XCD works through a contiguous range of one segment, and each workgroup
// finds its segment by a binary search over the segments' first blocks.
// Integer results (sums, histograms) are accumulated with integer atomics
// only, so they are exact and independent of scheduling and batching.
#include "device.hpp"
#include "host_wait.hpp"
#include "internal.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <mutex>
#include <type_traits>
#include <vector>

namespace tiplhip {
namespace {

// Work item = one wave tile column (16 x 4 voxels of a plane) marched
// through a run of zc planes.
struct BatchGeo {
    int fw, fh, fd;             // from grid
    uint32_t ntx, ncols;        // tiles per row, tile columns per plane
    uint32_t zc, nzc;           // planes per item, z runs
    uint32_t items;             // ncols * nzc
    uint32_t items_per_block;
    uint32_t K;
    uint32_t total_blocks;      // K * ceil(items / items_per_block)
    uint32_t from_bytes, to_bytes8;
    VolGeo to;
};

VolGeo make_geo(const int* dim)
{
    VolGeo g;
    g.w = dim[0];
    g.h = dim[1];
    g.d = dim[2];
    g.wh = uint32_t(dim[0]) * uint32_t(dim[1]);
    g.wm1 = float(dim[0] - 1);
    g.hm1 = float(dim[1] - 1);
    g.dm1 = float(dim[2] - 1);
    return g;
}

constexpr uint32_t kTileX = 16, kTileY = 4;  // one wave = 16 x 4 voxels of one plane

// Item geometry of one pair for K candidates, items of up to zc planes
// (blocks chosen by batch_blocks).
BatchGeo batch_geometry(const tiplhip_volume* from, const tiplhip_volume* to, int K, uint32_t zc)
{
    BatchGeo b;
    b.fw = from->dim[0];
    b.fh = from->dim[1];
    b.fd = from->dim[2];
    b.ntx = (uint32_t(b.fw) + kTileX - 1) / kTileX;
    b.ncols = b.ntx * ((uint32_t(b.fh) + kTileY - 1) / kTileY);
    b.zc = std::max<uint32_t>(1u, std::min<uint32_t>(zc, uint32_t(b.fd)));
    b.nzc = (uint32_t(b.fd) + b.zc - 1) / b.zc;
    b.items = b.ncols * b.nzc;
    b.K = uint32_t(K);
    b.items_per_block = 1;
    b.total_blocks = 0;
    b.from_bytes = uint32_t(from->n);
    b.to_bytes8 = uint32_t(to->n * 8u);
    b.to = make_geo(to->dim);
    return b;
}

// Chooses items per workgroup so that the pair's K candidates take about
// `target_blocks` workgroups; max_voxels_per_block bounds the voxels one
// workgroup adds (the uint16 LDS counters of the 256-bin histogram hold at
// most 65535).
void batch_blocks(BatchGeo& b, uint32_t max_voxels_per_block, uint32_t target_blocks)
{
    target_blocks = std::max<uint32_t>(1u, target_blocks);
    const uint32_t max_items = std::max<uint32_t>(1u, max_voxels_per_block / (64u * b.zc));
    uint64_t want = (uint64_t(b.items) * uint64_t(b.K) + target_blocks - 1) / target_blocks;
    uint32_t ipb = uint32_t(std::min<uint64_t>(want, max_items));
    ipb = std::max<uint32_t>(ipb, std::min<uint32_t>(4u, max_items));
    b.items_per_block = ipb;
    b.total_blocks = ((b.items + ipb - 1) / ipb) * b.K;
}

// One image pair of a launch.  Candidates are numbered globally: the
// segment's candidates are cand_begin .. cand_begin + g.K - 1, its
// workgroups block_begin .. block_begin + g.total_blocks - 1.
struct Segment {
    BatchGeo g;
    const uint8_t* from;
    const uint8_t* to;
    const uint64_t* to8;
    const uint32_t* from_hist;   // value histogram of `from` at the launch's bin count
    long long sx, sxx;           // exact sums of `from` (correlation)
    double n;                    // voxels of `from`
    uint32_t block_begin;
    uint32_t cand_begin;
};

// Largest i with key(segs[i]) <= v, for a wave-uniform v (scalar loads).
template <class Key>
__device__ inline uint32_t find_segment(const Segment* __restrict__ segs, uint32_t nseg, uint32_t v,
                                        Key key)
{
    uint32_t lo = 0, hi = nseg;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (key(segs[mid]) <= v) lo = mid;
        else hi = mid;
    }
    return lo;
}

struct ByBlock {
    __device__ uint32_t operator()(const Segment& s) const { return s.block_begin; }
};
struct ByCandidate {
    __device__ uint32_t operator()(const Segment& s) const { return s.cand_begin; }
};

// Wave-uniform decode of an item into this lane's (x, y) and the z run.
struct Item {
    uint32_t x, y;      // this lane's voxel column
    uint32_t z0, z1;    // planes [z0, z1)
    bool ok;
};
__device__ inline Item decode(const BatchGeo& g, uint32_t it, uint32_t lane)
{
    const uint32_t c = it % g.ncols;
    const uint32_t j = it / g.ncols;
    const uint32_t tx = c % g.ntx;
    const uint32_t ty = c / g.ntx;
    Item m;
    m.x = tx * kTileX + (lane & (kTileX - 1));
    m.y = ty * kTileY + (lane / kTileX);
    m.z0 = j * g.zc;
    m.z1 = min(m.z0 + g.zc, uint32_t(g.fd));
    m.ok = m.x < uint32_t(g.fw) && m.y < uint32_t(g.fh);
    return m;
}

// The x/y part of vector_transformation's sums, ((x*r0 + y*r1) + z*r2) + s0
// (numerical/transformation.hpp:58-73): x*r0 + y*r1 is the first rounded
// partial sum, so it is formed once per column and reused for every z.
struct ColumnSums {
    float hx, hy, hz;
};
__device__ inline ColumnSums column_sums(const float* T, float fx, float fy)
{
    ColumnSums s;
    s.hx = fx * T[0] + fy * T[1];
    s.hy = fx * T[3] + fy * T[4];
    s.hz = fx * T[6] + fy * T[7];
    return s;
}

// Position of voxel (x, y, z) of the from grid in the to volume:
// vector_transformation(.., rotation, shift, vdim<3>)
// (numerical/transformation.hpp:58-73): ((x*r0 + y*r1) + z*r2) + s0.
struct Pos {
    float x, y, z;
};
__device__ inline Pos apply(const float* T, float fx, float fy, float fz)
{
    Pos p;
    p.x = fx * T[0] + fy * T[1] + fz * T[2] + T[9];
    p.y = fx * T[3] + fy * T[4] + fz * T[5] + T[10];
    p.z = fx * T[6] + fy * T[7] + fz * T[8] + T[11];
    return p;
}

// fz * T[2], fz * T[5], fz * T[8] of plane z (the z terms of the transformed
// position).  Large launches take them from a table built once per
// candidate and plane by ztab_kernel -- the same FP32 products, so the
// positions are bit-identical -- and read it with scalar loads, since z is
// uniform across a wave: four fewer VALU instructions per sample.
template <bool ZT>
__device__ inline float4 z_products(const float4* __restrict__ ztab, uint32_t row, uint32_t z,
                                    const float* T)
{
    if constexpr (ZT) {
        return ztab[row + z];
    } else {
        const float fz = float(z);
        return make_float4(fz * T[2], fz * T[5], fz * T[8], 0.0f);
    }
}

// Whether every sample of this wave's columns lies inside the `to` volume,
// so the march can skip the per-sample bounds test (sample2_u8_packed_in).
// For fixed (x, y) each rounded step of ((hx + z*r2) + s) is monotone in z,
// so a column's positions lie between those of its first and last plane:
// the column is inside when both ends are.  Also requires w*h*d <= 2^24
// (small_to) for the exact FP32 cell index.  Wave-uniform.
template <bool ZT>
__device__ inline bool column_inside(const BatchGeo& g, const ColumnSums& cs, const float* T,
                                     const float4* __restrict__ ztab, uint32_t zrow, const Item& m,
                                     bool small_to)
{
    if (!small_to) return false;
    const float4 a = z_products<ZT>(ztab, zrow, m.z0, T), b = z_products<ZT>(ztab, zrow, m.z1 - 1, T);
    const bool in = inside_u8(g.to, (cs.hx + a.x) + T[9], (cs.hy + a.y) + T[10], (cs.hz + a.z) + T[11]) &&
                    inside_u8(g.to, (cs.hx + b.x) + T[9], (cs.hy + b.y) + T[10], (cs.hz + b.z) + T[11]);
    return __ballot(in) == __ballot(1);
}

__global__ __launch_bounds__(256) void ztab_kernel(const float* __restrict__ Ts, uint32_t K,
                                                   uint32_t zstride, float4* __restrict__ ztab)
{
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= K * zstride) return;
    const uint32_t k = i / zstride;
    const float fz = float(i % zstride);
    ztab[i] = make_float4(fz * Ts[12 * k + 2], fz * Ts[12 * k + 5], fz * Ts[12 * k + 8], 0.0f);
}

// ---------------------------------------------------------------------------
// Correlation: per candidate the exact sums Sy, Syy, Sxy (Sx, Sxx of `from`
// do not depend on the transform and are kept with the volume).
// crop[k] = {flag, sx, sy, sz}: flag 1 reproduces the CPU resample's exact
// crop for an integer-shift transform (numerical/resampling.hpp:1244-1262).
// ---------------------------------------------------------------------------
template <bool ZT>
__global__ __launch_bounds__(256) void corr_sums_kernel(const Segment* __restrict__ segs,
                                                        uint32_t nseg, uint32_t total_blocks,
                                                        const float* __restrict__ Ts,
                                                        const int4* __restrict__ crop,
                                                        const float4* __restrict__ ztab,
                                                        uint32_t zstride,
                                                        unsigned long long* __restrict__ sums)
{
    const uint32_t L = xcd_remap(blockIdx.x, total_blocks);
    const Segment& sg = segs[find_segment(segs, nseg, L, ByBlock{})];
    const BatchGeo g = sg.g;
    const uint32_t local = L - sg.block_begin;
    const uint32_t cand = sg.cand_begin + local % g.K;
    const uint32_t chunk = local / g.K;
    float T[12];
#pragma unroll
    for (int i = 0; i < 12; ++i) T[i] = Ts[12 * cand + i];
    const int4 cr = crop[cand];
    const uint32_t zrow = cand * zstride;
    const __amdgpu_buffer_rsrc_t rf = make_rsrc(sg.from, g.from_bytes);
    const __amdgpu_buffer_rsrc_t rt = make_rsrc(sg.to, g.to_bytes8 / 8u);
    const __amdgpu_buffer_rsrc_t r8 = make_rsrc(sg.to8, g.to_bytes8);
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const uint32_t nwaves = blockDim.x >> 6;
    const uint32_t it0 = chunk * g.items_per_block;
    const uint32_t it1 = min(it0 + g.items_per_block, g.items);
    const uint32_t fwh = uint32_t(g.fw) * uint32_t(g.fh);
    const bool small_to = g.to.wh * uint64_t(uint32_t(g.to.d)) <= (uint64_t(1) << 24);
    const float tw = float(g.to.w), twh = float(g.to.wh);
    uint32_t sy = 0, syy = 0, sxy = 0;
    for (uint32_t it = it0 + wave; it < it1; it += nwaves) {
        const Item m = decode(g, it, lane);
        if (!m.ok) continue;
        const ColumnSums cs = column_sums(T, float(m.x), float(m.y));
        uint32_t fidx = m.z0 * fwh + m.y * uint32_t(g.fw) + m.x;
        auto add = [&](uint32_t xv, uint32_t yv) {
            sy += yv;
            syy += yv * yv;
            sxy += xv * yv;
        };
        if (cr.x) {
            for (uint32_t z = m.z0; z < m.z1; ++z, fidx += fwh) {
                int sx = int(m.x) + cr.y, s_y = int(m.y) + cr.z, sz = int(z) + cr.w;
                bool in = sx >= 0 && sx < g.to.w && s_y >= 0 && s_y < g.to.h && sz >= 0 &&
                          sz < g.to.d;
                add(load_u8(rf, fidx, 0),
                    in ? load_u8(rt, uint32_t(sz) * g.to.wh + uint32_t(s_y) * uint32_t(g.to.w) +
                                         uint32_t(sx), 0)
                       : 0u);
            }
            continue;
        }
        // two planes per step: both samples' gathers are in flight together
        auto march = [&](auto inside_tag) {
            constexpr bool kIn = decltype(inside_tag)::value;
            uint32_t z = m.z0;
            for (; z + 1 < m.z1; z += 2, fidx += 2 * fwh) {
                const uint32_t xa = load_u8(rf, fidx, 0), xb = load_u8(rf, fidx + fwh, 0);
                const float4 za = z_products<ZT>(ztab, zrow, z, T), zb = z_products<ZT>(ztab, zrow, z + 1, T);
                uint32_t ya, yb;
                if constexpr (kIn)
                    sample2_u8_packed_in(r8, tw, twh, (cs.hx + za.x) + T[9], (cs.hy + za.y) + T[10],
                                         (cs.hz + za.z) + T[11], (cs.hx + zb.x) + T[9],
                                         (cs.hy + zb.y) + T[10], (cs.hz + zb.z) + T[11], ya, yb);
                else
                    sample2_u8_packed(r8, g.to, (cs.hx + za.x) + T[9], (cs.hy + za.y) + T[10],
                                      (cs.hz + za.z) + T[11], (cs.hx + zb.x) + T[9],
                                      (cs.hy + zb.y) + T[10], (cs.hz + zb.z) + T[11], ya, yb);
                add(xa, ya);
                add(xb, yb);
            }
            if (z < m.z1) {
                const float4 zp = z_products<ZT>(ztab, zrow, z, T);
                add(load_u8(rf, fidx, 0), sample_u8_packed(r8, g.to, (cs.hx + zp.x) + T[9],
                                                           (cs.hy + zp.y) + T[10], (cs.hz + zp.z) + T[11]));
            }
        };
        // measured at 182x218x182 with 64 candidates: 0.672-0.692 -> 0.657-0.663
        // ms (candidates up to 10 degrees and 8 voxels off), 0.617-0.633 ->
        // 0.601-0.616 ms (candidates within 1 degree and 1 voxel)
        if (column_inside<ZT>(g, cs, T, ztab, zrow, m, small_to))
            march(std::true_type{});
        else
            march(std::false_type{});
    }
    __shared__ unsigned long long red[3][4];
    unsigned long long a = wave_sum_u64(sy), b = wave_sum_u64(syy), c = wave_sum_u64(sxy);
    if (lane == 0) {
        red[0][wave] = a;
        red[1][wave] = b;
        red[2][wave] = c;
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        unsigned long long s = 0;
        for (uint32_t w = 0; w < nwaves; ++w) s += red[threadIdx.x][w];
        atomicAdd(&sums[cand * 4 + threadIdx.x], s);
    }
}

// tipl::correlation from exact sums (numerical/statistics.hpp:555-586,
// 407-443) in FP64, then `float c; 1.0f - c*c` (reg/linear.hpp:34-35).
__global__ void corr_finalize_kernel(const Segment* __restrict__ segs, uint32_t nseg,
                                     const unsigned long long* __restrict__ sums, int K,
                                     double* __restrict__ costs, double* __restrict__ host_costs,
                                     long long* __restrict__ sums_out)
{
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= K) return;
    const Segment& sg = segs[find_segment(segs, nseg, uint32_t(k), ByCandidate{})];
    const long long sx = sg.sx, sxx = sg.sxx;
    const double n = sg.n;
    long long sy = (long long)sums[k * 4 + 0];
    long long syy = (long long)sums[k * 4 + 1];
    long long sxy = (long long)sums[k * 4 + 2];
    double mean_x = double(sx) / n;
    double mean_y = double(sy) / n;
    double var_x = double(sxx) / n - mean_x * mean_x;
    double var_y = double(syy) / n - mean_y * mean_y;
    double sd1 = var_x > 0.0 ? sqrt(var_x) : 0.0;
    double sd2 = var_y > 0.0 ? sqrt(var_y) : 0.0;
    double r = 0.0;
    if (!(sd1 == 0.0 || sd2 == 0.0)) {
        double cov = double(sxy) / n - mean_x * mean_y;
        r = cov / sd1 / sd2;
    }
    float c = float(r);
    costs[k] = double(1.0f - c * c);
    if (host_costs) host_costs[k] = double(1.0f - c * c);
    if (sums_out) {
        sums_out[k * 5 + 0] = sx;
        sums_out[k * 5 + 1] = sxx;
        sums_out[k * 5 + 2] = sy;
        sums_out[k * 5 + 3] = syy;
        sums_out[k * 5 + 4] = sxy;
    }
}

// ---------------------------------------------------------------------------
// Mutual information: joint histogram in LDS.
//   BINS = 256: 65536 bins as packed uint16 pairs in 32768 words (128 KB),
//               so a workgroup may add at most 65535 voxels (items_per_block
//               <= 1023 segments of 64);
//   BINS = 64 : 4096 uint32 words, each row of 64 padded to 65 words
//               (16.25 KB), so that bins of one column in different rows sit
//               in different LDS banks: 0.80-0.81 -> 0.73 ms for 64
//               candidates at 182x218x182.  The same padding of the 256-bin
//               rows (129 words, 129 KB) was 1.4% slower.
// Every sample is one plain LDS atomic.  Aggregating a wave's lanes that
// share the first active lane's bin (one ballot, one atomic) paid off while
// the 64-bin rows were unpadded (0.97 against 1.18 ms); with padded rows
// plain atomics are faster (0.68-0.69 against 0.72-0.73 ms), and with 256
// bins they always were (1.08 against 1.22 ms).  At the end the non-zero
// words are flushed into the candidate's global int32 histogram.
// ---------------------------------------------------------------------------
template <int BINS>
struct MiTraits;
template <>
struct MiTraits<256> {
    static constexpr int kWords = 32768;
    static constexpr int kStride = 128;        // words per row in LDS (no padding)
    static constexpr int kLdsWords = 32768;
    static constexpr int kBand = 8;
    static constexpr int kThreads = 1024;
    __device__ static uint32_t word(uint32_t bin) { return bin >> 1; }
    __device__ static uint32_t inc(uint32_t bin) { return (bin & 1u) ? 0x10000u : 1u; }
};
template <>
struct MiTraits<64> {
    static constexpr int kWords = 4096;
    static constexpr int kStride = 65;         // words per row in LDS: one pad word
    static constexpr int kLdsWords = 64 * 65;
    static constexpr int kThreads = 256;
};

template <int BINS, bool ZT>
__global__ __launch_bounds__(MiTraits<BINS>::kThreads) void mi_hist_kernel(
    const Segment* __restrict__ segs, uint32_t nseg, uint32_t total_blocks,
    const float* __restrict__ Ts, const float4* __restrict__ ztab, uint32_t zstride,
    int* __restrict__ hist, uint32_t* __restrict__ to_hist)
{
    using Tr = MiTraits<BINS>;
    // cleared 16 bytes a store: the 256-bin kernel 0.8-1.4% faster at
    // 182x218x182 with 64 candidates (0.970 -> 0.963 ms), the 64-bin one unchanged
    static_assert(Tr::kLdsWords % 4 == 0, "the clear stores 16 bytes at a time");
    __shared__ __align__(16) uint32_t lds[Tr::kLdsWords];
    for (uint32_t i = threadIdx.x; i < uint32_t(Tr::kLdsWords) / 4u; i += blockDim.x)
        reinterpret_cast<uint4*>(lds)[i] = make_uint4(0u, 0u, 0u, 0u);
    __syncthreads();

    const uint32_t L = xcd_remap(blockIdx.x, total_blocks);
    const Segment& sg = segs[find_segment(segs, nseg, L, ByBlock{})];
    const BatchGeo g = sg.g;
    const uint32_t local = L - sg.block_begin;
    const uint32_t cand = sg.cand_begin + local % g.K;
    const uint32_t chunk = local / g.K;
    float T[12];
#pragma unroll
    for (int i = 0; i < 12; ++i) T[i] = Ts[12 * cand + i];
    const uint32_t zrow = cand * zstride;
    const __amdgpu_buffer_rsrc_t rf = make_rsrc(sg.from, g.from_bytes);
    const __amdgpu_buffer_rsrc_t r8 = make_rsrc(sg.to8, g.to_bytes8);
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const uint32_t nwaves = blockDim.x >> 6;
    const uint32_t it0 = chunk * g.items_per_block;
    const uint32_t it1 = min(it0 + g.items_per_block, g.items);
    const uint32_t fwh = uint32_t(g.fw) * uint32_t(g.fh);
    // the exact FP32 cell index needs w*h*d <= 2^24 (sample2_u8_packed_in)
    const bool small_to = g.to.wh * uint64_t(uint32_t(g.to.d)) <= (uint64_t(1) << 24);
    const float tw = float(g.to.w), twh = float(g.to.wh);
    for (uint32_t it = it0 + wave; it < it1; it += nwaves) {
        const Item m = decode(g, it, lane);
        if (!m.ok) continue;
        const ColumnSums cs = column_sums(T, float(m.x), float(m.y));
        uint32_t fidx = m.z0 * fwh + m.y * uint32_t(g.fw) + m.x;
        auto count = [&](uint32_t fv, uint32_t tv) {
            if constexpr (BINS == 64) {
                // the byte address of word (fv >> 2) * 65 + (tv >> 2)
                const uint32_t a = __umul24(fv & 0xFCu, 65u) + (tv & 0xFCu);
                atomicAdd(reinterpret_cast<uint32_t*>(reinterpret_cast<char*>(lds) + a), 1u);
            } else {
                const uint32_t bin = (fv << Tr::kBand) | tv;
                atomicAdd(&lds[Tr::word(bin)], Tr::inc(bin));
            }
        };
        // Two planes per step: both samples' gathers are in flight together.
        // (The correlation kernel's all-inside march does not pay here: 1-2%
        // slower with the wave-aggregated atomics, no faster with plain ones.)
        // The 256-bin kernel forms the cell index in FP32 when the `to`
        // volume allows it (sample2_u8_packed_fi).
        uint32_t z = m.z0;
        auto march = [&](auto fi_tag) {
            constexpr bool kFi = decltype(fi_tag)::value;
            for (; z + 1 < m.z1; z += 2, fidx += 2 * fwh) {
                const uint32_t fva = load_u8(rf, fidx, 0), fvb = load_u8(rf, fidx + fwh, 0);
                const float4 za = z_products<ZT>(ztab, zrow, z, T), zb = z_products<ZT>(ztab, zrow, z + 1, T);
                const float pxa = (cs.hx + za.x) + T[9];
                const float pya = (cs.hy + za.y) + T[10];
                const float pza = (cs.hz + za.z) + T[11];
                const float pxb = (cs.hx + zb.x) + T[9];
                const float pyb = (cs.hy + zb.y) + T[10];
                const float pzb = (cs.hz + zb.z) + T[11];
                uint32_t tva, tvb;
                if constexpr (kFi)
                    sample2_u8_packed_fi(r8, g.to, tw, twh, pxa, pya, pza, pxb, pyb, pzb, tva, tvb);
                else
                    sample2_u8_packed(r8, g.to, pxa, pya, pza, pxb, pyb, pzb, tva, tvb);
                count(fva, tva);
                count(fvb, tvb);
            }
        };
        if constexpr (BINS == 256) {
            if (small_to)
                march(std::true_type{});
            else
                march(std::false_type{});
        } else {
            march(std::false_type{});
        }
        if (z < m.z1) {
            const uint32_t fv = load_u8(rf, fidx, 0);
            const float4 zp = z_products<ZT>(ztab, zrow, z, T);
            count(fv, sample_u8_packed(r8, g.to, (cs.hx + zp.x) + T[9], (cs.hy + zp.y) + T[10],
                                       (cs.hz + zp.z) + T[11]));
        }
    }
    __syncthreads();
    // Flush the non-zero words into the candidate's histogram.  The thread
    // count is a multiple of the words per histogram row, so all of a
    // thread's words lie in one column (pair): it also sums that column over
    // its rows, and the partial column sums go to the candidate's to_hist
    // (the histogram of the sampled values, get_mutual_info's to_hist).
    constexpr uint32_t kWpr = uint32_t(Tr::kWords) / BINS;   // words per row
    constexpr uint32_t kBpw = BINS / kWpr;                   // bins per word
    constexpr uint32_t kQ = uint32_t(Tr::kThreads) / kWpr;   // threads per column word
    static_assert(uint32_t(Tr::kThreads) % kWpr == 0 && kQ * BINS <= uint32_t(Tr::kWords),
                  "flush layout");
    int* h = hist + size_t(cand) * BINS * BINS;
    uint32_t col_lo = 0, col_hi = 0;
    for (uint32_t i = threadIdx.x; i < uint32_t(Tr::kWords); i += blockDim.x) {
        const uint32_t v = lds[(i / kWpr) * uint32_t(Tr::kStride) + (i % kWpr)];
        if constexpr (BINS == 256) {
            col_lo += v & 0xffffu;
            col_hi += v >> 16;
        } else {
            col_lo += v;
        }
        if (!v) continue;
        if constexpr (BINS == 256) {
            // both bins of the word in one 64-bit atomic on the int32 pair:
            // a bin's total is below 2^31, so the low half never carries
            // (9-13% off the kernel at 48^3-96^3, 5% at 182x218x182)
            atomicAdd(reinterpret_cast<unsigned long long*>(h + 2 * i),
                      (static_cast<unsigned long long>(v >> 16) << 32) | (v & 0xffffu));
        } else {
            atomicAdd(&h[i], int(v));
        }
    }
    __syncthreads();  // every word has been read: lds becomes kQ x BINS column partials
    {
        const uint32_t j = threadIdx.x % kWpr, q = threadIdx.x / kWpr;
        lds[q * BINS + kBpw * j] = col_lo;
        if constexpr (kBpw == 2) lds[q * BINS + 2 * j + 1] = col_hi;
    }
    __syncthreads();
    if (threadIdx.x < uint32_t(BINS)) {
        uint32_t c = 0;
        for (uint32_t q = 0; q < kQ; ++q) c += lds[q * BINS + threadIdx.x];
        if (c) atomicAdd(&to_hist[size_t(cand) * BINS + threadIdx.x], c);
    }
}

// get_mutual_info_mean (reg/linear.hpp:87-104): the sum of
// mu*log(mu/from_h[row]/to_h[col]) in FP64 over non-zero bins, divided by
// float(BINS*BINS) and negated (reg/linear.hpp:201).  to_h comes from the
// histogram kernel.  Each candidate's bins are split into slices of 4096
// (16 per thread), one workgroup per slice, so even a single candidate
// fills many CUs; the slice sums of 256-bin histograms are combined by
// mi_entropy_final_kernel in a fixed tree (a 64-bin histogram is one slice).  Each wave first packs its non-zero bins (a ballot and a
// lane rank per bin slot) so the FP64 log terms run on dense lanes only.
// The result is deterministic (the reference's sum is sequential in
// row-major order; the two agree to rounding).
constexpr int kEntropyThreads = 256;
constexpr int kEntropySlice = 4096;

template <int BINS>
constexpr int entropy_slices() { return BINS * BINS / kEntropySlice; }

template <int BINS>
__global__ __launch_bounds__(kEntropyThreads) void mi_entropy_kernel(
    const Segment* __restrict__ segs, uint32_t nseg, int* hist, uint32_t* to_hist,
    double* __restrict__ part, double* __restrict__ costs, double* __restrict__ host_costs)
{
    constexpr int G = entropy_slices<BINS>();
    constexpr int kVec = kEntropySlice / 4 / kEntropyThreads;
    constexpr int kWaves = kEntropyThreads / 64;
    const uint32_t cand = blockIdx.x / G, slice = blockIdx.x % G;
    const uint32_t* __restrict__ from_hist =
        segs[find_segment(segs, nseg, cand, ByCandidate{})].from_hist;
    __shared__ uint32_t to_h[BINS];
    __shared__ double wpart[kWaves];
    // each wave's non-zero bins, packed in (slot, lane) order: at most 64 x 4 x kVec a wave
    constexpr int kPerWave = 64 * 4 * kVec;
    __shared__ int s_mu[kEntropySlice];
    __shared__ uint16_t s_bin[kEntropySlice];
    for (int i = threadIdx.x; i < BINS; i += kEntropyThreads) to_h[i] = to_hist[size_t(cand) * BINS + i];
    // four bins per 16-byte load, all of a thread's loads issued before use
    int4* h4 = reinterpret_cast<int4*>(hist + size_t(cand) * BINS * BINS + size_t(slice) * kEntropySlice);
    int4 q[kVec];
#pragma unroll
    for (int u = 0; u < kVec; ++u) q[u] = h4[threadIdx.x + u * kEntropyThreads];
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    int* w_mu = s_mu + wave * kPerWave;
    uint16_t* w_bin = s_bin + wave * kPerWave;
    uint32_t n = 0;  // wave-uniform count so far
#pragma unroll
    for (int u = 0; u < kVec; ++u) {
        const int mus[4] = {q[u].x, q[u].y, q[u].z, q[u].w};
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const bool nz = mus[c] != 0;
            const unsigned long long mk = __ballot(nz);
            if (nz) {
                const uint32_t r = __builtin_amdgcn_mbcnt_hi(uint32_t(mk >> 32),
                                                             __builtin_amdgcn_mbcnt_lo(uint32_t(mk), 0u));
                w_mu[n + r] = mus[c];
                w_bin[n + r] = uint16_t(4 * (threadIdx.x + u * kEntropyThreads) + c);
            }
            n += uint32_t(__popcll(mk));
        }
    }
    __syncthreads();
    // the slice is read (its values are in registers and LDS): leave it zero
    // for the next launch, which then needs no clearing pass
#pragma unroll
    for (int u = 0; u < kVec; ++u) h4[threadIdx.x + u * kEntropyThreads] = make_int4(0, 0, 0, 0);
    // the log terms of the wave's non-zero bins only, lane l taking entries l, l + 64, ...
    double acc = 0.0;
    for (uint32_t e = lane; e < n; e += 64) {
        const int i = int(slice) * kEntropySlice + int(w_bin[e]);
        const double m = double(w_mu[e]);
        acc += m * log(m / double(from_hist[i / BINS]) / double(to_h[i % BINS]));
    }
    acc = wave_sum_f64(acc);
    if ((threadIdx.x & 63) == 0) wpart[threadIdx.x >> 6] = acc;
    __syncthreads();
    // One slice per candidate (64 bins): the cost itself, and the candidate's
    // to_hist row cleared (this workgroup was its only reader).  Several
    // slices (256 bins): the slice sum, combined by mi_entropy_final_kernel.
    if (threadIdx.x == 0) {
        double t[kWaves];
        for (int i = 0; i < kWaves; ++i) t[i] = wpart[i];
        for (int w = kWaves / 2; w > 0; w /= 2)
            for (int i = 0; i < w; ++i) t[i] = t[2 * i] + t[2 * i + 1];
        if constexpr (G == 1) {
            const double c = -(t[0] / double(float(BINS * BINS)));
            costs[cand] = c;
            if (host_costs) host_costs[cand] = c;
        } else {
            part[blockIdx.x] = t[0];
        }
    }
    if constexpr (G == 1) {
        uint4* row4 = reinterpret_cast<uint4*>(to_hist + size_t(cand) * BINS);
        for (int i = int(threadIdx.x); i < BINS / 4; i += kEntropyThreads) row4[i] = make_uint4(0u, 0u, 0u, 0u);
    }
}

// Also clears the block's candidates' to_hist rows (every slice has read
// them), so the whole histogram buffer is zero again after this launch.
template <int BINS>
__global__ __launch_bounds__(256) void mi_entropy_final_kernel(const double* __restrict__ part,
                                                               uint32_t K, double* __restrict__ costs,
                                                               double* __restrict__ host_costs,
                                                               uint32_t* to_hist)
{
    constexpr int G = entropy_slices<BINS>();
    const uint32_t first = blockIdx.x * 256u, cand = first + threadIdx.x;
    if (cand < K) {
        double t[G];
#pragma unroll
        for (int i = 0; i < G; ++i) t[i] = part[size_t(cand) * G + i];
#pragma unroll
        for (int w = G / 2; w > 0; w /= 2)
#pragma unroll
            for (int i = 0; i < w; ++i) t[i] = t[2 * i] + t[2 * i + 1];
        const double c = -(t[0] / double(float(BINS * BINS)));
        costs[cand] = c;
        if (host_costs) host_costs[cand] = c;
    }
    const uint32_t words4 = (min(K - first, 256u) * uint32_t(BINS)) / 4u;
    uint4* row4 = reinterpret_cast<uint4*>(to_hist + size_t(first) * BINS);
    for (uint32_t i = threadIdx.x; i < words4; i += 256u) row4[i] = make_uint4(0u, 0u, 0u, 0u);
}

struct TransformArg {
    float t[12];
};

// tipl::resample for one transform into a device uint8 grid of shape
// (ow, oh, od): out(p) = src(T p), or the exact integer-shift crop when
// crop.x is set.
__global__ __launch_bounds__(256) void resample_kernel(const uint8_t* __restrict__ src,
                                                       uint32_t src_bytes, VolGeo sg, int ow,
                                                       int oh, int od, TransformArg T, int4 crop,
                                                       uint8_t* __restrict__ out)
{
    const __amdgpu_buffer_rsrc_t rs = make_rsrc(src, src_bytes);
    const size_t n = size_t(ow) * oh * od;
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += size_t(gridDim.x) * blockDim.x) {
        const uint32_t x = uint32_t(i % uint32_t(ow));
        const uint32_t r = uint32_t(i / uint32_t(ow));
        const uint32_t y = r % uint32_t(oh);
        const uint32_t z = r / uint32_t(oh);
        uint32_t v;
        if (crop.x) {
            int sx = int(x) + crop.y, s_y = int(y) + crop.z, sz = int(z) + crop.w;
            bool in = sx >= 0 && sx < sg.w && s_y >= 0 && s_y < sg.h && sz >= 0 && sz < sg.d;
            v = in ? load_u8(rs, uint32_t(sz) * sg.wh + uint32_t(s_y) * uint32_t(sg.w) + uint32_t(sx), 0)
                   : 0u;
        } else {
            Pos p = apply(T.t, float(x), float(y), float(z));
            v = sample_u8(rs, sg, p.x, p.y, p.z);
        }
        out[i] = uint8_t(v);
    }
}

__global__ void fill_u64_kernel(unsigned long long* p, size_t n)
{
    size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    size_t stride = size_t(gridDim.x) * blockDim.x;
    for (; i < n; i += stride) p[i] = 0ull;
}

__global__ void fill_u32x4_kernel(uint4* p, size_t n)
{
    size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    size_t stride = size_t(gridDim.x) * blockDim.x;
    for (; i < n; i += stride) p[i] = make_uint4(0u, 0u, 0u, 0u);
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------

constexpr int kMaxK = 4096;  // candidates per launch, over all pairs

void check_volume(const tiplhip_ctx* ctx, const tiplhip_volume* v, const char* what)
{
    TIPLHIP_REQUIRE(v != nullptr, std::string(what) + " is null");
    TIPLHIP_REQUIRE(v->ctx != nullptr && v->ctx->device == ctx->device,
                    std::string(what) + " lives on another device");
    TIPLHIP_REQUIRE(v->n > 0, std::string(what) + " is empty");
    TIPLHIP_REQUIRE(v->n < (size_t(1) << 29), std::string(what) + " exceeds 2^29 voxels");
    TIPLHIP_REQUIRE(size_t(v->dim[0]) * size_t(v->dim[1]) < (size_t(1) << 24),
                    std::string(what) + " plane exceeds 2^24 voxels");
}

// Kernel timing brackets each timed launch with a pair of events and does
// not wait for them, so timed runs keep their throughput; timing_flush
// resolves the pairs (at most kTimingPairs are held before a flush).
constexpr size_t kTimingPairs = 512;

void timing_begin(tiplhip_ctx* ctx)
{
    if (!ctx->kernel_timing) return;
    if (ctx->kt_pending == kTimingPairs) timing_flush(ctx);
    if (ctx->kt_events.size() < 2 * (ctx->kt_pending + 1)) {
        for (int i = 0; i < 2; ++i) {
            hipEvent_t e = nullptr;
            TIPLHIP_CHECK(hipEventCreate(&e));
            ctx->kt_events.push_back(e);
        }
    }
    TIPLHIP_CHECK(hipEventRecord(ctx->kt_events[2 * ctx->kt_pending], ctx->stream));
}

void timing_end(tiplhip_ctx* ctx)
{
    if (!ctx->kernel_timing) return;
    TIPLHIP_CHECK(hipEventRecord(ctx->kt_events[2 * ctx->kt_pending + 1], ctx->stream));
    ++ctx->kt_pending;
}

void zero_u64(tiplhip_ctx* ctx, void* p, size_t n)
{
    if (!n) return;
    int blocks = int(std::min<size_t>((n + 255) / 256, 1024));
    hipLaunchKernelGGL(fill_u64_kernel, dim3(blocks), dim3(256), 0, ctx->stream,
                       static_cast<unsigned long long*>(p), n);
    TIPLHIP_CHECK(hipGetLastError());
}

void zero_bytes16(tiplhip_ctx* ctx, void* p, size_t bytes)
{
    size_t n = bytes / 16;
    if (!n) return;
    int blocks = int(std::min<size_t>((n + 255) / 256, 4096));
    hipLaunchKernelGGL(fill_u32x4_kernel, dim3(blocks), dim3(256), 0, ctx->stream,
                       static_cast<uint4*>(p), n);
    TIPLHIP_CHECK(hipGetLastError());
}

// K candidate transforms of one image pair.
struct PairJob {
    const tiplhip_volume* from;
    const tiplhip_volume* to;
    int K;
};

// A launch's device inputs: transforms, crop flags and the segment table.
struct Staged {
    uint32_t nseg = 0, total_blocks = 0, K = 0;
    const Segment* segs = nullptr;
    const float* T = nullptr;
    const int4* crop = nullptr;
    double samples = 0.0;   // voxel-candidate samples of the launch
    uint32_t zstride = 0;   // the largest `from` depth: rows of the z-product table
    size_t bytes = 0;       // staged bytes (transforms, crop flags, segment table)
    bool grew = false;      // the device staging buffer was reallocated
};

// Launches of at least this many samples get a z-product table: its build
// (one small kernel) costs less than the four VALU instructions per sample
// it saves.
constexpr double kZTableSamples = double(1 << 25);

bool hist256(int cost, int bins) { return cost == TIPLHIP_COST_MI && bins == 256; }

// Cached cost-launch graphs per context (cleared when full).
constexpr size_t kMaxGraphs = 64;

// A quiet-NaN bit pattern no cost kernel writes (MI and correlation costs
// are finite), and how long the host spins for the costs before it
// synchronises the stream instead.
constexpr uint64_t kCostPending = 0x7FF4C0575EED0001ull;
constexpr int kCostSpinUs = 2000;

// Work items (16 x 4 voxel columns of up to zc planes) per workgroup above
// which a launch gets more workgroups: about 256 items, 256K samples at zc = 16.
constexpr double kMaxItemsPerBlock = 256.0;

// The launch shape of one pair: planes per work item and workgroups per CU.
// Measured on the device over 48^3 .. 182x218x182 pairs with 12 and 64
// candidates: the 256-bin histogram (128 KB of LDS, one 1024-thread
// workgroup per CU) is fastest with about one workgroup per CU -- fewer,
// fuller workgroups amortise the LDS clear and flush -- while the 64-bin and
// correlation kernels gain from finer slices up to 32-64 per CU (64-bin MI at
// 182x218x182, 64 candidates: 0.76-0.79 ms with 8, 0.65 with 16, 0.60 with
// 32; the same at 48^3-128^3 within noise).  (Splitting the 256-bin histogram into 2 or 4 ranges
// of `from` rows, one workgroup each with a half or a quarter of the LDS,
// was slower at every size from 64^3 with 24 or more candidates: 1.4x and
// 2.6x at 182x218x182, 64 candidates.)
LaunchShape default_shape(int cost, int bins)
{
    return LaunchShape{16u, hist256(cost, bins) ? 1u : 32u};
}

TuneKey tune_key(const tiplhip_volume* from, const tiplhip_volume* to, int K, int cost, int bins)
{
    return TuneKey{from->dim[0], from->dim[1], from->dim[2], to->dim[0], to->dim[1], to->dim[2], K,
                   cost, cost == TIPLHIP_COST_MI ? bins : 0};
}

// Builds the segment table of `jobs` (T holds their transforms, job after
// job; crop, when not null, one entry per candidate) and uploads all of it
// in one copy.  Each pair gets a share of about cu_count x per_cu workgroups
// in proportion to its voxel-candidate work.
Staged stage_jobs(tiplhip_ctx* ctx, const std::vector<PairJob>& jobs, const float* T,
                  const int4* crop, int cost, int bins, const LaunchShape* force = nullptr,
                  bool upload = true)
{
    Staged st;
    std::vector<Segment> segs(jobs.size());
    const uint32_t max_vox = hist256(cost, bins) ? 65535u : (1u << 23);
    // a single pair uses the autotuned shape for its key when there is one
    LaunchShape shape = default_shape(cost, bins);
    if (force) {
        shape = *force;
    } else if (jobs.size() == 1) {
        auto it = ctx->tuned.find(tune_key(jobs[0].from, jobs[0].to, jobs[0].K, cost, bins));
        if (it != ctx->tuned.end()) shape = it->second;
    }
    double work = 0.0;
    for (size_t j = 0; j < jobs.size(); ++j) {
        segs[j].g = batch_geometry(jobs[j].from, jobs[j].to, jobs[j].K, shape.zc);
        work += double(segs[j].g.items) * double(jobs[j].K);
    }
    // at least per_cu workgroups per CU, and no workgroup with more than
    // kMaxItemsPerBlock items (large batches get more workgroups, not longer ones:
    // concurrent workgroups then share z slabs of one pair in L2)
    // (per_cu 0 asks for the fewest workgroups: each as full as the counters allow)
    const double target = shape.per_cu == 0
                              ? 1.0
                              : std::max(double(ctx->cu_count) * double(shape.per_cu),
                                         std::ceil(work / double(kMaxItemsPerBlock)));
    uint64_t blocks = 0;
    for (size_t j = 0; j < jobs.size(); ++j) {
        Segment& sg = segs[j];
        const double share = double(sg.g.items) * double(jobs[j].K) / work;
        batch_blocks(sg.g, max_vox, uint32_t(std::min(std::ceil(share * target), 2147483647.0)));
        sg.from = jobs[j].from->d;
        sg.to = jobs[j].to->d;
        sg.to8 = jobs[j].to->d_pack;
        sg.from_hist = jobs[j].from->d_hist + (bins == 64 ? 256 : 0);
        sg.sx = (long long)jobs[j].from->sum;
        sg.sxx = (long long)jobs[j].from->sum_sq;
        sg.n = double(jobs[j].from->n);
        sg.block_begin = uint32_t(blocks);
        sg.cand_begin = st.K;
        blocks += sg.g.total_blocks;
        st.K += uint32_t(jobs[j].K);
        st.samples += double(jobs[j].from->n) * double(jobs[j].K);
        st.zstride = std::max(st.zstride, uint32_t(jobs[j].from->dim[2]));
    }
    TIPLHIP_REQUIRE(blocks < (uint64_t(1) << 31), "launch exceeds 2^31 workgroups");
    st.nseg = uint32_t(segs.size());
    st.total_blocks = uint32_t(blocks);
    const size_t tb = sizeof(float) * 12 * size_t(st.K);   // a multiple of 16
    const size_t cb = crop ? sizeof(int4) * size_t(st.K) : 0;
    const size_t sb = sizeof(Segment) * segs.size();
    ctx->h_T.reserve(tb + cb + sb);
    const size_t had = ctx->d_T.bytes;
    ctx->d_T.reserve(tb + cb + sb);
    st.grew = ctx->d_T.bytes != had;
    st.bytes = tb + cb + sb;
    char* h = static_cast<char*>(ctx->h_T.p);
    std::memcpy(h, T, tb);
    // a zero shift component is stored as +0.0 (the packed sampler's bounds
    // test needs coordinates that are never -0.0; the sample is unchanged)
    float* hT = reinterpret_cast<float*>(h);
    for (size_t k = 0; k < size_t(st.K); ++k)
        for (int i = 9; i < 12; ++i)
            if (hT[12 * k + size_t(i)] == 0.0f) hT[12 * k + size_t(i)] = 0.0f;
    if (crop) std::memcpy(h + tb, crop, cb);
    std::memcpy(h + tb + cb, segs.data(), sb);
    if (upload) TIPLHIP_CHECK(hipMemcpyAsync(ctx->d_T.p, h, st.bytes, hipMemcpyHostToDevice, ctx->stream));
    char* d = static_cast<char*>(ctx->d_T.p);
    st.T = reinterpret_cast<const float*>(d);
    st.crop = crop ? reinterpret_cast<const int4*>(d + tb) : nullptr;
    st.segs = reinterpret_cast<const Segment*>(d + tb + cb);
    return st;
}

// The staged bytes' copy to the device (when stage_jobs was told not to).
void upload_staged(tiplhip_ctx* ctx, const Staged& st)
{
    TIPLHIP_CHECK(hipMemcpyAsync(ctx->d_T.p, ctx->h_T.p, st.bytes, hipMemcpyHostToDevice, ctx->stream));
}

// The z-product table of a large launch (nullptr for a small one).
const float4* stage_ztab(tiplhip_ctx* ctx, const Staged& st)
{
    if (st.samples < kZTableSamples) return nullptr;
    const size_t n = size_t(st.K) * st.zstride;
    TIPLHIP_REQUIRE(n < (size_t(1) << 32), "z-product table too large");
    ctx->d_ztab.reserve(sizeof(float4) * n);
    hipLaunchKernelGGL(ztab_kernel, dim3(uint32_t((n + 255) / 256)), dim3(256), 0, ctx->stream,
                       st.T, st.K, st.zstride, ctx->d_ztab.as<float4>());
    TIPLHIP_CHECK(hipGetLastError());
    return ctx->d_ztab.as<float4>();
}

// Correlation sums Sy, Syy, Sxy of every staged candidate into ctx->d_sums (K x 4 u64).
void launch_corr_sums(tiplhip_ctx* ctx, const Staged& st)
{
    ctx->d_sums.reserve(sizeof(unsigned long long) * 4 * size_t(st.K));
    zero_u64(ctx, ctx->d_sums.p, 4 * size_t(st.K));
    const float4* zt = stage_ztab(ctx, st);
    // per-lane uint32 sums of squares stay below 2^32 for < 65536 voxels a lane
    timing_begin(ctx);
    if (zt)
        hipLaunchKernelGGL(corr_sums_kernel<true>, dim3(st.total_blocks), dim3(256), 0, ctx->stream,
                           st.segs, st.nseg, st.total_blocks, st.T, st.crop, zt, st.zstride,
                           ctx->d_sums.as<unsigned long long>());
    else
        hipLaunchKernelGGL(corr_sums_kernel<false>, dim3(st.total_blocks), dim3(256), 0, ctx->stream,
                           st.segs, st.nseg, st.total_blocks, st.T, st.crop, nullptr, 0u,
                           ctx->d_sums.as<unsigned long long>());
    TIPLHIP_CHECK(hipGetLastError());
    timing_end(ctx);
}

void launch_corr_finalize(tiplhip_ctx* ctx, const Staged& st, long long* sums_out,
                          double* host_costs = nullptr)
{
    ctx->d_costs.reserve(sizeof(double) * size_t(st.K));
    hipLaunchKernelGGL(corr_finalize_kernel, dim3((st.K + 63) / 64), dim3(64), 0, ctx->stream,
                       st.segs, st.nseg, ctx->d_sums.as<unsigned long long>(), int(st.K),
                       ctx->d_costs.as<double>(), host_costs, sums_out);
    TIPLHIP_CHECK(hipGetLastError());
}

// Joint histograms of every staged candidate into ctx->d_hist.
// d_hist holds K histograms of bins^2 int32, then K to_hist rows of bins.
uint32_t* to_hist_of(tiplhip_ctx* ctx, const Staged& st, int bins)
{
    return reinterpret_cast<uint32_t*>(ctx->d_hist.as<int>() + size_t(bins) * bins * size_t(st.K));
}

size_t hist_bytes(const Staged& st, int bins)
{
    return sizeof(int) * (size_t(bins) * bins + size_t(bins)) * size_t(st.K);
}

// The histogram kernel accumulates into a zeroed buffer.  The entropy
// kernels leave what they read zero again, so clearing passes run only over
// bytes never cleared since the buffer was allocated and after a launch
// whose histograms were not consumed by them (the raw histogram entry
// point, the autotuner).
// Returns true when the buffer was reallocated.
bool reserve_hist(tiplhip_ctx* ctx, size_t hb)
{
    const size_t had = ctx->d_hist.bytes;
    ctx->d_hist.reserve(hb);
    if (ctx->d_hist.bytes == had) return false;
    ctx->hist_dirty = ctx->hist_zero = 0;
    return true;
}

// Queues the clearing passes that make the first hb bytes zero.
void make_hist_clean(tiplhip_ctx* ctx, size_t hb)
{
    char* base = static_cast<char*>(ctx->d_hist.p);
    if (ctx->hist_dirty > 0) zero_bytes16(ctx, base, ctx->hist_dirty);
    ctx->hist_dirty = 0;
    if (ctx->hist_zero < hb) zero_bytes16(ctx, base + ctx->hist_zero, hb - ctx->hist_zero);
    ctx->hist_zero = std::max(ctx->hist_zero, hb);
}

void launch_mi_hist(tiplhip_ctx* ctx, const Staged& st, int bins)
{
    const size_t hb = hist_bytes(st, bins);  // a multiple of 16
    reserve_hist(ctx, hb);
    make_hist_clean(ctx, hb);
    ctx->hist_dirty = hb;
    const float4* zt = stage_ztab(ctx, st);
    int* h = ctx->d_hist.as<int>();
    uint32_t* th = to_hist_of(ctx, st, bins);
    timing_begin(ctx);
    if (bins == 256 && zt)
        hipLaunchKernelGGL((mi_hist_kernel<256, true>), dim3(st.total_blocks), dim3(1024), 0,
                           ctx->stream, st.segs, st.nseg, st.total_blocks, st.T, zt, st.zstride, h, th);
    else if (bins == 256)
        hipLaunchKernelGGL((mi_hist_kernel<256, false>), dim3(st.total_blocks), dim3(1024), 0,
                           ctx->stream, st.segs, st.nseg, st.total_blocks, st.T, nullptr, 0u, h, th);
    else if (zt)
        hipLaunchKernelGGL((mi_hist_kernel<64, true>), dim3(st.total_blocks), dim3(256), 0,
                           ctx->stream, st.segs, st.nseg, st.total_blocks, st.T, zt, st.zstride, h, th);
    else
        hipLaunchKernelGGL((mi_hist_kernel<64, false>), dim3(st.total_blocks), dim3(256), 0,
                           ctx->stream, st.segs, st.nseg, st.total_blocks, st.T, nullptr, 0u, h, th);
    TIPLHIP_CHECK(hipGetLastError());
    timing_end(ctx);
}

// d_costs holds the K costs, then K x slices partial sums.
template <int BINS>
void launch_mi_entropy_t(tiplhip_ctx* ctx, const Staged& st, double* host_costs)
{
    constexpr int G = entropy_slices<BINS>();
    ctx->d_costs.reserve(sizeof(double) * size_t(st.K) * (1 + G));
    double* costs = ctx->d_costs.as<double>();
    double* part = costs + st.K;
    hipLaunchKernelGGL(mi_entropy_kernel<BINS>, dim3(st.K * uint32_t(G)), dim3(kEntropyThreads), 0,
                       ctx->stream, st.segs, st.nseg, ctx->d_hist.as<int>(), to_hist_of(ctx, st, BINS),
                       part, costs, host_costs);
    TIPLHIP_CHECK(hipGetLastError());
    if constexpr (G > 1) {
        hipLaunchKernelGGL(mi_entropy_final_kernel<BINS>, dim3((st.K + 255) / 256), dim3(256), 0,
                           ctx->stream, part, st.K, costs, host_costs, to_hist_of(ctx, st, BINS));
        TIPLHIP_CHECK(hipGetLastError());
    }
    ctx->hist_dirty = 0;
}

// Costs into ctx->d_costs and, when host_costs is set (device-visible host
// memory), straight into host memory as well.
void launch_mi_entropy(tiplhip_ctx* ctx, const Staged& st, int bins, double* host_costs)
{
    if (bins == 256)
        launch_mi_entropy_t<256>(ctx, st, host_costs);
    else
        launch_mi_entropy_t<64>(ctx, st, host_costs);
}

// tipl::resample of src into a device grid of shape out_dim.
void launch_resample(tiplhip_ctx* ctx, const tiplhip_volume* src, const int* out_dim, const float* T,
                     bool crop_shortcut, uint8_t* out)
{
    TransformArg ta;
    std::memcpy(ta.t, T, sizeof(ta.t));
    int pos[3] = {0, 0, 0};
    bool c = crop_shortcut && is_integer_shift(T, pos);
    int4 crop = make_int4(c ? 1 : 0, pos[0], pos[1], pos[2]);
    const size_t n = vox(out_dim);
    int blocks = int(std::min<size_t>((n + 255) / 256, 65536));
    hipLaunchKernelGGL(resample_kernel, dim3(blocks), dim3(256), 0, ctx->stream, src->d,
                       uint32_t(src->n), make_geo(src->dim), out_dim[0], out_dim[1], out_dim[2], ta,
                       crop, out);
    TIPLHIP_CHECK(hipGetLastError());
}

// Costs of every candidate of every job in one launch (plus one finishing
// launch); costs[] receives them job after job.  T holds the jobs'
// transforms in the same order.
void cost_jobs(tiplhip_ctx* ctx, std::vector<PairJob> jobs, const float* T, int cost, int bins,
               double* costs, const LaunchShape* force = nullptr)
{
    TIPLHIP_REQUIRE(!jobs.empty(), "no image pairs");
    TIPLHIP_REQUIRE(T != nullptr && costs != nullptr, "null transform or cost buffer");
    TIPLHIP_REQUIRE(cost == TIPLHIP_COST_CORR || cost == TIPLHIP_COST_MI, "unknown cost type");
    if (cost == TIPLHIP_COST_MI) TIPLHIP_REQUIRE(bins == 256 || bins == 64, "bins must be 256 or 64");
    size_t K = 0;
    for (const PairJob& j : jobs) {
        check_volume(ctx, j.from, "from");
        check_volume(ctx, j.to, "to");
        TIPLHIP_REQUIRE(j.K >= 1 && j.K <= kMaxK, "K must be in [1, 4096]");
        K += size_t(j.K);
    }
    TIPLHIP_REQUIRE(K <= size_t(kMaxK), "more than 4096 candidates in one launch");
    DeviceGuard dg(ctx->device);
    std::vector<float> Tk(T, T + 12 * K);
    std::vector<int4> crop;
    if (cost == TIPLHIP_COST_CORR) {
        // reg::correlation swaps the roles when `from` is larger
        // (reg/linear.hpp:28-33); a singular inverse leaves T unchanged, as
        // transformation_matrix::inverse does.
        crop.resize(K);
        size_t base = 0;
        for (PairJob& j : jobs) {
            float* Tj = &Tk[12 * base];
            if (j.from->n > j.to->n) {
                std::swap(j.from, j.to);
                for (int k = 0; k < j.K; ++k) (void)matrix_inverse(&Tj[12 * size_t(k)]);
            }
            for (int k = 0; k < j.K; ++k) {
                int pos[3];
                bool c = is_integer_shift(&Tj[12 * size_t(k)], pos);
                crop[base + size_t(k)] = make_int4(c ? 1 : 0, pos[0], pos[1], pos[2]);
            }
            base += size_t(j.K);
        }
    }
    for (const PairJob& j : jobs) volume_ensure_packed(ctx, const_cast<tiplhip_volume*>(j.to));
    const Staged st = stage_jobs(ctx, jobs, Tk.data(), crop.empty() ? nullptr : crop.data(), cost,
                                 bins, force, /*upload=*/false);
    // every buffer the launch touches, reserved up front (a graph capture may
    // not allocate); a reallocation invalidates the cached graphs
    bool grew = st.grew;
    auto grow = [&](DeviceBuffer& b, size_t n) {
        const size_t had = b.bytes;
        b.reserve(n);
        grew = grew || b.bytes != had;
    };
    const bool mi = cost == TIPLHIP_COST_MI;
    if (mi) {
        grew = reserve_hist(ctx, hist_bytes(st, bins)) || grew;
        grow(ctx->d_costs, sizeof(double) * K * size_t(1 + (bins == 256 ? entropy_slices<256>()
                                                                        : entropy_slices<64>())));
    } else {
        grow(ctx->d_sums, sizeof(unsigned long long) * 4 * K);
        grow(ctx->d_costs, sizeof(double) * K);
    }
    if (st.samples >= kZTableSamples) grow(ctx->d_ztab, sizeof(float4) * K * st.zstride);
    // the finishing kernel writes the costs into coherent host memory: no
    // copy back, the host only waits for the stream
    const size_t had_hc = ctx->h_costs.bytes;
    ctx->h_costs.reserve(sizeof(double) * K);
    grew = grew || ctx->h_costs.bytes != had_hc;
    double* hc = static_cast<double*>(ctx->h_costs.dev);
    auto enqueue = [&] {
        upload_staged(ctx, st);
        if (mi) {
            launch_mi_hist(ctx, st, bins);
            launch_mi_entropy(ctx, st, bins, hc);
        } else {
            launch_corr_sums(ctx, st);
            launch_corr_finalize(ctx, st, nullptr, hc);
        }
    };
    if (grew && !ctx->graphs.empty()) {
        TIPLHIP_CHECK(hipStreamSynchronize(ctx->stream));
        clear_graphs(ctx);
    }
    // The host waits for the costs themselves: each slot starts as a NaN
    // pattern no cost kernel writes, and the finishing kernel's stores land
    // in coherent host memory.  Spinning on them returns sooner than a
    // stream synchronisation; after kCostSpinUs the wait falls back to one
    // (which also reports a failed launch).
    uint64_t* slots = static_cast<uint64_t*>(ctx->h_costs.p);
    std::fill(slots, slots + K, kCostPending);
    // from here on a kernel may be queued that stores into the slots: if a
    // later step throws, drain the stream before the caller moves on
    auto drain = [ctx] { (void)hipStreamSynchronize(ctx->stream); };
    DrainOnUnwind<decltype(drain)> unwind(drain);
    if (ctx->use_graphs && !ctx->kernel_timing) {
        // the same launch shape replays one graph; only the staged bytes
        // (transforms, crop flags, segment table) are new
        if (mi) make_hist_clean(ctx, hist_bytes(st, bins));  // outside the graph
        // (the buffers' addresses are part of the key: a graph is replayed
        // only on the buffers it was captured with, each reserved above for
        // this launch's sizes)
        const uint64_t key_fields[] = {uint64_t(cost), uint64_t(mi ? bins : 0), uint64_t(K),
                                       uint64_t(st.nseg), uint64_t(st.total_blocks), uint64_t(st.bytes),
                                       uint64_t(st.samples >= kZTableSamples), uint64_t(st.zstride),
                                       uint64_t(st.crop != nullptr),
                                       uint64_t(reinterpret_cast<uintptr_t>(ctx->d_T.p)),
                                       uint64_t(reinterpret_cast<uintptr_t>(ctx->h_T.p)),
                                       uint64_t(reinterpret_cast<uintptr_t>(ctx->d_hist.p)),
                                       uint64_t(reinterpret_cast<uintptr_t>(ctx->d_costs.p)),
                                       uint64_t(reinterpret_cast<uintptr_t>(ctx->d_sums.p)),
                                       uint64_t(reinterpret_cast<uintptr_t>(ctx->d_ztab.p)),
                                       uint64_t(reinterpret_cast<uintptr_t>(ctx->h_costs.dev))};
        const std::string key(reinterpret_cast<const char*>(key_fields), sizeof(key_fields));
        auto it = ctx->graphs.find(key);
        if (it == ctx->graphs.end()) {
            if (ctx->graphs.size() >= kMaxGraphs) {
                TIPLHIP_CHECK(hipStreamSynchronize(ctx->stream));
                clear_graphs(ctx);
            }
            TIPLHIP_CHECK(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
            hipGraph_t g = nullptr;
            try {
                enqueue();
            } catch (...) {
                (void)hipStreamEndCapture(ctx->stream, &g);
                if (g) (void)hipGraphDestroy(g);
                throw;
            }
            TIPLHIP_CHECK(hipStreamEndCapture(ctx->stream, &g));
            hipGraphExec_t ex = nullptr;
            const hipError_t e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            TIPLHIP_CHECK(e);
            it = ctx->graphs.emplace(key, ex).first;
        }
        TIPLHIP_CHECK(hipGraphLaunch(it->second, ctx->stream));
    } else {
        enqueue();
    }
    wait_for_slots(static_cast<const volatile uint64_t*>(slots), K, kCostPending,
                   std::chrono::microseconds(kCostSpinUs),
                   [ctx] { TIPLHIP_CHECK(hipStreamSynchronize(ctx->stream)); });
    std::memcpy(costs, ctx->h_costs.p, sizeof(double) * K);
}

// ---------------------------------------------------------------------------
// Cost evaluation that allocates nothing (tiplhip_cost_batch_noalloc).
//
// Every device buffer the launch needs beyond its inputs and outputs lives
// in `scratch`, a file-scope array each device holds for the life of the
// process.  Its address on a device is looked up once, with that device
// current, by hipGetSymbolAddress; host code uses only that address (the
// array's name in host code is a host-side stand-in, not device memory).
// Layout (byte offsets): transforms, crop flags, segment table, costs and
// entropy partials, correlation sums, then the K joint histograms and
// to_hist rows.  Calls on one device share the array: calls that overlap on
// one GPU race.
// ---------------------------------------------------------------------------
constexpr int kNoAllocMaxK = 256;
constexpr size_t kNoAllocScratchBytes = size_t(20) << 20;  // 20 MiB per device
constexpr size_t kNaT = 0;                                  // K x 12 floats
constexpr size_t kNaCrop = kNaT + sizeof(float) * 12 * kNoAllocMaxK;
constexpr size_t kNaSeg = kNaCrop + sizeof(int4) * kNoAllocMaxK;
constexpr size_t kNaCosts = kNaSeg + 1024;                  // K x (1 + 16) doubles
constexpr size_t kNaSums = kNaCosts + sizeof(double) * 17 * kNoAllocMaxK;
constexpr size_t kNaHist = (kNaSums + sizeof(unsigned long long) * 4 * kNoAllocMaxK + 255) & ~size_t(255);
static_assert(sizeof(Segment) <= 1024, "segment slot");

static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch[kMaxDevices] = {};

// The device address of `scratch` on ctx's device (current), looked up on
// the first call for that device.
unsigned char* scratch_of(tiplhip_ctx* ctx)
{
    TIPLHIP_REQUIRE(ctx->device >= 0 && ctx->device < kMaxDevices, "device index out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch[ctx->device]) {
        void* p = nullptr;
        TIPLHIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        TIPLHIP_REQUIRE(p != nullptr, "no device address for the static scratch buffer");
        g_scratch[ctx->device] = p;
    }
    return static_cast<unsigned char*>(g_scratch[ctx->device]);
}

void cost_noalloc(tiplhip_ctx* ctx, const tiplhip_volume* from, const tiplhip_volume* to,
                  const float* T, int K, int cost, int bins, double* costs)
{
    TIPLHIP_REQUIRE(T != nullptr && costs != nullptr, "null transform or cost buffer");
    TIPLHIP_REQUIRE(cost == TIPLHIP_COST_CORR || cost == TIPLHIP_COST_MI, "unknown cost type");
    const bool mi = cost == TIPLHIP_COST_MI;
    if (mi) TIPLHIP_REQUIRE(bins == 256 || bins == 64, "bins must be 256 or 64");
    TIPLHIP_REQUIRE(K >= 1 && K <= kNoAllocMaxK,
                    "K must be in [1, 256] for the allocation-free cost evaluation");
    check_volume(ctx, from, "from");
    check_volume(ctx, to, "to");
    const size_t hist_need = mi ? sizeof(int) * (size_t(bins) * bins + size_t(bins)) * size_t(K) : 0;
    TIPLHIP_REQUIRE(kNaHist + hist_need <= kNoAllocScratchBytes,
                    "the histograms of this many candidates exceed the 20 MiB static scratch "
                    "buffer (256 bins: at most 79 candidates)");
    DeviceGuard dg(ctx->device);
    unsigned char* scr = scratch_of(ctx);
    // host-side staging on the stack: transforms (with the correlation's
    // role swap), crop flags, the segment
    float Tk[12 * kNoAllocMaxK];
    int4 crop[kNoAllocMaxK];
    std::memcpy(Tk, T, sizeof(float) * 12 * size_t(K));
    if (!mi && from->n > to->n) {
        // reg::correlation swaps the roles when `from` is larger (reg/linear.hpp:28-33)
        std::swap(from, to);
        for (int k = 0; k < K; ++k) (void)matrix_inverse(&Tk[12 * k]);
    }
    for (int k = 0; k < K; ++k) {
        int pos[3] = {0, 0, 0};
        const bool c = !mi && is_integer_shift(&Tk[12 * k], pos);
        crop[k] = make_int4(c ? 1 : 0, pos[0], pos[1], pos[2]);
        for (int i = 9; i < 12; ++i)
            if (Tk[12 * k + i] == 0.0f) Tk[12 * k + i] = 0.0f;  // +0.0 (packed sampler's bounds test)
    }
    TIPLHIP_REQUIRE(to->d_pack != nullptr,
                    "the sampled volume has no packed copy yet: evaluate it once with "
                    "tiplhip_cost_batch (which builds it) before allocation-free calls");
    LaunchShape shape = default_shape(cost, bins);
    auto it = ctx->tuned.find(tune_key(from, to, K, cost, bins));
    if (it != ctx->tuned.end()) shape = it->second;
    Segment sg{};  // every field zero until set below
    sg.g = batch_geometry(from, to, K, shape.zc);
    const double work = double(sg.g.items) * double(K);
    const double target = shape.per_cu == 0
                              ? 1.0
                              : std::max(double(ctx->cu_count) * double(shape.per_cu),
                                         std::ceil(work / double(kMaxItemsPerBlock)));
    batch_blocks(sg.g, hist256(cost, bins) ? 65535u : (1u << 23),
                 uint32_t(std::min(std::ceil(target), 2147483647.0)));
    sg.from = from->d;
    sg.to = to->d;
    sg.to8 = to->d_pack;
    sg.from_hist = from->d_hist + (bins == 64 ? 256 : 0);
    sg.sx = (long long)from->sum;
    sg.sxx = (long long)from->sum_sq;
    sg.n = double(from->n);
    sg.block_begin = 0;
    sg.cand_begin = 0;
    const uint32_t blocks = sg.g.total_blocks;
    const Segment* d_seg = reinterpret_cast<const Segment*>(scr + kNaSeg);
    const float* d_T = reinterpret_cast<const float*>(scr + kNaT);
    const int4* d_crop = reinterpret_cast<const int4*>(scr + kNaCrop);
    double* d_costs = reinterpret_cast<double*>(scr + kNaCosts);
    hipStream_t st = ctx->stream;
    // the host arrays above stay alive until the stream is synchronised below
    auto drain = [st] { (void)hipStreamSynchronize(st); };
    DrainOnUnwind<decltype(drain)> unwind(drain);
    TIPLHIP_CHECK(hipMemcpyAsync(scr + kNaT, Tk, sizeof(float) * 12 * size_t(K), hipMemcpyHostToDevice, st));
    TIPLHIP_CHECK(hipMemcpyAsync(scr + kNaSeg, &sg, sizeof(Segment), hipMemcpyHostToDevice, st));
    if (mi) {
        // the array keeps the previous call's contents: clear the histograms
        // (a fill kernel, which a graph capture may also record)
        zero_bytes16(ctx, scr + kNaHist, hist_need);
        int* h = reinterpret_cast<int*>(scr + kNaHist);
        uint32_t* th = reinterpret_cast<uint32_t*>(h + size_t(bins) * bins * size_t(K));
        if (bins == 256) {
            hipLaunchKernelGGL((mi_hist_kernel<256, false>), dim3(blocks), dim3(1024), 0, st, d_seg, 1u,
                               blocks, d_T, nullptr, 0u, h, th);
            TIPLHIP_CHECK(hipGetLastError());
            constexpr int G = entropy_slices<256>();
            hipLaunchKernelGGL(mi_entropy_kernel<256>, dim3(uint32_t(K) * uint32_t(G)), dim3(kEntropyThreads),
                               0, st, d_seg, 1u, h, th, d_costs + K, d_costs, nullptr);
            TIPLHIP_CHECK(hipGetLastError());
            hipLaunchKernelGGL(mi_entropy_final_kernel<256>, dim3((uint32_t(K) + 255) / 256), dim3(256), 0,
                               st, d_costs + K, uint32_t(K), d_costs, nullptr, th);
        } else {
            hipLaunchKernelGGL((mi_hist_kernel<64, false>), dim3(blocks), dim3(256), 0, st, d_seg, 1u,
                               blocks, d_T, nullptr, 0u, h, th);
            TIPLHIP_CHECK(hipGetLastError());
            hipLaunchKernelGGL(mi_entropy_kernel<64>, dim3(uint32_t(K)), dim3(kEntropyThreads), 0, st, d_seg,
                               1u, h, th, d_costs + K, d_costs, nullptr);
        }
        TIPLHIP_CHECK(hipGetLastError());
    } else {
        TIPLHIP_CHECK(hipMemcpyAsync(scr + kNaCrop, crop, sizeof(int4) * size_t(K), hipMemcpyHostToDevice, st));
        unsigned long long* sums = reinterpret_cast<unsigned long long*>(scr + kNaSums);
        zero_u64(ctx, sums, 4 * size_t(K));
        hipLaunchKernelGGL(corr_sums_kernel<false>, dim3(blocks), dim3(256), 0, st, d_seg, 1u, blocks, d_T,
                           d_crop, nullptr, 0u, sums);
        TIPLHIP_CHECK(hipGetLastError());
        hipLaunchKernelGGL(corr_finalize_kernel, dim3((uint32_t(K) + 63) / 64), dim3(64), 0, st, d_seg, 1u,
                           sums, K, d_costs, nullptr, nullptr);
        TIPLHIP_CHECK(hipGetLastError());
    }
    TIPLHIP_CHECK(hipMemcpyAsync(costs, d_costs, sizeof(double) * size_t(K), hipMemcpyDeviceToHost, st));
    TIPLHIP_CHECK(hipStreamSynchronize(st));
}

}  // namespace
}  // namespace tiplhip

using namespace tiplhip;

extern "C" {

int tiplhip_resample(tiplhip_ctx* ctx, const tiplhip_volume* src, const int out_dim[3],
                     const float T[12], int crop_shortcut, uint8_t* out)
{
    return guarded([&] {
        TIPLHIP_REQUIRE(ctx != nullptr, "null context");
        std::lock_guard<std::recursive_mutex> lk(ctx->mu);
        check_volume(ctx, src, "src");
        TIPLHIP_REQUIRE(out_dim && out_dim[0] > 0 && out_dim[1] > 0 && out_dim[2] > 0,
                        "bad output shape");
        TIPLHIP_REQUIRE(T && out, "null transform or output");
        DeviceGuard dg(ctx->device);
        size_t n = vox(out_dim);
        ctx->d_tmp.reserve(n);
        launch_resample(ctx, src, out_dim, T, crop_shortcut != 0, ctx->d_tmp.as<uint8_t>());
        read_back(ctx, ctx->d_tmp.p, out, n);
    });
}

int tiplhip_volume_resample(tiplhip_ctx* ctx, const tiplhip_volume* src, const int out_dim[3],
                            const float T[12], int crop_shortcut, tiplhip_volume** out)
{
    return guarded([&] {
        TIPLHIP_REQUIRE(out != nullptr, "null output");
        *out = nullptr;
        TIPLHIP_REQUIRE(ctx != nullptr, "null context");
        std::lock_guard<std::recursive_mutex> lk(ctx->mu);
        check_volume(ctx, src, "src");
        TIPLHIP_REQUIRE(out_dim && out_dim[0] > 0 && out_dim[1] > 0 && out_dim[2] > 0,
                        "bad output shape");
        TIPLHIP_REQUIRE(T != nullptr, "null transform");
        DeviceGuard dg(ctx->device);
        tiplhip_volume* v = volume_alloc(ctx, out_dim);
        try {
            launch_resample(ctx, src, out_dim, T, crop_shortcut != 0, v->d);
            volume_compute_stats(ctx, v);
        } catch (...) {
            volume_release(v);
            throw;
        }
        *out = v;
    });
}

int tiplhip_corr_sums(tiplhip_ctx* ctx, const tiplhip_volume* from, const tiplhip_volume* to,
                      const float* T, int K, int64_t* sums)
{
    return guarded([&] {
        TIPLHIP_REQUIRE(ctx != nullptr, "null context");
        std::lock_guard<std::recursive_mutex> lk(ctx->mu);
        check_volume(ctx, from, "from");
        check_volume(ctx, to, "to");
        TIPLHIP_REQUIRE(T && sums, "null transform or output");
        TIPLHIP_REQUIRE(K >= 1 && K <= kMaxK, "K must be in [1, 4096]");
        DeviceGuard dg(ctx->device);
        volume_ensure_packed(ctx, const_cast<tiplhip_volume*>(to));
        std::vector<int4> crop(size_t(K), make_int4(0, 0, 0, 0));
        const Staged st = stage_jobs(ctx, {PairJob{from, to, K}}, T, crop.data(),
                                     TIPLHIP_COST_CORR, 256);
        launch_corr_sums(ctx, st);
        ctx->d_tmp2.reserve(sizeof(long long) * 5 * size_t(K));
        launch_corr_finalize(ctx, st, ctx->d_tmp2.as<long long>());
        read_back(ctx, ctx->d_tmp2.p, sums, sizeof(int64_t) * 5 * size_t(K));
    });
}

int tiplhip_mi_hist(tiplhip_ctx* ctx, const tiplhip_volume* from, const tiplhip_volume* to,
                    const float* T, int K, int bins, int32_t* hist)
{
    return guarded([&] {
        TIPLHIP_REQUIRE(ctx != nullptr, "null context");
        std::lock_guard<std::recursive_mutex> lk(ctx->mu);
        check_volume(ctx, from, "from");
        check_volume(ctx, to, "to");
        TIPLHIP_REQUIRE(T && hist, "null transform or output");
        TIPLHIP_REQUIRE(K >= 1 && K <= kMaxK, "K must be in [1, 4096]");
        TIPLHIP_REQUIRE(bins == 256 || bins == 64, "bins must be 256 or 64");
        DeviceGuard dg(ctx->device);
        volume_ensure_packed(ctx, const_cast<tiplhip_volume*>(to));
        const Staged st = stage_jobs(ctx, {PairJob{from, to, K}}, T, nullptr, TIPLHIP_COST_MI, bins);
        launch_mi_hist(ctx, st, bins);
        read_back(ctx, ctx->d_hist.p, hist, sizeof(int) * size_t(bins) * bins * size_t(K));
    });
}

int tiplhip_autotune(tiplhip_ctx* ctx, const tiplhip_volume* from, const tiplhip_volume* to,
                     const float* T, int K, int cost, int bins, int reps, int* zc, int* per_cu,
                     double* best_ms)
{
    return guarded([&] {
        TIPLHIP_REQUIRE(ctx != nullptr && T != nullptr, "null argument");
        std::lock_guard<std::recursive_mutex> lk(ctx->mu);
        check_volume(ctx, from, "from");
        check_volume(ctx, to, "to");
        TIPLHIP_REQUIRE(K >= 1 && K <= kMaxK, "K must be in [1, 4096]");
        TIPLHIP_REQUIRE(cost == TIPLHIP_COST_CORR || cost == TIPLHIP_COST_MI, "unknown cost type");
        if (cost == TIPLHIP_COST_MI) TIPLHIP_REQUIRE(bins == 256 || bins == 64, "bins must be 256 or 64");
        TIPLHIP_REQUIRE(reps >= 1 && reps <= 100, "reps must be in [1, 100]");
        DeviceGuard dg(ctx->device);
        // the correlation path evaluates with the roles swapped when `from`
        // is larger (reg/linear.hpp:28-33); tune what will run
        const tiplhip_volume *f = from, *t = to;
        std::vector<float> Tk(T, T + 12 * size_t(K));
        std::vector<int4> crop;
        if (cost == TIPLHIP_COST_CORR) {
            if (f->n > t->n) {
                std::swap(f, t);
                for (int k = 0; k < K; ++k) (void)matrix_inverse(&Tk[12 * size_t(k)]);
            }
            crop.resize(size_t(K));
            for (int k = 0; k < K; ++k) {
                int pos[3];
                const bool c = is_integer_shift(&Tk[12 * size_t(k)], pos);
                crop[size_t(k)] = make_int4(c ? 1 : 0, pos[0], pos[1], pos[2]);
            }
        }
        volume_ensure_packed(ctx, const_cast<tiplhip_volume*>(t));
        const bool wide = hist256(cost, bins);  // 1024-thread workgroups
        // measured at 182x218x182 with 64 candidates: the 64-bin MI and
        // correlation kernels keep getting faster up to 48-64 workgroups per
        // CU (finer slices of the work; 0.647 -> 0.584 ms from 16 to 64 for
        // MI), the 256-bin one is indifferent from 0 to 4
        const uint32_t zcs[] = {8u, 16u, 32u};
        const std::vector<uint32_t> pcs = wide ? std::vector<uint32_t>{1u, 2u, 4u, 8u, 0u}
                                               : std::vector<uint32_t>{2u, 4u, 8u, 16u, 32u, 48u, 64u, 0u};
        timing_flush(ctx);  // the caller's own pending launches count for the caller
        const bool keep_timing = ctx->kernel_timing;
        const double keep_ms = ctx->kernel_ms;
        const int64_t keep_n = ctx->kernel_launches;
        ctx->kernel_timing = true;
        // `n` timed launches of one shape (after one that warms up), appended
        // to `times`
        auto time_shape = [&](const LaunchShape& shape, int n, std::vector<double>& times) {
            const Staged st = stage_jobs(ctx, {PairJob{f, t, K}}, Tk.data(),
                                         crop.empty() ? nullptr : crop.data(), cost, bins, &shape);
            for (int r = 0; r <= n; ++r) {
                ctx->kernel_ms = 0.0;
                if (cost == TIPLHIP_COST_CORR) launch_corr_sums(ctx, st);
                else launch_mi_hist(ctx, st, bins);
                timing_flush(ctx);
                if (r > 0) times.push_back(ctx->kernel_ms);
            }
        };
        auto median = [](std::vector<double> v) {
            std::sort(v.begin(), v.end());
            return v[v.size() / 2];
        };
        // Two passes against the noise of single launches (clocks drift by a
        // few percent within a run): every shape by the median of `reps`
        // launches, then the four best again, interleaved, by the median of
        // all their launches.  A lone minimum over two dozen shapes picks
        // lucky samples: (8, 48) measured 0.58 ms once and ran at 0.67 ms.
        std::vector<std::pair<double, LaunchShape>> first;
        for (uint32_t z : zcs)
            for (uint32_t pc : pcs) {
                std::vector<double> v;
                time_shape(LaunchShape{z, pc}, reps, v);
                first.emplace_back(median(v), LaunchShape{z, pc});
            }
        std::sort(first.begin(), first.end(),
                  [](const auto& a, const auto& b) { return a.first < b.first; });
        const size_t finalists = std::min<size_t>(4, first.size());
        std::vector<std::vector<double>> again(finalists);
        for (int round = 0; round < 3; ++round)
            for (size_t i = 0; i < finalists; ++i) time_shape(first[i].second, reps, again[i]);
        LaunchShape best = first[0].second;
        double best_t = 1e30;
        for (size_t i = 0; i < finalists; ++i) {
            const double m = median(again[i]);
            if (m < best_t) {
                best_t = m;
                best = first[i].second;
            }
        }
        ctx->kernel_timing = keep_timing;
        ctx->kernel_ms = keep_ms;
        ctx->kernel_launches = keep_n;
        ctx->tuned[tune_key(f, t, K, cost, bins)] = best;  // the order stage_jobs sees
        if (zc) *zc = int(best.zc);
        if (per_cu) *per_cu = int(best.per_cu);
        if (best_ms) *best_ms = best_t;
    });
}

int tiplhip_set_launch_shape(tiplhip_ctx* ctx, const tiplhip_volume* from, const tiplhip_volume* to,
                             int K, int cost, int bins, int zc, int per_cu)
{
    return guarded([&] {
        TIPLHIP_REQUIRE(ctx != nullptr, "null context");
        std::lock_guard<std::recursive_mutex> lk(ctx->mu);
        check_volume(ctx, from, "from");
        check_volume(ctx, to, "to");
        TIPLHIP_REQUIRE(K >= 1 && K <= kMaxK, "K must be in [1, 4096]");
        TIPLHIP_REQUIRE(cost == TIPLHIP_COST_CORR || cost == TIPLHIP_COST_MI, "unknown cost type");
        if (cost == TIPLHIP_COST_MI) TIPLHIP_REQUIRE(bins == 256 || bins == 64, "bins must be 256 or 64");
        TIPLHIP_REQUIRE(zc >= 1 && zc <= 1024, "planes per item must be in [1, 1024]");
        TIPLHIP_REQUIRE(per_cu >= 0 && per_cu <= 64, "workgroups per CU must be in [0, 64]");
        // correlation evaluates with the roles swapped when `from` is larger
        const tiplhip_volume *f = from, *t = to;
        if (cost == TIPLHIP_COST_CORR && f->n > t->n) std::swap(f, t);
        ctx->tuned[tune_key(f, t, K, cost, bins)] = LaunchShape{uint32_t(zc), uint32_t(per_cu)};
    });
}

int tiplhip_cost_batch(tiplhip_ctx* ctx, const tiplhip_volume* from, const tiplhip_volume* to,
                       const float* T, int K, int cost, int bins, double* costs)
{
    return guarded([&] {
        TIPLHIP_REQUIRE(ctx != nullptr, "null context");
        std::lock_guard<std::recursive_mutex> lk(ctx->mu);
        cost_jobs(ctx, {PairJob{from, to, K}}, T, cost, bins, costs);
    });
}

int tiplhip_cost_batch_noalloc(tiplhip_ctx* ctx, const tiplhip_volume* from,
                               const tiplhip_volume* to, const float* T, int K, int cost, int bins,
                               double* costs)
{
    return guarded([&] {
        TIPLHIP_REQUIRE(ctx != nullptr, "null context");
        std::lock_guard<std::recursive_mutex> lk(ctx->mu);
        cost_noalloc(ctx, from, to, T, K, cost, bins, costs);
    });
}

int tiplhip_cost_batch_pairs(tiplhip_ctx* ctx, const tiplhip_pair_costs* pairs, int n, int cost,
                             int bins)
{
    return guarded([&] {
        TIPLHIP_REQUIRE(ctx != nullptr && pairs != nullptr, "null argument");
        TIPLHIP_REQUIRE(n >= 1, "need at least one pair");
        std::lock_guard<std::recursive_mutex> lk(ctx->mu);
        std::vector<PairJob> jobs;
        std::vector<float> T;
        size_t K = 0;
        for (int i = 0; i < n; ++i) {
            const tiplhip_pair_costs& p = pairs[i];
            TIPLHIP_REQUIRE(p.T != nullptr && p.costs != nullptr, "null transform or cost buffer");
            TIPLHIP_REQUIRE(p.K >= 1 && p.K <= kMaxK, "K must be in [1, 4096]");
            K += size_t(p.K);
            TIPLHIP_REQUIRE(K <= size_t(kMaxK), "more than 4096 candidates in one launch");
            jobs.push_back(PairJob{p.from, p.to, p.K});
            T.insert(T.end(), p.T, p.T + 12 * size_t(p.K));
        }
        std::vector<double> costs(K);
        // Multi-pair launches serve throughput (a batcher merging the candidate
        // sets of many registrations): 256-bin histograms then use the fewest,
        // fullest workgroups, which spend the least on clearing and flushing
        // their 128 KB LDS histograms per sample.
        const LaunchShape fullest{16u, 0u};
        cost_jobs(ctx, jobs, T.data(), cost, bins, costs.data(), hist256(cost, bins) ? &fullest : nullptr);
        size_t base = 0;
        for (int i = 0; i < n; ++i) {
            std::copy_n(costs.begin() + std::ptrdiff_t(base), pairs[i].K, pairs[i].costs);
            base += size_t(pairs[i].K);
        }
    });
}

int tiplhip_cost_batch_params(tiplhip_ctx* ctx, const tiplhip_volume* from, const float from_vs[3],
                              const tiplhip_volume* to, const float to_vs[3], const float* params,
                              int K, int cost, int bins, double* costs)
{
    return guarded([&] {
        TIPLHIP_REQUIRE(ctx != nullptr, "null context");
        std::lock_guard<std::recursive_mutex> lk(ctx->mu);
        check_volume(ctx, from, "from");
        check_volume(ctx, to, "to");
        TIPLHIP_REQUIRE(params && from_vs && to_vs, "null parameters");
        TIPLHIP_REQUIRE(K >= 1 && K <= kMaxK, "K must be in [1, 4096]");
        std::vector<float> T(12 * size_t(K));
        for (int k = 0; k < K; ++k)
            affine_matrix(params + 12 * size_t(k), from->dim, from_vs, to->dim, to_vs,
                          &T[12 * size_t(k)]);
        cost_jobs(ctx, {PairJob{from, to, K}}, T.data(), cost, bins, costs);
    });
}

}  // extern "C"
