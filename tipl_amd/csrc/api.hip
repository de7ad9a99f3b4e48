// api.hip -- errors, contexts, device-resident volumes, quantisation,
// value statistics, the uint8 pyramid and host geometry helpers.
#include "device.hpp"
#include "internal.hpp"

#include <algorithm>
#include <cmath>
#include <vector>

namespace tiplhip {

namespace {
thread_local std::string g_last_error;
}

void set_last_error(const std::string& msg) { g_last_error = msg; }
void clear_last_error() { g_last_error.clear(); }

// ---------------------------------------------------------------------------
// Stream-ordered device memory
// ---------------------------------------------------------------------------

namespace {
std::mutex g_pool_mu;
std::vector<hipMemPool_t> g_pools;  // one per device, kept for the process lifetime
}

hipMemPool_t device_pool(int device)
{
    std::lock_guard<std::mutex> lk(g_pool_mu);
    if (g_pools.size() <= size_t(device)) g_pools.resize(size_t(device) + 1, nullptr);
    if (!g_pools[size_t(device)]) {
        hipMemPoolProps props;
        std::memset(&props, 0, sizeof(props));
        props.allocType = hipMemAllocationTypePinned;
        props.handleTypes = hipMemHandleTypeNone;
        props.location.type = hipMemLocationTypeDevice;
        props.location.id = device;
        hipMemPool_t pool = nullptr;
        TIPLHIP_CHECK(hipMemPoolCreate(&pool, &props));
        // keep freed blocks for reuse instead of returning them at every sync
        uint64_t keep = ~uint64_t(0);
        TIPLHIP_CHECK(hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep));
        g_pools[size_t(device)] = pool;
    }
    return g_pools[size_t(device)];
}

void* pool_alloc(tiplhip_ctx* ctx, size_t bytes)
{
    if (!bytes) return nullptr;
    void* p = nullptr;
    TIPLHIP_CHECK(hipMallocFromPoolAsync(&p, bytes, device_pool(ctx->device), ctx->stream));
    return p;
}

void pool_free(tiplhip_ctx* ctx, void* p)
{
    if (p) (void)hipFreeAsync(p, ctx->stream);
}

void clear_graphs(tiplhip_ctx* ctx)
{
    for (auto& kv : ctx->graphs) (void)hipGraphExecDestroy(kv.second);
    ctx->graphs.clear();
}

// ---------------------------------------------------------------------------
// Host geometry, float arithmetic in the reference's order
// ---------------------------------------------------------------------------

// transformation_matrix(affine_param, from, from_vs, to, to_vs) with
// rotation_scaling_affine_matrix (numerical/transformation.hpp:141-216,
// 612-666).
void affine_matrix(const float* p, const int* fdim, const float* fvs, const int* tdim,
                   const float* tvs, float* T)
{
    const float* tr = p;
    const float* rot = p + 3;
    const float* sc = p + 6;
    const float* af = p + 9;
    const float sin_x = std::sin(rot[0]), cos_x = std::cos(rot[0]);
    const float sin_y = std::sin(rot[1]), cos_y = std::cos(rot[1]);
    const float sin_z = std::sin(rot[2]), cos_z = std::cos(rot[2]);
    const float cxsz = cos_x * sin_z, cxcz = cos_x * cos_z;
    const float sxsz = sin_x * sin_z, sxcz = sin_x * cos_z;
    float m[9] = {cos_y * cos_z,         cos_y * sin_z,         -sin_y,
                  -cxsz + sxcz * sin_y,  cxcz + sxsz * sin_y,   sin_x * cos_y,
                  sxsz + cxcz * sin_y,   -sxcz + cxsz * sin_y,  cos_x * cos_y};
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) m[r * 3 + c] *= sc[r];
    m[0] += m[3] * af[0] + m[6] * af[1];
    m[1] += m[4] * af[0] + m[7] * af[1];
    m[2] += m[5] * af[0] + m[8] * af[1];
    m[3] += m[6] * af[2];
    m[4] += m[7] * af[2];
    m[5] += m[8] * af[2];
    float t[3] = {float(fdim[0]), float(fdim[1]), float(fdim[2])};
    for (int i = 0; i < 3; ++i) {
        t[i] *= -0.5f;
        t[i] *= fvs[i];
        t[i] += tr[i];
    }
    float sh[3];
    for (int r = 0; r < 3; ++r) sh[r] = m[r * 3] * t[0] + m[r * 3 + 1] * t[1] + m[r * 3 + 2] * t[2];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) m[r * 3 + c] *= fvs[c];
    for (int r = 0; r < 3; ++r)
        if (tvs[r] != 1.0f) {
            for (int c = 0; c < 3; ++c) m[r * 3 + c] /= tvs[r];
            sh[r] /= tvs[r];
        }
    for (int r = 0; r < 3; ++r) sh[r] += float(tdim[r]) * 0.5f;
    std::copy(m, m + 9, T);
    std::copy(sh, sh + 3, T + 9);
}

// transformation_matrix::inverse with the 3x3 cofactor inverse
// (numerical/transformation.hpp:820-832, numerical/matrix.hpp:1346-1352,
// 1479-1506).  Leaves T unchanged and returns false when singular.
bool matrix_inverse(float* T)
{
    const float* a = T;
    const float det = a[0] * (a[4] * a[8] - a[5] * a[7]) + a[1] * (a[5] * a[6] - a[3] * a[8]) +
                      a[2] * (a[3] * a[7] - a[4] * a[6]);
    if (det + 1.0f == 1.0f) return false;
    float c[9] = {a[4] * a[8] - a[5] * a[7], a[2] * a[7] - a[1] * a[8], a[1] * a[5] - a[2] * a[4],
                  a[5] * a[6] - a[3] * a[8], a[0] * a[8] - a[2] * a[6], a[2] * a[3] - a[0] * a[5],
                  a[3] * a[7] - a[4] * a[6], a[1] * a[6] - a[0] * a[7], a[0] * a[4] - a[1] * a[3]};
    for (float& v : c) v /= det;
    float s[3];
    for (int r = 0; r < 3; ++r) s[r] = T[9] * c[r * 3] + T[10] * c[r * 3 + 1] + T[11] * c[r * 3 + 2];
    std::copy(c, c + 9, T);
    for (int r = 0; r < 3; ++r) T[9 + r] = -s[r];
    return true;
}

// The CPU resample's shortcut test (numerical/resampling.hpp:1244-1262):
// identity rotation and integer shift to within 1e-4.
bool is_integer_shift(const float* T, int* from_pos)
{
    const double eps = 1.0e-4;
    for (int i = 0; i < 3; ++i) {
        const long s = std::lround(T[9 + i]);
        if (std::fabs(T[9 + i] - float(s)) > eps) return false;
        from_pos[i] = int(s);
        for (int j = 0; j < 3; ++j)
            if (std::fabs(T[i * 3 + j] - float(i == j)) > eps) return false;
    }
    return true;
}

namespace {

// ---------------------------------------------------------------------------
// Volume kernels
// ---------------------------------------------------------------------------

// Exact sum, sum of squares and 256-bin histogram of a uint8 volume.  The
// LDS histogram adds are aggregated over lanes that share the first active
// lane's value (the background value in practice).
__global__ __launch_bounds__(256) void volume_stats_kernel(const uint8_t* __restrict__ v, size_t n,
                                                           unsigned long long* __restrict__ sums,
                                                           uint32_t* __restrict__ hist)
{
    __shared__ uint32_t lh[256];
    __shared__ unsigned long long red[2][4];
    lh[threadIdx.x] = 0u;
    __syncthreads();
    unsigned long long s = 0, s2 = 0;
    const size_t stride = size_t(gridDim.x) * blockDim.x;
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
        const uint32_t x = v[i];
        s += x;
        s2 += x * x;
        const uint32_t x0 = __builtin_amdgcn_readfirstlane(x);
        const unsigned long long m = __ballot(x == x0);
        if (x == x0) {
            if ((threadIdx.x & 63u) == uint32_t(__ffsll((long long)m) - 1))
                atomicAdd(&lh[x0], uint32_t(__popcll(m)));
        } else {
            atomicAdd(&lh[x], 1u);
        }
    }
    s = wave_sum_u64(s);
    s2 = wave_sum_u64(s2);
    if ((threadIdx.x & 63u) == 0) {
        red[0][threadIdx.x >> 6] = s;
        red[1][threadIdx.x >> 6] = s2;
    }
    __syncthreads();
    if (threadIdx.x < 2) {
        unsigned long long t = red[threadIdx.x][0] + red[threadIdx.x][1] + red[threadIdx.x][2] +
                               red[threadIdx.x][3];
        atomicAdd(&sums[threadIdx.x], t);
    }
    if (lh[threadIdx.x]) atomicAdd(&hist[threadIdx.x], lh[threadIdx.x]);
}

__device__ inline int f32_key(float f)
{
    int b = __float_as_int(f);
    return b >= 0 ? b : (b ^ 0x7fffffff);
}

template <class T>
__device__ inline int value_key(T v);
template <>
__device__ inline int value_key<float>(float v) { return f32_key(v); }
template <>
__device__ inline int value_key<int16_t>(int16_t v) { return int(v); }

// Min and max of the input (minmax_value, used by normalize_upper_lower2).
template <class T>
__global__ __launch_bounds__(256) void minmax_kernel(const T* __restrict__ in, size_t n,
                                                     int* __restrict__ mm)
{
    int lo = 0x7fffffff, hi = int(0x80000000);
    const size_t stride = size_t(gridDim.x) * blockDim.x;
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
        const int k = value_key<T>(in[i]);
        lo = min(lo, k);
        hi = max(hi, k);
    }
    for (int off = 32; off > 0; off >>= 1) {
        lo = min(lo, __shfl_xor(lo, off, 64));
        hi = max(hi, __shfl_xor(hi, off, 64));
    }
    if ((threadIdx.x & 63u) == 0) {
        atomicMin(&mm[0], lo);
        atomicMax(&mm[1], hi);
    }
}

// out = (in - min) * coef truncated to uint8 (normalize_upper_lower2,
// numerical/numerical.hpp:909-927 / 1085-1108).
template <class T>
__global__ __launch_bounds__(256) void normalize_kernel(const T* __restrict__ in, size_t n,
                                                        float mn, float coef,
                                                        uint8_t* __restrict__ out)
{
    const size_t stride = size_t(gridDim.x) * blockDim.x;
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
        const float v = (float(in[i]) - mn) * coef;
        out[i] = uint8_t(uint32_t(v));
    }
}

// downsample_with_padding for uint8 (numerical/resampling.hpp:575-613).
__global__ __launch_bounds__(256) void downsample_u8_kernel(const uint8_t* __restrict__ in, int w,
                                                            int h, int d, int ow, int oh, int od,
                                                            uint8_t* __restrict__ out)
{
    const size_t n = size_t(ow) * oh * od;
    const size_t wh = size_t(w) * h;
    for (size_t o = size_t(blockIdx.x) * blockDim.x + threadIdx.x; o < n;
         o += size_t(gridDim.x) * blockDim.x) {
        const int x = int(o % uint32_t(ow));
        const uint32_t r = uint32_t(o / uint32_t(ow));
        const int y = int(r % uint32_t(oh));
        const int z = int(r / uint32_t(oh));
        const int x2 = 2 * x, y2 = 2 * y, z2 = 2 * z;
        const size_t base = (size_t(z2) * h + y2) * w + x2;
        int has = 0;
        if (x2 + 1 < w) has += 1;
        if (y2 + 1 < h) has += 2;
        if (z2 + 1 < d) has += 4;
        const size_t shift[8] = {0, 1, size_t(w), size_t(w) + 1, wh, wh + 1, wh + w, wh + w + 1};
        uint32_t buf[8];
        uint32_t sum = buf[0] = in[base];
        for (int i = 1; i < 8; ++i) {
            const int hh = has & i;
            buf[i] = (hh == i) ? uint32_t(in[base + shift[i]]) : buf[hh];
            sum += buf[i];
        }
        out[o] = uint8_t((sum & 0xffffu) >> 3);
    }
}

// 8-tap packing of a uint8 volume: word v = bytes at v, v+1, v+w, v+w+1,
// v+wh, v+wh+1, v+wh+w, v+wh+w+1 (0 past the end of the array).  Only cells
// whose far corner is inside the volume are ever sampled, and for those the
// word holds exactly the 8 trilinear taps.
__global__ __launch_bounds__(256) void pack8_kernel(const uint8_t* __restrict__ v, size_t n,
                                                    uint32_t w, uint32_t wh,
                                                    uint64_t* __restrict__ out)
{
    const size_t stride = size_t(gridDim.x) * blockDim.x;
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
        const size_t off[8] = {0, 1, w, size_t(w) + 1, wh, size_t(wh) + 1, size_t(wh) + w,
                               size_t(wh) + w + 1};
        uint64_t word = 0;
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const size_t j = i + off[k];
            const uint64_t b = j < n ? uint64_t(v[j]) : 0ull;
            word |= b << (8 * k);
        }
        out[i] = word;
    }
}

// Kernels of known HBM traffic (bandwidth ceiling and memory-counter
// calibration).  Each block owns a contiguous chunk; every thread keeps
// kProbeUnroll 16-byte accesses in flight; loads and stores are
// non-temporal (streamed data is not kept in the caches).
constexpr int kProbeThreads = 256;
constexpr int kProbeUnroll = 8;
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

// kind 0: copy, 16 bytes a lane.
__global__ __launch_bounds__(kProbeThreads) void stream_copy_kernel(const u32x4* __restrict__ in,
                                                                    u32x4* __restrict__ out, size_t n,
                                                                    size_t per_block)
{
    const size_t b0 = size_t(blockIdx.x) * per_block, b1 = min(b0 + per_block, n);
    constexpr size_t step = size_t(kProbeThreads) * kProbeUnroll;
    size_t i = b0 + threadIdx.x;
    for (; i + (kProbeUnroll - 1) * kProbeThreads < b1; i += step) {
        u32x4 v[kProbeUnroll];
#pragma unroll
        for (int u = 0; u < kProbeUnroll; ++u) v[u] = __builtin_nontemporal_load(in + i + u * kProbeThreads);
#pragma unroll
        for (int u = 0; u < kProbeUnroll; ++u) __builtin_nontemporal_store(v[u], out + i + u * kProbeThreads);
    }
    for (; i < b1; i += kProbeThreads) __builtin_nontemporal_store(__builtin_nontemporal_load(in + i), out + i);
}

// kind 1: read-only, 16 bytes a lane; kind 2: read-only as the field
// kernels read packed float3 voxels (three 4-byte loads a lane at a 12-byte
// stride).  The xor of everything read goes to `sink` (one word a block),
// so no load is dropped.
template <int KIND>
__global__ __launch_bounds__(kProbeThreads) void stream_read_kernel(const u32x4* __restrict__ in, size_t n,
                                                                    size_t per_block, uint32_t* __restrict__ sink)
{
    const size_t b0 = size_t(blockIdx.x) * per_block, b1 = min(b0 + per_block, n);
    uint32_t acc = 0;
    if constexpr (KIND == 1) {
        constexpr size_t step = size_t(kProbeThreads) * kProbeUnroll;
        size_t i = b0 + threadIdx.x;
        for (; i + (kProbeUnroll - 1) * kProbeThreads < b1; i += step) {
#pragma unroll
            for (int u = 0; u < kProbeUnroll; ++u) {
                const u32x4 v = __builtin_nontemporal_load(in + i + u * kProbeThreads);
                acc ^= v.x ^ v.y ^ v.z ^ v.w;
            }
        }
        for (; i < b1; i += kProbeThreads) {
            const u32x4 v = in[i];
            acc ^= v.x ^ v.y ^ v.z ^ v.w;
        }
    } else {
        // the chunk as 12-byte voxels: voxel j at words 3j, 3j+1, 3j+2
        const uint32_t* w = reinterpret_cast<const uint32_t*>(in);
        const size_t v0 = b0 * 4 / 3, v1 = b1 * 4 / 3;  // 16-byte units -> 12-byte voxels
        size_t j = v0 + threadIdx.x;
        for (; j + (kProbeUnroll - 1) * kProbeThreads < v1; j += size_t(kProbeThreads) * kProbeUnroll) {
#pragma unroll
            for (int u = 0; u < kProbeUnroll; ++u) {
                const size_t q = 3 * (j + size_t(u) * kProbeThreads);
                acc ^= w[q] ^ w[q + 1] ^ w[q + 2];
            }
        }
        for (; j < v1; j += kProbeThreads) acc ^= w[3 * j] ^ w[3 * j + 1] ^ w[3 * j + 2];
    }
    for (int off = 32; off > 0; off >>= 1) acc ^= __shfl_xor(acc, off, 64);
    if ((threadIdx.x & 63u) == 0) atomicXor(&sink[blockIdx.x], acc);
}

int grid_for(size_t n, int cap = 16384)
{
    return int(std::max<size_t>(1, std::min<size_t>((n + 255) / 256, size_t(cap))));
}

}  // namespace

void launch_downsample_u8(tiplhip_ctx* ctx, const uint8_t* in, const int* dim, uint8_t* out)
{
    const int od[3] = {(dim[0] + 1) / 2, (dim[1] + 1) / 2, (dim[2] + 1) / 2};
    hipLaunchKernelGGL(downsample_u8_kernel, dim3(grid_for(vox(od))), dim3(256), 0, ctx->stream, in,
                       dim[0], dim[1], dim[2], od[0], od[1], od[2], out);
    TIPLHIP_CHECK(hipGetLastError());
}

tiplhip_volume* volume_alloc(tiplhip_ctx* ctx, const int* dim)
{
    TIPLHIP_REQUIRE(dim && dim[0] > 0 && dim[1] > 0 && dim[2] > 0, "volume dims must be positive");
    const size_t n = vox(dim);
    TIPLHIP_REQUIRE(n < (size_t(1) << 31), "volume exceeds 2^31 voxels");
    auto* v = new tiplhip_volume();
    v->ctx = ctx;
    std::copy(dim, dim + 3, v->dim);
    v->n = n;
    try {
        v->d = static_cast<uint8_t*>(pool_alloc(ctx, n));
        v->d_hist = static_cast<uint32_t*>(pool_alloc(ctx, sizeof(uint32_t) * (256 + 64)));
    } catch (...) {
        volume_release(v);
        throw;
    }
    return v;
}

void volume_release(tiplhip_volume* v)
{
    if (!v) return;
    pool_free(v->ctx, v->d);
    pool_free(v->ctx, v->d_hist);
    pool_free(v->ctx, v->d_pack);
    delete v;
}

void volume_ensure_packed(tiplhip_ctx* ctx, tiplhip_volume* vol)
{
    // A volume may be used from several contexts of its device: the first
    // user builds the packed copy and waits for it, so every later user, on
    // any stream, reads a finished copy.
    std::lock_guard<std::mutex> lk(vol->pack_mu);
    if (vol->d_pack) return;
    TIPLHIP_REQUIRE(vol->n * 8u < (size_t(1) << 32), "volume too large for the packed sampler");
    void* p = pool_alloc(ctx, vol->n * 8u);
    hipLaunchKernelGGL(pack8_kernel, dim3(grid_for(vol->n)), dim3(256), 0, ctx->stream, vol->d,
                       vol->n, uint32_t(vol->dim[0]), uint32_t(vol->dim[0]) * uint32_t(vol->dim[1]),
                       static_cast<uint64_t*>(p));
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) {
        pool_free(ctx, p);
        TIPLHIP_CHECK(e);
    }
    vol->d_pack = static_cast<uint64_t*>(p);
}

void timing_flush(tiplhip_ctx* ctx)
{
    if (ctx->kt_pending == 0) return;
    const size_t n = ctx->kt_pending;
    ctx->kt_pending = 0;
    TIPLHIP_CHECK(hipEventSynchronize(ctx->kt_events[2 * n - 1]));
    for (size_t i = 0; i < n; ++i) {
        float ms = 0.0f;
        TIPLHIP_CHECK(hipEventElapsedTime(&ms, ctx->kt_events[2 * i], ctx->kt_events[2 * i + 1]));
        ctx->kernel_ms += ms;
        ctx->kernel_launches += 1;
    }
}

void read_back(tiplhip_ctx* ctx, const void* dev, void* host, size_t bytes)
{
    ctx->h_out.reserve(bytes);
    TIPLHIP_CHECK(hipMemcpyAsync(ctx->h_out.p, dev, bytes, hipMemcpyDeviceToHost, ctx->stream));
    TIPLHIP_CHECK(hipStreamSynchronize(ctx->stream));
    std::memcpy(host, ctx->h_out.p, bytes);
}

tiplhip_ctx* aux_context(tiplhip_ctx* ctx)
{
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    if (!ctx->aux) {
        tiplhip_ctx* a = nullptr;
        if (tiplhip_ctx_create(ctx->device, &a) != TIPLHIP_OK)
            throw Error(TIPLHIP_ERR_DEVICE, tiplhip_last_error());
        a->cdm_batch = ctx->cdm_batch;
        ctx->aux = a;
    }
    return ctx->aux;
}

void volume_invalidate(tiplhip_ctx* ctx, tiplhip_volume* vol)
{
    // the packed copy is rebuilt on its next use; the old one is released
    // once ctx's stream (which wrote the new voxels) has passed this point
    {
        std::lock_guard<std::mutex> lk(vol->pack_mu);
        if (vol->d_pack) {
            TIPLHIP_CHECK(hipStreamSynchronize(ctx->stream));
            pool_free(vol->ctx, vol->d_pack);
            vol->d_pack = nullptr;
        }
    }
    volume_compute_stats(ctx, vol);
}

void volume_compute_stats(tiplhip_ctx* ctx, tiplhip_volume* vol)
{
    ctx->d_tmp2.reserve(sizeof(unsigned long long) * 2 + sizeof(uint32_t) * 256);
    auto* sums = ctx->d_tmp2.as<unsigned long long>();
    auto* hist = reinterpret_cast<uint32_t*>(sums + 2);
    TIPLHIP_CHECK(hipMemsetAsync(sums, 0, sizeof(unsigned long long) * 2 + sizeof(uint32_t) * 256,
                                 ctx->stream));
    hipLaunchKernelGGL(volume_stats_kernel, dim3(grid_for(vol->n, 2048)), dim3(256), 0, ctx->stream,
                       vol->d, vol->n, sums, hist);
    TIPLHIP_CHECK(hipGetLastError());
    struct {
        unsigned long long s[2];
        uint32_t h[256];
    } host;
    read_back(ctx, sums, &host, sizeof(host));
    vol->sum = int64_t(host.s[0]);
    vol->sum_sq = int64_t(host.s[1]);
    std::copy(host.h, host.h + 256, vol->hist256);
    for (int i = 0; i < 64; ++i)
        vol->hist64[i] = host.h[4 * i] + host.h[4 * i + 1] + host.h[4 * i + 2] + host.h[4 * i + 3];
    uint32_t all[256 + 64];
    std::copy(vol->hist256, vol->hist256 + 256, all);
    std::copy(vol->hist64, vol->hist64 + 64, all + 256);
    TIPLHIP_CHECK(hipMemcpyAsync(vol->d_hist, all, sizeof(all), hipMemcpyHostToDevice, ctx->stream));
    TIPLHIP_CHECK(hipStreamSynchronize(ctx->stream));
}

namespace {

template <class T>
void quantise(tiplhip_ctx* ctx, const T* d_in, size_t n, float upper, uint8_t* d_out)
{
    ctx->d_tmp2.reserve(sizeof(int) * 2);
    int init[2] = {0x7fffffff, int(0x80000000)};
    TIPLHIP_CHECK(hipMemcpyAsync(ctx->d_tmp2.p, init, sizeof(init), hipMemcpyHostToDevice, ctx->stream));
    hipLaunchKernelGGL(minmax_kernel<T>, dim3(grid_for(n, 2048)), dim3(256), 0, ctx->stream, d_in, n,
                       ctx->d_tmp2.as<int>());
    TIPLHIP_CHECK(hipGetLastError());
    int mm[2];
    read_back(ctx, ctx->d_tmp2.p, mm, sizeof(mm));
    float mn, mx;
    if constexpr (std::is_same<T, float>::value) {
        auto unkey = [](int k) {
            int b = k >= 0 ? k : (k ^ 0x7fffffff);
            float f;
            std::memcpy(&f, &b, 4);
            return f;
        };
        mn = unkey(mm[0]);
        mx = unkey(mm[1]);
    } else {
        mn = float(mm[0]);
        mx = float(mm[1]);
    }
    TIPLHIP_CHECK(hipMemsetAsync(d_out, 0, n, ctx->stream));
    const float range = mx - mn;
    if (range == 0.0f) return;  // the reference leaves the output untouched
    const float coef = upper / range;
    hipLaunchKernelGGL(normalize_kernel<T>, dim3(grid_for(n)), dim3(256), 0, ctx->stream, d_in, n,
                       mn, coef, d_out);
    TIPLHIP_CHECK(hipGetLastError());
}

tiplhip_volume* volume_create(tiplhip_ctx* ctx, const void* src, bool src_on_device, int dtype,
                              const int* dim, float upper)
{
    TIPLHIP_REQUIRE(ctx != nullptr, "null context");
    TIPLHIP_REQUIRE(src != nullptr, "null volume data");
    TIPLHIP_REQUIRE(dtype == TIPLHIP_U8 || dtype == TIPLHIP_I16 || dtype == TIPLHIP_F32,
                    "unknown dtype");
    DeviceGuard dg(ctx->device);
    tiplhip_volume* v = volume_alloc(ctx, dim);
    try {
        const size_t n = v->n;
        const hipMemcpyKind kind = src_on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
        if (dtype == TIPLHIP_U8) {
            TIPLHIP_CHECK(hipMemcpyAsync(v->d, src, n, kind, ctx->stream));
        } else {
            const size_t es = dtype == TIPLHIP_I16 ? 2 : 4;
            const void* d_in = src;
            if (!src_on_device) {
                ctx->d_tmp.reserve(n * es);
                TIPLHIP_CHECK(hipMemcpyAsync(ctx->d_tmp.p, src, n * es, kind, ctx->stream));
                d_in = ctx->d_tmp.p;
            }
            if (dtype == TIPLHIP_I16)
                quantise(ctx, static_cast<const int16_t*>(d_in), n, upper, v->d);
            else
                quantise(ctx, static_cast<const float*>(d_in), n, upper, v->d);
        }
        volume_compute_stats(ctx, v);
    } catch (...) {
        volume_release(v);
        throw;
    }
    return v;
}

}  // namespace
}  // namespace tiplhip

using namespace tiplhip;

extern "C" {

const char* tiplhip_last_error(void) { return g_last_error.c_str(); }

int tiplhip_device_count(int* count)
{
    return guarded([&] {
        TIPLHIP_REQUIRE(count != nullptr, "null count");
        int n = 0;
        hipError_t e = hipGetDeviceCount(&n);
        *count = (e == hipSuccess) ? n : 0;
    });
}

int tiplhip_ctx_create(int device, tiplhip_ctx** out)
{
    return guarded([&] {
        TIPLHIP_REQUIRE(out != nullptr, "null output");
        *out = nullptr;
        int n = 0;
        if (hipGetDeviceCount(&n) != hipSuccess || n == 0)
            throw Error(TIPLHIP_ERR_NO_DEVICE, "no HIP device available");
        TIPLHIP_REQUIRE(device >= 0 && device < n, "device index out of range");
        auto* c = new tiplhip_ctx();
        c->device = device;
        // costs written by kernels straight into host memory: uncached on the
        // device side, so a finished kernel's writes are already visible
        c->h_costs.flags = hipHostMallocMapped | hipHostMallocCoherent;
        c->h_cdm.flags = hipHostMallocMapped | hipHostMallocCoherent;
        try {
            DeviceGuard dg(device);
            hipDeviceProp_t prop;
            TIPLHIP_CHECK(hipGetDeviceProperties(&prop, device));
            c->cu_count = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
            TIPLHIP_CHECK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
            TIPLHIP_CHECK(hipEventCreate(&c->t0));
            TIPLHIP_CHECK(hipEventCreate(&c->t1));
            TIPLHIP_CHECK(hipEventCreateWithFlags(&c->e_read, hipEventDisableTiming));
        } catch (...) {
            if (c->stream) (void)hipStreamDestroy(c->stream);
            delete c;
            throw;
        }
        *out = c;
    });
}

int tiplhip_ctx_destroy(tiplhip_ctx* ctx)
{
    return guarded([&] {
        if (!ctx) return;
        if (ctx->aux) (void)tiplhip_ctx_destroy(ctx->aux);
        {
            DeviceGuard dg(ctx->device);
            (void)hipStreamSynchronize(ctx->stream);
            for (hipEvent_t e : {ctx->t0, ctx->t1, ctx->e_read})
                if (e) (void)hipEventDestroy(e);
            for (hipEvent_t e : ctx->kt_events) (void)hipEventDestroy(e);
            ctx->kt_events.clear();
            clear_graphs(ctx);
            ctx->d_T.release();
            ctx->d_hist.release();
            ctx->d_sums.release();
            ctx->d_costs.release();
            ctx->d_ztab.release();
            ctx->d_tmp.release();
            ctx->d_tmp2.release();
            ctx->h_T.release();
            ctx->h_out.release();
            ctx->h_costs.release();
            ctx->h_cdm.release();
            (void)hipStreamDestroy(ctx->stream);
        }
        delete ctx;
    });
}

int tiplhip_ctx_synchronize(tiplhip_ctx* ctx)
{
    return guarded([&] {
        TIPLHIP_REQUIRE(ctx != nullptr, "null context");
        DeviceGuard dg(ctx->device);
        TIPLHIP_CHECK(hipStreamSynchronize(ctx->stream));
    });
}

int tiplhip_ctx_set_graphs(tiplhip_ctx* ctx, int enable)
{
    return guarded([&] {
        TIPLHIP_REQUIRE(ctx != nullptr, "null context");
        std::lock_guard<std::recursive_mutex> lk(ctx->mu);
        DeviceGuard dg(ctx->device);
        ctx->use_graphs = enable != 0;
        if (!ctx->use_graphs) {
            TIPLHIP_CHECK(hipStreamSynchronize(ctx->stream));
            clear_graphs(ctx);
        }
    });
}

int tiplhip_ctx_set_kernel_timing(tiplhip_ctx* ctx, int enable)
{
    return guarded([&] {
        TIPLHIP_REQUIRE(ctx != nullptr, "null context");
        std::lock_guard<std::recursive_mutex> lk(ctx->mu);
        ctx->kernel_timing = enable != 0;
        ctx->kt_pending = 0;  // pairs not yet read are dropped with the totals
        ctx->kernel_ms = 0.0;
        ctx->kernel_launches = 0;
    });
}

int tiplhip_ctx_kernel_time(tiplhip_ctx* ctx, double* total_ms, int64_t* launches)
{
    return guarded([&] {
        TIPLHIP_REQUIRE(ctx != nullptr, "null context");
        std::lock_guard<std::recursive_mutex> lk(ctx->mu);
        DeviceGuard dg(ctx->device);
        timing_flush(ctx);
        if (total_ms) *total_ms = ctx->kernel_ms;
        if (launches) *launches = ctx->kernel_launches;
    });
}

int tiplhip_timer_start(tiplhip_ctx* ctx)
{
    return guarded([&] {
        TIPLHIP_REQUIRE(ctx != nullptr, "null context");
        DeviceGuard dg(ctx->device);
        TIPLHIP_CHECK(hipEventRecord(ctx->t0, ctx->stream));
    });
}

int tiplhip_timer_stop(tiplhip_ctx* ctx, float* ms)
{
    return guarded([&] {
        TIPLHIP_REQUIRE(ctx != nullptr && ms != nullptr, "null argument");
        DeviceGuard dg(ctx->device);
        TIPLHIP_CHECK(hipEventRecord(ctx->t1, ctx->stream));
        TIPLHIP_CHECK(hipEventSynchronize(ctx->t1));
        TIPLHIP_CHECK(hipEventElapsedTime(ms, ctx->t0, ctx->t1));
    });
}

int tiplhip_traffic_probe(tiplhip_ctx* ctx, int kind, size_t bytes, int reps, double* gbps)
{
    return guarded([&] {
        TIPLHIP_REQUIRE(ctx && gbps, "null argument");
        TIPLHIP_REQUIRE(kind >= 0 && kind <= 2, "kind must be 0 (copy), 1 or 2 (reads)");
        TIPLHIP_REQUIRE(bytes >= 192 && reps >= 1, "need bytes >= 192 and reps >= 1");
        std::lock_guard<std::recursive_mutex> lk(ctx->mu);
        DeviceGuard dg(ctx->device);
        // whole 48-byte groups (four 12-byte voxels = three 16-byte units)
        const size_t n = bytes / 48 * 3;
        // about 8 blocks per CU, each a contiguous chunk (a multiple of 3 units)
        const size_t blocks = std::max<size_t>(1, std::min<size_t>(size_t(ctx->cu_count) * 8,
                                                                   n / (3 * kProbeThreads)));
        const size_t per_block = (n / blocks + 2) / 3 * 3;
        const size_t nb = (n + per_block - 1) / per_block;
        void *a = nullptr, *b = nullptr;
        TIPLHIP_CHECK(hipMalloc(&a, n * 16));
        hipError_t e = hipMalloc(&b, kind == 0 ? n * 16 : nb * sizeof(uint32_t));
        if (e != hipSuccess) {
            (void)hipFree(a);
            TIPLHIP_CHECK(e);
        }
        auto launch = [&] {
            if (kind == 0)
                hipLaunchKernelGGL(stream_copy_kernel, dim3(uint32_t(nb)), dim3(kProbeThreads), 0, ctx->stream,
                                   static_cast<const u32x4*>(a), static_cast<u32x4*>(b), n, per_block);
            else if (kind == 1)
                hipLaunchKernelGGL(stream_read_kernel<1>, dim3(uint32_t(nb)), dim3(kProbeThreads), 0,
                                   ctx->stream, static_cast<const u32x4*>(a), n, per_block,
                                   static_cast<uint32_t*>(b));
            else
                hipLaunchKernelGGL(stream_read_kernel<2>, dim3(uint32_t(nb)), dim3(kProbeThreads), 0,
                                   ctx->stream, static_cast<const u32x4*>(a), n, per_block,
                                   static_cast<uint32_t*>(b));
        };
        e = hipMemsetAsync(a, 1, n * 16, ctx->stream);
        if (e == hipSuccess && kind != 0) e = hipMemsetAsync(b, 0, nb * sizeof(uint32_t), ctx->stream);
        if (e != hipSuccess) {
            (void)hipFree(a);
            (void)hipFree(b);
            TIPLHIP_CHECK(e);
        }
        launch();  // warm up
        e = hipGetLastError();
        if (e == hipSuccess) e = hipEventRecord(ctx->t0, ctx->stream);
        for (int r = 0; r < reps && e == hipSuccess; ++r) launch();
        if (e == hipSuccess) e = hipEventRecord(ctx->t1, ctx->stream);
        if (e == hipSuccess) e = hipEventSynchronize(ctx->t1);
        float ms = 0.0f;
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, ctx->t0, ctx->t1);
        (void)hipStreamSynchronize(ctx->stream);
        (void)hipFree(a);
        (void)hipFree(b);
        TIPLHIP_CHECK(e);
        *gbps = double(kind == 0 ? 2 : 1) * double(n * 16) * reps / (double(ms) * 1e-3) / 1e9;
    });
}

int tiplhip_stream_copy(tiplhip_ctx* ctx, size_t bytes, int reps, double* gbps)
{
    return tiplhip_traffic_probe(ctx, 0, bytes, reps, gbps);
}

int tiplhip_volume_create(tiplhip_ctx* ctx, const void* host, int dtype, const int dim[3],
                          float upper_limit, tiplhip_volume** out)
{
    return guarded([&] {
        TIPLHIP_REQUIRE(out != nullptr, "null output");
        *out = nullptr;
        TIPLHIP_REQUIRE(ctx != nullptr, "null context");
        std::lock_guard<std::recursive_mutex> lk(ctx->mu);
        *out = volume_create(ctx, host, false, dtype, dim, upper_limit);
    });
}

int tiplhip_volume_create_device(tiplhip_ctx* ctx, const void* device_ptr, int dtype,
                                 const int dim[3], float upper_limit, tiplhip_volume** out)
{
    return guarded([&] {
        TIPLHIP_REQUIRE(out != nullptr, "null output");
        *out = nullptr;
        TIPLHIP_REQUIRE(ctx != nullptr, "null context");
        std::lock_guard<std::recursive_mutex> lk(ctx->mu);
        *out = volume_create(ctx, device_ptr, true, dtype, dim, upper_limit);
    });
}

int tiplhip_volume_destroy(tiplhip_volume* vol)
{
    return guarded([&] {
        if (!vol) return;
        tiplhip_ctx* ctx = vol->ctx;
        std::lock_guard<std::recursive_mutex> lk(ctx->mu);
        DeviceGuard dg(ctx->device);
        // stream-ordered: the memory is reused only after the owner's
        // queued work has finished; no device-wide synchronisation
        volume_release(vol);
    });
}

int tiplhip_volume_dims(const tiplhip_volume* vol, int dim[3])
{
    return guarded([&] {
        TIPLHIP_REQUIRE(vol && dim, "null argument");
        std::copy(vol->dim, vol->dim + 3, dim);
    });
}

int tiplhip_volume_read(tiplhip_ctx* ctx, const tiplhip_volume* vol, uint8_t* host)
{
    return guarded([&] {
        TIPLHIP_REQUIRE(ctx && vol && host, "null argument");
        TIPLHIP_REQUIRE(vol->ctx == ctx, "volume belongs to another context");
        std::lock_guard<std::recursive_mutex> lk(ctx->mu);
        DeviceGuard dg(ctx->device);
        TIPLHIP_CHECK(hipMemcpyAsync(host, vol->d, vol->n, hipMemcpyDeviceToHost, ctx->stream));
        TIPLHIP_CHECK(hipStreamSynchronize(ctx->stream));
    });
}

int tiplhip_volume_stats(const tiplhip_volume* vol, int64_t* sum, int64_t* sum_sq,
                         uint32_t hist256[256])
{
    return guarded([&] {
        TIPLHIP_REQUIRE(vol != nullptr, "null volume");
        if (sum) *sum = vol->sum;
        if (sum_sq) *sum_sq = vol->sum_sq;
        if (hist256) std::copy(vol->hist256, vol->hist256 + 256, hist256);
    });
}

int tiplhip_volume_downsample(tiplhip_ctx* ctx, const tiplhip_volume* in, tiplhip_volume** out)
{
    return guarded([&] {
        TIPLHIP_REQUIRE(out != nullptr, "null output");
        *out = nullptr;
        TIPLHIP_REQUIRE(ctx && in, "null argument");
        TIPLHIP_REQUIRE(in->ctx && in->ctx->device == ctx->device, "volume lives on another device");
        std::lock_guard<std::recursive_mutex> lk(ctx->mu);
        DeviceGuard dg(ctx->device);
        int od[3] = {(in->dim[0] + 1) / 2, (in->dim[1] + 1) / 2, (in->dim[2] + 1) / 2};
        tiplhip_volume* v = volume_alloc(ctx, od);
        try {
            launch_downsample_u8(ctx, in->d, in->dim, v->d);
            volume_compute_stats(ctx, v);
        } catch (...) {
            volume_release(v);
            throw;
        }
        *out = v;
    });
}

int tiplhip_affine_matrix(const float param[12], const int from_dim[3], const float from_vs[3],
                          const int to_dim[3], const float to_vs[3], float T[12])
{
    return guarded([&] {
        TIPLHIP_REQUIRE(param && from_dim && from_vs && to_dim && to_vs && T, "null argument");
        affine_matrix(param, from_dim, from_vs, to_dim, to_vs, T);
    });
}

int tiplhip_matrix_inverse(float T[12])
{
    return guarded([&] {
        TIPLHIP_REQUIRE(T != nullptr, "null argument");
        if (!matrix_inverse(T)) throw Error(TIPLHIP_ERR_INVALID, "singular transform");
    });
}

}  // extern "C"
