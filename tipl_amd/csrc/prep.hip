// prep.hip -- subject preprocessing on gfx950: reg::subject_image_pre
// (reg/mm_reg.hpp:14-20) = filter::gaussian(I) (bandwidth 1), then
// segmentation::normalize_otsu_median(I, 255.99f) (numerical/otsu.hpp:96-136),
// then the conversion to uint8.
//
// Steps (one pass each over the float image in HBM):
//   min / max                  per-block partials, final on the host
//   256-bin histogram          LDS, exact counts (numerical/basic_op.hpp:1294-1311)
//   Otsu threshold             on the host: 256 floats, the reference's float scan
//   values > threshold         per-block counts, host prefix, order-preserving
//                              compaction (the reference pushes them in index order)
//   median                     tipl::median (numerical/statistics.hpp:327-363):
//                              below 2048 values rank n/2, otherwise rank 128 of the
//                              256 contiguous chunks' rank-len/2 elements; each rank
//                              by an in-workgroup radix select on order-preserving keys
//   scale, offset, clamp, u8   one fused pass (plus one count of positives)
#include "device.hpp"
#include "internal.hpp"

#include <algorithm>
#include <cmath>
#include <numeric>
#include <vector>

namespace tiplhip {
namespace {

constexpr int kPB = 256;         // threads per block of the streaming passes
constexpr int kPItems = 16;      // items per thread
constexpr int kPSpan = kPB * kPItems;

__global__ __launch_bounds__(kPB) void to_float_kernel(const void* __restrict__ in, int dtype,
                                                       size_t n, float* __restrict__ out)
{
    for (size_t i = size_t(blockIdx.x) * kPB + threadIdx.x; i < n; i += size_t(gridDim.x) * kPB) {
        float v;
        if (dtype == TIPLHIP_I16) v = float(static_cast<const int16_t*>(in)[i]);
        else if (dtype == TIPLHIP_U8) v = float(static_cast<const uint8_t*>(in)[i]);
        else v = static_cast<const float*>(in)[i];
        out[i] = v;
    }
}

__global__ __launch_bounds__(kPB) void minmax_part_kernel(const float* __restrict__ v, size_t n,
                                                          float2* __restrict__ parts)
{
    float mn = INFINITY, mx = -INFINITY;
    const size_t b0 = size_t(blockIdx.x) * kPSpan;
    for (int k = 0; k < kPItems; ++k) {
        const size_t i = b0 + size_t(k) * kPB + threadIdx.x;
        if (i < n) {
            mn = fminf(mn, v[i]);
            mx = fmaxf(mx, v[i]);
        }
    }
    for (int off = 32; off > 0; off >>= 1) {
        mn = fminf(mn, __shfl_xor(mn, off, kWave));
        mx = fmaxf(mx, __shfl_xor(mx, off, kWave));
    }
    __shared__ float2 w[kPB / 64];
    if ((threadIdx.x & 63) == 0) w[threadIdx.x >> 6] = make_float2(mn, mx);
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int i = 1; i < kPB / 64; ++i) {
            mn = fminf(mn, w[i].x);
            mx = fmaxf(mx, w[i].y);
        }
        parts[blockIdx.x] = make_float2(fminf(mn, w[0].x), fmaxf(mx, w[0].y));
    }
}

// histogram(src, hist, min, max) with 256 bins: bin = clamp(floor((v-min)*r), 0, 255)
__global__ __launch_bounds__(kPB) void hist256_kernel(const float* __restrict__ v, size_t n,
                                                      float mn, float r,
                                                      unsigned int* __restrict__ hist)
{
    __shared__ unsigned int h[256];
    h[threadIdx.x] = 0u;
    __syncthreads();
    const size_t b0 = size_t(blockIdx.x) * kPSpan;
    for (int k = 0; k < kPItems; ++k) {
        const size_t i = b0 + size_t(k) * kPB + threadIdx.x;
        if (i < n) {
            const float value = (v[i] - mn) * r;
            const int b = min(max(int(floorf(value)), 0), 255);
            atomicAdd(&h[b], 1u);
        }
    }
    __syncthreads();
    if (h[threadIdx.x]) atomicAdd(&hist[threadIdx.x], h[threadIdx.x]);
}

// Per block: how many of its values are > t (mode 0), or have v * t > 0 (mode 1).
__global__ __launch_bounds__(kPB) void count_kernel(const float* __restrict__ v, size_t n, float t,
                                                    int mode, unsigned int* __restrict__ counts)
{
    unsigned int c = 0;
    const size_t b0 = size_t(blockIdx.x) * kPSpan + size_t(threadIdx.x) * kPItems;
    for (int k = 0; k < kPItems; ++k) {
        const size_t i = b0 + size_t(k);
        if (i < n) c += mode == 0 ? (v[i] > t ? 1u : 0u) : (v[i] * t > 0.0f ? 1u : 0u);
    }
    for (int off = 32; off > 0; off >>= 1) c += __shfl_xor(c, off, kWave);
    __shared__ unsigned int w[kPB / 64];
    if ((threadIdx.x & 63) == 0) w[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned int s = 0;
        for (int i = 0; i < kPB / 64; ++i) s += w[i];
        counts[blockIdx.x] = s;
    }
}

// Values > t in index order: thread k of block b owns items
// [b*kPSpan + k*kPItems, +kPItems); block offsets come from the host prefix.
__global__ __launch_bounds__(kPB) void compact_kernel(const float* __restrict__ v, size_t n, float t,
                                                      const unsigned int* __restrict__ offsets,
                                                      float* __restrict__ out)
{
    const size_t b0 = size_t(blockIdx.x) * kPSpan + size_t(threadIdx.x) * kPItems;
    unsigned int c = 0;
    for (int k = 0; k < kPItems; ++k) {
        const size_t i = b0 + size_t(k);
        if (i < n && v[i] > t) ++c;
    }
    __shared__ unsigned int pre[kPB];
    pre[threadIdx.x] = c;
    __syncthreads();
    for (int off = 1; off < kPB; off <<= 1) {  // inclusive Hillis-Steele scan
        const unsigned int add = threadIdx.x >= unsigned(off) ? pre[threadIdx.x - off] : 0u;
        __syncthreads();
        pre[threadIdx.x] += add;
        __syncthreads();
    }
    unsigned int pos = offsets[blockIdx.x] + pre[threadIdx.x] - c;
    for (int k = 0; k < kPItems; ++k) {
        const size_t i = b0 + size_t(k);
        if (i < n && v[i] > t) out[pos++] = v[i];
    }
}

// Order-preserving key of a float (non-NaN): larger float, larger key.
__device__ inline uint32_t fkey(float f)
{
    const uint32_t b = __float_as_uint(f);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__device__ inline float fkey_inv(uint32_t k)
{
    return __uint_as_float((k & 0x80000000u) ? (k & 0x7fffffffu) : ~k);
}

// Block s: the element of rank len/2 of buf[start_s, start_s + len_s), with
// chunks laid out as tipl::median's (nchunks = 256 or 1 for the whole).
__global__ __launch_bounds__(1024) void chunk_select_kernel(const float* __restrict__ buf, size_t m,
                                                            int nchunks, float* __restrict__ out)
{
    const size_t cs = m / size_t(nchunks), rem = m % size_t(nchunks), i = blockIdx.x;
    const size_t start = i * cs + min(i, rem);
    const size_t len = cs + (i < rem ? 1 : 0);
    __shared__ unsigned int h[256];
    __shared__ uint32_t s_prefix, s_rank;
    uint32_t prefix = 0, mask = 0;
    if (threadIdx.x == 0) s_rank = uint32_t(len / 2);
    for (int shift = 24; shift >= 0; shift -= 8) {
        if (threadIdx.x < 256) h[threadIdx.x] = 0u;
        __syncthreads();
        for (size_t j = threadIdx.x; j < len; j += blockDim.x) {
            const uint32_t k = fkey(buf[start + j]);
            if ((k & mask) == prefix) atomicAdd(&h[(k >> shift) & 255u], 1u);
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            uint32_t r = s_rank, cum = 0, sel = 255;
            for (uint32_t b = 0; b < 256; ++b) {
                if (cum + h[b] > r) {
                    sel = b;
                    break;
                }
                cum += h[b];
            }
            s_rank = r - cum;
            s_prefix = prefix | (sel << shift);
        }
        __syncthreads();
        prefix = s_prefix;
        mask |= 255u << shift;
    }
    if (threadIdx.x == 0) out[blockIdx.x] = fkey_inv(prefix);
}

// v * c (then - s when sub), clamped to [0, upper], truncated to uint8; or,
// when the median is 0 (no normalisation), the plain truncating conversion.
__global__ __launch_bounds__(kPB) void finalize_kernel(const float* __restrict__ v, size_t n,
                                                       int normalise, float c, int sub, float s,
                                                       float upper, uint8_t* __restrict__ out)
{
    for (size_t i = size_t(blockIdx.x) * kPB + threadIdx.x; i < n; i += size_t(gridDim.x) * kPB) {
        float t = v[i];
        if (normalise) {
            t = t * c;
            if (sub) t = t - s;
            t = t > upper ? upper : (t < 0.0f ? 0.0f : t);
        }
        out[i] = uint8_t(uint32_t(int32_t(t)));
    }
}

uint32_t span_blocks(size_t n) { return uint32_t((n + kPSpan - 1) / kPSpan); }
uint32_t stream_blocks(size_t n) { return uint32_t(std::min<size_t>((n + kPB - 1) / kPB, 16384)); }

template <class T>
void read_sync(tiplhip_ctx* ctx, const void* dev, std::vector<T>& host)
{
    read_back(ctx, dev, host.data(), sizeof(T) * host.size());
}

// otsu_threshold_from_hist (numerical/otsu.hpp:15-65), the float scan.
float otsu_from_hist(std::vector<unsigned int> hist, float mn, float mx)
{
    hist[0] = 0;  // skip = 1 (numerical/otsu.hpp:78)
    float total_w = 0.0f, total_sum = 0.0f;
    for (size_t i = 0; i < hist.size(); ++i) {
        const float val = float(hist[i]);
        total_w += val;
        total_sum += val * float(i);
    }
    if (total_w == 0.0f) return mn;
    float max_sig_b = 0.0f, w_b = 0.0f, sum_b = 0.0f;
    size_t opt = 0;
    for (size_t i = 0; i + 1 < hist.size(); ++i) {
        w_b += float(hist[i]);
        if (w_b == 0.0f) continue;
        const float w_f = total_w - w_b;
        if (w_f == 0.0f) break;
        sum_b += float(hist[i]) * float(i);
        const float d = (sum_b * total_w) - (total_sum * w_b);
        const float sig_b = (d * d) / (w_b * w_f);
        if (sig_b > max_sig_b) {
            max_sig_b = sig_b;
            opt = i;
        }
    }
    return mn + float(opt) * (mx - mn) / float(hist.size());
}

}  // namespace

void subject_preprocess(tiplhip_ctx* ctx, const float* d_in, const int* dim, uint8_t* d_out)
{
    const size_t n = vox(dim);
    TIPLHIP_REQUIRE(n < (size_t(1) << 31), "volume exceeds 2^31 voxels");
    const size_t fb = sizeof(float) * n;
    void* g = pool_alloc(ctx, fb);   // the smoothed image
    void* t = pool_alloc(ctx, fb);   // gaussian scratch, then the compacted values
    struct Release {
        tiplhip_ctx* c;
        void* a;
        void* b;
        ~Release()
        {
            pool_free(c, a);
            pool_free(c, b);
        }
    } release{ctx, g, t};
    float* I = static_cast<float*>(g);
    launch_gaussian_to(ctx, d_in, I, static_cast<float*>(t), dim, 1, 1);

    const uint32_t nb = span_blocks(n);
    ctx->d_tmp2.reserve(std::max<size_t>(sizeof(float2), sizeof(unsigned int)) * (nb + 512));
    // min / max
    hipLaunchKernelGGL(minmax_part_kernel, dim3(nb), dim3(kPB), 0, ctx->stream, I, n,
                       ctx->d_tmp2.as<float2>());
    TIPLHIP_CHECK(hipGetLastError());
    std::vector<float2> mm(nb);
    read_sync(ctx, ctx->d_tmp2.p, mm);
    float mn = mm[0].x, mx = mm[0].y;
    for (const float2& p : mm) {
        mn = std::min(mn, p.x);
        mx = std::max(mx, p.y);
    }
    // Otsu threshold
    float ot = mn;
    if (mn < mx) {
        float range = mx - mn;
        if (range == 0.0f) range = 1.0f;
        const float r = 256.0f / range;
        unsigned int* dh = ctx->d_tmp2.as<unsigned int>();
        TIPLHIP_CHECK(hipMemsetAsync(dh, 0, sizeof(unsigned int) * 256, ctx->stream));
        hipLaunchKernelGGL(hist256_kernel, dim3(nb), dim3(kPB), 0, ctx->stream, I, n, mn, r, dh);
        TIPLHIP_CHECK(hipGetLastError());
        std::vector<unsigned int> hist(256);
        read_sync(ctx, dh, hist);
        ot = otsu_from_hist(hist, mn, mx);
    }
    // values > ot, compacted in index order
    unsigned int* counts = ctx->d_tmp2.as<unsigned int>();
    hipLaunchKernelGGL(count_kernel, dim3(nb), dim3(kPB), 0, ctx->stream, I, n, ot, 0, counts);
    TIPLHIP_CHECK(hipGetLastError());
    std::vector<unsigned int> cnt(nb);
    read_sync(ctx, counts, cnt);
    std::vector<unsigned int> off(nb);
    size_t m = 0;
    for (uint32_t b = 0; b < nb; ++b) {
        off[b] = unsigned(m);
        m += cnt[b];
    }
    float med = ot;
    if (m > 0) {
        TIPLHIP_CHECK(hipMemcpyAsync(counts, off.data(), sizeof(unsigned int) * nb,
                                     hipMemcpyHostToDevice, ctx->stream));
        float* buf = static_cast<float*>(t);
        hipLaunchKernelGGL(compact_kernel, dim3(nb), dim3(kPB), 0, ctx->stream, I, n, ot, counts, buf);
        TIPLHIP_CHECK(hipGetLastError());
        const int nchunks = m < 2048 ? 1 : 256;
        float* sel = reinterpret_cast<float*>(counts);  // nb >= 1 slots; 256 more are reserved
        hipLaunchKernelGGL(chunk_select_kernel, dim3(nchunks), dim3(1024), 0, ctx->stream, buf, m,
                           nchunks, sel);
        TIPLHIP_CHECK(hipGetLastError());
        std::vector<float> meds(static_cast<size_t>(nchunks));
        read_sync(ctx, sel, meds);
        if (nchunks == 1) {
            med = meds[0];
        } else {
            std::nth_element(meds.begin(), meds.begin() + 128, meds.end());
            med = meds[128];
        }
    }
    const float upper = 255.99f;
    int normalise = 0, sub = 0;
    float c = 0.0f;
    const float s = upper * 0.1f;
    if (med != 0.0f) {
        normalise = 1;
        c = upper * 0.5f / med;
        hipLaunchKernelGGL(count_kernel, dim3(nb), dim3(kPB), 0, ctx->stream, I, n, c, 1, counts);
        TIPLHIP_CHECK(hipGetLastError());
        read_sync(ctx, counts, cnt);
        const size_t pos = std::accumulate(cnt.begin(), cnt.end(), size_t(0));
        sub = double(pos) > 0.5 * double(n) ? 1 : 0;
    }
    hipLaunchKernelGGL(finalize_kernel, dim3(stream_blocks(n)), dim3(kPB), 0, ctx->stream, I, n,
                       normalise, c, sub, s, upper, d_out);
    TIPLHIP_CHECK(hipGetLastError());
}

}  // namespace tiplhip

using namespace tiplhip;

extern "C" {

int tiplhip_volume_create_subject(tiplhip_ctx* ctx, const void* host, int dtype, const int dim[3],
                                  tiplhip_volume** out)
{
    return guarded([&] {
        TIPLHIP_REQUIRE(out != nullptr, "null output");
        *out = nullptr;
        TIPLHIP_REQUIRE(ctx && host && dim, "null argument");
        TIPLHIP_REQUIRE(dtype == TIPLHIP_U8 || dtype == TIPLHIP_I16 || dtype == TIPLHIP_F32,
                        "unknown dtype");
        std::lock_guard<std::recursive_mutex> lk(ctx->mu);
        DeviceGuard dg(ctx->device);
        tiplhip_volume* v = volume_alloc(ctx, dim);
        try {
            const size_t n = v->n;
            const size_t es = dtype == TIPLHIP_F32 ? 4 : dtype == TIPLHIP_I16 ? 2 : 1;
            void* raw = pool_alloc(ctx, n * es);
            void* f = pool_alloc(ctx, sizeof(float) * n);
            struct Release {
                tiplhip_ctx* c;
                void* a;
                void* b;
                ~Release()
                {
                    pool_free(c, a);
                    pool_free(c, b);
                }
            } release{ctx, raw, f};
            TIPLHIP_CHECK(hipMemcpyAsync(raw, host, n * es, hipMemcpyHostToDevice, ctx->stream));
            hipLaunchKernelGGL(to_float_kernel, dim3(stream_blocks(n)), dim3(kPB), 0, ctx->stream,
                               raw, dtype, n, static_cast<float*>(f));
            TIPLHIP_CHECK(hipGetLastError());
            subject_preprocess(ctx, static_cast<const float*>(f), dim, v->d);
            volume_compute_stats(ctx, v);
        } catch (...) {
            volume_release(v);
            throw;
        }
        *out = v;
    });
}

}  // extern "C"
