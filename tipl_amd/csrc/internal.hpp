// internal.hpp -- shared host/device definitions of libtiplhip (gfx950 only).
#pragma once

#include <hip/hip_runtime.h>

#include <array>
#include <cstdint>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/tiplhip.h"
#include "errors.hpp"

// Every FP32 operation on the bit-exact paths must be a separately rounded
// IEEE operation in source order (the reference's uint8 truncation after
// trilinear interpolation depends on it).  The build also passes
// -ffp-contract=off; the pragma keeps that true for anyone compiling these
// sources by hand.
#pragma clang fp contract(off)

namespace tiplhip {

class CdmBatcher;

#define TIPLHIP_CHECK(expr)                                                                      \
    do {                                                                                         \
        hipError_t e_ = (expr);                                                                  \
        if (e_ != hipSuccess)                                                                    \
            throw ::tiplhip::Error(e_ == hipErrorOutOfMemory ? TIPLHIP_ERR_MEMORY                \
                                                             : TIPLHIP_ERR_DEVICE,               \
                                   std::string(#expr) + ": " + hipGetErrorString(e_));           \
    } while (0)

// Device allocation that only grows.
struct DeviceBuffer {
    void* p = nullptr;
    size_t bytes = 0;
    DeviceBuffer() = default;
    DeviceBuffer(const DeviceBuffer&) = delete;
    DeviceBuffer& operator=(const DeviceBuffer&) = delete;
    ~DeviceBuffer() { release(); }
    void reserve(size_t n)
    {
        if (n <= bytes) return;
        release();
        TIPLHIP_CHECK(hipMalloc(&p, n));
        bytes = n;
    }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    template <class T>
    T* as() const { return static_cast<T*>(p); }
};

// Pinned host allocation that only grows.
struct PinnedBuffer {
    void* p = nullptr;
    void* dev = nullptr;  // the same memory as kernels address it
    size_t bytes = 0;
    unsigned flags = hipHostMallocDefault;
    PinnedBuffer() = default;
    PinnedBuffer(const PinnedBuffer&) = delete;
    PinnedBuffer& operator=(const PinnedBuffer&) = delete;
    ~PinnedBuffer() { release(); }
    void reserve(size_t n)
    {
        if (n <= bytes) return;
        release();
        TIPLHIP_CHECK(hipHostMalloc(&p, n, flags));
        bytes = n;
        TIPLHIP_CHECK(hipHostGetDevicePointer(&dev, p, 0));
    }
    void release()
    {
        if (p) (void)hipHostFree(p);
        p = nullptr;
        dev = nullptr;
        bytes = 0;
    }
    template <class T>
    T* as() const { return static_cast<T*>(p); }
};

// Makes ctx's device current for the lifetime of the guard.
struct DeviceGuard {
    int prev = -1;
    explicit DeviceGuard(int dev)
    {
        TIPLHIP_CHECK(hipGetDevice(&prev));
        if (prev != dev) TIPLHIP_CHECK(hipSetDevice(dev));
    }
    ~DeviceGuard()
    {
        int cur = -1;
        if (hipGetDevice(&cur) == hipSuccess && prev >= 0 && cur != prev) (void)hipSetDevice(prev);
    }
};

inline size_t vox(const int* d) { return size_t(d[0]) * size_t(d[1]) * size_t(d[2]); }

// Host restatements of the reference's float geometry (same operation order).
void affine_matrix(const float* p, const int* fdim, const float* fvs, const int* tdim,
                   const float* tvs, float* T);
bool matrix_inverse(float* T);
bool is_integer_shift(const float* T, int* from_pos);

// Launch shape of the affine cost kernels for one image pair: planes per
// work item and workgroups per CU (chosen by tiplhip_autotune per key).
struct LaunchShape {
    uint32_t zc;
    uint32_t per_cu;
};
// from dims, to dims, K, cost, bins
using TuneKey = std::array<int, 9>;

}  // namespace tiplhip

struct tiplhip_ctx {
    int device = 0;
    int cu_count = 256;
    hipStream_t stream = nullptr;
    std::recursive_mutex mu;
    // scratch for the affine batch path
    tiplhip::DeviceBuffer d_T;       // K x 12 floats, crop flags, segment table
    tiplhip::DeviceBuffer d_hist;    // K x bins^2 int32
    tiplhip::DeviceBuffer d_sums;    // K x 4 uint64
    tiplhip::DeviceBuffer d_costs;   // K doubles
    tiplhip::DeviceBuffer d_ztab;    // K x depth z products of large cost launches
    tiplhip::DeviceBuffer d_tmp;     // volume output / misc
    tiplhip::DeviceBuffer d_tmp2;
    tiplhip::PinnedBuffer h_T;
    tiplhip::PinnedBuffer h_out;
    tiplhip::PinnedBuffer h_costs;   // coherent: cost kernels write the K costs here directly
    tiplhip::PinnedBuffer h_cdm;     // coherent: a CDM iteration's per-modality costs
    // d_hist once queued work ends: [0, hist_dirty) holds counts no entropy
    // pass consumed, [hist_dirty, hist_zero) is zero
    size_t hist_dirty = 0, hist_zero = 0;
    // autotuned launch shapes of single-pair cost launches
    std::map<tiplhip::TuneKey, tiplhip::LaunchShape> tuned;
    // timing
    hipEvent_t t0 = nullptr, t1 = nullptr;
    hipEvent_t e_read = nullptr;     // marks a device-to-host copy the host waits for
    bool kernel_timing = false;
    bool use_graphs = false;         // see tiplhip_ctx_set_graphs
    std::map<std::string, hipGraphExec_t> graphs;  // by launch shape
    tiplhip_ctx* aux = nullptr;      // see aux_context
    // when set, cdm() level loops on this context run as batched steps
    // (cdm.hip), shared with the other contexts attached to the batcher
    std::shared_ptr<tiplhip::CdmBatcher> cdm_batch;
    double kernel_ms = 0.0;
    int64_t kernel_launches = 0;
    // kernel timing without host synchronisation: event pairs recorded
    // around each timed launch, resolved into kernel_ms by timing_flush
    std::vector<hipEvent_t> kt_events;
    size_t kt_pending = 0;
};

struct tiplhip_volume {
    tiplhip_ctx* ctx = nullptr;
    int dim[3] = {0, 0, 0};
    size_t n = 0;
    uint8_t* d = nullptr;            // device voxels
    uint32_t* d_hist = nullptr;      // device value histograms: 256 bins then 64 bins
    // 8-tap packed copy (built on first use as a resampling source): word v
    // holds the bytes at v, v+1, v+w, v+w+1, v+wh, v+wh+1, v+wh+w, v+wh+w+1
    // (0 past the array end), so one 8-byte gather fetches a trilinear cell.
    uint64_t* d_pack = nullptr;
    std::mutex pack_mu;              // guards the lazy build of d_pack
    int64_t sum = 0, sum_sq = 0;     // exact
    uint32_t hist256[256];
    uint32_t hist64[64];
};

namespace tiplhip {

// Stream-ordered device memory from a per-device pool (hipMallocFromPoolAsync /
// hipFreeAsync on ctx->stream).  Unlike hipFree, a free does not synchronise
// the device, so registrations running on other streams are not stalled.
// Memory allocated here and used on another context's stream must be
// finished there before it is freed (every blocking entry point is).
void* pool_alloc(tiplhip_ctx* ctx, size_t bytes);
void pool_free(tiplhip_ctx* ctx, void* p);
// Destroys the context's cached cost-launch graphs.
void clear_graphs(tiplhip_ctx* ctx);

// Allocates a volume object with device storage for `dim` voxels.
tiplhip_volume* volume_alloc(tiplhip_ctx* ctx, const int* dim);
// Frees a volume's device storage (stream-ordered, on its owner's stream) and the object.
void volume_release(tiplhip_volume* v);
// downsample_with_padding of a device uint8 volume into `out` ((dim+1)/2).
void launch_downsample_u8(tiplhip_ctx* ctx, const uint8_t* in, const int* dim, uint8_t* out);
// Builds vol->d_pack if it does not exist yet.
void volume_ensure_packed(tiplhip_ctx* ctx, tiplhip_volume* vol);
// Computes sum / sum_sq / histograms of vol->d on the device and stores them.
void volume_compute_stats(tiplhip_ctx* ctx, tiplhip_volume* vol);
// After vol->d was rewritten on ctx's stream: drops the packed copy and
// recomputes the statistics.
void volume_invalidate(tiplhip_ctx* ctx, tiplhip_volume* vol);

// Resolves the context's recorded kernel-timing event pairs into kernel_ms
// and kernel_launches (waits for the last one).
void timing_flush(tiplhip_ctx* ctx);

// cdm() / the fixed-schedule extension into the device field d_field
// (vox(It[0]->dim) x 3 floats, zeroed here) on ctx's stream.  Returns
// TIPLHIP_OK, TIPLHIP_ERR_FAILED (zero cost; d_field undefined) or
// TIPLHIP_ERR_TERMINATED (d_field holds the best field so far).
int cdm_device(tiplhip_ctx* ctx, const tiplhip_volume* const* It, const tiplhip_volume* const* Is,
               int n, const tiplhip_cdm_param* param, float* d_field,
               const volatile uint8_t* terminated, int* iterations);

// Copies `bytes` of device memory to host memory through ctx's pinned
// staging buffer and waits for it (ctx's stream is synchronised).
void read_back(tiplhip_ctx* ctx, const void* dev, void* host, size_t bytes);

// A second context on ctx's device (own stream), created on first use and
// destroyed with ctx: the reverse CDM of a registration runs on it.
tiplhip_ctx* aux_context(tiplhip_ctx* ctx);

// The CDM batch executor of a device (cdm.hip), created on first use and
// shared while any context holds it.  Contexts given it in cdm_batch run
// their cdm() level loops (one modality, 'r' gradient) as batched steps.
std::shared_ptr<CdmBatcher> shared_cdm_batcher(int device);
// Batched steps the batcher has launched, and fields over them.
void cdm_batcher_counts(const CdmBatcher& b, int64_t* steps, int64_t* fields);
// Whether servers batch CDM iterations: unless TIPLHIP_SERVER_CDM_BATCH=0
// when the server is created.
bool cdm_batching_enabled();

// filter::gaussian (filter/gaussian.hpp:18-89) of a device float image with
// `channels` interleaved channels (1 or 3), ping-ponging between `a` and
// `tmp`; returns whichever of the two holds the result.
constexpr unsigned kMaxGaussianBandwidth = 16;
float* launch_gaussian(tiplhip_ctx* ctx, float* a, float* tmp, const int* dim, int channels,
                       unsigned bandwidth);
// The same from src into dst (src is only read; tmp is clobbered).
void launch_gaussian_to(tiplhip_ctx* ctx, const float* src, float* dst, float* tmp,
                        const int* dim, int channels, unsigned bandwidth);

}  // namespace tiplhip
