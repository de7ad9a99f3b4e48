/*
 * tiplhip.h -- C ABI of the MI355X (gfx950) registration hot path.
 *
 * This is the drop-in boundary under TIPL's registration entry points.  The
 * reference plugs its GPU path in through two C++-mangled explicit template
 * instantiations compiled in a CUDA translation unit:
 *
 *   linear_reg_imp<out,mr,uint8,3>(std::true_type, shared_ptr<linear_reg_runner>,
 *                                  cost_type, bool&)
 *       declared reg/linear.hpp:606-609, instantiated reg/linear_cu.hpp:7-16;
 *       its cost functors are correlation_cuda (reg/linear.hpp:40-78) and
 *       mutual_information_cuda (reg/linear.hpp:207-252);
 *   cdm_cuda<out,uint8,dim>(It, Is, d, terminated, cdm_param)
 *       declared reg/cdm.hpp:525-531, instantiated reg/cdm_cu.hpp:25-36.
 *
 * This library replaces both with plain C entry points (no C++ types, no
 * exceptions across the boundary).  The header-only C++ mirror in
 * include/tipl/ calls them from tipl::reg::linear and tipl::reg::cdm_common.
 *
 * Conventions
 *   - Volumes are x-fastest uint8 arrays, dim = {w, h, d}
 *     (tipl::image<3,unsigned char>, utility/basic_image.hpp:136-317).
 *   - A transform is 12 floats: sr[9] (row-major 3x3) then shift[3]
 *     (tipl::transformation_matrix<float,3>, numerical/transformation.hpp:589-604);
 *     it maps a voxel of the `from` grid to a position in the `to` volume.
 *   - An affine parameter vector is 12 floats: translocation[3], rotation[3],
 *     scaling[3], affine[3] (tipl::affine_param<float,3>,
 *     numerical/transformation.hpp:421-437).
 *   - Displacement fields are packed float3 per voxel
 *     (tipl::image<3,tipl::vector<3>>).
 *   - Every function returns TIPLHIP_OK (0) or an error code; the message of
 *     the last error on the calling thread is tiplhip_last_error().
 *   - A context owns one HIP stream on one device.  Calls on one context are
 *     serialised by the context (the reference's cost functors are re-entered
 *     from several threads, reg/linear.hpp:431-436); use one context per
 *     thread for concurrency.
 */
#ifndef TIPLHIP_H
#define TIPLHIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TIPLHIP_OK 0
#define TIPLHIP_ERR_INVALID 1     /* bad argument */
#define TIPLHIP_ERR_DEVICE 2      /* HIP runtime error */
#define TIPLHIP_ERR_MEMORY 3      /* device allocation failed */
#define TIPLHIP_ERR_NO_DEVICE 4   /* no usable GPU */
#define TIPLHIP_ERR_TERMINATED 5  /* the caller's terminate flag was raised */
#define TIPLHIP_ERR_FAILED 6      /* the algorithm returned false (e.g. zero cost) */

#define TIPLHIP_U8 0
#define TIPLHIP_I16 1
#define TIPLHIP_F32 2

#define TIPLHIP_COST_CORR 0       /* tipl::reg::corr */
#define TIPLHIP_COST_MI 1         /* tipl::reg::mutual_info */

typedef struct tiplhip_ctx tiplhip_ctx;
typedef struct tiplhip_volume tiplhip_volume;

/* ---- errors, devices, contexts ------------------------------------------ */

/* Message of the last failed call on this thread ("" if none).
 * Replaces the std::runtime_error(cudaGetErrorName(..)) throws of cu.hpp:288-322. */
const char* tiplhip_last_error(void);

/* Number of visible HIP devices (tipl::has_gpu, def.hpp:7-13). */
int tiplhip_device_count(int* count);

/* A context: one device, one stream, reusable scratch memory.  Replaces the
 * per-functor device state of correlation_cuda / mutual_information_cuda
 * (reg/linear.hpp:43-44, 211-213) and the legacy default stream. */
int tiplhip_ctx_create(int device, tiplhip_ctx** out);
int tiplhip_ctx_destroy(tiplhip_ctx* ctx);
int tiplhip_ctx_synchronize(tiplhip_ctx* ctx);

/* Device time accounting: when enabled, the affine cost entry points record
 * HIP events around their main kernel on the context stream; the host does
 * not wait for them, so a timed run keeps its launch rate.  The pairs are
 * resolved (waiting for the last one) when tiplhip_ctx_kernel_time reads the
 * totals, and every 512 launches.  Enabling or disabling resets the totals
 * and drops pairs not yet read. */
int tiplhip_ctx_set_kernel_timing(tiplhip_ctx* ctx, int enable);
int tiplhip_ctx_kernel_time(tiplhip_ctx* ctx, double* total_ms, int64_t* launches);

/* HIP graph replay of cost evaluations (default off).  When enabled, a
 * cost launch (tiplhip_cost_batch, _params, _pairs) whose launch shape --
 * cost, bins, candidate count, segment and workgroup counts, staged bytes --
 * was seen before on this context replays a captured graph of its upload and
 * kernels instead of enqueuing them one by one; only the staged transforms
 * and segment table change between replays.  Results are bit-identical to
 * the eager path.  Ignored while kernel timing is on.  No counterpart in the
 * reference (the launch pattern of its line search, reg/linear.hpp:255-338,
 * re-evaluates the same shapes many times). */
int tiplhip_ctx_set_graphs(tiplhip_ctx* ctx, int enable);

/* Stream-level timer: events recorded on the context stream. */
int tiplhip_timer_start(tiplhip_ctx* ctx);
int tiplhip_timer_stop(tiplhip_ctx* ctx, float* ms);

/* Streaming device copy of `bytes` (rounded down to 16), repeated `reps`
 * times on the context stream: the achievable HBM bandwidth (read + write
 * bytes / time, GB/s) and a kernel of known traffic for calibrating memory
 * counters.  No counterpart in the reference (measurement support). */
int tiplhip_stream_copy(tiplhip_ctx* ctx, size_t bytes, int reps, double* gbps);

/* Kernels of known HBM traffic over a `bytes` buffer (rounded down to 48),
 * `reps` launches: kind 0 the copy above (read + write), kind 1 a read of
 * every byte with 16-byte loads, kind 2 a read of every byte as packed
 * 12-byte float3 voxels (three 4-byte loads a lane, the field kernels'
 * pattern).  gbps = bytes moved / time.  With memory counters they give the
 * factor between a counter and the bytes a kernel really moved, per access
 * pattern.  No counterpart in the reference (measurement support). */
int tiplhip_traffic_probe(tiplhip_ctx* ctx, int kind, size_t bytes, int reps, double* gbps);

/* ---- volumes ------------------------------------------------------------- */

/* Upload a volume and keep it resident on the device as uint8.
 * dtype TIPLHIP_U8: copied as is (reg/linear.hpp:58-62, 222-229).
 * dtype TIPLHIP_I16 / TIPLHIP_F32: quantised on the device with
 * normalize_upper_lower2(in, out, upper_limit) (numerical/numerical.hpp:909-927,
 * 1085-1108); the reference uses upper_limit 255.99 for correlation
 * (reg/linear.hpp:69-70) and 255.99 for MI (mi_his_bandwidth - 0.01,
 * reg/linear.hpp:195-196).  A constant volume is left all zero. */
int tiplhip_volume_create(tiplhip_ctx* ctx, const void* host, int dtype, const int dim[3],
                          float upper_limit, tiplhip_volume** out);
/* Upload from a device pointer (no host copy). */
int tiplhip_volume_create_device(tiplhip_ctx* ctx, const void* device_ptr, int dtype,
                                 const int dim[3], float upper_limit, tiplhip_volume** out);
/* reg::subject_image_pre (reg/mm_reg.hpp:14-20) on the device: the image
 * (uint8, int16 or float32, converted to float) smoothed by
 * filter::gaussian (bandwidth 1), normalised by
 * segmentation::normalize_otsu_median(I, 255.99f) (numerical/otsu.hpp:96-136)
 * and truncated to uint8 -- bit-exact.  (Template images use
 * tiplhip_volume_create with upper_limit 255.999, reg/mm_reg.hpp:28-34.) */
int tiplhip_volume_create_subject(tiplhip_ctx* ctx, const void* host, int dtype, const int dim[3],
                                  tiplhip_volume** out);
int tiplhip_volume_destroy(tiplhip_volume* vol);
int tiplhip_volume_dims(const tiplhip_volume* vol, int dim[3]);
int tiplhip_volume_read(tiplhip_ctx* ctx, const tiplhip_volume* vol, uint8_t* host);
/* Exact value statistics kept with the volume: sum, sum of squares and the
 * 256-bin value histogram (histogram(), numerical/basic_op.hpp:1293-1311). */
int tiplhip_volume_stats(const tiplhip_volume* vol, int64_t* sum, int64_t* sum_sq,
                         uint32_t hist256[256]);

/* downsample_with_padding (numerical/resampling.hpp:575-613): a new volume of
 * shape (dim+1)/2. */
int tiplhip_volume_downsample(tiplhip_ctx* ctx, const tiplhip_volume* in, tiplhip_volume** out);

/* ---- affine resampling and costs ---------------------------------------- */

/* tipl::resample(from, to, trans) (numerical/resampling.hpp:1239-1299):
 * out (host, out_dim voxels) receives src(T * p) by trilinear interpolation
 * truncated to uint8 (0 outside).  crop_shortcut = 1 reproduces the CPU
 * reference's exact crop when T is an integer shift (:1244-1262);
 * 0 reproduces the reference's device resample (no shortcut). */
int tiplhip_resample(tiplhip_ctx* ctx, const tiplhip_volume* src, const int out_dim[3],
                     const float T[12], int crop_shortcut, uint8_t* out);

/* Exact correlation sums for K transforms: for candidate k,
 * sums[5k..5k+4] = {Sx, Sxx, Sy, Syy, Sxy} with x = from and
 * y = resample(to, from grid, T_k) (tipl::correlation on uint8,
 * numerical/statistics.hpp:186-322, 491-625).  No role swap is applied. */
int tiplhip_corr_sums(tiplhip_ctx* ctx, const tiplhip_volume* from, const tiplhip_volume* to,
                      const float* T, int K, int64_t* sums);

/* Joint histogram of get_mutual_info (reg/linear.hpp:138-169): for
 * candidate k, hist[k*bins*bins + (from << b) + to(T_k * p)] over the from
 * grid.  bins = 256 is the reference's (mi_band_width 8, reg/linear.hpp:84-85);
 * bins = 64 bins both values by >> 2. */
int tiplhip_mi_hist(tiplhip_ctx* ctx, const tiplhip_volume* from, const tiplhip_volume* to,
                    const float* T, int K, int bins, int32_t* hist);

/* K cost evaluations in one batched launch -- the optimiser's candidate sets
 * (numerical/optimization.hpp:16-28, 59-96, 272-512).
 * cost = TIPLHIP_COST_CORR: reg::correlation()(from, to, T) (reg/linear.hpp:23-37),
 *        1 - float(r)^2, with the from/to role swap of :28-33 and the
 *        CPU resample's integer-shift crop;
 * cost = TIPLHIP_COST_MI: mutual_information<3>()(from, to, T) (reg/linear.hpp:173-203),
 *        the negated FP64 entropy mean.
 * T is host memory (K x 12 floats); costs is host memory (K doubles). */
int tiplhip_cost_batch(tiplhip_ctx* ctx, const tiplhip_volume* from, const tiplhip_volume* to,
                       const float* T, int K, int cost, int bins, double* costs);

/* tiplhip_cost_batch that allocates nothing (no counterpart in the reference;
 * for callers that cannot allocate device memory: code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Every
 * device buffer beyond the inputs and outputs -- staged transforms, segment
 * table, joint histograms, correlation sums, costs -- lives in a 20 MiB
 * static device array that each GPU the library runs on holds until the
 * process ends.  Limits: K <= 256; with 256 bins, K <= 79 (the histograms
 * must fit the array).  The sampled volume (`to`, or `from` when correlation
 * swaps the roles of a larger `from`) must already hold its packed copy,
 * which its first use in tiplhip_cost_batch builds.  Costs equal
 * tiplhip_cost_batch's bit for bit.  Calls on one GPU share the array:
 * calls that overlap on one GPU race (the context's own lock serialises
 * calls on one context). */
int tiplhip_cost_batch_noalloc(tiplhip_ctx* ctx, const tiplhip_volume* from,
                               const tiplhip_volume* to, const float* T, int K, int cost, int bins,
                               double* costs);

/* Autotuner of the cost kernels' launch shape (no reference counterpart):
 * times the histogram (MI) or sums (correlation) kernel for these K
 * transforms over a grid of work-item depths (8, 16, 32 planes) and
 * workgroups per CU (per_cu 0: the fewest workgroups, each as full as the
 * histogram counters allow), `reps` launches each, and keeps the fastest for
 * every later single-pair launch on ctx with the same shapes, K, cost and
 * bins.  Results do not depend on the shape (integer accumulation).  Outputs
 * may be NULL. */
int tiplhip_autotune(tiplhip_ctx* ctx, const tiplhip_volume* from, const tiplhip_volume* to,
                     const float* T, int K, int cost, int bins, int reps, int* zc, int* per_cu,
                     double* best_ms);

/* One image pair's candidates in a multi-pair launch. */
typedef struct tiplhip_pair_costs {
    const tiplhip_volume* from;
    const tiplhip_volume* to;
    const float* T;          /* K transforms, 12 floats each, host memory */
    int K;
    double* costs;           /* K costs out, host memory */
} tiplhip_pair_costs;

/* Pins the launch shape of later single-pair cost launches on ctx with these
 * volumes' shapes, K, cost and bins, as the autotuner would (for example a
 * shape found offline): zc planes per work item, per_cu workgroups per CU
 * (0: the fewest).  Results do not depend on the shape.  No reference
 * counterpart. */
int tiplhip_set_launch_shape(tiplhip_ctx* ctx, const tiplhip_volume* from, const tiplhip_volume* to,
                             int K, int cost, int bins, int zc, int per_cu);

/* tiplhip_cost_batch for n image pairs of any sizes in one launch (all
 * candidates of all pairs, at most 4096, share one histogram or sum kernel
 * and one finishing kernel).  Each cost equals what tiplhip_cost_batch
 * returns for that pair alone.  Volumes may come from any context of the
 * calling context's device.  No reference counterpart: the reference
 * evaluates one candidate of one pair per call (reg/linear.hpp:40-78,
 * 205-252). */
int tiplhip_cost_batch_pairs(tiplhip_ctx* ctx, const tiplhip_pair_costs* pairs, int n, int cost,
                             int bins);

/* The same with the K transforms computed on the host from affine
 * parameters (transformation_matrix(arg, from, from_vs, to, to_vs),
 * numerical/transformation.hpp:612-666). */
int tiplhip_cost_batch_params(tiplhip_ctx* ctx, const tiplhip_volume* from, const float from_vs[3],
                              const tiplhip_volume* to, const float to_vs[3], const float* params,
                              int K, int cost, int bins, double* costs);

/* ---- host geometry helpers (same float arithmetic as the reference) ---- */

/* transformation_matrix(affine_param, from, from_vs, to, to_vs). */
int tiplhip_affine_matrix(const float param[12], const int from_dim[3], const float from_vs[3],
                          const int to_dim[3], const float to_vs[3], float T[12]);
/* transformation_matrix::inverse (numerical/transformation.hpp:820-832);
 * TIPLHIP_ERR_INVALID when singular. */
int tiplhip_matrix_inverse(float T[12]);

/* ---- CDM ------------------------------------------------------------------ */

/* cdm_param (reg/cdm.hpp:36-44) plus the iteration cap that cdm() hard-codes
 * as 200 (reg/cdm.hpp:418), plus a fixed schedule (an extension with no
 * reference counterpart, off by default): with levels > 0 the pyramid has
 * exactly `levels` levels, every level runs exactly `level_iterations`
 * iterations of the loop body with no early exit and keeps its running
 * field, and the field is smoothed each iteration by
 * filter::gaussian(field_bandwidth) instead of cdm_smooth when
 * field_bandwidth > 0 (bandwidth b has variance b/2: 8 is sigma 2).
 * resolution, min_dimension and max_iterations do not apply then. */
typedef struct {
    float resolution;       /* 2.0 */
    float speed;            /* 0.3 */
    float smoothing;        /* 0.05 (in the fixed schedule: cdm_smooth's weight) */
    int min_dimension;      /* 16 */
    int max_iterations;     /* 200 */
    char gradient_type[8];  /* per modality: 'r' (default) or 'd' */
    int levels;             /* 0: the reference's pyramid rule */
    int level_iterations;   /* iterations per level in the fixed schedule */
    int field_bandwidth;    /* 0: cdm_smooth; 1..16: gaussian smoothing of the field */
} tiplhip_cdm_param;

void tiplhip_cdm_param_default(tiplhip_cdm_param* p);

/* tipl::reg::cdm / cdm_cuda (reg/cdm.hpp:373-532) for n modality pairs of
 * equal shape.  field (host) receives vox(dim) float3.  terminated may be
 * NULL; it is a one-byte flag (a C++ bool) polled once per iteration, as
 * the reference polls its bool& (reg/cdm.hpp:418).  Returns
 * TIPLHIP_ERR_FAILED where the reference returns false, and
 * TIPLHIP_ERR_TERMINATED (with the best field so far written) when the flag
 * stopped the run. */
int tiplhip_cdm(tiplhip_ctx* ctx, const tiplhip_volume* const* It, const tiplhip_volume* const* Is,
                int n, const tiplhip_cdm_param* param, float* field,
                const volatile uint8_t* terminated, int* iterations);

/* One resolution level of the CDM loop with no pyramid: runs exactly
 * `iterations` loop iterations (no early exit) from a zero field, for
 * timing.  field may be NULL. */
int tiplhip_cdm_level_iterations(tiplhip_ctx* ctx, const tiplhip_volume* It,
                                 const tiplhip_volume* Is, const tiplhip_cdm_param* param,
                                 int iterations, float* field);

/* Single CDM steps on host buffers (upload, one kernel, download), for
 * parity checks of each kernel against the reference routine it restates. */
int tiplhip_compose_displacement(tiplhip_ctx* ctx, const uint8_t* Is, const int dim[3],
                                 const float* d, uint8_t* Js);            /* numerical/dif.hpp:90-107 */
int tiplhip_cdm_gradient(tiplhip_ctx* ctx, const uint8_t* Js, const uint8_t* It,
                         const int dim[3], char type, float* new_d,
                         float* cost);                                   /* reg/cdm.hpp:46-106 */
int tiplhip_cdm_poisson(tiplhip_ctx* ctx, float* new_d, const int dim[3], int sweeps,
                        float* max_length);                              /* reg/cdm.hpp:158-296 */
int tiplhip_accumulate_displacement(tiplhip_ctx* ctx, const float* cur, float* new_d,
                                    const int dim[3], float scale);      /* numerical/dif.hpp:166-186 */
int tiplhip_cdm_smooth(tiplhip_ctx* ctx, const float* d, float* dd, const int dim[3],
                       float smoothing);                                 /* reg/cdm.hpp:302-342 */
int tiplhip_upsample_field(tiplhip_ctx* ctx, const float* in, const int in_dim[3], float* out,
                           const int out_dim[3]);                        /* numerical/resampling.hpp:849-858 */

/* ---- 2-D CDM -------------------------------------------------------------- */

/* tipl::reg::cdm for 2-D uint8 images: the reference's second GPU plug,
 * cdm_cuda<out,uint8,2> (declared reg/cdm.hpp:525-531, instantiated
 * reg/cdm_cu.hpp:12-17), which runs the cdm() template (reg/cdm.hpp:373-474)
 * at dimension 2.  It and Is are n host images of dim[0] x dim[1] pixels
 * (x-fastest); field (host) receives dim[0]*dim[1] float2.  param as for
 * tiplhip_cdm (the fixed schedule, levels > 0, is 3-D only); terminated and
 * the return codes as for tiplhip_cdm. */
int tiplhip_cdm2d(tiplhip_ctx* ctx, const uint8_t* const* It, const uint8_t* const* Is, int n,
                  const int dim[2], const tiplhip_cdm_param* param, float* field,
                  const volatile uint8_t* terminated, int* iterations);

/* Single 2-D CDM steps on host buffers, for parity checks of each kernel
 * against the reference routine it restates at dimension 2. */
int tiplhip_cdm2d_compose(tiplhip_ctx* ctx, const uint8_t* Is, const int dim[2], const float* d,
                          uint8_t* Js);                                   /* numerical/dif.hpp:90-107 */
int tiplhip_cdm2d_gradient(tiplhip_ctx* ctx, const uint8_t* Js, const uint8_t* It,
                           const int dim[2], char type, float* new_d,
                           float* cost);                                  /* reg/cdm.hpp:46-106 */
int tiplhip_cdm2d_poisson(tiplhip_ctx* ctx, float* new_d, const int dim[2], int sweeps,
                          float* max_length);                             /* reg/cdm.hpp:158-296 */
int tiplhip_cdm2d_accumulate(tiplhip_ctx* ctx, const float* cur, float* new_d, const int dim[2],
                             float scale);                                /* numerical/dif.hpp:166-186 */
int tiplhip_cdm2d_smooth(tiplhip_ctx* ctx, const float* d, float* dd, const int dim[2],
                         float smoothing);                                /* reg/cdm.hpp:302-342 */
int tiplhip_cdm2d_downsample(tiplhip_ctx* ctx, const uint8_t* in, const int dim[2],
                             uint8_t* out);                               /* numerical/resampling.hpp:575-613 */
int tiplhip_cdm2d_upsample(tiplhip_ctx* ctx, const float* in, const int in_dim[2], float* out,
                           const int out_dim[2]);                         /* numerical/resampling.hpp:849-858 */

/* tipl::filter::gaussian(image, bandwidth) (filter/gaussian.hpp:18-105), in
 * place on a host float image of dim[0]*dim[1]*dim[2] voxels with `channels`
 * interleaved floats each (1: scalar image, 3: vector<3> field).  Bit-exact:
 * each [1 2 1] pass is rounded to float as the reference's pass arrays are.
 * bandwidth 0 leaves the image unchanged; bandwidth above 16 is rejected. */
int tiplhip_gaussian(tiplhip_ctx* ctx, float* data, const int dim[3], int channels,
                     unsigned bandwidth);

/* tipl::resample (numerical/resampling.hpp:1239-1299) of a device volume
 * into a new device volume of shape out_dim: out(p) = src(T p), trilinear,
 * uint8 truncation; with crop_shortcut, an integer-shift T is an exact crop
 * as on the CPU path.  The result lives on ctx. */
int tiplhip_volume_resample(tiplhip_ctx* ctx, const tiplhip_volume* src, const int out_dim[3],
                            const float T[12], int crop_shortcut, tiplhip_volume** out);

/* ---------------------------------------------------------------------------
 * Post-registration warping (numerical/dif.hpp) and mm_reg's r.  Fields and
 * mappings are host arrays of vox(dim) packed float3 (tipl::image<3,
 * tipl::vector<3>>), x fastest.  A transform T is 12 floats: sr[9] row-major,
 * then shift[3] (tipl::transformation_matrix).  All results are bit-exact.
 * ------------------------------------------------------------------------- */

/* displacement_to_mapping (numerical/dif.hpp:9-28; with T :41-50):
 * mapping(x) = T(x + dis(x)); T == NULL is the identity.  mapping may equal dis. */
int tiplhip_displacement_to_mapping(tiplhip_ctx* ctx, const float* dis, const int dim[3],
                                    const float* T, float* mapping);

/* mapping_to_displacement (numerical/dif.hpp:30-38): dis(x) = mapping(x) - x. */
int tiplhip_mapping_to_displacement(tiplhip_ctx* ctx, const float* mapping, const int dim[3],
                                    float* dis);

/* inv_displacement_to_mapping<linear> (numerical/dif.hpp:52-64): over a grid
 * of out_dim, mapping(x) = p + inv_dis(p) with p = T^-1(x); inv_dis (dis_dim)
 * is sampled trilinearly and counts as zero outside its grid. */
int tiplhip_inv_displacement_to_mapping(tiplhip_ctx* ctx, const float* inv_dis,
                                        const int dis_dim[3], const float T[12],
                                        const int out_dim[3], float* mapping);

/* compose_mapping<linear> (numerical/dif.hpp:66-86): out(x) = from(mapping(x))
 * over mapping's grid dim, trilinear, 0 outside `from`.  dtype TIPLHIP_U8
 * (interp_type<uint8> truncation) or TIPLHIP_F32; out has vox(dim) voxels. */
int tiplhip_compose_mapping(tiplhip_ctx* ctx, const void* from, int dtype, const int from_dim[3],
                            const float* mapping, const int dim[3], void* out);

/* jacobian_determinant_dis<3> (numerical/dif.hpp:399-465): per voxel of a
 * displacement field, in the reference's mixed float/double arithmetic,
 * stored as float; 1 on the volume's edge voxels. */
int tiplhip_jacobian_determinant_dis(tiplhip_ctx* ctx, const float* dis, const int dim[3],
                                     float* out);

/* jacobian_determinant (numerical/dif.hpp:356-397) of a mapping, float; 0 on
 * edge voxels (the value-returning overload's fresh image). */
int tiplhip_jacobian_determinant(tiplhip_ctx* ctx, const float* mapping, const int dim[3],
                                 float* out);

/* mm_reg::It_match_fov (reg/mm_reg.hpp:313-328) in place on a device volume:
 * voxels x with T(x) outside [0, moving_dim) are set to 0. */
int tiplhip_match_fov(tiplhip_ctx* ctx, tiplhip_volume* It, const float T[12],
                      const int moving_dim[3]);

/* r of two device volumes of one shape (reg/mm_reg.hpp:302-305): with masked,
 * correlation_ygz (only voxels where y > 0, numerical/statistics.hpp:599-619),
 * else tipl::correlation (:569-597); exact integer sums, FP64 tail. */
int tiplhip_correlation(tiplhip_ctx* ctx, const tiplhip_volume* x, const tiplhip_volume* y,
                        int masked, double* r);

/* ---------------------------------------------------------------------------
 * NIfTI files on the host (io/nifti.hpp), so real images can be handed to
 * tiplhip_register.  NIfTI-1 (.nii, n+1 / ni1 + .img), NIfTI-2, Analyze
 * .hdr/.img, either byte order, gzip-compressed or not.  Reading follows
 * tipl::io::nifti's `>>`: qform handled when there is no sform, the axes
 * swapped and flipped into LPS (toLPS), voxels converted to the output type,
 * then scl_slope / scl_inter.
 * ------------------------------------------------------------------------- */
typedef struct tiplhip_nifti_info {
    int dim[8];          /* after the LPS reorientation: dim[0] rank, dim[1..3] x, y, z,
                            dim[4] volumes */
    float pixdim[8];     /* voxel sizes in pixdim[1..3], after the reorientation */
    float srow[12];      /* srow_x, srow_y, srow_z after qform handling and reorientation
                            (the reference's image transformation, last row 0 0 0 1) */
    int datatype, bitpix;
    float scl_slope, scl_inter;
    int sform_code, qform_code;
    int is_mni;          /* nifti_base::is_mni (io/nifti.hpp:807-814) */
    int version;         /* 1 (n+1), 2 (n+2), 0 (ni1 or Analyze pair) */
    int big_endian;
    int flip_swap[6];    /* the toLPS sequence: 0-2 flip x/y/z, 3 swap xy, 4 swap yz, 5 swap xz */
    int flip_swap_n;
    char descrip[80];
} tiplhip_nifti_info;

/* nifti_base::open + header (io/nifti.hpp:601-723). */
int tiplhip_nifti_read_info(const char* path, tiplhip_nifti_info* info);

/* `>> image` of volume `volume` (io/nifti.hpp:1126-1160) into out: dtype
 * TIPLHIP_U8, TIPLHIP_I16 or TIPLHIP_F32, vox of the reoriented dim[1..3]
 * voxels, x fastest.  info may be NULL. */
int tiplhip_nifti_read(const char* path, int volume, int dtype, void* out, tiplhip_nifti_info* info);

/* `<< vs << R << image` (io/nifti.hpp:937-985) on a fresh header: NIfTI-1
 * n+1, sform_code 1 (4 when is_mni), qform_code 0, srow = the 12 floats;
 * gzip-compressed when path ends in .gz. */
int tiplhip_nifti_write(const char* path, const void* data, int dtype, const int dim[3],
                        const float vs[3], const float srow[12], int is_mni);

/* ---------------------------------------------------------------------------
 * Registration of one image pair in the reference's mm_reg order
 * (reg/mm_reg.hpp:285-416), every step on the device:
 *   linear_reg: linear(fixed, moving) (or the given arg with skip_linear);
 *   calculate_linear_r: J = resample(moving) onto the fixed grid, r_linear;
 *   nonlinear requests: It_match_fov on the fixed image (match_fov), then
 *   cdm(fixed, J) -> t2f and, with reverse, cdm(J, fixed) -> f2t on a
 *   second stream at the same time; to2from = displacement_to_mapping(t2f, T),
 *   from2to = inv_displacement_to_mapping(f2t, T); calculate_nonlinear_r:
 *   J = compose_mapping(moving, to2from), r_nonlinear.
 * fixed plays `from` / It (template), moving plays `to` / I (subject).  Inputs
 * of type int16 or float32 are quantised on the device (preprocess).
 * ------------------------------------------------------------------------- */
typedef struct tiplhip_reg_request {
    const void* fixed;        /* host voxels, x fastest */
    int fixed_dtype;          /* TIPLHIP_U8 / TIPLHIP_I16 / TIPLHIP_F32 */
    int fixed_dim[3];
    float fixed_vs[3];
    const void* moving;
    int moving_dtype;
    int moving_dim[3];
    float moving_vs[3];
    int cost;                 /* TIPLHIP_COST_MI (default) or TIPLHIP_COST_CORR */
    int bins;                 /* MI bins: 256 (reference) or 64 */
    int nonlinear;            /* 0: affine only; 1: affine then CDM */
    tiplhip_cdm_param cdm;
    float* field;             /* host, vox(fixed_dim) x 3 floats, or NULL */
    int preprocess;           /* 0: both images by normalize_upper_lower2(255.99), the
                                 linear cost functor's rule (reg/linear.hpp:65-71);
                                 1: mm_reg's rule -- fixed by template_image_pre,
                                 moving by subject_image_pre (reg/mm_reg.hpp:14-34) */
    int reverse;              /* nonlinear: also cdm(J, fixed) -> f2t, concurrently
                                 (mm_reg's skip_reverse = false); default 0 */
    int match_fov;            /* nonlinear: It_match_fov before cdm; default 1 (mm_reg) */
    int masked_r;             /* r over fixed > 0 (correlation_ygz); default 1 (mm_reg) */
    int skip_linear;          /* 1: no linear(); `arg` is used as given */
    float arg[12];            /* affine_param for skip_linear: t, r, s, a */
    float* f2t;               /* host vox(fixed) x 3: reverse field (needs reverse), or NULL */
    float* to2from;           /* host vox(fixed) x 3: fixed voxel -> moving position, or NULL */
    float* from2to;           /* host vox(moving) x 3: moving voxel -> fixed position
                                 (needs reverse), or NULL */
    uint8_t* warped;          /* host vox(fixed): the last J -- moving warped onto the fixed
                                 grid (affine or nonlinear), or NULL */
} tiplhip_reg_request;

typedef struct tiplhip_reg_result {
    int status;               /* TIPLHIP_OK or the error code of the failing step */
    float affine[12];         /* affine_param found by linear(): t, r, s, a */
    float T[12];              /* transformation_matrix: fixed voxel -> moving voxel */
    float cost;               /* linear()'s return value */
    int cdm_iterations;
    int64_t evaluations;      /* affine cost evaluations */
    double queue_ms;          /* submit to start (served requests; 0 on the plain path) */
    double run_ms;            /* start to done */
    double prep_ms;           /* upload, quantisation, host mask / moment estimate */
    double affine_ms;         /* the linear() optimisation */
    double cdm_ms;            /* match_fov, cdm (both directions), mappings, r_nonlinear */
    float r_linear;           /* mm_reg's r after the linear step (reg/mm_reg.hpp:302-305) */
    float r_nonlinear;        /* mm_reg's r after the nonlinear step (:341-343); 0 if none */
} tiplhip_reg_result;

void tiplhip_reg_request_default(tiplhip_reg_request* r);

/* The plain single-pair path: the whole registration on the calling thread
 * with ctx, each optimiser candidate set in its own launch.  The caller's
 * input and field buffers are used in place. */
int tiplhip_register(tiplhip_ctx* ctx, const tiplhip_reg_request* req, tiplhip_reg_result* res);

/* ---------------------------------------------------------------------------
 * Serving loop (no reference counterpart; the reference registers one pair
 * per call).  `slots` worker threads, each with its own context (stream), run
 * registrations concurrently; a request is admitted when a slot is free and
 * its estimated device memory fits in hbm_budget bytes (FIFO).  The affine
 * cost evaluations of all running registrations are merged by one batcher
 * into multi-pair launches (tiplhip_cost_batch_pairs) as they arrive, so
 * requests join and leave the batch at every optimiser step.  Each result
 * equals what tiplhip_register returns for that request alone.
 * A request's buffers must stay valid until tiplhip_server_wait returns it.
 * ------------------------------------------------------------------------- */
typedef struct tiplhip_server tiplhip_server;

typedef struct tiplhip_server_stats {
    int64_t submitted, completed, failed, running, queued;
    int64_t batch_launches;   /* multi-pair cost launches */
    int64_t batched_pairs;    /* pair segments over those launches */
    size_t hbm_in_use, hbm_peak;
    double latency_p50_ms, latency_p95_ms, latency_p99_ms;  /* submit to done */
    double wall_s;            /* first submit to last completion */
    double regs_per_s;        /* completed / wall_s */
    /* batched CDM iterations: steps launched since the server started and
       the fields over them (counted on the device's batcher, which servers
       on one device share; both 0 when batching is off) */
    int64_t cdm_steps, cdm_fields;
} tiplhip_server_stats;

int tiplhip_server_create(int device, int slots, size_t hbm_budget, tiplhip_server** out);
/* Estimated device bytes a request holds while it runs (what admission counts). */
/* The same with an explicit number of batch launcher threads (each with its
 * own stream): -1 picks the default (one per four slots, 1 to 4), 0 runs no
 * batcher -- every slot then launches its own candidate sets on its own
 * stream. */
int tiplhip_server_create_ex(int device, int slots, int launchers, size_t hbm_budget,
                             tiplhip_server** out);
size_t tiplhip_request_bytes(const tiplhip_reg_request* req);
int tiplhip_server_submit(tiplhip_server* s, const tiplhip_reg_request* req, uint64_t* ticket);
/* Blocks until the ticket is done; returns its status and forgets the ticket. */
int tiplhip_server_wait(tiplhip_server* s, uint64_t ticket, tiplhip_reg_result* res);
int tiplhip_server_stats_get(tiplhip_server* s, tiplhip_server_stats* st);
/* Finishes every submitted request, then stops the workers. */
int tiplhip_server_destroy(tiplhip_server* s);

#ifdef __cplusplus
}
#endif

#endif /* TIPLHIP_H */
