/*
 * oracle.c -- CPU restatement of TIPL's registration hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker that the HIP product
 * path is compared against.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it; the product library never links
 * or calls it.
 *
 * Build: gcc -std=c11 -O2 -ffp-contract=off -fPIC -shared (see oracle/Makefile).
 * -ffp-contract=off and no -march=native keep every FP32 operation a separate
 * IEEE-rounded operation in exactly the order the reference writes it; that
 * is what makes the uint8 truncation after trilinear interpolation, and hence
 * the integer sums and joint histograms, reproducible bit for bit.
 *
 * Conventions (all follow the reference's containers):
 *   - volumes are x-fastest, then y, then z; dim[3] = {w, h, d}
 *     (utility/basic_image.hpp:136-317, utility/shape.hpp);
 *   - a transform is 12 floats: sr[9] row-major then shift[3]
 *     (tipl::transformation_matrix, numerical/transformation.hpp:589-604);
 *   - an affine parameter vector is 12 floats: translocation[3], rotation[3],
 *     scaling[3], affine[3] (tipl::affine_param, numerical/transformation.hpp:421-437);
 *   - displacement fields are packed float3 per voxel, AoS
 *     (tipl::image<3,tipl::vector<3>>, utility/pixel_index.hpp:410-640).
 *
 * Pinning status of each function is recorded in tests/ (test_oracle_*.py);
 * every function here is checked bit-exactly against the compiled reference
 * (oracle/_ref, built from the reference headers by oracle/Makefile) and
 * against committed golden vectors in tests/golden/.
 */
#include <math.h>
#include <stddef.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ORC_API __attribute__((visibility("default")))

static size_t vox(const int* dim) { return (size_t)dim[0] * (size_t)dim[1] * (size_t)dim[2]; }

/* ------------------------------------------------------------------------ */
/* Affine parameters -> transformation matrix                               */
/* ------------------------------------------------------------------------ */

/* rotation_matrix(theta, m, vdim<3>) numerical/transformation.hpp:141-162,
 * rotation_scaling_matrix :176-190, rotation_scaling_affine_matrix :203-214,
 * transformation_matrix(affine_param, from, from_vs, to, to_vs) :612-666.
 * All arithmetic is float, as the reference instantiates it with float. */
ORC_API void orc_affine_matrix(const float* p, const int* fdim, const float* fvs,
                               const int* tdim, const float* tvs, float* T)
{
    const float* tr = p;        /* translocation */
    const float* rot = p + 3;   /* rotation */
    const float* sc = p + 6;    /* scaling */
    const float* af = p + 9;    /* affine (shear) */
    float m[9];
    float sin_x = sinf(rot[0]);
    float cos_x = cosf(rot[0]);
    float sin_y = sinf(rot[1]);
    float cos_y = cosf(rot[1]);
    float sin_z = sinf(rot[2]);
    float cos_z = cosf(rot[2]);
    float cxsz = cos_x * sin_z;
    float cxcz = cos_x * cos_z;
    float sxsz = sin_x * sin_z;
    float sxcz = sin_x * cos_z;
    m[0] = cos_y * cos_z;
    m[1] = cos_y * sin_z;
    m[2] = -sin_y;
    m[3] = -cxsz + sxcz * sin_y;
    m[4] = cxcz + sxsz * sin_y;
    m[5] = sin_x * cos_y;
    m[6] = sxsz + cxcz * sin_y;
    m[7] = -sxcz + cxsz * sin_y;
    m[8] = cos_x * cos_y;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            m[i * 3 + j] *= sc[i];
    /* m[0] += m[3]*a0 + m[6]*a1 : the right-hand side is formed first */
    m[0] += m[3] * af[0] + m[6] * af[1];
    m[1] += m[4] * af[0] + m[7] * af[1];
    m[2] += m[5] * af[0] + m[8] * af[1];
    m[3] += m[6] * af[2];
    m[4] += m[7] * af[2];
    m[5] += m[8] * af[2];

    float t[3];
    t[0] = (float)fdim[0];
    t[1] = (float)fdim[1];
    t[2] = (float)fdim[2];
    for (int i = 0; i < 3; ++i) t[i] *= -0.5f;
    for (int i = 0; i < 3; ++i) t[i] *= fvs[i];
    for (int i = 0; i < 3; ++i) t[i] += tr[i];
    float sh[3];
    sh[0] = m[0] * t[0] + m[1] * t[1] + m[2] * t[2];
    sh[1] = m[3] * t[0] + m[4] * t[1] + m[5] * t[2];
    sh[2] = m[6] * t[0] + m[7] * t[1] + m[8] * t[2];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            m[i * 3 + j] *= fvs[j];
    for (int i = 0; i < 3; ++i)
        if (tvs[i] != 1.0f) {
            m[i * 3 + 0] /= tvs[i];
            m[i * 3 + 1] /= tvs[i];
            m[i * 3 + 2] /= tvs[i];
            sh[i] /= tvs[i];
        }
    for (int i = 0; i < 3; ++i)
        sh[i] += (float)tdim[i] * 0.5f;
    memcpy(T, m, sizeof(m));
    memcpy(T + 9, sh, sizeof(sh));
}

/* transformation_matrix::inverse numerical/transformation.hpp:820-832 with
 * mat::inverse(iter, dim<3,3>) numerical/matrix.hpp:1479-1506 and
 * determinant numerical/matrix.hpp:1346-1352.  Returns 0 when singular. */
ORC_API int orc_matrix_inverse(float* T)
{
    const float* a = T;
    float det = a[0] * (a[4] * a[8] - a[5] * a[7]) +
                a[1] * (a[5] * a[6] - a[3] * a[8]) +
                a[2] * (a[3] * a[7] - a[4] * a[6]);
    if (det + 1.0f == 1.0f)
        return 0;
    float t[9];
    t[0] = a[4] * a[8] - a[5] * a[7];
    t[1] = a[2] * a[7] - a[1] * a[8];
    t[2] = a[1] * a[5] - a[2] * a[4];
    t[3] = a[5] * a[6] - a[3] * a[8];
    t[4] = a[0] * a[8] - a[2] * a[6];
    t[5] = a[2] * a[3] - a[0] * a[5];
    t[6] = a[3] * a[7] - a[4] * a[6];
    t[7] = a[1] * a[6] - a[0] * a[7];
    t[8] = a[0] * a[4] - a[1] * a[3];
    for (int i = 0; i < 9; ++i) t[i] /= det;
    float s[3];
    /* vector_rotation(shift, new_shift, iT, vdim<3>) :80-91 */
    s[0] = T[9] * t[0] + T[10] * t[1] + T[11] * t[2];
    s[1] = T[9] * t[3] + T[10] * t[4] + T[11] * t[5];
    s[2] = T[9] * t[6] + T[10] * t[7] + T[11] * t[8];
    for (int i = 0; i < 3; ++i) T[9 + i] = -s[i];
    memcpy(T, t, sizeof(t));
    return 1;
}

/* ------------------------------------------------------------------------ */
/* Trilinear interpolation                                                  */
/* ------------------------------------------------------------------------ */

/* interpolator::linear<3>::get_location numerical/interpolation.hpp:337-383.
 * Returns 0 when the sample is "outside": any coordinate < 0, or
 * ix+1 >= w (resp. h, d).  A NaN coordinate is outside too (the x86 build of
 * the reference converts NaN to 2^63, which fails the ix+1 < w test). */
typedef struct {
    size_t idx[8];
    float w[8];
} orc_tri;

static int orc_tri_locate(const int* dim, float x, float y, float z, orc_tri* t)
{
    if (x < 0.0f || y < 0.0f || z < 0.0f)
        return 0;
    if (isnan(x) || isnan(y) || isnan(z))
        return 0;
    if (x >= 16777216.0f || y >= 16777216.0f || z >= 16777216.0f)
        return 0;
    size_t ix = (size_t)x, iy = (size_t)y, iz = (size_t)z;
    size_t w = (size_t)dim[0];
    if (ix + 1 >= w || iy + 1 >= (size_t)dim[1] || iz + 1 >= (size_t)dim[2])
        return 0;
    float p0 = x - (float)ix;
    float p1 = y - (float)iy;
    float p2 = z - (float)iz;
    float n0 = 1.0f - p0;
    float n1 = 1.0f - p1;
    float n2 = 1.0f - p2;
    size_t wh = w * (size_t)dim[1];
    t->idx[0] = iz * wh + iy * w + ix;
    t->idx[1] = t->idx[0] + 1;
    t->idx[2] = t->idx[0] + w;
    t->idx[3] = t->idx[2] + 1;
    t->idx[4] = t->idx[0] + wh;
    t->idx[5] = t->idx[1] + wh;
    t->idx[6] = t->idx[2] + wh;
    t->idx[7] = t->idx[3] + wh;
    float n0n1 = n0 * n1;
    float p0n1 = p0 * n1;
    float n0p1 = n0 * p1;
    float p0p1 = p0 * p1;
    t->w[0] = n0n1 * n2;
    t->w[1] = p0n1 * n2;
    t->w[2] = n0p1 * n2;
    t->w[3] = p0p1 * n2;
    t->w[4] = n0n1 * p2;
    t->w[5] = p0n1 * p2;
    t->w[6] = n0p1 * p2;
    t->w[7] = p0p1 * p2;
    return 1;
}

/* weighted_sum numerical/interpolation.hpp:61-73 for uint8 data, then
 * interp_type<uint8>::assign :25-35 (v <= 0 -> 0, else truncate). */
static uint8_t orc_tri_u8(const uint8_t* I, const orc_tri* t)
{
    float r = (float)I[t->idx[0]] * t->w[0];
    for (int k = 1; k < 8; ++k) {
        float v = (float)I[t->idx[k]];
        v *= t->w[k];
        r += v;
    }
    return r <= 0.0f ? 0 : (uint8_t)r;
}

/* weighted_sum for tipl::vector<3,float> data (no truncation). */
static void orc_tri_f3(const float* F, const orc_tri* t, float* out)
{
    for (int c = 0; c < 3; ++c) {
        float r = F[t->idx[0] * 3 + c] * t->w[0];
        for (int k = 1; k < 8; ++k) {
            float v = F[t->idx[k] * 3 + c];
            v *= t->w[k];
            r += v;
        }
        out[c] = r;
    }
}

/* transformation_matrix::operator()(pixel_index) numerical/transformation.hpp:834-839
 * with vector_transformation(.., rotation, shift, vdim<3>) :58-73. */
static void orc_apply(const float* T, float x, float y, float z, float* o)
{
    o[0] = x * T[0] + y * T[1] + z * T[2] + T[9];
    o[1] = x * T[3] + y * T[4] + z * T[5] + T[10];
    o[2] = x * T[6] + y * T[7] + z * T[8] + T[11];
}

/* tipl::estimate<linear>(I, pos) for uint8 I: 0 when outside. */
ORC_API uint8_t orc_sample_u8(const uint8_t* I, const int* dim, float x, float y, float z)
{
    orc_tri t;
    if (!orc_tri_locate(dim, x, y, z, &t))
        return 0;
    return orc_tri_u8(I, &t);
}

/* ------------------------------------------------------------------------ */
/* Resample (numerical/resampling.hpp:1239-1270)                            */
/* ------------------------------------------------------------------------ */

/* draw()/crop() numerical/basic_op.hpp:470-520: copy the overlap of `from`
 * shifted by -pos into `to`; everything else stays 0. */
static void orc_crop(const uint8_t* from, const int* fdim, uint8_t* to, const int* tdim,
                     const int* from_pos)
{
    memset(to, 0, vox(tdim));
    for (int z = 0; z < tdim[2]; ++z) {
        int sz = z + from_pos[2];
        if (sz < 0 || sz >= fdim[2]) continue;
        for (int y = 0; y < tdim[1]; ++y) {
            int sy = y + from_pos[1];
            if (sy < 0 || sy >= fdim[1]) continue;
            for (int x = 0; x < tdim[0]; ++x) {
                int sx = x + from_pos[0];
                if (sx < 0 || sx >= fdim[0]) continue;
                to[((size_t)z * tdim[1] + y) * tdim[0] + x] =
                    from[((size_t)sz * fdim[1] + sy) * fdim[0] + sx];
            }
        }
    }
}

/* Returns 1 and the integer offset when T is an identity rotation with an
 * integer shift to within 1e-4 (numerical/resampling.hpp:1244-1262). */
ORC_API int orc_is_integer_shift(const float* T, int* from_pos)
{
    const double eps = 1.0e-4;
    for (int i = 0; i < 3; ++i) {
        long s = lroundf(T[9 + i]);
        if (fabsf(T[9 + i] - (float)s) > eps)
            return 0;
        from_pos[i] = (int)s;
        for (int j = 0; j < 3; ++j)
            if (fabsf(T[i * 3 + j] - (float)(i == j)) > eps)
                return 0;
    }
    return 1;
}

/* tipl::resample(from, to, trans) CPU branch: `to` has shape tdim and
 * receives from(T * p) for every voxel p of `to`. */
ORC_API void orc_resample_u8(const uint8_t* from, const int* fdim, uint8_t* to, const int* tdim,
                             const float* T)
{
    int from_pos[3];
    if (orc_is_integer_shift(T, from_pos)) {
        orc_crop(from, fdim, to, tdim, from_pos);
        return;
    }
    size_t i = 0;
    for (int z = 0; z < tdim[2]; ++z)
        for (int y = 0; y < tdim[1]; ++y)
            for (int x = 0; x < tdim[0]; ++x, ++i) {
                float p[3];
                orc_apply(T, (float)x, (float)y, (float)z, p);
                to[i] = orc_sample_u8(from, fdim, p[0], p[1], p[2]);
            }
}

/* Same, without the integer-shift crop shortcut (the reference's device
 * resample, numerical/resampling.hpp:1277-1299, has no shortcut). */
ORC_API void orc_resample_u8_nocrop(const uint8_t* from, const int* fdim, uint8_t* to,
                                    const int* tdim, const float* T)
{
    size_t i = 0;
    for (int z = 0; z < tdim[2]; ++z)
        for (int y = 0; y < tdim[1]; ++y)
            for (int x = 0; x < tdim[0]; ++x, ++i) {
                float p[3];
                orc_apply(T, (float)x, (float)y, (float)z, p);
                to[i] = orc_sample_u8(from, fdim, p[0], p[1], p[2]);
            }
}

/* ------------------------------------------------------------------------ */
/* Correlation cost                                                          */
/* ------------------------------------------------------------------------ */

/* Exact integer sums for tipl::correlation on uint8 data:
 * sum, square_sum, inner_product with int64 accumulation
 * (numerical/statistics.hpp:186-322, 491-536).  s = {Sx, Sxx, Sy, Syy, Sxy}. */
ORC_API void orc_corr_sums(const uint8_t* x, const uint8_t* y, size_t n, int64_t* s)
{
    int64_t sx = 0, sxx = 0, sy = 0, syy = 0, sxy = 0;
    for (size_t i = 0; i < n; ++i) {
        int64_t a = x[i], b = y[i];
        sx += a;
        sxx += a * a;
        sy += b;
        syy += b * b;
        sxy += a * b;
    }
    s[0] = sx; s[1] = sxx; s[2] = sy; s[3] = syy; s[4] = sxy;
}

/* tipl::correlation(x, y) (numerical/statistics.hpp:579-586, 407-443,
 * 555-560) from the exact sums, then reg::correlation's
 * `float c = ...; return 1.0f - c*c;` (reg/linear.hpp:34-35). */
ORC_API double orc_corr_cost_from_sums(const int64_t* s, size_t n)
{
    double dn = (double)n;
    double mean_x = (double)s[0] / dn;
    double mean_y = (double)s[2] / dn;
    double var_x = (double)s[1] / dn - mean_x * mean_x;
    double var_y = (double)s[3] / dn - mean_y * mean_y;
    double sd1 = var_x > 0.0 ? sqrt(var_x) : 0.0;
    double sd2 = var_y > 0.0 ? sqrt(var_y) : 0.0;
    double r;
    if (sd1 == 0.0 || sd2 == 0.0)
        r = 0.0;
    else {
        double cov = (double)s[4] / dn - mean_x * mean_y;
        r = cov / sd1 / sd2;
    }
    float c = (float)r;
    float cost = 1.0f - c * c;
    return (double)cost;
}

/* reg::correlation::operator() reg/linear.hpp:23-36: swaps the roles when
 * `from` is larger than `to` (inverting the transform), resamples `to`
 * onto `from`'s grid (with the integer-shift crop shortcut) and returns
 * 1 - r^2 as float.  Returns NAN if the swap needs a singular inverse. */
ORC_API double orc_corr_cost(const uint8_t* from, const int* fdim, const uint8_t* to,
                             const int* tdim, const float* T_in)
{
    float T[12];
    memcpy(T, T_in, sizeof(T));
    if (vox(fdim) > vox(tdim)) {
        orc_matrix_inverse(T);   /* the reference ignores the return value */
        const uint8_t* tp = from; from = to; to = tp;
        const int* dp = fdim; fdim = tdim; tdim = dp;
    }
    size_t n = vox(fdim);
    uint8_t* r = (uint8_t*)malloc(n ? n : 1);
    if (!r) return NAN;
    orc_resample_u8(to, tdim, r, fdim, T);
    int64_t s[5];
    orc_corr_sums(from, r, n, s);
    free(r);
    return orc_corr_cost_from_sums(s, n);
}

/* ------------------------------------------------------------------------ */
/* Mutual information                                                        */
/* ------------------------------------------------------------------------ */

/* get_mutual_info reg/linear.hpp:138-146: hist[(from << b) + to(T*p)] over the
 * `from` grid, with to(T*p)=0 when outside.  bins is 256 (the reference's
 * mi_band_width = 8, reg/linear.hpp:84-85) or 64; for 64 bins both values are
 * reduced by >> 2 (the 64-bin extension). */
ORC_API void orc_mi_hist(const uint8_t* from, const int* fdim, const uint8_t* to,
                         const int* tdim, const float* T, int bins, int32_t* hist)
{
    int shift = bins == 256 ? 0 : 2;
    int band = bins == 256 ? 8 : 6;
    memset(hist, 0, sizeof(int32_t) * (size_t)bins * bins);
    size_t i = 0;
    for (int z = 0; z < fdim[2]; ++z)
        for (int y = 0; y < fdim[1]; ++y)
            for (int x = 0; x < fdim[0]; ++x, ++i) {
                float p[3];
                orc_apply(T, (float)x, (float)y, (float)z, p);
                uint32_t t = orc_sample_u8(to, tdim, p[0], p[1], p[2]);
                uint32_t f = from[i];
                hist[((f >> shift) << band) + (t >> shift)]++;
            }
}

/* histogram(from, from_hist, 0, 255, 256) numerical/basic_op.hpp:1293-1311
 * for uint8 data is a plain count per value; with 64 bins, per value>>2. */
ORC_API void orc_value_hist(const uint8_t* v, size_t n, int bins, uint32_t* h)
{
    int shift = bins == 256 ? 0 : 2;
    memset(h, 0, sizeof(uint32_t) * (size_t)bins);
    for (size_t i = 0; i < n; ++i)
        h[v[i] >> shift]++;
}

/* get_mutual_info_mean reg/linear.hpp:87-104: column sums of the joint
 * histogram, then the FP64 sum of mu*log(mu/from_h[row]/to_h[col]) in
 * row-major order, divided by float(bins*bins). */
ORC_API double orc_mi_mean(const int32_t* hist, const uint32_t* from_hist, int bins)
{
    uint32_t to_hist[256];
    memset(to_hist, 0, sizeof(to_hist));
    for (int r = 0; r < bins; ++r)
        for (int c = 0; c < bins; ++c)
            to_hist[c] += (uint32_t)hist[r * bins + c];
    double sum = 0.0;
    for (int r = 0; r < bins; ++r)
        for (int c = 0; c < bins; ++c) {
            double mu = hist[r * bins + c];
            if (mu == 0.0)
                continue;
            sum += mu * log(mu / (double)from_hist[r] / (double)to_hist[c]);
        }
    return sum / (double)(float)(bins * bins);
}

/* mutual_information<3>::operator() reg/linear.hpp:173-203: -MI mean. */
ORC_API double orc_mi_cost(const uint8_t* from, const int* fdim, const uint8_t* to,
                           const int* tdim, const float* T, int bins)
{
    int32_t* hist = (int32_t*)malloc(sizeof(int32_t) * (size_t)bins * bins);
    if (!hist) return NAN;
    uint32_t fh[256];
    orc_value_hist(from, vox(fdim), bins, fh);
    orc_mi_hist(from, fdim, to, tdim, T, bins, hist);
    double v = -orc_mi_mean(hist, fh, bins);
    free(hist);
    return v;
}

/* ------------------------------------------------------------------------ */
/* Quantisation (numerical/numerical.hpp:909-927)                            */
/* ------------------------------------------------------------------------ */

/* normalize_upper_lower2(in, out, upper_limit) for float input: min/max in
 * float, out = (in - min) * (upper/range) truncated to uint8.  Returns 0 and
 * leaves `out` untouched when the range is 0. */
ORC_API int orc_normalize_f32(const float* in, size_t n, uint8_t* out, float upper_limit)
{
    if (n == 0) return 0;
    float mn = in[0], mx = in[0];
    for (size_t i = 1; i < n; ++i) {
        if (in[i] < mn) mn = in[i];
        if (in[i] > mx) mx = in[i];
    }
    float range = mx - mn;
    if (range == 0.0f) return 0;
    upper_limit /= range;
    for (size_t i = 0; i < n; ++i)
        out[i] = (uint8_t)((in[i] - mn) * upper_limit);
    return 1;
}

ORC_API int orc_normalize_i16(const int16_t* in, size_t n, uint8_t* out, float upper_limit)
{
    if (n == 0) return 0;
    float mn = in[0], mx = in[0];
    for (size_t i = 1; i < n; ++i) {
        if ((float)in[i] < mn) mn = in[i];
        if ((float)in[i] > mx) mx = in[i];
    }
    float range = mx - mn;
    if (range == 0.0f) return 0;
    upper_limit /= range;
    for (size_t i = 0; i < n; ++i)
        out[i] = (uint8_t)(((float)in[i] - mn) * upper_limit);
    return 1;
}

/* ------------------------------------------------------------------------ */
/* Pyramid (numerical/resampling.hpp:575-613, 849-858)                       */
/* ------------------------------------------------------------------------ */

/* downsample_with_padding for uint8: out shape (s+1)/2; each output is the
 * uint16 sum of the 2x2x2 block, with missing far-edge taps replaced by the
 * tap that drops the missing axes, shifted right by 3. */
ORC_API void orc_downsample_u8(const uint8_t* in, const int* dim, uint8_t* out, int* odim)
{
    int ow = (dim[0] + 1) / 2, oh = (dim[1] + 1) / 2, od = (dim[2] + 1) / 2;
    odim[0] = ow; odim[1] = oh; odim[2] = od;
    size_t w = dim[0], wh = (size_t)dim[0] * dim[1];
    size_t shift[8] = {0, 1, w, 1 + w, wh, wh + 1, wh + w, wh + 1 + w};
    size_t o = 0;
    for (int z = 0; z < od; ++z)
        for (int y = 0; y < oh; ++y)
            for (int x = 0; x < ow; ++x, ++o) {
                int x2 = 2 * x, y2 = 2 * y, z2 = 2 * z;
                size_t base = ((size_t)z2 * dim[1] + y2) * dim[0] + x2;
                int has = 0;
                if (x2 + 1 < dim[0]) has += 1;
                if (y2 + 1 < dim[1]) has += 2;
                if (z2 + 1 < dim[2]) has += 4;
                uint8_t buf[8];
                uint16_t sum = buf[0] = in[base];
                for (int i = 1; i < 8; ++i) {
                    int h = has & i;
                    buf[i] = (h == i) ? in[base + shift[i]] : buf[h];
                    sum += buf[i];
                }
                out[o] = (uint8_t)(sum >> 3);
            }
}

/* upsample_with_padding<linear> for a float3 field: out(p) = in(p/2) by
 * trilinear interpolation; outside samples stay 0. */
ORC_API void orc_upsample_f3(const float* in, const int* idim, float* out, const int* odim)
{
    memset(out, 0, sizeof(float) * 3 * vox(odim));
    size_t o = 0;
    for (int z = 0; z < odim[2]; ++z)
        for (int y = 0; y < odim[1]; ++y)
            for (int x = 0; x < odim[0]; ++x, ++o) {
                float p0 = (float)x, p1 = (float)y, p2 = (float)z;
                p0 *= 0.5f; p1 *= 0.5f; p2 *= 0.5f;
                orc_tri t;
                if (orc_tri_locate(idim, p0, p1, p2, &t))
                    orc_tri_f3(in, &t, out + o * 3);
            }
}

/* ------------------------------------------------------------------------ */
/* CDM iteration pieces                                                      */
/* ------------------------------------------------------------------------ */

/* compose_displacement numerical/dif.hpp:90-107: Js(x) = Is(x + d(x)), with
 * an exact copy where d(x) == 0 and 0 where the sample is outside. */
ORC_API void orc_compose_displacement(const uint8_t* Is, const int* dim, const float* d,
                                      uint8_t* Js)
{
    size_t i = 0;
    for (int z = 0; z < dim[2]; ++z)
        for (int y = 0; y < dim[1]; ++y)
            for (int x = 0; x < dim[0]; ++x, ++i) {
                const float* v = d + i * 3;
                if (v[0] == 0.0f && v[1] == 0.0f && v[2] == 0.0f) {
                    Js[i] = Is[i];
                    continue;
                }
                float p0 = (float)x, p1 = (float)y, p2 = (float)z;
                p0 += v[0]; p1 += v[1]; p2 += v[2];
                Js[i] = orc_sample_u8(Is, dim, p0, p1, p2);
            }
}

/* gradient_at numerical/numerical.hpp:102-109: central differences, 0 on the
 * first and last slab of each axis. */
static void orc_gradient_at(const uint8_t* I, const int* dim, int x, int y, int z, size_t pos,
                            float* g)
{
    size_t w = dim[0], wh = (size_t)dim[0] * dim[1];
    g[0] = (x + 1 < dim[0] && x) ? (float)((int)I[pos + 1] - (int)I[pos - 1]) : 0.0f;
    g[1] = (y + 1 < dim[1] && y) ? (float)((int)I[pos + w] - (int)I[pos - w]) : 0.0f;
    g[2] = (z + 1 < dim[2] && z) ? (float)((int)I[pos + wh] - (int)I[pos - wh]) : 0.0f;
}

/* cdm_get_gradient_r_imp reg/cdm.hpp:46-66 with get_window_at_width<2>
 * numerical/window.hpp:92-110 and linear_regression(.., a, b, r2)
 * numerical/statistics.hpp:694-713.  Window sums are exact integers, so the
 * traversal order of the window does not matter. */
static void orc_grad_r_voxel(const uint8_t* Js, const uint8_t* It, const int* dim, int x, int y,
                             int z, size_t pos, float* nd, float* cost)
{
    if (Js[pos] == 0 || It[pos] == 0)
        return;
    int x0 = x - 2 < 0 ? 0 : x - 2, x1 = x + 2 >= dim[0] ? dim[0] - 1 : x + 2;
    int y0 = y - 2 < 0 ? 0 : y - 2, y1 = y + 2 >= dim[1] ? dim[1] - 1 : y + 2;
    int z0 = z - 2 < 0 ? 0 : z - 2, z1 = z + 2 >= dim[2] ? dim[2] - 1 : z + 2;
    int64_t sx = 0, sxx = 0, sy = 0, syy = 0, sxy = 0, n = 0;
    for (int zz = z0; zz <= z1; ++zz)
        for (int yy = y0; yy <= y1; ++yy)
            for (int xx = x0; xx <= x1; ++xx) {
                size_t q = ((size_t)zz * dim[1] + yy) * dim[0] + xx;
                int64_t a = Js[q], b = It[q];
                sx += a; sxx += a * a; sy += b; syy += b * b; sxy += a * b; ++n;
            }
    double dn = (double)n;
    double mean_x = (double)sx / dn;
    double mean_y = (double)sy / dn;
    double x_var = (double)sxx / dn - mean_x * mean_x;
    double y_var = (double)syy / dn - mean_y * mean_y;
    float a, b, r2;
    if (x_var == 0.0 || y_var == 0.0) {
        a = 0.0f;
        b = (float)mean_y;
        r2 = 0.0f;
    } else {
        double cov = (double)sxy / dn - mean_x * mean_y;
        a = (float)(cov / x_var);
        b = (float)(mean_y - (double)a * mean_x);
        r2 = (float)(cov * cov / x_var / y_var);
    }
    if (a > 0.0f) {
        float g[3];
        orc_gradient_at(Js, dim, x, y, z, pos, g);
        float s = (float)Js[pos] * a;
        s = s + b;
        s = s - (float)It[pos];
        s = r2 * s;
        nd[pos * 3 + 0] = 0.0f + g[0] * s;
        nd[pos * 3 + 1] = 0.0f + g[1] * s;
        nd[pos * 3 + 2] = 0.0f + g[2] * s;
        cost[pos] = -r2;
    }
}

/* cdm_get_gradient_d_imp reg/cdm.hpp:68-78 (SSD gradient). */
static void orc_grad_d_voxel(const uint8_t* Js, const uint8_t* It, const int* dim, int x, int y,
                             int z, size_t pos, float* nd, float* cost)
{
    float g[3];
    orc_gradient_at(Js, dim, x, y, z, pos, g);
    int dif = (int)Js[pos] - (int)It[pos];
    nd[pos * 3 + 0] = 0.0f + g[0] * (float)dif;
    nd[pos * 3 + 1] = 0.0f + g[1] * (float)dif;
    nd[pos * 3 + 2] = 0.0f + g[2] * (float)dif;
    cost[pos] = (float)(dif * dif);
}

/* cdm_get_gradient reg/cdm.hpp:84-106: new_d is overwritten (zero where the
 * voxel contributes nothing); the return value is float(mean(cost_map)),
 * the mean taken in FP64 (numerical/statistics.hpp:316-321).  The sum runs
 * sequentially in voxel order here; the reference's threads add partial
 * sums in a schedule-dependent order, so the last bits of the double mean
 * may differ from the reference's. */
ORC_API float orc_cdm_gradient(const uint8_t* Js, const uint8_t* It, const int* dim, char type,
                               float* new_d)
{
    size_t n = vox(dim);
    memset(new_d, 0, sizeof(float) * 3 * n);
    float* cost = (float*)calloc(n ? n : 1, sizeof(float));
    if (!cost) return NAN;
    size_t i = 0;
    for (int z = 0; z < dim[2]; ++z)
        for (int y = 0; y < dim[1]; ++y)
            for (int x = 0; x < dim[0]; ++x, ++i) {
                if (type == 'd')
                    orc_grad_d_voxel(Js, It, dim, x, y, z, i, new_d, cost);
                else
                    orc_grad_r_voxel(Js, It, dim, x, y, z, i, new_d, cost);
            }
    double s = 0.0;
    for (size_t k = 0; k < n; ++k) s += (double)cost[k];
    free(cost);
    return n ? (float)(s / (double)n) : 0.0f;
}

/* cdm_solve_poisson reg/cdm.hpp:158-198: solve = new_d * (-1/6), then
 * `sweeps` Jacobi sweeps of the 7-point stencil with neighbours taken by
 * flat index (x +- 1 wraps across rows, y +- 1 across planes) and missing
 * neighbours dropped at the ends of the array.  Result replaces new_d. */
ORC_API void orc_cdm_poisson(float* new_d, const int* dim, int sweeps)
{
    size_t n = vox(dim);
    float* a = (float*)malloc(sizeof(float) * 3 * (n ? n : 1));
    float* b = (float*)malloc(sizeof(float) * 3 * (n ? n : 1));
    if (!a || !b) { free(a); free(b); return; }
    const float inv_d2 = -0.5f / 3.0f;
    for (size_t i = 0; i < 3 * n; ++i) a[i] = new_d[i] * inv_d2;
    long w = dim[0], wh = (long)dim[0] * dim[1], size = (long)n;
    for (int it = 0; it < sweeps; ++it) {
        for (long p = 0; p < size; ++p)
            for (int c = 0; c < 3; ++c) {
                float v = new_d[p * 3 + c];
                if (p >= 1) v -= a[(p - 1) * 3 + c];
                if (p < size - 1) v -= a[(p + 1) * 3 + c];
                if (p >= w) v -= a[(p - w) * 3 + c];
                if (p < size - w) v -= a[(p + w) * 3 + c];
                if (p >= wh) v -= a[(p - wh) * 3 + c];
                if (p < size - wh) v -= a[(p + wh) * 3 + c];
                v *= inv_d2;
                b[p * 3 + c] = v;
            }
        float* t = a; a = b; b = t;
    }
    memcpy(new_d, a, sizeof(float) * 3 * n);
    free(a);
    free(b);
}

/* cdm_max_displacement_length reg/cdm.hpp:269-284 with
 * vector<3>::length() utility/pixel_index.hpp:558: float(sqrt(double(x*x+y*y+z*z))). */
ORC_API float orc_cdm_max_length(const float* d, size_t n)
{
    float m = 0.0f;
    for (size_t i = 0; i < n; ++i) {
        float s = d[i * 3] * d[i * 3] + d[i * 3 + 1] * d[i * 3 + 1];
        s = s + d[i * 3 + 2] * d[i * 3 + 2];
        float l = (float)sqrt((double)s);
        if (l > m) m = l;
    }
    return m;
}

/* accumulate_displacement numerical/dif.hpp:166-186: with the mapping
 * m(x) = cur(x) + x, new(x) = m(x + new(x)) - x by trilinear interpolation;
 * where new(x) == 0 or the sample is outside, new(x) = cur(x). */
ORC_API void orc_accumulate_displacement(const float* cur, float* nd, const int* dim)
{
    size_t n = vox(dim);
    float* map = (float*)malloc(sizeof(float) * 3 * (n ? n : 1));
    if (!map) return;
    size_t i = 0;
    for (int z = 0; z < dim[2]; ++z)
        for (int y = 0; y < dim[1]; ++y)
            for (int x = 0; x < dim[0]; ++x, ++i) {
                map[i * 3 + 0] = cur[i * 3 + 0] + (float)x;
                map[i * 3 + 1] = cur[i * 3 + 1] + (float)y;
                map[i * 3 + 2] = cur[i * 3 + 2] + (float)z;
            }
    i = 0;
    for (int z = 0; z < dim[2]; ++z)
        for (int y = 0; y < dim[1]; ++y)
            for (int x = 0; x < dim[0]; ++x, ++i) {
                float* v = nd + i * 3;
                int moved = 0;
                if (v[0] != 0.0f || v[1] != 0.0f || v[2] != 0.0f) {
                    float p0 = v[0] + (float)x, p1 = v[1] + (float)y, p2 = v[2] + (float)z;
                    orc_tri t;
                    if (orc_tri_locate(dim, p0, p1, p2, &t)) {
                        float r[3];
                        orc_tri_f3(map, &t, r);
                        v[0] = r[0] - (float)x;
                        v[1] = r[1] - (float)y;
                        v[2] = r[2] - (float)z;
                        moved = 1;
                    }
                }
                if (!moved) {
                    v[0] = cur[i * 3 + 0];
                    v[1] = cur[i * 3 + 1];
                    v[2] = cur[i * 3 + 2];
                }
            }
    free(map);
}

/* cdm_smooth reg/cdm.hpp:302-342: dd = d*(1-s) + (s/6)*(sum of the 6
 * flat-index neighbours, added in the order +1, -1, +w, -w, +wh, -wh).
 * Does nothing when smoothing == 0. */
ORC_API void orc_cdm_smooth(const float* d, float* dd, const int* dim, float smoothing)
{
    if (smoothing == 0.0f)
        return;
    size_t n = vox(dim), w = dim[0], wh = (size_t)dim[0] * dim[1];
    float w_6 = smoothing / (2.0f * 3.0f);
    float w_1 = 1.0f - smoothing;
    for (size_t i = 0; i < n; ++i)
        for (int c = 0; c < 3; ++c) {
            float v = 0.0f;
            if (i + 1 < n) v += d[(i + 1) * 3 + c];
            if (i >= 1) v += d[(i - 1) * 3 + c];
            if (i + w < n) v += d[(i + w) * 3 + c];
            if (i >= w) v += d[(i - w) * 3 + c];
            if (i + wh < n) v += d[(i + wh) * 3 + c];
            if (i >= wh) v += d[(i - wh) * 3 + c];
            v *= w_6;
            dd[i * 3 + c] = d[i * 3 + c] * w_1 + v;
        }
}

/* cdm_improved reg/cdm.hpp:19-34 with linear_regression(int*, float*, a, b, r2)
 * numerical/statistics.hpp:694-713.  `cost` holds *n floats; the oldest entry
 * is dropped (in place) when more than 7 are held.  Returns 0 when the cost
 * trend is rising (a > 0). */
ORC_API int orc_cdm_improved(float* cost, int* n)
{
    int sz = *n;
    if (sz > 5) {
        int64_t sx = 0, sxx = 0;
        double sy = 0.0, syy = 0.0, sxy = 0.0;
        for (int i = 0; i < sz; ++i) {
            sx += i;
            sxx += (int64_t)i * i;
            sy += (double)cost[i];
            syy += (double)cost[i] * (double)cost[i];
            sxy += (double)((float)i * cost[i]);
        }
        double dn = (double)sz;
        double mean_x = (double)sx / dn;
        double mean_y = sy / dn;
        double x_var = (double)sxx / dn - mean_x * mean_x;
        double y_var = syy / dn - mean_y * mean_y;
        float a;
        if (x_var == 0.0 || y_var == 0.0)
            a = 0.0f;
        else {
            double cov = sxy / dn - mean_x * mean_y;
            a = (float)(cov / x_var);
        }
        if (a > 0.0f)
            return 0;
        if (sz > 7) {
            memmove(cost, cost + 1, sizeof(float) * (size_t)(sz - 1));
            *n = sz - 1;
        }
    }
    return 1;
}

/* ------------------------------------------------------------------------ */
/* Whole CDM (reg/cdm.hpp:373-474)                                           */
/* ------------------------------------------------------------------------ */

typedef struct {
    float resolution;     /* 2.0 */
    float speed;          /* 0.3 */
    float smoothing;      /* 0.05 */
    int min_dimension;    /* 16 */
    int max_iterations;   /* 200 */
    char gradient_type;   /* 'r' or 'd' */
} orc_cdm_param;

static int min3(const int* d)
{
    int m = d[0];
    if (d[1] < m) m = d[1];
    if (d[2] < m) m = d[2];
    return m;
}

/* One call of cdm() for `nmod` modality pairs at shape dim.  best_d holds
 * vox(dim) float3 and is zero on entry (the reference resizes an empty image).
 * *iterations accumulates the loop iterations run at every level. */
static int orc_cdm_rec(const uint8_t* const* It, const uint8_t* const* Is, int nmod,
                       const int* dim, float* best_d, orc_cdm_param p, int* iterations)
{
    size_t n = vox(dim);
    if (min3(dim) > p.min_dimension) {
        int rdim[3] = {(dim[0] + 1) / 2, (dim[1] + 1) / 2, (dim[2] + 1) / 2};
        size_t rn = vox(rdim);
        uint8_t* buf = (uint8_t*)malloc(2 * (size_t)nmod * rn);
        float* rd = (float*)calloc(3 * rn, sizeof(float));
        const uint8_t* rIt[16];
        const uint8_t* rIs[16];
        if (!buf || !rd || nmod > 16) { free(buf); free(rd); return 0; }
        for (int i = 0; i < nmod; ++i) {
            int od[3];
            orc_downsample_u8(It[i], dim, buf + (size_t)(2 * i) * rn, od);
            orc_downsample_u8(Is[i], dim, buf + (size_t)(2 * i + 1) * rn, od);
            rIt[i] = buf + (size_t)(2 * i) * rn;
            rIs[i] = buf + (size_t)(2 * i + 1) * rn;
        }
        orc_cdm_param p2 = p;
        p2.resolution /= 2.0f;
        int ok = orc_cdm_rec(rIt, rIs, nmod, rdim, rd, p2, iterations);
        free(buf);
        if (!ok) { free(rd); return 0; }
        for (size_t k = 0; k < 3 * rn; ++k) rd[k] *= 2.0f;
        orc_upsample_f3(rd, rdim, best_d, dim);
        free(rd);
        if (p.resolution > 1.0f)
            return 1;
    }

    float* cur = (float*)malloc(sizeof(float) * 3 * (n ? n : 1));
    float* nd = (float*)malloc(sizeof(float) * 3 * (n ? n : 1));
    float* dd = (float*)malloc(sizeof(float) * 3 * (n ? n : 1));
    uint8_t* Js = (uint8_t*)malloc(n ? n : 1);
    if (!cur || !nd || !dd || !Js) { free(cur); free(nd); free(dd); free(Js); return 0; }
    memcpy(cur, best_d, sizeof(float) * 3 * n);
    float theta = 0.0f;
    float cost[16];
    int ncost = 0;
    float best_cost = 3.402823466e+38f;
    float cur_smoothing = 0.8f;
    int ok = 1;
    for (int index = 0; index < p.max_iterations; ++index) {
        ++*iterations;
        float sum_cost = 0.0f;
        for (int i = 0; i < nmod; ++i) {
            orc_compose_displacement(Is[i], dim, cur, Js);
            float c = orc_cdm_gradient(Js, It[i], dim, p.gradient_type, i == 0 ? nd : dd);
            sum_cost += c;
            if (sum_cost == 0.0f) { ok = 0; goto done; }
            if (i)
                for (size_t k = 0; k < 3 * n; ++k) nd[k] += dd[k];
        }
        cost[ncost++] = sum_cost / (float)nmod;
        if (cost[ncost - 1] < best_cost) {
            best_cost = cost[ncost - 1];
            if (index)
                memcpy(best_d, cur, sizeof(float) * 3 * n);
        }
        if (!orc_cdm_improved(cost, &ncost)) {
            if (cur_smoothing <= p.smoothing)
                break;
            memcpy(cur, best_d, sizeof(float) * 3 * n);
            cur_smoothing *= 0.5f;
            ncost = 0;
        }
        orc_cdm_poisson(nd, dim, 12);
        if (theta == 0.0f)
            theta = orc_cdm_max_length(nd, n);
        if (theta == 0.0f)
            break;
        float k = p.speed / theta;
        for (size_t j = 0; j < 3 * n; ++j) nd[j] *= k;
        orc_accumulate_displacement(cur, nd, dim);
        orc_cdm_smooth(nd, cur, dim, cur_smoothing);
    }
done:
    free(cur); free(nd); free(dd); free(Js);
    return ok;
}

/* tipl::reg::cdm(It, Is, best_d, terminated, param) reg/cdm.hpp:373-474 for
 * nmod uint8 modality pairs of shape dim.  `field` receives vox(dim) float3.
 * Returns 1 on success, 0 when the reference returns false. */
ORC_API int orc_cdm(const uint8_t* const* It, const uint8_t* const* Is, int nmod, const int* dim,
                    float resolution, float speed, float smoothing, int min_dimension,
                    int max_iterations, char gradient_type, float* field, int* iterations)
{
    orc_cdm_param p;
    p.resolution = resolution;
    p.speed = speed;
    p.smoothing = smoothing;
    p.min_dimension = min_dimension;
    p.max_iterations = max_iterations;
    p.gradient_type = gradient_type;
    memset(field, 0, sizeof(float) * 3 * vox(dim));
    int it = 0;
    int ok = orc_cdm_rec(It, Is, nmod, dim, field, p, &it);
    if (iterations) *iterations = it;
    return ok;
}

/* ------------------------------------------------------------------------ */
/* 2-D CDM (cdm<out,uint8,2>, cdm_cuda<out,uint8,2>: reg/cdm_cu.hpp:12-17)   */
/* ------------------------------------------------------------------------ */
/* Images are x-fastest, dim[2] = {w, h}; fields are packed float2 per pixel
 * (tipl::image<2,tipl::vector<2>>).  Each function is the 2-D case of the
 * same reference template as its 3-D counterpart above. */

static size_t pix(const int* dim) { return (size_t)dim[0] * (size_t)dim[1]; }

/* interpolator::linear<2>::get_location numerical/interpolation.hpp:283-306
 * (the same outside rules as the 3-D locate). */
typedef struct {
    size_t idx[4];
    float w[4];
} orc_bil;

static int orc_bil_locate(const int* dim, float x, float y, orc_bil* t)
{
    if (x < 0.0f || y < 0.0f)
        return 0;
    if (isnan(x) || isnan(y))
        return 0;
    if (x >= 16777216.0f || y >= 16777216.0f)
        return 0;
    size_t ix = (size_t)x, iy = (size_t)y, w = (size_t)dim[0];
    if (ix + 1 >= w || iy + 1 >= (size_t)dim[1])
        return 0;
    t->idx[0] = iy * w + ix;
    t->idx[1] = t->idx[0] + 1;
    t->idx[2] = t->idx[0] + w;
    t->idx[3] = t->idx[2] + 1;
    float p0 = x - (float)ix;
    float p1 = y - (float)iy;
    float n0 = 1.0f - p0;
    float n1 = 1.0f - p1;
    t->w[0] = n0 * n1;
    t->w[1] = p0 * n1;
    t->w[2] = n0 * p1;
    t->w[3] = p0 * p1;
    return 1;
}

/* weighted_sum (numerical/interpolation.hpp:61-73) of uint8 data, then the
 * uint8 assignment (:25-35). */
static uint8_t orc_bil_u8(const uint8_t* I, const orc_bil* t)
{
    float r = (float)I[t->idx[0]] * t->w[0];
    for (int k = 1; k < 4; ++k) {
        float v = (float)I[t->idx[k]];
        v *= t->w[k];
        r += v;
    }
    return r <= 0.0f ? 0 : (uint8_t)r;
}

/* weighted_sum of vector<2,float> data. */
static void orc_bil_f2(const float* F, const orc_bil* t, float* out)
{
    for (int c = 0; c < 2; ++c) {
        float r = F[t->idx[0] * 2 + c] * t->w[0];
        for (int k = 1; k < 4; ++k) {
            float v = F[t->idx[k] * 2 + c];
            v *= t->w[k];
            r += v;
        }
        out[c] = r;
    }
}

/* tipl::estimate<linear>(I, pos) for a 2-D uint8 image: 0 when outside. */
ORC_API uint8_t orc2_sample_u8(const uint8_t* I, const int* dim, float x, float y)
{
    orc_bil t;
    if (!orc_bil_locate(dim, x, y, &t))
        return 0;
    return orc_bil_u8(I, &t);
}

/* downsample_with_padding numerical/resampling.hpp:575-613 in 2-D: shape
 * (s+1)/2, the uint16 sum of the 2x2 block (a missing far-edge tap replaced by
 * the tap that drops the missing axes) shifted right by 2. */
ORC_API void orc2_downsample_u8(const uint8_t* in, const int* dim, uint8_t* out, int* odim)
{
    int ow = (dim[0] + 1) / 2, oh = (dim[1] + 1) / 2;
    odim[0] = ow; odim[1] = oh;
    size_t w = dim[0];
    size_t shift[4] = {0, 1, w, 1 + w};
    size_t o = 0;
    for (int y = 0; y < oh; ++y)
        for (int x = 0; x < ow; ++x, ++o) {
            int x2 = 2 * x, y2 = 2 * y;
            size_t base = (size_t)y2 * w + x2;
            int has = 0;
            if (x2 + 1 < dim[0]) has += 1;
            if (y2 + 1 < dim[1]) has += 2;
            uint8_t buf[4];
            uint16_t sum = buf[0] = in[base];
            for (int i = 1; i < 4; ++i) {
                int h = has & i;
                buf[i] = (h == i) ? in[base + shift[i]] : buf[h];
                sum += buf[i];
            }
            out[o] = (uint8_t)(sum >> 2);
        }
}

/* upsample_with_padding<linear> (numerical/resampling.hpp:849-858) of a
 * float2 field: out(p) = in(p * 0.5) bilinearly; outside samples stay 0. */
ORC_API void orc2_upsample_f2(const float* in, const int* idim, float* out, const int* odim)
{
    memset(out, 0, sizeof(float) * 2 * pix(odim));
    size_t o = 0;
    for (int y = 0; y < odim[1]; ++y)
        for (int x = 0; x < odim[0]; ++x, ++o) {
            float p0 = (float)x, p1 = (float)y;
            p0 *= 0.5f; p1 *= 0.5f;
            orc_bil t;
            if (orc_bil_locate(idim, p0, p1, &t))
                orc_bil_f2(in, &t, out + o * 2);
        }
}

/* compose_displacement numerical/dif.hpp:90-107 in 2-D. */
ORC_API void orc2_compose_displacement(const uint8_t* Is, const int* dim, const float* d,
                                       uint8_t* Js)
{
    size_t i = 0;
    for (int y = 0; y < dim[1]; ++y)
        for (int x = 0; x < dim[0]; ++x, ++i) {
            const float* v = d + i * 2;
            if (v[0] == 0.0f && v[1] == 0.0f) {
                Js[i] = Is[i];
                continue;
            }
            float p0 = (float)x, p1 = (float)y;
            p0 += v[0]; p1 += v[1];
            Js[i] = orc2_sample_u8(Is, dim, p0, p1);
        }
}

/* gradient_at numerical/numerical.hpp:112-117 (the 2-D overload). */
static void orc2_gradient_at(const uint8_t* I, const int* dim, int x, int y, size_t pos, float* g)
{
    size_t w = dim[0];
    g[0] = (x + 1 < dim[0] && x) ? (float)((int)I[pos + 1] - (int)I[pos - 1]) : 0.0f;
    g[1] = (y + 1 < dim[1] && y) ? (float)((int)I[pos + w] - (int)I[pos - w]) : 0.0f;
}

/* cdm_get_gradient_r_imp reg/cdm.hpp:46-66 with the 2-D
 * get_window_at_width<2> (numerical/window.hpp:73-84, a 5x5 window clipped
 * at the edges) and linear_regression (numerical/statistics.hpp:694-713). */
static void orc2_grad_r_pixel(const uint8_t* Js, const uint8_t* It, const int* dim, int x, int y,
                              size_t pos, float* nd, float* cost)
{
    if (Js[pos] == 0 || It[pos] == 0)
        return;
    int x0 = x - 2 < 0 ? 0 : x - 2, x1 = x + 2 >= dim[0] ? dim[0] - 1 : x + 2;
    int y0 = y - 2 < 0 ? 0 : y - 2, y1 = y + 2 >= dim[1] ? dim[1] - 1 : y + 2;
    int64_t sx = 0, sxx = 0, sy = 0, syy = 0, sxy = 0, n = 0;
    for (int yy = y0; yy <= y1; ++yy)
        for (int xx = x0; xx <= x1; ++xx) {
            size_t q = (size_t)yy * dim[0] + xx;
            int64_t a = Js[q], b = It[q];
            sx += a; sxx += a * a; sy += b; syy += b * b; sxy += a * b; ++n;
        }
    double dn = (double)n;
    double mean_x = (double)sx / dn;
    double mean_y = (double)sy / dn;
    double x_var = (double)sxx / dn - mean_x * mean_x;
    double y_var = (double)syy / dn - mean_y * mean_y;
    float a, b, r2;
    if (x_var == 0.0 || y_var == 0.0) {
        a = 0.0f;
        b = (float)mean_y;
        r2 = 0.0f;
    } else {
        double cov = (double)sxy / dn - mean_x * mean_y;
        a = (float)(cov / x_var);
        b = (float)(mean_y - (double)a * mean_x);
        r2 = (float)(cov * cov / x_var / y_var);
    }
    if (a > 0.0f) {
        float g[2];
        orc2_gradient_at(Js, dim, x, y, pos, g);
        float s = (float)Js[pos] * a;
        s = s + b;
        s = s - (float)It[pos];
        s = r2 * s;
        nd[pos * 2 + 0] = 0.0f + g[0] * s;
        nd[pos * 2 + 1] = 0.0f + g[1] * s;
        cost[pos] = -r2;
    }
}

/* cdm_get_gradient_d_imp reg/cdm.hpp:68-78 in 2-D. */
static void orc2_grad_d_pixel(const uint8_t* Js, const uint8_t* It, const int* dim, int x, int y,
                              size_t pos, float* nd, float* cost)
{
    float g[2];
    orc2_gradient_at(Js, dim, x, y, pos, g);
    int dif = (int)Js[pos] - (int)It[pos];
    nd[pos * 2 + 0] = 0.0f + g[0] * (float)dif;
    nd[pos * 2 + 1] = 0.0f + g[1] * (float)dif;
    cost[pos] = (float)(dif * dif);
}

/* cdm_get_gradient reg/cdm.hpp:84-106 in 2-D (the FP64 mean summed in pixel
 * order, as orc_cdm_gradient). */
ORC_API float orc2_cdm_gradient(const uint8_t* Js, const uint8_t* It, const int* dim, char type,
                                float* new_d)
{
    size_t n = pix(dim);
    memset(new_d, 0, sizeof(float) * 2 * n);
    float* cost = (float*)calloc(n ? n : 1, sizeof(float));
    if (!cost) return NAN;
    size_t i = 0;
    for (int y = 0; y < dim[1]; ++y)
        for (int x = 0; x < dim[0]; ++x, ++i) {
            if (type == 'd')
                orc2_grad_d_pixel(Js, It, dim, x, y, i, new_d, cost);
            else
                orc2_grad_r_pixel(Js, It, dim, x, y, i, new_d, cost);
        }
    double s = 0.0;
    for (size_t k = 0; k < n; ++k) s += (double)cost[k];
    free(cost);
    return n ? (float)(s / (double)n) : 0.0f;
}

/* cdm_solve_poisson reg/cdm.hpp:158-198 in 2-D: solve = new_d * (-1/4), then
 * `sweeps` Jacobi sweeps of the 5-point flat-index stencil (neighbours in the
 * order -1, +1, -w, +w; missing ones dropped at the ends of the array). */
ORC_API void orc2_cdm_poisson(float* new_d, const int* dim, int sweeps)
{
    size_t n = pix(dim);
    float* a = (float*)malloc(sizeof(float) * 2 * (n ? n : 1));
    float* b = (float*)malloc(sizeof(float) * 2 * (n ? n : 1));
    if (!a || !b) { free(a); free(b); return; }
    const float inv_d2 = -0.5f / 2.0f;
    for (size_t i = 0; i < 2 * n; ++i) a[i] = new_d[i] * inv_d2;
    long w = dim[0], size = (long)n;
    for (int it = 0; it < sweeps; ++it) {
        for (long p = 0; p < size; ++p)
            for (int c = 0; c < 2; ++c) {
                float v = new_d[p * 2 + c];
                if (p >= 1) v -= a[(p - 1) * 2 + c];
                if (p < size - 1) v -= a[(p + 1) * 2 + c];
                if (p >= w) v -= a[(p - w) * 2 + c];
                if (p < size - w) v -= a[(p + w) * 2 + c];
                v *= inv_d2;
                b[p * 2 + c] = v;
            }
        float* t = a; a = b; b = t;
    }
    memcpy(new_d, a, sizeof(float) * 2 * n);
    free(a);
    free(b);
}

/* cdm_max_displacement_length reg/cdm.hpp:269-284 with vector<2>::length()
 * utility/pixel_index.hpp:340: std::sqrt of the float x*x+y*y. */
ORC_API float orc2_cdm_max_length(const float* d, size_t n)
{
    float m = 0.0f;
    for (size_t i = 0; i < n; ++i) {
        float s = d[i * 2] * d[i * 2] + d[i * 2 + 1] * d[i * 2 + 1];
        float l = sqrtf(s);
        if (l > m) m = l;
    }
    return m;
}

/* accumulate_displacement numerical/dif.hpp:166-186 in 2-D. */
ORC_API void orc2_accumulate_displacement(const float* cur, float* nd, const int* dim)
{
    size_t n = pix(dim);
    float* map = (float*)malloc(sizeof(float) * 2 * (n ? n : 1));
    if (!map) return;
    size_t i = 0;
    for (int y = 0; y < dim[1]; ++y)
        for (int x = 0; x < dim[0]; ++x, ++i) {
            map[i * 2 + 0] = cur[i * 2 + 0] + (float)x;
            map[i * 2 + 1] = cur[i * 2 + 1] + (float)y;
        }
    i = 0;
    for (int y = 0; y < dim[1]; ++y)
        for (int x = 0; x < dim[0]; ++x, ++i) {
            float* v = nd + i * 2;
            int moved = 0;
            if (v[0] != 0.0f || v[1] != 0.0f) {
                float p0 = v[0] + (float)x, p1 = v[1] + (float)y;
                orc_bil t;
                if (orc_bil_locate(dim, p0, p1, &t)) {
                    float r[2];
                    orc_bil_f2(map, &t, r);
                    v[0] = r[0] - (float)x;
                    v[1] = r[1] - (float)y;
                    moved = 1;
                }
            }
            if (!moved) {
                v[0] = cur[i * 2 + 0];
                v[1] = cur[i * 2 + 1];
            }
        }
    free(map);
}

/* cdm_smooth reg/cdm.hpp:302-342 in 2-D: dd = d*(1-s) + (s/4)*(the four
 * flat-index neighbours, added in the order +1, -1, +w, -w).  Does nothing
 * when smoothing == 0. */
ORC_API void orc2_cdm_smooth(const float* d, float* dd, const int* dim, float smoothing)
{
    if (smoothing == 0.0f)
        return;
    size_t n = pix(dim), w = dim[0];
    float w_6 = smoothing / (2.0f * 2.0f);
    float w_1 = 1.0f - smoothing;
    for (size_t i = 0; i < n; ++i)
        for (int c = 0; c < 2; ++c) {
            float v = 0.0f;
            if (i + 1 < n) v += d[(i + 1) * 2 + c];
            if (i >= 1) v += d[(i - 1) * 2 + c];
            if (i + w < n) v += d[(i + w) * 2 + c];
            if (i >= w) v += d[(i - w) * 2 + c];
            v *= w_6;
            dd[i * 2 + c] = d[i * 2 + c] * w_1 + v;
        }
}

/* One call of the 2-D cdm() (reg/cdm.hpp:373-474), as orc_cdm_rec. */
static int orc2_cdm_rec(const uint8_t* const* It, const uint8_t* const* Is, int nmod,
                        const int* dim, float* best_d, orc_cdm_param p, int* iterations)
{
    size_t n = pix(dim);
    int mn = dim[0] < dim[1] ? dim[0] : dim[1];
    if (mn > p.min_dimension) {
        int rdim[2] = {(dim[0] + 1) / 2, (dim[1] + 1) / 2};
        size_t rn = pix(rdim);
        uint8_t* buf = (uint8_t*)malloc(2 * (size_t)nmod * rn);
        float* rd = (float*)calloc(2 * rn, sizeof(float));
        const uint8_t* rIt[16];
        const uint8_t* rIs[16];
        if (!buf || !rd || nmod > 16) { free(buf); free(rd); return 0; }
        for (int i = 0; i < nmod; ++i) {
            int od[2];
            orc2_downsample_u8(It[i], dim, buf + (size_t)(2 * i) * rn, od);
            orc2_downsample_u8(Is[i], dim, buf + (size_t)(2 * i + 1) * rn, od);
            rIt[i] = buf + (size_t)(2 * i) * rn;
            rIs[i] = buf + (size_t)(2 * i + 1) * rn;
        }
        orc_cdm_param p2 = p;
        p2.resolution /= 2.0f;
        int ok = orc2_cdm_rec(rIt, rIs, nmod, rdim, rd, p2, iterations);
        free(buf);
        if (!ok) { free(rd); return 0; }
        for (size_t k = 0; k < 2 * rn; ++k) rd[k] *= 2.0f;
        orc2_upsample_f2(rd, rdim, best_d, dim);
        free(rd);
        if (p.resolution > 1.0f)
            return 1;
    }

    float* cur = (float*)malloc(sizeof(float) * 2 * (n ? n : 1));
    float* nd = (float*)malloc(sizeof(float) * 2 * (n ? n : 1));
    float* dd = (float*)malloc(sizeof(float) * 2 * (n ? n : 1));
    uint8_t* Js = (uint8_t*)malloc(n ? n : 1);
    if (!cur || !nd || !dd || !Js) { free(cur); free(nd); free(dd); free(Js); return 0; }
    memcpy(cur, best_d, sizeof(float) * 2 * n);
    float theta = 0.0f;
    float cost[16];
    int ncost = 0;
    float best_cost = 3.402823466e+38f;
    float cur_smoothing = 0.8f;
    int ok = 1;
    for (int index = 0; index < p.max_iterations; ++index) {
        ++*iterations;
        float sum_cost = 0.0f;
        for (int i = 0; i < nmod; ++i) {
            orc2_compose_displacement(Is[i], dim, cur, Js);
            float c = orc2_cdm_gradient(Js, It[i], dim, p.gradient_type, i == 0 ? nd : dd);
            sum_cost += c;
            if (sum_cost == 0.0f) { ok = 0; goto done; }
            if (i)
                for (size_t k = 0; k < 2 * n; ++k) nd[k] += dd[k];
        }
        cost[ncost++] = sum_cost / (float)nmod;
        if (cost[ncost - 1] < best_cost) {
            best_cost = cost[ncost - 1];
            if (index)
                memcpy(best_d, cur, sizeof(float) * 2 * n);
        }
        if (!orc_cdm_improved(cost, &ncost)) {
            if (cur_smoothing <= p.smoothing)
                break;
            memcpy(cur, best_d, sizeof(float) * 2 * n);
            cur_smoothing *= 0.5f;
            ncost = 0;
        }
        orc2_cdm_poisson(nd, dim, 12);
        if (theta == 0.0f)
            theta = orc2_cdm_max_length(nd, n);
        if (theta == 0.0f)
            break;
        float k = p.speed / theta;
        for (size_t j = 0; j < 2 * n; ++j) nd[j] *= k;
        orc2_accumulate_displacement(cur, nd, dim);
        orc2_cdm_smooth(nd, cur, dim, cur_smoothing);
    }
done:
    free(cur); free(nd); free(dd); free(Js);
    return ok;
}

/* tipl::reg::cdm for 2-D uint8 images (reg/cdm.hpp:373-474; the reference's
 * GPU instantiation cdm_cuda<out,uint8,2>, reg/cdm_cu.hpp:12-17, runs the same
 * template).  `field` receives pix(dim) float2.  Returns 1 on success, 0
 * when the reference returns false. */
ORC_API int orc2_cdm(const uint8_t* const* It, const uint8_t* const* Is, int nmod, const int* dim,
                     float resolution, float speed, float smoothing, int min_dimension,
                     int max_iterations, char gradient_type, float* field, int* iterations)
{
    orc_cdm_param p;
    p.resolution = resolution;
    p.speed = speed;
    p.smoothing = smoothing;
    p.min_dimension = min_dimension;
    p.max_iterations = max_iterations;
    p.gradient_type = gradient_type;
    memset(field, 0, sizeof(float) * 2 * pix(dim));
    int it = 0;
    int ok = orc2_cdm_rec(It, Is, nmod, dim, field, p, &it);
    if (iterations) *iterations = it;
    return ok;
}

/* filter::gaussian filter/gaussian.hpp:18-89 on a float image of `ch`
 * interleaved channels (1 = scalar image, 3 = vector<3> field): `bandwidth`
 * rounds of the separable [1 2 1] stencil along x, then y, then z, each pass
 * written to its own buffer as b = (a[left] + (c + c)) + a[right] with the
 * edge neighbour replaced by the centre (replicate padding); finally every
 * value is scaled by 2^(-6*bandwidth).  In place on `img`. */
ORC_API void orc_gaussian_f32(float* img, const int* dim, int ch, unsigned bandwidth)
{
    size_t n = vox(dim);
    if (n == 0 || bandwidth == 0)
        return;
    size_t w = dim[0], h = dim[1], d = dim[2], wh = w * h;
    float* a = (float*)malloc(sizeof(float) * n * ch);
    float* b = (float*)malloc(sizeof(float) * n * ch);
    memcpy(a, img, sizeof(float) * n * ch);
    for (unsigned k = 0; k < bandwidth; ++k) {
        for (int axis = 0; axis < 3; ++axis) {
            for (size_t i = 0; i < n; ++i) {
                size_t lo, hi;
                if (axis == 0) {
                    size_t x = i % w;
                    lo = x ? i - 1 : i;
                    hi = x + 1 < w ? i + 1 : i;
                } else if (axis == 1) {
                    size_t y = (i / w) % h;
                    lo = y ? i - w : i;
                    hi = y + 1 < h ? i + w : i;
                } else {
                    size_t z = i / wh;
                    lo = z ? i - wh : i;
                    hi = z + 1 < d ? i + wh : i;
                }
                for (int c = 0; c < ch; ++c) {
                    float cc = a[i * ch + c];
                    cc += cc;
                    b[i * ch + c] = (a[lo * ch + c] + cc) + a[hi * ch + c];
                }
            }
            float* t = a;
            a = b;
            b = t;
        }
    }
    double scale = ldexp(1.0, -(int)(6 * bandwidth));
    for (size_t j = 0; j < n * (size_t)ch; ++j)
        img[j] = (float)((double)a[j] * scale);
    free(a);
    free(b);
}

/* A fixed-schedule CDM (an extension; the reference has no such mode):
 * `levels` pyramid levels, each the downsample_with_padding of the one
 * above, all run, coarsest first, for exactly `level_iterations` iterations
 * of the reference loop body (reg/cdm.hpp:418-472: compose, gradient,
 * 12-sweep Poisson, theta on a level's first iteration, scale,
 * accumulate), keeping the running field (no cost-driven control).  Each
 * iteration smooths the field with filter::gaussian(bandwidth) of new_d, or
 * with cdm_smooth(new_d, smoothing) when bandwidth is 0.  A level's result
 * is doubled and upsample_with_padding'ed into the next.  `field` (vox(dim)
 * float3) receives the finest level's field. */
static void orc_cdm_fixed_rec(const uint8_t* const* It, const uint8_t* const* Is, int nmod,
                              const int* dim, float* field, int levels, int level_iterations,
                              float speed, float smoothing, unsigned bandwidth, char gradient_type,
                              int* iterations)
{
    size_t n = vox(dim);
    memset(field, 0, sizeof(float) * 3 * n);
    if (levels > 1) {
        int rdim[3] = {(dim[0] + 1) / 2, (dim[1] + 1) / 2, (dim[2] + 1) / 2};
        size_t rn = vox(rdim);
        uint8_t* buf = (uint8_t*)malloc(2 * (size_t)nmod * rn);
        float* rd = (float*)calloc(3 * rn, sizeof(float));
        const uint8_t* rIt[16];
        const uint8_t* rIs[16];
        for (int i = 0; i < nmod && i < 16; ++i) {
            int od[3];
            orc_downsample_u8(It[i], dim, buf + (size_t)(2 * i) * rn, od);
            orc_downsample_u8(Is[i], dim, buf + (size_t)(2 * i + 1) * rn, od);
            rIt[i] = buf + (size_t)(2 * i) * rn;
            rIs[i] = buf + (size_t)(2 * i + 1) * rn;
        }
        orc_cdm_fixed_rec(rIt, rIs, nmod, rdim, rd, levels - 1, level_iterations, speed, smoothing,
                          bandwidth, gradient_type, iterations);
        free(buf);
        for (size_t k = 0; k < 3 * rn; ++k) rd[k] *= 2.0f;
        orc_upsample_f3(rd, rdim, field, dim);
        free(rd);
    }
    float* nd = (float*)malloc(sizeof(float) * 3 * n);
    float* dd = (float*)malloc(sizeof(float) * 3 * n);
    uint8_t* Js = (uint8_t*)malloc(n);
    float theta = 0.0f;
    for (int index = 0; index < level_iterations; ++index) {
        ++*iterations;
        for (int i = 0; i < nmod; ++i) {
            orc_compose_displacement(Is[i], dim, field, Js);
            (void)orc_cdm_gradient(Js, It[i], dim, gradient_type, i == 0 ? nd : dd);
            if (i)
                for (size_t k = 0; k < 3 * n; ++k) nd[k] += dd[k];
        }
        orc_cdm_poisson(nd, dim, 12);
        if (theta == 0.0f)
            theta = orc_cdm_max_length(nd, n);
        if (theta == 0.0f)
            break;
        float k = speed / theta;
        for (size_t j = 0; j < 3 * n; ++j) nd[j] *= k;
        orc_accumulate_displacement(field, nd, dim);
        if (bandwidth) {
            orc_gaussian_f32(nd, dim, 3, bandwidth);
            memcpy(field, nd, sizeof(float) * 3 * n);
        } else {
            orc_cdm_smooth(nd, field, dim, smoothing);
        }
    }
    free(nd); free(dd); free(Js);
}

ORC_API void orc_cdm_fixed(const uint8_t* const* It, const uint8_t* const* Is, int nmod,
                           const int* dim, int levels, int level_iterations, float speed,
                           float smoothing, unsigned bandwidth, char gradient_type, float* field,
                           int* iterations)
{
    int it = 0;
    orc_cdm_fixed_rec(It, Is, nmod, dim, field, levels, level_iterations, speed, smoothing,
                      bandwidth, gradient_type, &it);
    if (iterations) *iterations = it;
}

/* --- subject preprocessing ------------------------------------------------
 * reg::subject_image_pre (reg/mm_reg.hpp:14-20): filter::gaussian(I) with
 * bandwidth 1, segmentation::normalize_otsu_median(I, 255.99f)
 * (numerical/otsu.hpp:96-136), then image<3,uint8>(I) (std::copy, so a
 * truncating conversion). */

static int orc_cmp_float(const void* a, const void* b)
{
    float x = *(const float*)a, y = *(const float*)b;
    return (x > y) - (x < y);
}

/* The value of rank k (0-based) of v[0..n): std::nth_element's element. */
static float orc_rank(const float* v, size_t n, size_t k)
{
    float* t = (float*)malloc(sizeof(float) * n);
    memcpy(t, v, sizeof(float) * n);
    qsort(t, n, sizeof(float), orc_cmp_float);
    float r = t[k];
    free(t);
    return r;
}

/* tipl::median (numerical/statistics.hpp:327-363): below 2048 values the
 * element of rank n/2; otherwise the rank-128 element of the 256 medians of
 * contiguous chunks (chunk i: i*(n/256) + min(i, n%256), one longer for
 * i < n%256), each the element of rank len/2 of its chunk. */
ORC_API float orc_median(const float* v, size_t n)
{
    if (n < 2048)
        return orc_rank(v, n, n / 2);
    float chunk[256];
    size_t cs = n / 256, rem = n % 256;
    for (size_t i = 0; i < 256; ++i) {
        size_t start = i * cs + (i < rem ? i : rem);
        size_t len = cs + (i < rem ? 1 : 0);
        chunk[i] = orc_rank(v + start, len, len / 2);
    }
    return orc_rank(chunk, 256, 128);
}

/* otsu_threshold (numerical/otsu.hpp:15-81) of a float image: 256-bin
 * histogram (numerical/basic_op.hpp:1294-1311) between min and max, bin 0
 * cleared, then the between-class variance scan in float. */
ORC_API float orc_otsu_threshold(const float* v, size_t n)
{
    if (n == 0) return 0.0f;
    float mn = v[0], mx = v[0];
    for (size_t i = 1; i < n; ++i) {
        if (v[i] > mx) mx = v[i];
        else if (v[i] < mn) mn = v[i];
    }
    if (mn >= mx) return mn;  /* histogram() leaves the histogram empty */
    unsigned hist[256] = {0};
    float range = mx - mn;
    if (range == 0.0f) range = 1.0f;
    range = 256.0f / range;
    for (size_t i = 0; i < n; ++i) {
        float value = (v[i] - mn) * range;
        int b = (int)floorf(value);
        b = b < 0 ? 0 : (b > 255 ? 255 : b);
        ++hist[b];
    }
    hist[0] = 0;
    float total_w = 0.0f, total_sum = 0.0f;
    for (int i = 0; i < 256; ++i) {
        float val = (float)hist[i];
        total_w += val;
        total_sum += val * (float)i;
    }
    if (total_w == 0.0f) return mn;
    float max_sig_b = 0.0f, w_b = 0.0f, sum_b = 0.0f;
    int opt = 0;
    for (int i = 0; i < 255; ++i) {
        w_b += (float)hist[i];
        if (w_b == 0.0f) continue;
        float w_f = total_w - w_b;
        if (w_f == 0.0f) break;
        sum_b += (float)hist[i] * (float)i;
        float d = (sum_b * total_w) - (total_sum * w_b);
        float sig_b = (d * d) / (w_b * w_f);
        if (sig_b > max_sig_b) {
            max_sig_b = sig_b;
            opt = i;
        }
    }
    return mn + (float)opt * (mx - mn) / 256.0f;
}

/* otsu_median (numerical/otsu.hpp:96-108). */
ORC_API float orc_otsu_median(const float* v, size_t n)
{
    float ot = orc_otsu_threshold(v, n);
    float* buf = (float*)malloc(sizeof(float) * (n ? n : 1));
    size_t m = 0;
    for (size_t i = 0; i < n; ++i)
        if (v[i] > ot) buf[m++] = v[i];
    float r = m ? orc_median(buf, m) : ot;
    free(buf);
    return r;
}

/* The float -> uint8 conversion of std::copy on the reference's x86 build:
 * truncation to int32, low byte kept (values are in [0, 255.99] after the
 * normalisation; other values only arise when the median is 0). */
static uint8_t orc_to_u8(float v)
{
    return (uint8_t)(uint32_t)(int32_t)v;
}

ORC_API void orc_subject_pre(const float* in, const int* dim, uint8_t* out)
{
    size_t n = vox(dim);
    float* I = (float*)malloc(sizeof(float) * (n ? n : 1));
    memcpy(I, in, sizeof(float) * n);
    orc_gaussian_f32(I, dim, 1, 1);
    const float upper = 255.99f;
    float ot = orc_otsu_median(I, n);
    if (ot != 0.0f) {
        float c = upper * 0.5f / ot;
        size_t pos = 0;
        for (size_t i = 0; i < n; ++i) {
            I[i] *= c;
            if (I[i] > 0.0f) ++pos;
        }
        if ((double)pos > 0.5 * (double)n) {
            float s = upper * 0.1f;
            for (size_t i = 0; i < n; ++i) I[i] -= s;
        }
        for (size_t i = 0; i < n; ++i) {
            float v = I[i];
            I[i] = v > upper ? upper : (v < 0.0f ? 0.0f : v);
        }
    }
    for (size_t i = 0; i < n; ++i) out[i] = orc_to_u8(I[i]);
    free(I);
}

/* ------------------------------------------------------------------------ */
/* Post-registration warping (numerical/dif.hpp) and mm_reg's r values       */
/* ------------------------------------------------------------------------ */

/* displacement_to_mapping(dis, mapping) numerical/dif.hpp:19-28 (T == NULL)
 * and displacement_to_mapping(dis, mapping, T) :42-50:
 * mapping(x) = dis(x) + vector(x), then T applied in place
 * (transformation_matrix::operator()(vector&) numerical/transformation.hpp:846-850). */
ORC_API void orc_disp_to_mapping(const float* dis, const int* dim, const float* T, float* out)
{
    size_t i = 0;
    for (int z = 0; z < dim[2]; ++z)
        for (int y = 0; y < dim[1]; ++y)
            for (int x = 0; x < dim[0]; ++x, ++i) {
                float m[3] = {dis[3 * i] + (float)x, dis[3 * i + 1] + (float)y,
                              dis[3 * i + 2] + (float)z};
                if (T) {
                    float o[3];
                    orc_apply(T, m[0], m[1], m[2], o);
                    memcpy(m, o, sizeof(m));
                }
                memcpy(out + 3 * i, m, sizeof(m));
            }
}

/* mapping_to_displacement numerical/dif.hpp:30-38. */
ORC_API void orc_mapping_to_disp(const float* mapping, const int* dim, float* out)
{
    size_t i = 0;
    for (int z = 0; z < dim[2]; ++z)
        for (int y = 0; y < dim[1]; ++y)
            for (int x = 0; x < dim[0]; ++x, ++i) {
                out[3 * i] = mapping[3 * i] - (float)x;
                out[3 * i + 1] = mapping[3 * i + 1] - (float)y;
                out[3 * i + 2] = mapping[3 * i + 2] - (float)z;
            }
}

/* inv_displacement_to_mapping<linear>(inv_dis, inv_mapping, T)
 * numerical/dif.hpp:52-64: p = T^-1(x) over the mapping's grid, then
 * p += estimate<linear>(inv_dis, p) (zero vector when outside,
 * numerical/interpolation.hpp:726-740). */
ORC_API void orc_inv_disp_to_mapping(const float* inv_dis, const int* ddim, const float* T_in,
                                     const int* odim, float* out)
{
    float iT[12];
    memcpy(iT, T_in, sizeof(iT));
    orc_matrix_inverse(iT);   /* the reference ignores the return value */
    size_t i = 0;
    for (int z = 0; z < odim[2]; ++z)
        for (int y = 0; y < odim[1]; ++y)
            for (int x = 0; x < odim[0]; ++x, ++i) {
                float p[3], v[3] = {0.0f, 0.0f, 0.0f};
                orc_apply(iT, (float)x, (float)y, (float)z, p);
                orc_tri t;
                if (orc_tri_locate(ddim, p[0], p[1], p[2], &t))
                    orc_tri_f3(inv_dis, &t, v);
                out[3 * i] = p[0] + v[0];
                out[3 * i + 1] = p[1] + v[1];
                out[3 * i + 2] = p[2] + v[2];
            }
}

/* compose_mapping<linear>(from, mapping, to) numerical/dif.hpp:66-86 for a
 * uint8 image: to(x) = from(mapping(x)), 0 where the sample is outside. */
ORC_API void orc_compose_mapping_u8(const uint8_t* from, const int* fdim, const float* mapping,
                                    const int* dim, uint8_t* out)
{
    size_t n = vox(dim);
    for (size_t i = 0; i < n; ++i)
        out[i] = orc_sample_u8(from, fdim, mapping[3 * i], mapping[3 * i + 1], mapping[3 * i + 2]);
}

/* The same for a float image (weighted_sum in float, no truncation). */
ORC_API void orc_compose_mapping_f32(const float* from, const int* fdim, const float* mapping,
                                     const int* dim, float* out)
{
    size_t n = vox(dim);
    for (size_t i = 0; i < n; ++i) {
        orc_tri t;
        float r = 0.0f;
        if (orc_tri_locate(fdim, mapping[3 * i], mapping[3 * i + 1], mapping[3 * i + 2], &t)) {
            r = from[t.idx[0]] * t.w[0];
            for (int k = 1; k < 8; ++k) {
                float v = from[t.idx[k]];
                v *= t.w[k];
                r += v;
            }
        }
        out[i] = r;
    }
}

static int orc_is_edge(const int* dim, int x, int y, int z)
{
    return x == 0 || y == 0 || z == 0 || x + 1 == dim[0] || y + 1 == dim[1] || z + 1 == dim[2];
}

/* jacobian_determinant_dis(displacement, dest) numerical/dif.hpp:451-465 with
 * jacobian_determinant_dis_at<3> :399-423: central differences (not halved)
 * with +1.0 on the diagonal.  The mixed float/double expression types of the
 * reference are kept: differences are float, the diagonal terms and every
 * product with one of them are double, the other products are float.  The
 * double result is stored as float; edge voxels are 1. */
ORC_API void orc_jacobian_det_dis(const float* d, const int* dim, float* out)
{
    const size_t w = (size_t)dim[0], wh = w * (size_t)dim[1];
    size_t i = 0;
    for (int z = 0; z < dim[2]; ++z)
        for (int y = 0; y < dim[1]; ++y)
            for (int x = 0; x < dim[0]; ++x, ++i) {
                if (orc_is_edge(dim, x, y, z)) {
                    out[i] = 1.0f;
                    continue;
                }
                const float* v1_0 = d + 3 * (i + 1);
                const float* v1_1 = d + 3 * (i - 1);
                const float* v2_0 = d + 3 * (i + w);
                const float* v2_1 = d + 3 * (i - w);
                const float* v3_0 = d + 3 * (i + wh);
                const float* v3_1 = d + 3 * (i - wh);
                float d2_0 = v2_0[0] - v2_1[0];
                double d2_1 = (double)(v2_0[1] - v2_1[1]) + 1.0;
                float d2_2 = v2_0[2] - v2_1[2];
                float d3_0 = v3_0[0] - v3_1[0];
                float d3_1 = v3_0[1] - v3_1[1];
                double d3_2 = (double)(v3_0[2] - v3_1[2]) + 1.0;
                double a = (double)(v1_0[0] - v1_1[0]) + 1.0;
                float b = v1_0[1] - v1_1[1];
                float c = v1_0[2] - v1_1[2];
                float p22_31 = d2_2 * d3_1;
                float p22_30 = d2_2 * d3_0;
                float p20_31 = d2_0 * d3_1;
                double r = a * (d2_1 * d3_2 - (double)p22_31) +
                           (double)b * ((double)p22_30 - (double)d2_0 * d3_2) +
                           (double)c * ((double)p20_31 - d2_1 * (double)d3_0);
                out[i] = (float)r;
            }
}

/* jacobian_determinant(src, dest) numerical/dif.hpp:356-388 of a mapping:
 * float central differences; edge voxels are left untouched (0 here, as in
 * a freshly sized dest).  The reference's expression applies its 0.125f
 * factor to the third product only; that is kept. */
ORC_API void orc_jacobian_det(const float* m, const int* dim, float* out)
{
    const size_t w = (size_t)dim[0], wh = w * (size_t)dim[1];
    size_t i = 0;
    for (int z = 0; z < dim[2]; ++z)
        for (int y = 0; y < dim[1]; ++y)
            for (int x = 0; x < dim[0]; ++x, ++i) {
                if (orc_is_edge(dim, x, y, z)) {
                    out[i] = 0.0f;
                    continue;
                }
                float dx[3], dy[3], dz[3];
                for (int c = 0; c < 3; ++c) {
                    dx[c] = m[3 * (i + 1) + c] - m[3 * (i - 1) + c];
                    dy[c] = m[3 * (i + w) + c] - m[3 * (i - w) + c];
                    dz[c] = m[3 * (i + wh) + c] - m[3 * (i - wh) + c];
                }
                float t0 = dx[0] * (dy[1] * dz[2] - dy[2] * dz[1]);
                float t1 = dx[1] * (dy[2] * dz[0] - dy[0] * dz[2]);
                float t2 = dx[2] * (dy[0] * dz[1] - dy[1] * dz[0]) * 0.125f;
                out[i] = t0 + t1 + t2;
            }
}

/* tipl::correlation(x, y) as r (numerical/statistics.hpp:569-597) from the
 * exact uint8 sums s = {Sx, Sxx, Sy, Syy, Sxy} over n voxels; 0 for n = 0. */
ORC_API double orc_corr_r_from_sums(const int64_t* s, size_t n)
{
    if (n == 0) return 0.0;
    double dn = (double)n;
    double mean_x = (double)s[0] / dn;
    double mean_y = (double)s[2] / dn;
    double var_x = (double)s[1] / dn - mean_x * mean_x;
    double var_y = (double)s[3] / dn - mean_y * mean_y;
    double sd1 = var_x > 0.0 ? sqrt(var_x) : 0.0;
    double sd2 = var_y > 0.0 ? sqrt(var_y) : 0.0;
    if (sd1 == 0.0 || sd2 == 0.0) return 0.0;
    double cov = (double)s[4] / dn - mean_x * mean_y;
    return cov / sd1 / sd2;
}

/* mm_reg's r (reg/mm_reg.hpp:302-305, 341-343): correlation_ygz
 * (numerical/statistics.hpp:599-619, only voxels with y > 0) when masked,
 * else tipl::correlation over every voxel. */
ORC_API double orc_correlation(const uint8_t* x, const uint8_t* y, size_t n, int masked)
{
    int64_t s[5] = {0, 0, 0, 0, 0};
    size_t m = 0;
    for (size_t i = 0; i < n; ++i) {
        if (masked && y[i] == 0) continue;
        int64_t a = x[i], b = y[i];
        s[0] += a;
        s[1] += a * a;
        s[2] += b;
        s[3] += b * b;
        s[4] += a * b;
        ++m;
    }
    return orc_corr_r_from_sums(s, m);
}

/* mm_reg::It_match_fov reg/mm_reg.hpp:313-328: zero every template voxel whose
 * T(x) lies outside the subject grid; "inside" is shape::is_valid
 * (utility/shape.hpp:582-587): 0 <= p < dim as float comparisons. */
ORC_API void orc_match_fov(uint8_t* It, const int* tdim, const float* T, const int* sdim)
{
    size_t i = 0;
    for (int z = 0; z < tdim[2]; ++z)
        for (int y = 0; y < tdim[1]; ++y)
            for (int x = 0; x < tdim[0]; ++x, ++i) {
                float p[3];
                orc_apply(T, (float)x, (float)y, (float)z, p);
                int valid = p[0] >= 0.0f && p[1] >= 0.0f && p[2] >= 0.0f &&
                            p[0] < (float)(unsigned)sdim[0] && p[1] < (float)(unsigned)sdim[1] &&
                            p[2] < (float)(unsigned)sdim[2];
                if (!valid) It[i] = 0;
            }
}
