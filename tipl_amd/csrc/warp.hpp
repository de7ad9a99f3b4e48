// warp.hpp -- device-resident warping steps shared by warp.hip and the
// registration pipeline (server.cpp).  All pointers are device pointers;
// every launch is asynchronous on ctx->stream.
#pragma once

#include "internal.hpp"

namespace tiplhip {

// A transformation_matrix: sr[9] row-major, then shift[3].
struct Affine {
    float m[12];
};

// displacement_to_mapping: out(x) = T(x + dis(x)); T may be null (identity).
void launch_mapping(tiplhip_ctx* ctx, const float* dis, const int* dim, const float* T, float* out);
// inv_displacement_to_mapping: over out_dim, out(x) = p + inv_dis(p), p = T^-1(x).
void launch_inv_mapping(tiplhip_ctx* ctx, const float* inv_dis, const int* dis_dim, const float* T,
                        const int* out_dim, float* out);
// compose_mapping of a uint8 image: out(x) = from(mapping(x)), 0 outside.
void launch_compose_mapping_u8(tiplhip_ctx* ctx, const uint8_t* from, const int* from_dim,
                               const float* mapping, const int* dim, uint8_t* out);
// mm_reg::It_match_fov in place on a uint8 template image.
void launch_match_fov(tiplhip_ctx* ctx, uint8_t* It, const int* dim, const float* T,
                      const int* moving_dim);
// tipl::correlation (or correlation_ygz when masked) of two device uint8
// arrays of n voxels; synchronises ctx->stream.
double device_correlation(tiplhip_ctx* ctx, const uint8_t* x, const uint8_t* y, size_t n,
                          bool masked);

}  // namespace tiplhip
