/*
 * Synthetic Gaussian-splat scenes and camera matrices for the render path's workloads.
 *
 * Scenes come out as a SplatCloud in structure-of-arrays form (SplatCloud.h:18-39 of the
 * reference): positions xyz, covariance upper triangle xx xy xz yy yz zz, sRGB colour, opacity,
 * and 3*((d+1)^2-1) SH floats per splat, coefficient-major with the channel fastest.
 * Bakes follow the reference's decoder (SpzDecoder.cpp:301-319): colour clamp(0.5 + 0.282095
 * DC, 0, 1), sigmoid opacity, exp scales, covariance R diag(s)^2 R^T (Covariance.cpp:7-17).
 *
 * Randomness: each splat i draws from its own xoshiro256** stream seeded by SplitMix64 over
 * (seed, i), normals by Box-Muller, attributes in the order position, log-scale, quaternion,
 * opacity logit, DC, SH. The output depends only on the spec, not on the thread count.
 */
#pragma once
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum skm_scene_layout {
  SKM_SCENE_GAUSSIAN = 0, /* positions iid N(0, spread^2) per axis */
  SKM_SCENE_SURFACES = 1  /* surface_share near a ground square and boxes, the rest iid */
} skm_scene_layout;

typedef struct skm_scene_spec {
  uint32_t count;
  int32_t sh_degree;     /* 0..3 */
  uint64_t seed;
  int32_t layout;        /* skm_scene_layout */
  float spread;          /* sigma of the iid positions, metres */
  float area;            /* side of the ground square, metres (SURFACES) */
  float surface_share;   /* fraction of splats within N(0, surface_sigma) of a surface */
  float surface_sigma;   /* metres */
  float log_scale_mean;  /* per-axis log scale ~ N(mean, sigma) */
  float log_scale_sigma;
  float opacity_sigma;   /* opacity = sigmoid(N(0, sigma)) */
  float dc_sigma;        /* DC ~ N(0, sigma) */
  float sh_sigma;        /* higher bands ~ N(0, sigma) */
} skm_scene_spec;

/* Fills caller-allocated arrays; sh may be NULL for degree 0. Returns 0 on success. */
int skm_scene_fill(const skm_scene_spec* spec, float* positions, float* covariances,
                   float* colors, float* alphas, float* sh, uint32_t threads);

/* Column-major view matrix of a camera at eye looking at target (RUB: looks down -Z, +Y up). */
void skm_look_at(const float* eye3, const float* target3, const float* up3, float* view16);
/* Mat4::perspective conventions (Mat4.h:75-86): z in [-near, -far] maps to depth [0, 1]. */
void skm_perspective(float fovy_radians, float aspect, float near_plane, float far_plane,
                     float* proj16);

#ifdef __cplusplus
}
#endif
