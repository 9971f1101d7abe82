/*
 * CPU oracle for the per-frame splat render path (test infrastructure only).
 *
 * This library is a checker. Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it; the product (splatkit-ios_amd/) never links or calls it.
 *
 * It restates, in scalar C++, the reference's packing (GpuLayout.cpp, Half.h), camera
 * uniform derivation (MetalFrameCompute.mm), visibility/projection kernel
 * (SplatVisibility.metal, SplatProjection.metalh), the radix sort contract
 * (MetalRadixSort.h:13-19, checked like MetalRadixSortTest.mm:58-62 with std::stable_sort),
 * the tile binning this build adds, and the front-to-back composite (SplatRaster.metal,
 * MetalSplatPipeline.mm:46-63,153-161) in FP64 and FP32.
 *
 * Every layout here is the reference's: 32-byte GpuSplat records (GpuLayout.h:50-56),
 * per-splat SH words (GpuLayout.cpp:92-113) and 32-byte ProjectedSplat records
 * (GpuLayout.h:60-70).
 */
#pragma once
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Camera uniform, byte-identical to CameraUniform (MetalShaderTypes.h:14-30). */
typedef struct orc_camera {
  float view[16];
  float proj[16];
  float focal[2];
  float tan_half_fov[2];
  float screen_size[2];
  uint32_t output_linear;
  float min_pixel_radius;
  float camera_position[4];
  float reveal[4];
  float reveal_tint[4];
} orc_camera;

/* Half conversions. */
uint16_t orc_to_half(float v);       /* saturating RNE, Half.h:12-38 */
uint16_t orc_half_rne(float v);      /* IEEE RNE, overflow to infinity: the shader's half(float) */
float orc_from_half(uint16_t h);     /* Half.h:40-65 */

/* Packing (AoS, the reference layout). out: 8 words per splat. */
void orc_pack_splats(const float* positions, const float* covariances, const float* colors,
                     const float* alphas, uint32_t n, uint32_t* out);
/* SH words per splat for `degree` (GpuLayout.cpp:66-70, float16 storage). */
uint32_t orc_sh_stride(int degree);
/* src holds 3*((src_degree+1)^2-1) floats per splat; out holds orc_sh_stride(degree) words. */
void orc_pack_sh(const float* sh, uint32_t n, int src_degree, int degree, uint32_t* out);

void orc_make_camera(const float* view16, const float* proj16, uint32_t width, uint32_t height,
                     int output_linear, float min_pixel_radius, const float* camera_position3,
                     const float* reveal8, orc_camera* out);

/* Visibility: cull, project, key and compact in source order. active == NULL for ranges.
 * Returns the survivor count; keys[] and projected[] (8 words each) are written for it. */
uint32_t orc_cull(const orc_camera* cam, const uint32_t* splats, const uint32_t* sh,
                  int sh_degree, const uint32_t* ranges, uint32_t range_count,
                  const uint32_t* active, uint32_t active_count, int key16, uint32_t* keys,
                  uint32_t* projected);
/* Single-splat projection, for tests: returns 1 when visible. */
int orc_project(const orc_camera* cam, const uint32_t* splat8, const uint32_t* sh_words,
                int sh_degree, uint32_t index, uint32_t* projected8);
uint32_t orc_depth_key16(const orc_camera* cam, float depth);

/* Stable ascending argsort of u32 keys (std::stable_sort). */
void orc_argsort(const uint32_t* keys, uint32_t n, uint32_t* order);

/* mortonCode (SpatialOrder.cpp:39-44) over the box [lo, hi]. */
uint32_t orc_morton_code(const float* p3, const float* lo3, const float* hi3);
/* reorderSpatially's permutation (SpatialOrder.cpp:48-84): order[i] is the source index of splat i
 * after three stable 10-bit counting passes over the Morton codes. */
void orc_spatial_order(const float* positions, uint32_t n, const float* lo3, const float* hi3, uint32_t* order);

/* Tile binning in depth order, then stable by tile. ranges: 2 words per tile.
 * Returns the pair count; pairs are written only when it fits `capacity`. */
uint64_t orc_bin(const uint32_t* projected, const uint32_t* order, uint32_t n, uint32_t width,
                 uint32_t height, uint32_t tile, uint32_t* pair_tiles, uint32_t* pair_values,
                 uint64_t capacity, uint32_t* ranges);
/* Tile rectangle of one projected record: rect[4] = tx0, ty0, tx1, ty1; returns tile count. */
uint32_t orc_tile_rect(const uint32_t* projected8, uint32_t width, uint32_t height,
                       uint32_t tile, int32_t* rect);

/* Brute-force check of the binning: counts (pixel, record) pairs the coverage rule accepts whose
 * record is missing from the pixel's tile list. Scans every pixel of the quad's loose bounding box
 * c +- r(|a1| + |a2|) plus two pixels per record. Test infrastructure. */
uint64_t orc_missed_coverage(const uint32_t* projected, uint32_t count, uint32_t width,
                             uint32_t height, uint32_t tile, const uint32_t* pair_values,
                             const uint32_t* ranges, uint64_t* accepted);

/* The raster's per-region record test over pixel centres [x0, x0+span] x [y0, y0+span], and a brute
 * force count of accepted (pixel, record) pairs whose `region`-aligned square it drops. Test
 * infrastructure: checks that the kernel's region lists are conservative. */
int orc_region_keep(const uint32_t* projected8, uint32_t width, uint32_t height, uint32_t x0, uint32_t y0,
                    uint32_t span);
uint64_t orc_region_misses(const uint32_t* projected, uint32_t count, uint32_t width, uint32_t height,
                           uint32_t region, uint64_t* accepted);

/* FP64 composite over the tile lists (no early termination). out: RGBA float, top-down rows. */
void orc_raster64(const uint32_t* projected, const uint32_t* pair_values, const uint32_t* ranges,
                  uint32_t width, uint32_t height, uint32_t tile, const float* background3,
                  float* out);
/* FP32 composite with per-pixel termination at transmittance t_stop (the CPU baseline). */
void orc_raster32(const uint32_t* projected, const uint32_t* pair_values, const uint32_t* ranges,
                  uint32_t width, uint32_t height, uint32_t tile, const float* background3,
                  float t_stop, uint32_t threads, float* out);
/* Per-pixel coverage decision of one record at one pixel (the fp32 rule shared with the kernel):
 * returns 1 when the pixel takes the splat; power_out gets the fp32 log2 alpha. */
int orc_decide(const uint32_t* projected8, uint32_t width, uint32_t height, uint32_t px,
               uint32_t py, float* power_out);

/* The whole CPU frame (cull, sort, bin, raster32) for the baseline timing; returns seconds of
 * the timed stages in stage_seconds[4] (cull, sort, bin, raster). threads >= 1. */
int orc_render_frame(const orc_camera* cam, const uint32_t* splats, const uint32_t* sh,
                     int sh_degree, uint32_t n, int key16, uint32_t tile, const float* background3,
                     uint32_t threads, float* out_rgba, double* stage_seconds,
                     uint32_t* visible_out, uint64_t* pairs_out);

/* LOD tree cut (SplatLOD.metal:26-220, MetalLodSelection.mm:109-192) restated as sequential rounds.
 * clusters: 16 words each (LodCluster, LodTree.h:28-35). Writes the selected node indices in the
 * reference's emission order and returns their count; stats[6] = count, active, packets, accepted,
 * limited (denied refinements), evaluated (interior clusters evaluated). */
uint32_t orc_lod_select(const orc_camera* cam, const uint32_t* clusters, uint32_t cluster_count,
                        const uint32_t* leaves, uint32_t rounds, uint32_t capacity, float pixel_limit,
                        float color_weight, int cull, uint32_t* indices, uint32_t* stats);
/* The fp32 decisions of one cluster: returns 1 when outside the view; error_out gets the
 * screen-space error in pixels. */
int orc_lod_cluster(const orc_camera* cam, const uint32_t* cluster16, float color_weight,
                    float* error_out);

/* LodThreshold (LodThreshold.h:20-36) over `steps` completed selections: out[i] is the threshold
 * after the i-th (selected, limited, limit), starting from the policy's. */
void orc_lod_threshold(float policy, uint32_t steps, const uint32_t* selected, const uint32_t* limited,
                       const uint32_t* limit, float* out);

/* Transcendentals built from IEEE basic operations only (argument reduction plus series), so
 * any two IEEE machines compute identical bits. */
double orc_log_d(double x);
float orc_log_f(float x);
double orc_exp_d(double x);

#ifdef __cplusplus
}
#endif
