/*
 * skm: the per-frame Gaussian-splat render path on one AMD Instinct MI355X (gfx950), behind a C ABI
 * that mirrors the splatkit-engine renderer boundary:
 *
 *   class SplatRenderer        packages/splatkit-engine/include/splatkit/rendering/SplatRenderer.h:36-160
 *   struct Frame, Range        SplatRenderer.h:100-131
 *   struct GpuWorldInfo        SplatRenderer.h:25-28
 *   struct RenderPolicy et al. packages/splatkit-engine/include/splatkit/rendering/RenderPolicy.h:22-77
 *   captureNextFrame           packages/splatkit-ios/include/SplatKit/metal/MetalRenderer.h:29
 *
 * One entry point per virtual of SplatRenderer, plain pointers and sizes only. Every call returns a
 * status (0 = ok) or a value; failures set a per-renderer message (skm_last_error) and, like the
 * reference, leave the previous world, policy and buffers in place. Nothing throws across the ABI.
 * A renderer is used from one thread at a time (SplatRenderer.h:32), except the stat getters.
 */
#pragma once
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct skm_renderer skm_renderer;
struct skm_lod_tree_view; /* include/skm_lod.h */

typedef enum skm_status {
  SKM_OK = 0,
  SKM_ERR_INVALID = -1,     /* bad argument: the call did nothing */
  SKM_ERR_DEVICE = -2,      /* HIP runtime or kernel failure; the renderer latches gpu_failed */
  SKM_ERR_MEMORY = -3,      /* device allocation failed; previous buffers kept */
  SKM_ERR_STATE = -4,       /* no world / no extent / wrong order source */
  SKM_ERR_UNSUPPORTED = -5, /* a feature this backend does not implement */
  SKM_ERR_CAPACITY = -6     /* the frame's tile pairs exceeded capacity; it was dropped */
} skm_status;

typedef struct skm_config {
  int32_t device;          /* HIP device ordinal */
  uint32_t tile_size;      /* raster tile edge in pixels: 8, 16 or 32 (0 = 16) */
  uint32_t pair_factor;    /* initial tile-pair capacity per splat slot (grows on overflow); 0 = 8 */
  uint32_t sort_kernel;    /* radix pass: 0 = reduce then scan of tile counts, 1 = onesweep look-back */
  uint32_t raster_px;      /* raster pixels per lane side: 1 or 2 (0 = 1, or 2 at 32-pixel tiles);
                            * supported (tile, px): (8, 1), (16, 1), (16, 2), (32, 2) */
  uint32_t frames_in_flight; /* frames a draw may queue before it waits, each with its own HIP stream
                              * and scratch: 1 to 3 (0 = 2, the reference's count) */
} skm_config;

/* SplatRenderer::Range (SplatRenderer.h:100-103). */
typedef struct skm_range {
  uint32_t offset;
  uint32_t count;
} skm_range;

/* RevealUniform (packages/splatkit-engine/include/splatkit/reveal/Reveal.h:25-35); level = +inf is idle. */
typedef struct skm_reveal {
  float level;
  float band;
  float tint_amount;
  float brightness_boost;
  float tint[4];
} skm_reveal;

typedef enum skm_order_source { SKM_ORDER_CPU = 0, SKM_ORDER_GPU = 1 } skm_order_source;

/* SplatRenderer::Frame (SplatRenderer.h:111-131). view/proj are column-major Mat4 (Mat4.h:10-14). */
typedef struct skm_frame {
  int32_t order_source;     /* skm_order_source; this backend sorts on the GPU only */
  const uint32_t* order;    /* CPU order: unsupported, must be NULL */
  uint32_t order_count;
  uint32_t draw_count;
  const skm_range* ranges;  /* GPU order: the ranges to draw this frame */
  uint32_t range_count;
  int32_t sh_degree;        /* capped by the world's degree */
  float view[16];
  float proj[16];
  float camera_position[3];
  const void* label_styles; /* part labels: unsupported, must be NULL */
  skm_reveal reveal;
} skm_frame;

/* GpuWorldInfo (SplatRenderer.h:25-28). */
typedef struct skm_world_info {
  uint32_t count;
  int32_t sh_degree;
} skm_world_info;

/* A SplatCloud (packages/splat-core/include/splat/formats/SplatCloud.h:18-39) as borrowed arrays:
 * positions xyz, covariance xx xy xz yy yz zz, colour rgb in [0,1], opacity (LOD nodes up to 1000),
 * and 3*((sh_degree+1)^2-1) SH floats per splat, coefficient-major, channel fastest. */
typedef struct skm_cloud_view {
  uint32_t count;
  int32_t sh_degree;
  const float* positions;
  const float* covariances;
  const float* colors;
  const float* alphas;
  const float* sh;
} skm_cloud_view;

/* RenderPolicy (RenderPolicy.h:22-40), same field order; bools as bytes. */
typedef struct skm_render_policy {
  float lod_error_pixels;
  uint32_t lod_splat_limit;
  float alpha_threshold;
  float subpixel_threshold;
  uint8_t enable_frustum_culling;
  uint8_t enable_hiz_occlusion;
  uint8_t enable_early_termination;
  uint8_t reserved0;
  uint32_t sort_depth; /* 16 or 32 */
  double background_color[3];
} skm_render_policy;

/* RenderPolicySupport (RenderPolicy.h:45-60). */
typedef struct skm_policy_support {
  uint8_t lod_error_pixels;
  uint8_t lod_splat_limit;
  uint8_t alpha_threshold;
  uint8_t subpixel_threshold;
  uint8_t enable_frustum_culling;
  uint8_t enable_hiz_occlusion;
  uint8_t enable_early_termination;
  uint8_t sort_depth;
  uint8_t background_color;
  uint8_t reserved[3];
  float min_lod_error_pixels;
  float max_lod_error_pixels;
  float min_subpixel_threshold;
  float max_subpixel_threshold;
  skm_render_policy fallback;
} skm_policy_support;

/* DeviceCapabilities + SplatLimits (RenderPolicy.h:64-77). */
typedef struct skm_device_capabilities {
  uint32_t max_lod_capacity_splats;
  uint32_t min_residency_capacity_splats;
  uint32_t max_residency_capacity_splats;
  uint8_t supports_hiz_occlusion;
  uint8_t supports_subgroups;
  uint8_t reserved[2];
  uint32_t max_texture_dimension;
  skm_policy_support policy;
} skm_device_capabilities;

typedef enum skm_pixel_format {
  SKM_PIXELS_BGRA8 = 0,   /* MetalFrameCapture's BGRA, top-down rows (MetalFrameCapture.h:18) */
  SKM_PIXELS_RGBA32F = 1  /* the composite as computed, top-down rows */
} skm_pixel_format;

/* ---- lifecycle ---- */
skm_status skm_renderer_create(const skm_config* config, skm_renderer** out);
void skm_renderer_destroy(skm_renderer* r);
const char* skm_last_error(const skm_renderer* r);

/* ---- presentation controls (SplatRenderer.h:42-52) ---- */
void skm_set_render_scale(skm_renderer* r, float scale);
float skm_render_scale(const skm_renderer* r);
void skm_set_linear_blending(skm_renderer* r, int linear);
int skm_linear_blending(const skm_renderer* r);
void skm_set_vsync(skm_renderer* r, int vsync);
/* Replaces the platform surface attach/resize: the size frames are drawn at. */
skm_status skm_set_draw_extent(skm_renderer* r, uint32_t width, uint32_t height);

/* ---- capabilities and policy (SplatRenderer.h:56-70) ---- */
skm_status skm_get_device_capabilities(const skm_renderer* r, skm_device_capabilities* out);
skm_status skm_apply_render_policy(skm_renderer* r, const skm_render_policy* policy, char* reason,
                                   size_t reason_size);
/* Returns 1 and sets *bytes when known (device memory still free), else 0. */
int skm_available_memory(const skm_renderer* r, uint64_t* bytes);

/* ---- state (SplatRenderer.h:73-83) ---- */
int skm_ready(const skm_renderer* r);
int skm_gpu_failed(skm_renderer* r);
int skm_has_completed_world_frame(skm_renderer* r);
void skm_draw_extent(const skm_renderer* r, uint32_t* width, uint32_t* height);
uint32_t skm_generation(const skm_renderer* r);
int skm_surface_rebuild_pending(const skm_renderer* r);

/* ---- worlds (SplatRenderer.h:87-107) ---- */
skm_status skm_upload_world(skm_renderer* r, const skm_cloud_view* cloud, int max_sh_degree);
int skm_selects_lod_on_gpu(const skm_renderer* r);
/* uploadLodWorld (SplatRenderer.h:90): the tree's nodes become the world; every frame selects a cut
 * of at most `budget` nodes on the GPU (0 = the leaf count) at the policy's screen-space error,
 * adapted like LodThreshold (LodThreshold.h:27-36), and draws only that cut. */
skm_status skm_upload_lod_world(skm_renderer* r, const struct skm_lod_tree_view* tree, int max_sh_degree,
                                uint32_t budget);
skm_status skm_create_slab(skm_renderer* r, uint32_t capacity, int sh_degree);
skm_status skm_upload_tile(skm_renderer* r, uint32_t offset, const skm_cloud_view* cloud);
/* Returns 1 and fills *out when a world is resident. */
int skm_world(const skm_renderer* r, skm_world_info* out);
int skm_sorts_on_gpu(const skm_renderer* r);

/* ---- frames (SplatRenderer.h:134-159) ---- */
skm_status skm_draw(skm_renderer* r, const skm_frame* frame);
/* Waits until every submitted frame completed (the host side of a completion handler). Returns
 * SKM_ERR_CAPACITY when any frame or view batch since the previous skm_finish was dropped because
 * its tile pairs overflowed (the capacity has grown: drawing it again fits), and clears that report. */
skm_status skm_finish(skm_renderer* r);
uint64_t skm_submitted_frames(const skm_renderer* r);
uint64_t skm_completed_frames(skm_renderer* r);
double skm_last_gpu_millis(const skm_renderer* r);
double skm_last_sort_millis(const skm_renderer* r);
double skm_last_cull_millis(const skm_renderer* r);
uint32_t skm_last_draw_count(const skm_renderer* r);
uint32_t skm_last_selected_count(const skm_renderer* r);
double skm_last_select_millis(const skm_renderer* r);
const char* skm_device_description(const skm_renderer* r);
/* captureNextFrame (MetalRenderer.h:29): the last completed frame's pixels; SKM_ERR_CAPACITY when
 * that frame was dropped. */
skm_status skm_read_frame(skm_renderer* r, int format, void* dst, size_t dst_bytes);

/* ---- inspection of the last completed frame (parity tests and tools) ---- */
/* Survivor count, depth keys and 32-byte ProjectedSplat records (GpuLayout.h:60-70), compaction order. */
skm_status skm_read_visible(skm_renderer* r, uint32_t* count, uint32_t* keys, uint32_t* projected,
                            uint32_t capacity);
/* The depth sort's permutation of the survivors (MetalRadixSort::order, MetalRadixSort.h:27). */
skm_status skm_read_order(skm_renderer* r, uint32_t* order, uint32_t capacity);
/* Tile pairs in (tile | depth) order and [start, end) per tile. */
skm_status skm_read_tiles(skm_renderer* r, uint64_t* pairs, uint32_t* pair_tiles,
                          uint32_t* pair_values, uint64_t capacity, uint32_t* ranges,
                          uint32_t tile_capacity);
/* GPU milliseconds of the last completed frame: LOD cut, cull, depth sort, bin+tile sort+ranges,
 * raster, total. */
skm_status skm_read_stage_millis(skm_renderer* r, double* ms6);
/* The last completed frame's LOD cut: node indices in emission order and
 * stats6 = count, final frontier, packets, accepted extra, denied refinements, evaluated clusters. */
skm_status skm_read_lod_cut(skm_renderer* r, uint32_t* count, uint32_t* indices, uint32_t capacity,
                            uint32_t* stats6);

/* Sums over the frames completed since the last reset: GPU milliseconds per stage (as above),
 * survivors, tile pairs and, while skm_count_raster_tests is on, the raster's (pixel, record)
 * coverage tests of the frames drawn (every record a wave walked before its pixels saturated, times
 * the wave's pixels; 0 otherwise), and the frames and view batches dropped because their tile pairs
 * overflowed (not in the other sums). For throughput reports over many frames. */
typedef struct skm_frame_totals {
  uint64_t frames;
  double stage_ms[6];
  uint64_t visible;
  uint64_t pairs;
  uint64_t dropped;
  uint64_t raster_tests;
} skm_frame_totals;
skm_status skm_read_frame_totals(skm_renderer* r, skm_frame_totals* out, int reset);
/* Counting of the raster's coverage tests for the frames drawn from now on (off by default: the count
 * costs each raster block one device-wide atomic add, about 5% of a light frame's raster). No
 * counterpart in the reference: a profiling statistic of this backend. */
void skm_count_raster_tests(skm_renderer* r, int enable);
/* The tree cut's leading rounds whose frontier cannot exceed max_frontier clusters (the clusters at
 * that depth of the tree) run in one single-block kernel instead of two grid-wide launches per round;
 * the cut is the same either way. Default 4096; 0 runs every round grid-wide. No counterpart in the
 * reference: a launch shape of this backend's cut. */
void skm_set_lod_head(skm_renderer* r, uint32_t max_frontier);

/* ---- the sort on its own (MetalRadixSort, MetalRadixSort.h:13-19) ---- */
/* Stable ascending argsort of host keys below 2^key_bits, key_bits 1 to 32: ceil(key_bits / 8) LSD
 * passes with the bits split evenly (16 bits: 2 passes of 8; 13 bits, the tile ids of a 1080p frame:
 * 7 + 6; 32 bits: 4 passes of 8). */
skm_status skm_argsort_u32(skm_renderer* r, const uint32_t* keys, uint32_t n, int key_bits,
                           uint32_t* order);
/* Same on device memory already resident; *ms gets the GPU time of the sort. The key buffer is
 * used as ping-pong scratch: its contents are overwritten. */
skm_status skm_argsort_u32_device(skm_renderer* r, void* keys_dev, uint32_t n, int key_bits,
                                  void* order_dev, double* ms);

/* The same sort with no allocation of any kind, for callers that cannot allocate (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc). No counterpart in the
 * reference. Its scratch is one static device array of the library (8.7 MB per GPU, held until the
 * process ends), so n may be at most 1048576 (SKM_ERR_INVALID beyond); the caller passes the keys
 * (overwritten, as above) and the order. Enqueued on the renderer's stream and waited for. Every call
 * on one GPU shares that GPU's array: calls that overlap on one GPU, from any renderer, race. */
skm_status skm_argsort_u32_device_static(skm_renderer* r, void* keys_dev, uint32_t n, int key_bits,
                                         void* order_dev);

/* ---- the loader's spatial order (reorderSpatially, packages/splat-core/src/sorting/SpatialOrder.cpp:48-92) ---- */
/* mortonCode (SpatialOrder.cpp:39-44): 30-bit Morton code of a point quantised to 1024 cells per
 * axis of the box [lo, hi] (an axis where the box is flat quantises to 0). */
uint32_t skm_morton_code(const float position[3], const float lo[3], const float hi[3]);
/* Permutes a host cloud in place along the Morton curve over [lo, hi] (the cloud's bounds), on the
 * GPU: a stable sort of the codes, so equal codes keep their input order, then every attribute array
 * gathered in that order (positions 3, covariances 6, colours 3, alphas 1 and sh_floats SH floats per
 * splat; sh may be NULL when sh_floats is 0). order (may be NULL) receives the source index of each
 * splat. */
skm_status skm_reorder_spatially(skm_renderer* r, float* positions, float* covariances, float* colors,
                                 float* alphas, float* sh, uint32_t sh_floats, uint32_t count, const float lo[3],
                                 const float hi[3], uint32_t* order);

/* ---- the LOD threshold (LodThreshold, packages/splatkit-engine/include/splatkit/rendering/LodThreshold.h:9-45) ----
 * The screen-space error a tree world's cut runs at: the policy's, raised 10% per completed frame
 * whose refinements the budget denied (up to max(64 px, policy)), relaxed 3% per frame whose cut
 * stays under 90% of its limit, reset by a policy change (MetalFrameCompute.mm:166-172). The
 * renderer adapts one of these per frame; the value type is exported for callers and tests. */
typedef struct skm_lod_threshold {
  float policy_pixels;
  float pixels;
} skm_lod_threshold;
void skm_lod_threshold_reset(skm_lod_threshold* t, float policy_pixels);
/* Takes one completed selection (nodes selected, refinements denied, the cut's limit) and returns
 * the threshold for the next. */
float skm_lod_threshold_update(skm_lod_threshold* t, uint32_t selected, uint32_t limited, uint32_t limit);
/* The renderer's threshold after the frames completed so far: what its next frame's cut runs at. */
float skm_lod_threshold_pixels(skm_renderer* r);

/* ---- the LOD cut on its own (MetalLodSelection, MetalLodSelection.h:13-48) ---- */
/* Uploads the tree's clusters and leaf references, selects one cut for this camera and reads it
 * back: capacity is the most nodes the cut may hold, pixel_limit 0 refines every cluster. */
skm_status skm_lod_select(skm_renderer* r, const struct skm_lod_tree_view* tree, const float* view16,
                          const float* proj16, uint32_t width, uint32_t height, uint32_t capacity,
                          float pixel_limit, float color_weight, int cull, uint32_t* indices,
                          uint32_t* stats6);

/* ---- view batches: many cameras over the resident world in one graph launch ----
 * No SplatRenderer counterpart: this extends draw() (SplatRenderer.h:134) to many views per launch,
 * beside the per-slot cameras and counts of the frame slots (MetalVisibilityPass.h:89-92). Every
 * view draws the same ranges with the renderer's policy; each view's survivors, depth order, tile
 * pairs and pixels are those of a single draw() of that view at that extent, bit for bit. */
typedef struct skm_view {
  float view[16];
  float proj[16];
  float camera_position[3];
  uint32_t width;
  uint32_t height;
  uint32_t reserved;
} skm_view;

typedef struct skm_batch_tuning {
  uint32_t tile_size;       /* raster tile edge in pixels, as skm_config (0 = 16) */
  uint32_t tile_radix_bits; /* digit width of the tile sort's passes: 6, 7 or 8 (0 = 8); keys of k bits
                             * take ceil(k / width) passes of even widths */
  uint32_t use_graph;       /* 1: capture the batch's kernels in a hipGraph and replay it */
  uint32_t sort_kernel;     /* radix pass variant, as skm_config.sort_kernel */
  uint32_t raster_px;       /* raster pixels per lane side, as skm_config.raster_px */
  uint32_t reserved;
} skm_batch_tuning;

typedef struct skm_batch_stats {
  uint32_t views;
  uint32_t graph;     /* the batch ran as a graph replay */
  double gpu_ms;      /* GPU time of the whole batch */
  uint64_t visible;   /* survivors over all views */
  uint64_t pairs;     /* tile pairs over all views */
  uint32_t tile_passes;      /* radix passes of the tile sort over the pairs */
  uint32_t view_sort_passes; /* radix passes on the view over the survivors (view-local tile keys), or 0 */
} skm_batch_stats;

/* Enqueues the batch (waits for the previous batch first). ranges NULL draws the whole world.
 * tuning NULL uses the renderer's current batch tuning (skm_autotune_views). A batch whose tile
 * pairs overflowed is dropped: skm_finish then returns SKM_ERR_CAPACITY and the next draw fits. */
skm_status skm_draw_views(skm_renderer* r, const skm_view* views, uint32_t count, const skm_range* ranges,
                          uint32_t range_count, const skm_batch_tuning* tuning);
skm_status skm_read_batch_stats(skm_renderer* r, skm_batch_stats* out);

/* The batch autotuner: times every supported configuration (raster tile and block shape, radix pass
 * variant, tile-sort digit width of 8, 7 or 6 bits) on this batch of views, `repeats` timed launches each after one untimed launch, and makes
 * the fastest the renderer's batch tuning (used when skm_draw_views gets tuning NULL). results gets
 * up to `capacity` candidates with their best GPU milliseconds; *best the choice. Every candidate
 * gives the same pixels: the composite of a pixel does not depend on tiles or block shape. */
/* Applies a tuning (tile size, raster pixels per lane, radix pass variant, tile-sort digit width) to this renderer's
 * single-view draws as well; the next frame uses it. */
skm_status skm_set_tuning(skm_renderer* r, const skm_batch_tuning* tuning);

typedef struct skm_tuning_result {
  skm_batch_tuning tuning;
  double gpu_ms;
} skm_tuning_result;
skm_status skm_autotune_views(skm_renderer* r, const skm_view* views, uint32_t count, const skm_range* ranges,
                              uint32_t range_count, uint32_t repeats, skm_tuning_result* results,
                              uint32_t capacity, uint32_t* result_count, skm_batch_tuning* best);
/* Read-backs of one view of the last completed batch, in skm_read_frame / skm_read_visible /
 * skm_read_order / skm_read_tiles terms (indices local to the view). */
skm_status skm_read_view_frame(skm_renderer* r, uint32_t view, int format, void* dst, size_t dst_bytes);
skm_status skm_read_view_visible(skm_renderer* r, uint32_t view, uint32_t* count, uint32_t* keys,
                                 uint32_t* projected, uint32_t capacity);
skm_status skm_read_view_order(skm_renderer* r, uint32_t view, uint32_t* order, uint32_t capacity);
skm_status skm_read_view_tiles(skm_renderer* r, uint32_t view, uint64_t* pairs, uint32_t* pair_tiles,
                               uint32_t* pair_values, uint64_t capacity, uint32_t* ranges,
                               uint32_t tile_capacity);

#ifdef __cplusplus
}
#endif
