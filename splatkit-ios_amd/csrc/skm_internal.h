// Layouts shared by the HIP kernels (kernels.hip) and the host renderer (renderer.cpp).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace skm {

// Camera uniform, byte-identical to CameraUniform (MetalShaderTypes.h:14-30) and the shader's
// Camera (SplatTypes.metalh:7-22).
struct Camera {
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
};
static_assert(sizeof(Camera) == 208, "camera layout");

// The resident world in HBM, structure of arrays: every field of the reference's 32-byte GpuSplat
// record (GpuLayout.h:50-56) in its own array so a wave's 64 lanes read 256 contiguous bytes per
// field, and the half-packed SH words (GpuLayout.cpp:92-113) word-major: word w of splat i at
// sh[w * capacity + i].
struct World {
  const float* px;
  const float* py;
  const float* pz;
  const uint32_t* rgba8;
  const uint32_t* cov0;
  const uint32_t* cov1;
  const uint32_t* cov2;
  const uint32_t* lod_alpha;
  const uint32_t* sh;
  uint32_t capacity;
  int32_t sh_degree;
};

struct WorldOut {
  float* px;
  float* py;
  float* pz;
  uint32_t* rgba8;
  uint32_t* cov0;
  uint32_t* cov1;
  uint32_t* cov2;
  uint32_t* lod_alpha;
  uint32_t* sh;
  uint32_t capacity;
  int32_t sh_degree;
};

// Device-side control block of one frame pipeline. The prologue kernel bumps `epoch` and clears
// the rest; look-back status words carry a tag derived from the epoch, so the status arrays are
// never cleared.
struct Ctrl {
  uint32_t epoch;
  uint32_t visible;   // survivors of the cull
  uint32_t pairs;     // tile pairs, saturated at 2^32 - 1
  uint32_t overflow;  // nonzero when the pairs exceeded the pair capacity
  uint32_t sort_pairs;  // pairs the tile sort orders: `pairs`, or 0 after an overflow
  uint32_t cull_lanes;  // cull tiles x kCullTile: the extent of the cull's gapped survivor layout
  uint32_t reserved[2];
  uint32_t ticket[16];
  uint32_t lod[16];   // LOD selection counters
  // (pixel, record) coverage tests the raster performed: every record a wave walked, times its
  // pixels (records past a wave's or tile's termination are not walked). Block b adds its count to
  // slot b % 8, so the adds of concurrent blocks spread over several addresses.
  unsigned long long raster_tests[8];
};

enum Ticket : uint32_t {
  kTicketCull = 0,
  kTicketBin = 1,
  kTicketDepthSort = 2,  // + pass (4)
  kTicketTileSort = 6,   // + pass (4)
};

// How a radix pass finds where each tile's digits go: a reduce-then-scan of per-tile digit counts
// (three kernels, no inter-block waiting), or onesweep's decoupled look-back (one kernel).
enum SortVariant : int { kSortReduceScan = 0, kSortOnesweep = 1 };

// The digits of an LSD radix sort: pass p sorts by the `bits[p]` bits at `shift[p]`, low passes
// first; the widths sum to the key width. At most 4 passes of at most 8 bits.
struct DigitPlan {
  int passes;
  uint32_t shift[4];
  uint32_t bits[4];
};

// Even digits of at most `width` bits (6, 7 or 8) for keys of `key_bits` bits.
inline DigitPlan make_digit_plan(uint32_t key_bits, uint32_t width) {
  key_bits = key_bits == 0 ? 1 : key_bits;
  width = width < 6 ? 6 : (width > 8 ? 8 : width);
  DigitPlan p{};
  p.passes = static_cast<int>((key_bits + width - 1) / width);
  if (p.passes > 4) p.passes = 4;
  const uint32_t base = key_bits / static_cast<uint32_t>(p.passes);
  const uint32_t extra = key_bits % static_cast<uint32_t>(p.passes);
  uint32_t shift = 0;
  for (int k = 0; k < p.passes; ++k) {
    p.shift[k] = shift;
    p.bits[k] = base + (static_cast<uint32_t>(k) < extra ? 1u : 0u);
    shift += p.bits[k];
  }
  return p;
}

// Bits needed to write `max_key`.
inline uint32_t key_bits_for(uint32_t max_key) {
  uint32_t b = 0;
  while (b < 32 && (max_key >> b) != 0) ++b;
  return b;
}

constexpr uint32_t kThreads = 256;
// Lanes per thread of the cull (a kCullTile-lane tile per block); a build may set it.
#ifndef SKM_CULL_ITEMS
#define SKM_CULL_ITEMS 2
#endif
constexpr uint32_t kCullItems = SKM_CULL_ITEMS;
constexpr uint32_t kCullTile = kThreads * kCullItems;
#ifndef SKM_BIN_ITEMS
#define SKM_BIN_ITEMS 4
#endif
constexpr uint32_t kBinItems = SKM_BIN_ITEMS;  // survivors per thread of the binning (1024 per block)
// Keys per thread of a radix pass (a block sorts 256 x kSortItems keys); a build may set it.
#ifndef SKM_SORT_ITEMS
#define SKM_SORT_ITEMS 12
#endif
constexpr uint32_t kSortItems = SKM_SORT_ITEMS;
constexpr uint32_t kSortTile = kThreads * kSortItems;  // keys per radix-pass block (3072 at 12)
// Binning blocks per radix-pass block of survivors (the binning's block totals take one word each).
constexpr uint32_t kBinBlocksPerSortTile = (kSortTile + kThreads * kBinItems - 1u) / (kThreads * kBinItems);
constexpr uint32_t kRadix = 256;
constexpr uint32_t kMaxRanges = 65536;  // MetalVisibilityPass.h:72
constexpr uint32_t kOrderCounters = 2u * 4u * 256u;  // offset of the tile-order counters in hist
// The tile order's counters: tiles per bucket [0, 32), cursors [32, 64), and the raster's split
// tiles [64].
constexpr uint32_t kHistWords = kOrderCounters + 65u;
// The raster splits the heaviest whole buckets of its tile order, up to this many tiles, when frames
// do not overlap (one frame slot): a split tile's four quadrant blocks finish its long walk in about a
// third of the time, at about three times the VALU work, which overlapped frames would rather have.
constexpr uint32_t kSplitTiles = 1024;
constexpr float kTransmittanceStop = 1.0f / 4096.0f;

// One view of a batch: its extent, its tiles' place in the batch's global tile ids, and its first
// pixel in the batch image. A single-view frame has none (its extent sits in FrameArgs).
struct ViewInfo {
  uint32_t width, height, tiles_x, tiles_y;
  uint32_t tile_base;
  uint32_t reserved;
  uint64_t pixel_base;
};
static_assert(sizeof(ViewInfo) == 32, "view layout");

struct FrameArgs {
  const Camera* cam;
  World world;
  const uint32_t* ranges;  // pairs (offset, count)
  const uint32_t* range_starts;
  uint32_t range_count;
  uint32_t total;          // lanes for range mode
  const uint32_t* active;  // LOD indices, or null
  const uint32_t* active_count;
  int key16;
  uint32_t width, height, tile, tiles_x, tiles_y;
  float background[3];
  // scratch
  Ctrl* ctrl;
  uint32_t* cull_keys;  // depth keys in compaction order, as the cull wrote them
  uint32_t* keys[2];    // sort ping-pong
  uint32_t* order[2];
  uint32_t* projected;  // 8 words per survivor
  uint64_t* rects;      // per survivor: its packed tile rectangle (written by the cull, read by the binning)
  // Survivors sit in a gapped layout: cull tile b writes its survivors, in lane order, from position
  // b * kCullTile and its count to cull_counts[b]. Ordering by position is compaction order; the
  // depth sort's first pass skips the gaps, so the sorted order holds each survivor's position.
  uint32_t* cull_counts;
  uint64_t* scan_status;
  uint64_t* sort_status;
  uint32_t* hist;       // [2][4][256]: depth digits, tile digits
  uint32_t* pair_keys[2];
  uint32_t* pair_vals[2];
  uint32_t pair_capacity;
  uint32_t* tile_ranges;  // 2 words per tile
  // The raster's launch order of the tiles (one word per tile): the tiles with the most pairs first,
  // so the longest blocks start earliest. Built per frame from the ranges, with 64 counter words
  // at hist + kOrderCounters. Null: tile order.
  uint32_t* tile_order;
  float* image;           // RGBA float, top-down rows
  uint32_t capacity;      // survivor capacity
  // Batches (view_count > 1): one camera per view at cam[v]; every view culls the same lanes
  // [0, total). Cull tiles go tile-major: block b culls lane tile b / views for view b % views, so
  // the views of one lane tile run as neighbouring blocks and its records come from HBM once and
  // from cache for the other views. Ordering survivors by position is (tile, view, lane) order,
  // which is lane order within each view. Null for a single view.
  uint32_t view_count;
  const ViewInfo* views;
  uint32_t* view_visible;      // survivors per view
  const uint16_t* tile_view;   // the view of each global tile
  uint32_t total_tiles;        // tiles of all views
  // Views in one batch (view_count > 1): nonzero, the tile keys of the binning are
  // view << local_tile_bits | the tile's index within its view, and the depth sort ends with a pass
  // on the view, so that the pairs come out grouped by view and the tile sort needs only the local
  // bits. Zero: tile keys are global tile ids.
  uint32_t local_tile_bits;
  // Expected cull lanes, survivors and tile pairs, from earlier frames (0: the capacity): they size the
  // grids of the radix passes, binning and ranges, whose blocks loop past an estimate.
  uint32_t lane_estimate, survivor_estimate, pair_estimate;
  // Radix sort variant and its tile-count scratch (256 words per kSortTile-key tile of the larger of the
  // survivor and pair capacities).
  int sort_variant;
  uint32_t* sort_counts;
  // The binning's survivor-block totals, then starts (one word per 256 kBinItems survivors), kept
  // apart from sort_counts: the emission pass reads them while it writes the tile sort's first counts.
  uint32_t* bin_starts;
  // Raster block shape: pixels per lane side (1 or 2) at tile size `tile` (8, 16 or 32); the
  // supported pairs are (8, 1), (16, 1), (16, 2), (32, 2).
  uint32_t raster_px;
  // Nonzero: the raster adds its (pixel, record) coverage tests to ctrl->raster_tests (one atomic
  // per block, a few percent of a light frame's raster time).
  uint32_t count_tests;
  // At most this many of the heaviest tiles are drawn split (0: none).
  uint32_t split_cap;
};

// Whether the raster has a block shape for (tile, pixels per lane side).
inline bool raster_shape_supported(uint32_t tile, uint32_t px) {
  return (tile == 8 && px == 1) || (tile == 16 && (px == 1 || px == 2)) || (tile == 32 && px == 2);
}

// Launchers (kernels.hip).
hipError_t launch_pack(const WorldOut& out, uint32_t offset, uint32_t count, const float* positions,
                       const float* covariances, const float* colors, const float* alphas,
                       const float* sh, int src_degree, hipStream_t stream);
struct LodArgs;
// marks: 6 events recorded at the start, after the LOD cut, the cull, the depth sort, the tile
// binning/sort/ranges and the raster.
hipError_t launch_frame(const FrameArgs& a, const LodArgs* lod, const DigitPlan& depth_plan,
                        const DigitPlan& tile_plan, hipEvent_t* marks, hipStream_t stream);
// BGRA8 bytes (one u32 per pixel) of an RGBA float image.
hipError_t launch_to_bgra8(const float* rgba, uint32_t* out, uint64_t pixels, hipStream_t stream);
hipError_t launch_argsort(Ctrl* ctrl, uint32_t* keys[2], uint32_t* vals[2], const uint32_t* n_dev,
                          uint32_t capacity, const DigitPlan& plan, uint32_t* hist, uint64_t* status,
                          uint32_t* counts, int variant, hipStream_t stream);
// The 30-bit Morton code of each of n points (xyz interleaved) over the box [lo, hi]
// (mortonCode, SpatialOrder.cpp:39-44).
hipError_t launch_morton(const float* positions, uint32_t n, const float lo[3], const float hi[3], uint32_t* codes,
                         hipStream_t stream);
// out row i = in row order[i], rows of `stride` floats.
hipError_t launch_gather_rows(const float* in, float* out, const uint32_t* order, uint32_t n, uint32_t stride,
                              hipStream_t stream);

// mortonCode (SpatialOrder.cpp:10-44): each coordinate quantised to 1024 cells of the box (0 when the
// box is flat on that axis), bits interleaved x, y, z from the lowest.
__host__ __device__ inline uint32_t morton_spread(uint32_t v) {
  v = (v | (v << 16)) & 0x030000ffu;
  v = (v | (v << 8)) & 0x0300f00fu;
  v = (v | (v << 4)) & 0x030c30c3u;
  v = (v | (v << 2)) & 0x09249249u;
  return v;
}
__host__ __device__ inline uint32_t morton_cell(float value, float lo, float hi) {
  const float extent = hi - lo;
  if (!(extent > 0.0f)) return 0u;
  const float scaled = floorf(((value - lo) / extent) * 1024.0f);
  return static_cast<uint32_t>(scaled < 0.0f ? 0.0f : (scaled > 1023.0f ? 1023.0f : scaled));
}
__host__ __device__ inline uint32_t morton_code(const float p[3], const float lo[3], const float hi[3]) {
  return morton_spread(morton_cell(p[0], lo[0], hi[0])) | (morton_spread(morton_cell(p[1], lo[1], hi[1])) << 1) |
         (morton_spread(morton_cell(p[2], lo[2], hi[2])) << 2);
}

}  // namespace skm
