// Batches of views over one resident world: every stage of the frame pipeline runs once for all
// views (cull tiles tile-major over the views, one depth sort over every view's survivors, tile
// pairs keyed by batch-global tile ids, one raster grid over every view's tiles), captured into a
// hipGraph and replayed while the batch's shape stays the same. Per view, the survivors, their depth
// order, the (tile | depth) pairs and the pixels are the single-view frame's, bit for bit: survivor
// order is lane order within a view, the sorts are stable, and global tile ids keep each view's
// tiles contiguous and in order.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "skm.h"
#include "skm_host.h"
#include "skm_internal.h"

namespace skm {

struct BatchPolicy {
  bool key16 = true;
  bool linear = false;
  float subpixel = 1.0f;
  float background[3] = {0.05f, 0.05f, 0.08f};
};

class Batch {
 public:
  Batch() = default;
  Batch(const Batch&) = delete;
  Batch& operator=(const Batch&) = delete;
  ~Batch();

  // Enqueues one batch on `stream`: waits for this object's previous batch first (its buffers are
  // reused). `ranges` are (offset, count) pairs over the world, validated by the caller; every view
  // draws them. Fails without launching when a buffer cannot be allocated.
  hipError_t launch(const World& world, uint32_t world_count, int sh_degree, const skm_view* views,
                    uint32_t count, const skm_range* ranges, uint32_t range_count, const BatchPolicy& policy,
                    const skm_batch_tuning& tuning, hipStream_t stream, std::string* error);
  // Waits for the last launched batch. *dropped is set when its tile pairs overflowed the pair
  // capacity: the capacity has then been marked to grow and the batch must be launched again.
  hipError_t wait(bool* dropped);
  bool in_flight() const { return in_flight_; }
  // The pair capacity cannot grow further (32-bit pair indices): a batch that still overflows can
  // never be drawn whole.
  bool at_pair_limit() const;

  uint32_t view_count() const { return static_cast<uint32_t>(info_.size()); }
  const ViewInfo& info(uint32_t v) const { return info_[v]; }
  uint32_t view_visible(uint32_t v) const { return counts_host()[v]; }
  uint64_t visible() const { return ctrl_host()->visible; }
  uint64_t pairs() const { return ctrl_host()->pairs; }
  double gpu_ms() const { return gpu_ms_; }
  int tile_passes() const { return tile_passes_; }
  int view_sort_passes() const { return survivor_passes_ - depth_passes_; }
  const float* image_device() const { return image_.as<float>(); }
  uint64_t pixels() const { return info_.empty() ? 0 : info_.back().pixel_base + uint64_t{info_.back().width} * info_.back().height; }
  // Device bytes held now, and what a batch of this shape would hold (scratch grows, never shrinks).
  uint64_t device_bytes() const;
  uint64_t bytes_for(uint32_t views, uint64_t lanes, uint32_t tiles, uint64_t pixels) const;
  // Frees every device buffer (no batch may be in flight).
  void release();
  bool graph_used() const { return graph_used_; }

  // Read-backs of the last completed batch, in the single-view frame's terms.
  hipError_t read_image(uint32_t v, float* rgba) const;
  hipError_t read_visible(uint32_t v, uint32_t* keys, uint32_t* projected, uint32_t capacity) const;
  hipError_t read_order(uint32_t v, uint32_t* order, uint32_t capacity) const;
  hipError_t read_tiles(uint32_t v, uint64_t* pairs, uint32_t* pair_tiles, uint32_t* pair_values,
                        uint64_t capacity, uint32_t* ranges, uint32_t tile_capacity) const;

 private:
  const Ctrl* ctrl_host() const { return host_.as<Ctrl>(); }
  const uint32_t* counts_host() const { return reinterpret_cast<const uint32_t*>(host_.as<char>() + sizeof(Ctrl)); }
  // The cull's gapped layout of the last batch (read back on demand): survivors per cull tile, and
  // the index among its view's survivors of each tile's first survivor. Cull tile b belongs to view
  // b % views (tile-major tickets).
  hipError_t survivor_tiles() const;
  uint32_t tile_view_of(uint32_t b) const { return info_.size() > 1 ? b % static_cast<uint32_t>(info_.size()) : 0u; }
  // A survivor's index among its view's survivors, from its position in the gapped layout.
  uint32_t local_index(uint32_t pos) const { return st_local_[pos / kCullTile] + pos % kCullTile; }
  mutable std::vector<uint32_t> st_count_, st_local_;
  mutable bool st_ready_ = false;
  hipError_t ensure(uint32_t views, uint64_t lanes, uint64_t pairs, uint32_t tiles, uint64_t pixels,
                    uint32_t ranges, hipStream_t stream);
  void release_graph();

  // Device scratch.
  DeviceBuffer ctrl_, cams_, views_, tile_view_, view_visible_;
  DeviceBuffer cull_keys_, keys_[2], order_[2], projected_, rects_, scan_status_, sort_status_, sort_counts_, hist_;
  DeviceBuffer pair_keys_[2], pair_vals_[2], tile_ranges_, image_, ranges_, starts_;
  PinnedBuffer staging_, host_;
  size_t staging_bytes_ = 0, host_bytes_ = 0;
  uint64_t lane_capacity_ = 0, pair_capacity_ = 0, pair_want_ = 0, pixel_capacity_ = 0;
  uint32_t view_capacity_ = 0, tile_capacity_ = 0, range_capacity_ = 0;
  size_t scan_words_ = 0, sort_words_ = 0;
  uint64_t pipelines_ = 0;

  std::vector<ViewInfo> info_;
  std::vector<uint16_t> tile_view_host_;
  int depth_passes_ = 4, tile_passes_ = 2;
  int survivor_passes_ = 4;        // depth passes, then the passes on the view (view-local tile keys)
  uint32_t local_tile_bits_ = 0;   // nonzero: the tile keys are view << bits | the tile's index in the view
  hipStream_t stream_ = nullptr;
  hipEvent_t start_ = nullptr, end_ = nullptr;
  bool in_flight_ = false;
  double gpu_ms_ = 0.0;

  // The captured pipeline and what it was captured for.
  hipGraphExec_t exec_ = nullptr;
  std::vector<uint64_t> graph_key_;
  bool graph_used_ = false;
  // The lanes and pixels of the launch in flight, and the counts of the latest finished launch with
  // its lanes and pixels: an eager launch's grids are estimated from them.
  uint64_t launch_lanes_ = 0, launch_pixels_ = 0;
  uint64_t last_lanes_ = 0, last_pixels_ = 0, last_visible_ = 0, last_pairs_ = 0;
  bool counts_known_ = false;
};

// The renderer's batch and stream, for the server (renderer.cpp).
Batch* renderer_batch(skm_renderer* r);
hipStream_t renderer_stream(skm_renderer* r);
uint64_t renderer_world_bytes(const skm_renderer* r);

}  // namespace skm
