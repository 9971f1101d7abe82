// View batches (skm_batch.h): scratch sizing, camera and view uploads, hipGraph capture and
// replay of the frame pipeline over all views, and per-view read-backs in single-frame terms.
#include "skm_batch.h"

#include <algorithm>
#include <cstring>

namespace skm {

namespace {

constexpr uint32_t kMaxViews = 4096;
constexpr uint64_t kLaneLimit = 0xffff0000ull;  // survivor and pair indices are 32-bit

hipError_t grow(DeviceBuffer& b, size_t bytes) {
  if (b.bytes >= bytes && b.ptr != nullptr) return hipSuccess;
  return b.allocate(bytes);
}

// A capacity that has to grow grows by at least half: batches of changing shape (a served stream)
// then reallocate a few times rather than at every new maximum, and every reallocation frees the old
// buffers, which waits for the whole device.
uint64_t grown(uint64_t want, uint64_t have, uint64_t limit) {
  if (want <= have) return want;
  return std::min(std::max(want, have + have / 2), std::max(want, limit));
}

}  // namespace

Batch::~Batch() {
  if (in_flight_) (void)hipEventSynchronize(end_);
  release_graph();
  if (start_) (void)hipEventDestroy(start_);
  if (end_) (void)hipEventDestroy(end_);
}

void Batch::release_graph() {
  if (exec_ != nullptr) (void)hipGraphExecDestroy(exec_);
  exec_ = nullptr;
  graph_key_.clear();
}

hipError_t Batch::ensure(uint32_t views, uint64_t lanes, uint64_t pairs, uint32_t tiles, uint64_t pixels,
                         uint32_t ranges, hipStream_t stream) {
  lanes = grown(lanes, lane_capacity_, kLaneLimit) / kCullTile * kCullTile;
  pairs = grown(pairs, pair_capacity_, kLaneLimit);
  tiles = static_cast<uint32_t>(grown(tiles, tile_capacity_, 1u << 24));
  pixels = grown(pixels, pixel_capacity_, ~0ull);
  hipError_t e = hipSuccess;
  bool moved = false;
  auto need = [&](DeviceBuffer& b, size_t bytes) {
    if (e != hipSuccess || (b.ptr != nullptr && b.bytes >= bytes)) return;
    e = grow(b, bytes);
    moved = true;
  };
  if (ctrl_.ptr == nullptr) {
    need(ctrl_, sizeof(Ctrl));
    need(hist_, kHistWords * 4);
    if (e == hipSuccess) e = hipMemsetAsync(ctrl_.ptr, 0, sizeof(Ctrl), stream);
  }
  if (views > view_capacity_) {
    need(cams_, size_t{views} * sizeof(Camera));
    need(views_, size_t{views} * sizeof(ViewInfo));
    need(view_visible_, size_t{views} * 4);
    if (e == hipSuccess) view_capacity_ = views;
  }
  if (lanes > lane_capacity_) {
    const size_t n = static_cast<size_t>(lanes);
    need(cull_keys_, n * 4);
    need(keys_[0], n * 4);
    need(keys_[1], n * 4);
    need(order_[0], n * 4);
    need(order_[1], n * 4);
    need(projected_, n * 32);
    need(rects_, n * 8);
    if (e == hipSuccess) lane_capacity_ = lanes;
  }
  if (pairs > pair_capacity_) {
    const size_t n = static_cast<size_t>(pairs);
    need(pair_keys_[0], n * 4);
    need(pair_keys_[1], n * 4);
    need(pair_vals_[0], n * 4);
    need(pair_vals_[1], n * 4);
    if (e == hipSuccess) pair_capacity_ = pairs;
  }
  if (tiles > tile_capacity_) {
    need(tile_ranges_, size_t{tiles} * 12);  // ranges, then the raster's tile order
    need(tile_view_, size_t{tiles} * 2);
    if (e == hipSuccess) tile_capacity_ = tiles;
  }
  if (pixels > pixel_capacity_) {
    need(image_, static_cast<size_t>(pixels) * 16);
    if (e == hipSuccess) pixel_capacity_ = pixels;
  }
  if (ranges > range_capacity_ || ranges_.ptr == nullptr) {
    need(ranges_, std::max<size_t>(ranges, 1) * 8);
    need(starts_, (std::max<size_t>(ranges, 1) + 1) * 4);
    if (e == hipSuccess) range_capacity_ = std::max(ranges, 1u);
  }
  // Look-back status words: one per cull tile and per binning block; 256 per 4096-key sort tile.
  const size_t scan = static_cast<size_t>(lane_capacity_ / std::min(kCullTile, 256u * kBinItems) + 2);
  if (e == hipSuccess && scan > scan_words_) {
    need(scan_status_, scan * 8);
    if (e == hipSuccess) e = hipMemsetAsync(scan_status_.ptr, 0, scan_status_.bytes, stream);
    if (e == hipSuccess) scan_words_ = scan;
  }
  const size_t sort = static_cast<size_t>(std::max(lane_capacity_, pair_capacity_) / kSortTile + 2) * kRadix;
  if (e == hipSuccess && sort > sort_words_) {
    need(sort_status_, sort * 8);
    // The tile counts, then the binning's block totals and starts (kBinBlocksPerSortTile per sort tile
    // of survivors, and 2 spare).
    need(sort_counts_, (sort + sort / kRadix * kBinBlocksPerSortTile + 2) * 4);
    if (e == hipSuccess) e = hipMemsetAsync(sort_status_.ptr, 0, sort_status_.bytes, stream);
    if (e == hipSuccess) sort_words_ = sort;
  }
  // Staged inputs: cameras, view descriptors, ranges and their starts, then the view of each tile.
  const size_t stage = size_t{views} * (sizeof(Camera) + sizeof(ViewInfo)) + size_t{ranges} * 8 +
                       (size_t{ranges} + 1) * 4 + 64 + size_t{tiles} * 2 + 16;
  if (e == hipSuccess && stage > staging_bytes_) {
    e = staging_.allocate(stage * 2);
    if (e == hipSuccess) staging_bytes_ = stage * 2;
  }
  const size_t host = sizeof(Ctrl) + (size_t{views} + 1) * 4;
  if (e == hipSuccess && host > host_bytes_) {
    e = host_.allocate(host * 2);
    if (e == hipSuccess) host_bytes_ = host * 2;
  }
  if (e == hipSuccess && start_ == nullptr) e = hipEventCreate(&start_);
  if (e == hipSuccess && end_ == nullptr) e = hipEventCreate(&end_);
  if (moved) release_graph();  // the captured kernels hold the old addresses
  return e;
}

hipError_t Batch::launch(const World& world, uint32_t world_count, int sh_degree, const skm_view* views,
                         uint32_t count, const skm_range* ranges, uint32_t range_count,
                         const BatchPolicy& policy, const skm_batch_tuning& tuning, hipStream_t stream,
                         std::string* error) {
  if (in_flight_) {
    bool dropped = false;
    if (const hipError_t e = wait(&dropped); e != hipSuccess) return e;
  }
  const uint32_t kTile = tuning.tile_size ? tuning.tile_size : 16u;
  const uint32_t px = tuning.raster_px ? tuning.raster_px : (kTile == 32u ? 2u : 1u);
  const uint32_t radix_bits = tuning.tile_radix_bits ? tuning.tile_radix_bits : 8u;
  if (count == 0 || count > kMaxViews || !raster_shape_supported(kTile, px) || radix_bits < 6 || radix_bits > 8 ||
      tuning.sort_kernel > 1) {
    *error = "batch: 1 to 4096 views, a supported raster shape and 6- to 8-bit tile digits";
    return hipErrorInvalidValue;
  }
  uint64_t total = 0;
  for (uint32_t i = 0; i < range_count; ++i) total += ranges[i].count;
  const uint64_t view_lanes = (total + kCullTile - 1) / kCullTile * kCullTile;
  const uint64_t lanes = view_lanes * count;
  info_.assign(count, ViewInfo{});
  uint64_t tiles = 0, pixels = 0;
  for (uint32_t v = 0; v < count; ++v) {
    const uint32_t W = views[v].width, H = views[v].height;
    if (W == 0 || H == 0 || W > 16384 || H > 16384) {
      *error = "batch: view extent out of range";
      return hipErrorInvalidValue;
    }
    ViewInfo& vi = info_[v];
    vi.width = W;
    vi.height = H;
    vi.tiles_x = (W + kTile - 1) / kTile;
    vi.tiles_y = (H + kTile - 1) / kTile;
    vi.tile_base = static_cast<uint32_t>(tiles);
    vi.pixel_base = pixels;
    tiles += uint64_t{vi.tiles_x} * vi.tiles_y;
    pixels += uint64_t{W} * H;
  }
  if (lanes > kLaneLimit || tiles >= (1ull << 24)) {
    *error = "batch: too many lanes or tiles for 32-bit indices";
    return hipErrorInvalidValue;
  }
  const uint64_t want_pairs = std::min<uint64_t>(
      std::max<uint64_t>({pair_want_, lanes * 3, uint64_t{1} << 22}), kLaneLimit);
  if (hipError_t e = ensure(count, lanes, want_pairs, static_cast<uint32_t>(tiles), pixels, range_count, stream);
      e != hipSuccess) {
    *error = "batch: scratch allocation failed";
    return e;
  }
  stream_ = stream;

  // Inputs: cameras, view descriptors, ranges (one staged copy each).
  char* st = staging_.as<char>();
  auto* cams = reinterpret_cast<Camera*>(st);
  for (uint32_t v = 0; v < count; ++v)
    fill_camera(cams[v], views[v].view, views[v].proj, views[v].width, views[v].height, policy.linear,
                policy.subpixel, views[v].camera_position, nullptr);
  auto* vinfo = reinterpret_cast<ViewInfo*>(st + size_t{count} * sizeof(Camera));
  std::memcpy(vinfo, info_.data(), size_t{count} * sizeof(ViewInfo));
  auto* rh = reinterpret_cast<uint32_t*>(st + size_t{count} * (sizeof(Camera) + sizeof(ViewInfo)));
  uint32_t* sh = rh + 2 * size_t{range_count};
  sh[0] = 0;
  for (uint32_t i = 0; i < range_count; ++i) {
    rh[2 * i] = ranges[i].offset;
    rh[2 * i + 1] = ranges[i].count;
    sh[i + 1] = sh[i] + ranges[i].count;
  }
  hipError_t e = hipMemcpyAsync(cams_.ptr, cams, size_t{count} * sizeof(Camera), hipMemcpyHostToDevice, stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(views_.ptr, vinfo, size_t{count} * sizeof(ViewInfo), hipMemcpyHostToDevice, stream);
  if (e == hipSuccess && range_count > 0) {
    e = hipMemcpyAsync(ranges_.ptr, rh, size_t{range_count} * 8, hipMemcpyHostToDevice, stream);
    if (e == hipSuccess)
      e = hipMemcpyAsync(starts_.ptr, sh, (size_t{range_count} + 1) * 4, hipMemcpyHostToDevice, stream);
  }
  // The view of each global tile, uploaded on the stream (no device-wide wait) when the extents change.
  std::vector<uint16_t> tv(static_cast<size_t>(tiles));
  for (uint32_t v = 0; v < count; ++v)
    std::fill(tv.begin() + info_[v].tile_base, tv.begin() + info_[v].tile_base + info_[v].tiles_x * info_[v].tiles_y,
              static_cast<uint16_t>(v));
  if (e == hipSuccess && tv != tile_view_host_) {
    const size_t at = align_up(size_t{count} * (sizeof(Camera) + sizeof(ViewInfo)) + size_t{range_count} * 8 +
                                   (size_t{range_count} + 1) * 4,
                               16);
    std::memcpy(st + at, tv.data(), tv.size() * 2);
    e = hipMemcpyAsync(tile_view_.ptr, st + at, tv.size() * 2, hipMemcpyHostToDevice, stream);
    tile_view_host_ = std::move(tv);
  }
  if (e == hipSuccess && ++pipelines_ >= (1ull << 24)) {
    // Look-back tags carry 26 bits of epoch: clear the status words long before they repeat.
    pipelines_ = 1;
    e = hipMemsetAsync(scan_status_.ptr, 0, scan_status_.bytes, stream);
    if (e == hipSuccess) e = hipMemsetAsync(sort_status_.ptr, 0, sort_status_.bytes, stream);
  }
  if (e != hipSuccess) {
    *error = "batch: input upload failed";
    return e;
  }

  FrameArgs a{};
  a.cam = cams_.as<Camera>();
  a.world = world;
  a.ranges = ranges_.as<uint32_t>();
  a.range_starts = starts_.as<uint32_t>();
  a.range_count = range_count;
  a.total = static_cast<uint32_t>(total);
  a.key16 = policy.key16 ? 1 : 0;
  a.width = info_[0].width;
  a.height = info_[0].height;
  a.tile = kTile;
  a.tiles_x = info_[0].tiles_x;
  a.tiles_y = info_[0].tiles_y;
  for (int c = 0; c < 3; ++c) a.background[c] = policy.background[c];
  a.ctrl = ctrl_.as<Ctrl>();
  a.cull_keys = cull_keys_.as<uint32_t>();
  for (int k = 0; k < 2; ++k) {
    a.keys[k] = keys_[k].as<uint32_t>();
    a.order[k] = order_[k].as<uint32_t>();
    a.pair_keys[k] = pair_keys_[k].as<uint32_t>();
    a.pair_vals[k] = pair_vals_[k].as<uint32_t>();
  }
  a.projected = projected_.as<uint32_t>();
  a.rects = rects_.as<uint64_t>();
  a.scan_status = scan_status_.as<uint64_t>();
  a.cull_counts = reinterpret_cast<uint32_t*>(scan_status_.ptr);
  a.sort_status = sort_status_.as<uint64_t>();
  a.hist = hist_.as<uint32_t>();
  a.pair_capacity = static_cast<uint32_t>(pair_capacity_);
  a.tile_ranges = tile_ranges_.as<uint32_t>();
  a.tile_order = tile_ranges_.as<uint32_t>() + size_t{tile_capacity_} * 2;
  a.image = image_.as<float>();
  a.capacity = static_cast<uint32_t>(lane_capacity_);
  a.view_count = count;
  a.views = views_.as<ViewInfo>();
  a.view_visible = view_visible_.as<uint32_t>();
  a.tile_view = tile_view_.as<uint16_t>();
  a.total_tiles = static_cast<uint32_t>(tiles);
  a.sort_variant = tuning.sort_kernel == 1 ? kSortOnesweep : kSortReduceScan;
  a.raster_px = px;
  a.sort_counts = sort_counts_.as<uint32_t>();
  a.bin_starts = a.sort_counts + sort_words_;
  const DigitPlan depth_plan = make_digit_plan(policy.key16 ? 16u : 32u, 8u);
  DigitPlan tile_plan = make_digit_plan(key_bits_for(static_cast<uint32_t>(tiles - 1)), radix_bits);
  // View-local tile keys (view << bits | tile index in the view) where the tile sort then takes fewer
  // passes over the pairs than the sort on the view adds over the survivors (which are fewer).
  uint32_t view_tiles = 0;
  for (const ViewInfo& vi : info_) view_tiles = std::max(view_tiles, vi.tiles_x * vi.tiles_y);
  const uint32_t local_bits = key_bits_for(view_tiles - 1), view_bits = key_bits_for(count - 1);
  const DigitPlan local_plan = make_digit_plan(local_bits, radix_bits);
  const int view_passes = make_digit_plan(view_bits, 8u).passes;
  const bool local = count > 1 && local_bits + view_bits <= 32 && local_plan.passes < tile_plan.passes &&
                     view_passes <= tile_plan.passes - local_plan.passes;
  if (local) tile_plan = local_plan;
  a.local_tile_bits = local ? local_bits : 0u;
  local_tile_bits_ = a.local_tile_bits;
  // The cull's layout spans exactly `lanes`. An eager launch (a replayed graph keeps the grids it was
  // captured with) estimates survivors from the latest finished launch per lane and pairs per pixel.
  a.lane_estimate = static_cast<uint32_t>(lanes);
  if (tuning.use_graph == 0 && counts_known_ && last_lanes_ != 0 && last_pixels_ != 0) {
    auto ahead = [](uint64_t last, uint64_t now, uint64_t then, uint64_t slack) {
      const double v = static_cast<double>(last) * static_cast<double>(now) / static_cast<double>(then);
      return static_cast<uint32_t>(std::min(v * 1.5 + static_cast<double>(slack), 4294967295.0));
    };
    a.survivor_estimate = ahead(last_visible_, lanes, last_lanes_, 8u * kSortTile);
    a.pair_estimate = ahead(last_pairs_, pixels, last_pixels_, 8u * kSortTile);
  }
  launch_lanes_ = lanes;
  launch_pixels_ = pixels;
  depth_passes_ = depth_plan.passes;
  survivor_passes_ = depth_plan.passes + (local ? view_passes : 0);
  tile_passes_ = tile_plan.passes;
  (void)world_count;
  (void)sh_degree;

  graph_used_ = tuning.use_graph != 0;
  if (graph_used_) {
    // What the captured kernels bake in: sizes, passes, addresses (reset on reallocation), the
    // background and the world. Cameras, views and ranges are read from device memory.
    std::vector<uint64_t> key = {count, total, range_count, uint64_t(world.sh_degree), uint64_t(policy.key16),
                                 uint64_t(a.sort_variant), kTile, px, radix_bits,
                                 tiles, uint64_t(depth_passes_), uint64_t(tile_passes_), local_tile_bits_,
                                 reinterpret_cast<uint64_t>(world.px), world.capacity, pair_capacity_,
                                 lane_capacity_};
    for (int c = 0; c < 3; ++c) {
      uint32_t b;
      std::memcpy(&b, &policy.background[c], 4);
      key.push_back(b);
    }
    for (uint32_t v = 0; v < count; ++v) key.push_back((uint64_t{info_[v].width} << 32) | info_[v].height);
    if (exec_ == nullptr || key != graph_key_) {
      release_graph();
      hipGraph_t g = nullptr;
      e = hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal);
      if (e == hipSuccess) {
        const hipError_t le = launch_frame(a, nullptr, depth_plan, tile_plan, nullptr, stream);
        e = hipStreamEndCapture(stream, &g);
        if (e == hipSuccess) e = le;
      }
      if (e == hipSuccess) e = hipGraphInstantiateWithFlags(&exec_, g, 0);
      if (g != nullptr) (void)hipGraphDestroy(g);
      if (e != hipSuccess) {
        exec_ = nullptr;
        *error = "batch: graph capture failed";
        return e;
      }
      graph_key_ = std::move(key);
    }
    e = hipEventRecord(start_, stream);
    if (e == hipSuccess) e = hipGraphLaunch(exec_, stream);
  } else {
    e = hipEventRecord(start_, stream);
    if (e == hipSuccess) e = launch_frame(a, nullptr, depth_plan, tile_plan, nullptr, stream);
  }
  if (e == hipSuccess) e = hipEventRecord(end_, stream);
  if (e == hipSuccess) e = hipMemcpyAsync(host_.ptr, ctrl_.ptr, sizeof(Ctrl), hipMemcpyDeviceToHost, stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(host_.as<char>() + sizeof(Ctrl), view_visible_.ptr, size_t{count} * 4,
                       hipMemcpyDeviceToHost, stream);
  if (e != hipSuccess) {
    *error = "batch: launch failed";
    return e;
  }
  in_flight_ = true;
  st_ready_ = false;
  return hipSuccess;
}

hipError_t Batch::wait(bool* dropped) {
  *dropped = false;
  if (!in_flight_) return hipSuccess;
  in_flight_ = false;
  const hipError_t e = hipStreamSynchronize(stream_);
  if (e != hipSuccess) return e;
  float ms = 0.0f;
  (void)hipEventElapsedTime(&ms, start_, end_);
  gpu_ms_ = ms;
  last_lanes_ = launch_lanes_;
  last_pixels_ = launch_pixels_;
  last_visible_ = ctrl_host()->visible;
  last_pairs_ = ctrl_host()->pairs;
  counts_known_ = true;
  if (ctrl_host()->overflow != 0) {
    *dropped = true;
    pair_want_ = std::min<uint64_t>(uint64_t{ctrl_host()->pairs} + ctrl_host()->pairs / 4 + 4096, kLaneLimit);
  }
  return hipSuccess;
}

bool Batch::at_pair_limit() const { return pair_want_ >= kLaneLimit && pair_capacity_ >= kLaneLimit; }

uint64_t Batch::device_bytes() const {
  uint64_t b = 0;
  for (const DeviceBuffer* d : {&ctrl_, &cams_, &views_, &tile_view_, &view_visible_, &cull_keys_,
                                &keys_[0], &keys_[1], &order_[0], &order_[1], &projected_, &rects_, &scan_status_,
                                &sort_status_, &sort_counts_, &hist_, &pair_keys_[0], &pair_keys_[1], &pair_vals_[0],
                                &pair_vals_[1], &tile_ranges_, &image_, &ranges_, &starts_})
    b += d->bytes;
  return b;
}

uint64_t Batch::bytes_for(uint32_t views, uint64_t lanes, uint32_t tiles, uint64_t pixels) const {
  lanes = std::max(grown(lanes, lane_capacity_, kLaneLimit) / kCullTile * kCullTile, lane_capacity_);
  const uint64_t pairs = std::max(
      grown(std::max<uint64_t>({pair_want_, lanes * 3, uint64_t{1} << 22}), pair_capacity_, kLaneLimit),
      pair_capacity_);
  tiles = std::max(static_cast<uint32_t>(grown(tiles, tile_capacity_, 1u << 24)), tile_capacity_);
  pixels = std::max(grown(pixels, pixel_capacity_, ~0ull), pixel_capacity_);
  views = std::max(views, view_capacity_);
  const uint64_t status = (lanes / std::min(kCullTile, 256u * kBinItems) + 2) * 8 +
                          (std::max(lanes, pairs) / kSortTile + 2) * kRadix * 12;
  return lanes * (5 * 4 + 32 + 8) + pairs * 16 + uint64_t{tiles} * 14 + pixels * 16 +
         uint64_t{views} * (sizeof(Camera) + sizeof(ViewInfo) + 4) + status + (1u << 20);
}

void Batch::release() {
  release_graph();
  for (DeviceBuffer* d : {&ctrl_, &cams_, &views_, &tile_view_, &view_visible_, &cull_keys_,
                          &keys_[0], &keys_[1], &order_[0], &order_[1], &projected_, &rects_, &scan_status_, &sort_status_,
                          &sort_counts_, &hist_, &pair_keys_[0], &pair_keys_[1], &pair_vals_[0], &pair_vals_[1], &tile_ranges_,
                          &image_, &ranges_, &starts_})
    d->reset();
  lane_capacity_ = pair_capacity_ = pixel_capacity_ = 0;
  view_capacity_ = tile_capacity_ = range_capacity_ = 0;
  scan_words_ = sort_words_ = 0;
  tile_view_host_.clear();
}

hipError_t Batch::read_image(uint32_t v, float* rgba) const {
  const ViewInfo& vi = info_[v];
  return hipMemcpy(rgba, image_.as<float>() + vi.pixel_base * 4, size_t{vi.width} * vi.height * 16,
                   hipMemcpyDeviceToHost);
}

hipError_t Batch::survivor_tiles() const {
  if (st_ready_) return hipSuccess;
  const uint32_t blocks = ctrl_host()->cull_lanes / kCullTile;
  st_count_.assign(blocks, 0u);
  st_local_.assign(blocks, 0u);
  if (blocks) {
    if (hipError_t e = hipMemcpy(st_count_.data(), scan_status_.ptr, size_t{blocks} * 4, hipMemcpyDeviceToHost);
        e != hipSuccess)
      return e;
  }
  std::vector<uint32_t> next(info_.size(), 0);
  for (uint32_t b = 0; b < blocks; ++b) {
    st_local_[b] = next[tile_view_of(b)];
    next[tile_view_of(b)] += st_count_[b];
  }
  st_ready_ = true;
  return hipSuccess;
}

hipError_t Batch::read_visible(uint32_t v, uint32_t* keys, uint32_t* projected, uint32_t capacity) const {
  if (hipError_t e = survivor_tiles(); e != hipSuccess) return e;
  // View v's survivors: the segments of its cull tiles, in tile order, read in chunks of tiles.
  constexpr uint32_t kChunkTiles = (1u << 22) / kCullTile;
  const uint32_t blocks = static_cast<uint32_t>(st_count_.size());
  std::vector<uint32_t> kc, pc;
  uint32_t k = 0;
  for (uint32_t b0 = 0; b0 < blocks && k < capacity; b0 += kChunkTiles) {
    const uint32_t b1 = std::min(blocks, b0 + kChunkTiles);
    bool any = false;
    for (uint32_t b = b0; b < b1 && !any; ++b) any = tile_view_of(b) == v && st_count_[b] != 0u;
    if (!any) continue;
    const size_t lanes = size_t{b1 - b0} * kCullTile, first = size_t{b0} * kCullTile;
    hipError_t e = hipSuccess;
    if (keys) {
      kc.resize(lanes);
      e = hipMemcpy(kc.data(), cull_keys_.as<uint32_t>() + first, lanes * 4, hipMemcpyDeviceToHost);
    }
    if (e == hipSuccess && projected) {
      pc.resize(lanes * 8);
      e = hipMemcpy(pc.data(), projected_.as<uint32_t>() + first * 8, lanes * 32, hipMemcpyDeviceToHost);
    }
    if (e != hipSuccess) return e;
    for (uint32_t b = b0; b < b1; ++b) {
      if (tile_view_of(b) != v) continue;
      for (uint32_t j = 0; j < st_count_[b] && k < capacity; ++j, ++k) {
        const size_t i = size_t{b - b0} * kCullTile + j;
        if (keys) keys[k] = kc[i];
        if (projected) std::memcpy(projected + size_t{k} * 8, pc.data() + i * 8, 32);
      }
    }
  }
  return hipSuccess;
}

hipError_t Batch::read_order(uint32_t v, uint32_t* order, uint32_t capacity) const {
  if (hipError_t e = survivor_tiles(); e != hipSuccess) return e;
  const uint32_t total = ctrl_host()->visible, n = view_visible(v);
  std::vector<uint32_t> all(total);
  const uint32_t* src = order_[(survivor_passes_ - 1) & 1].as<uint32_t>();
  if (total) {
    if (hipError_t e = hipMemcpy(all.data(), src, size_t{total} * 4, hipMemcpyDeviceToHost); e != hipSuccess)
      return e;
  }
  uint32_t k = 0;
  for (uint32_t i = 0; i < total && k < std::min(n, capacity); ++i)
    if (tile_view_of(all[i] / kCullTile) == v) order[k++] = local_index(all[i]);
  return hipSuccess;
}

hipError_t Batch::read_tiles(uint32_t v, uint64_t* pairs, uint32_t* pair_tiles, uint32_t* pair_values,
                             uint64_t capacity, uint32_t* ranges, uint32_t tile_capacity) const {
  const ViewInfo& vi = info_[v];
  const uint32_t tiles = vi.tiles_x * vi.tiles_y;
  std::vector<uint32_t> rg(size_t{tiles} * 2);
  hipError_t e = hipMemcpy(rg.data(), tile_ranges_.as<uint32_t>() + size_t{vi.tile_base} * 2, rg.size() * 4,
                           hipMemcpyDeviceToHost);
  if (e != hipSuccess) return e;
  // The view's tiles in order, each tile's pairs a run of the sorted pairs (contiguous for the view
  // with global tile keys; interleaved with the other views' runs with view-local keys).
  uint32_t lo = ~0u, hi = 0;
  uint64_t total = 0;
  for (uint32_t t = 0; t < tiles; ++t) {
    if (rg[2 * t + 1] <= rg[2 * t]) continue;
    lo = std::min(lo, rg[2 * t]);
    hi = std::max(hi, rg[2 * t + 1]);
    total += rg[2 * t + 1] - rg[2 * t];
  }
  *pairs = total;
  if (ranges != nullptr) {
    uint32_t off = 0;
    for (uint32_t t = 0; t < std::min(tiles, tile_capacity); ++t) {
      const bool full = rg[2 * t + 1] > rg[2 * t];
      ranges[2 * t] = full ? off : 0;
      off += full ? rg[2 * t + 1] - rg[2 * t] : 0;
      ranges[2 * t + 1] = full ? off : 0;
    }
  }
  const uint64_t n = std::min<uint64_t>(total, capacity);
  if (n == 0 || (pair_tiles == nullptr && pair_values == nullptr)) return e;
  const int fin = tile_passes_ & 1;
  // Gather the view's runs from one copy of the span they lie in.
  auto gather = [&](const DeviceBuffer& src, uint32_t* out) -> hipError_t {
    std::vector<uint32_t> span(hi - lo);
    if (hipError_t ce = hipMemcpy(span.data(), src.as<uint32_t>() + lo, span.size() * 4, hipMemcpyDeviceToHost);
        ce != hipSuccess)
      return ce;
    uint64_t k = 0;
    for (uint32_t t = 0; t < tiles && k < n; ++t)
      for (uint32_t i = rg[2 * t]; i < rg[2 * t + 1] && k < n; ++i) out[k++] = span[i - lo];
    return hipSuccess;
  };
  if (pair_tiles != nullptr) {
    e = gather(pair_keys_[fin], pair_tiles);
    const uint32_t mask = (1u << local_tile_bits_) - 1u;
    for (uint64_t i = 0; e == hipSuccess && i < n; ++i)
      pair_tiles[i] = local_tile_bits_ != 0 ? pair_tiles[i] & mask : pair_tiles[i] - vi.tile_base;
  }
  if (e == hipSuccess && pair_values != nullptr) {
    if (e = survivor_tiles(); e != hipSuccess) return e;
    e = gather(pair_vals_[fin], pair_values);
    for (uint64_t i = 0; e == hipSuccess && i < n; ++i) pair_values[i] = local_index(pair_values[i]);
  }
  return e;
}

}  // namespace skm
