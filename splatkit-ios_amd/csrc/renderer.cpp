// The C-ABI renderer (include/skm.h): resident worlds in HBM, two frames in flight (each frame
// slot with its own HIP stream and scratch, so consecutive frames overlap on the GPU), the frame
// pipeline of kernels.hip, per-stage HIP-event timing and readback.
//
// Mirrors MetalSplatRenderer (packages/splatkit-ios/Sources/SplatKitCore/rendering/metal/
// MetalSplatRenderer.mm): transactional uploads (:166-208), range validation before any GPU work
// (MetalVisibilityPass.mm:135-168), a sticky GPU-failure latch (MetalFrameLoop.h:51-55) and
// completion-time counters (MetalFrameCompute.mm:210-237).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <memory>
#include <vector>

#include "skm.h"
#include "skm_batch.h"
#include "skm_host.h"
#include "skm_internal.h"
#include "skm_lod.h"
#include "skm_lod_internal.h"

using skm::align_up;
using skm::Camera;
using skm::Ctrl;
using skm::DeviceBuffer;
using skm::fill_camera;
using skm::PinnedBuffer;
using skm::sh_stride;

static_assert(sizeof(skm_view) == 152 && sizeof(skm_batch_tuning) == 24 && sizeof(skm_batch_stats) == 40 &&
                  sizeof(skm_config) == 24,
              "batch ABI layouts");

namespace {

// Frames in flight: the reference's two (kFramesInFlight, MetalFrameSlots.h:8) unless the config
// asks for one or three.
constexpr uint32_t kDefaultSlots = 2;
constexpr uint32_t kMaxSlots = 3;  // with the renderer's own stream, the process's 4 hardware queues
constexpr uint32_t kMaxExtent = 16384;     // per axis
constexpr uint32_t kUploadChunk = 1u << 20;  // splats per staging copy

// The resident world: every GpuSplat field as its own array, SH words word-major.
struct WorldBuffers {
  DeviceBuffer mem;
  uint32_t capacity = 0;
  uint32_t count = 0;
  int sh_degree = 0;
  bool slab = false;
  size_t off[9] = {};

  hipError_t allocate(uint32_t cap, int degree) {
    size_t o = 0;
    for (int k = 0; k < 8; ++k) {
      off[k] = o;
      o = align_up(o + size_t{cap} * 4u, 256);
    }
    off[8] = o;
    o += size_t{cap} * sh_stride(degree) * 4u + 256;
    capacity = cap;
    sh_degree = degree;
    return mem.allocate(o);
  }
  skm::World view() const {
    return {mem.as<float>(off[0]),    mem.as<float>(off[1]),    mem.as<float>(off[2]),
            mem.as<uint32_t>(off[3]), mem.as<uint32_t>(off[4]), mem.as<uint32_t>(off[5]),
            mem.as<uint32_t>(off[6]), mem.as<uint32_t>(off[7]), mem.as<uint32_t>(off[8]),
            capacity,                 sh_degree};
  }
  skm::WorldOut out() const {
    return {mem.as<float>(off[0]),    mem.as<float>(off[1]),    mem.as<float>(off[2]),
            mem.as<uint32_t>(off[3]), mem.as<uint32_t>(off[4]), mem.as<uint32_t>(off[5]),
            mem.as<uint32_t>(off[6]), mem.as<uint32_t>(off[7]), mem.as<uint32_t>(off[8]),
            capacity,                 sh_degree};
  }
};

// A frame's inputs travel in one copy: the camera at offset 0, the ranges at kInputRanges, their
// prefix starts right after the ranges.
constexpr size_t kInputRanges = 256;
constexpr size_t kInputBytes = kInputRanges + size_t{skm::kMaxRanges} * 8 + size_t{skm::kMaxRanges + 1} * 4 + 16;

// Device memory the frame pipeline writes: control block, survivor arrays, tile pairs and ranges.
// Each frame slot owns one, so frames on the slots' streams overlap; the standalone sort owns
// another.
struct Scratch {
  uint32_t capacity = 0;       // survivor lanes
  uint32_t pair_capacity = 0;  // tile pairs
  uint32_t tile_capacity = 0;  // tiles
  DeviceBuffer ctrl, hist, cull_keys, keys[2], order[2], projected, rects, scan_status, sort_status, sort_counts;
  DeviceBuffer pair_keys[2], pair_vals[2], tile_ranges;
  size_t sort_status_blocks = 0;
  uint64_t pipelines = 0;  // prologue launches since the status arrays were last cleared
};

// One cut's working state: the selection's control block, per-round states, frontiers, costs,
// packets, look-back status words and the selected indices. Each frame slot owns one, so the cuts of
// consecutive frames overlap on the GPU.
struct LodCut {
  DeviceBuffer state, ctrl, frontier[2], cost, packets, indices, status[3];
};

struct Slot {
  hipStream_t stream = nullptr;
  Scratch x;
  LodCut cut;  // a tree world's cut for this slot's frames
  DeviceBuffer inputs, image;
  PinnedBuffer inputs_host, ctrl_host;
  hipEvent_t marks[6] = {};
  hipEvent_t done = nullptr;
  bool in_flight = false;
  bool with_world = false;
  bool with_lod = false;
  uint32_t lod_limit = 0;
  uint32_t width = 0, height = 0, tile = 16;
  int depth_passes = 4, tile_passes = 2;
  skm::DigitPlan depth_plan{}, tile_plan{};
};

struct Policy {
  float lod_error_pixels = 1.0f;
  uint32_t lod_splat_limit = 0;
  float subpixel = 1.0f;
  bool key16 = true;  // RenderPolicy.h:36 default low16
  double background[3] = {0.05, 0.05, 0.08};
};

// The selection threshold of LodThreshold (LodThreshold.h:9-45), on the exported value type.
struct LodThreshold {
  skm_lod_threshold t{1.0f, 1.0f};
  void reset(float p) { skm_lod_threshold_reset(&t, p); }
  float update(uint32_t selected, uint32_t limited, uint32_t limit) {
    return skm_lod_threshold_update(&t, selected, limited, limit);
  }
  float pixels() const { return t.pixels; }
};

// Device buffers of one tree (MetalLodSelection::upload, MetalLodSelection.mm:60-108): clusters and
// leaf references, read by every cut.
struct LodBuffers {
  DeviceBuffer clusters, leaves;
  uint32_t cluster_count = 0, rounds = 0, capacity = 0, frontier_capacity = 0;
  std::vector<uint32_t> depth_clusters;  // clusters at each depth: the largest frontier of each round

  hipError_t upload(const skm_lod_tree_view* t, uint32_t cap, uint32_t n_rounds, hipStream_t stream) {
    cluster_count = static_cast<uint32_t>(t->cluster_count);
    rounds = n_rounds;
    // Round r evaluates children of round r - 1's clusters, i.e. clusters at depth r (breadth-first:
    // a child's index exceeds its parent's).
    std::vector<uint32_t> depth(t->cluster_count, 0);
    depth_clusters.assign(rounds + 1, 0);
    for (uint64_t k = 0; k < t->cluster_count; ++k) {
      const skm_lod_cluster& c = t->clusters[k];
      if (depth[k] < depth_clusters.size()) ++depth_clusters[depth[k]];
      for (uint32_t j = 0; j < c.child_count; ++j) depth[c.child_start + j] = depth[k] + 1;
    }
    capacity = cap;
    frontier_capacity = std::max(1u, std::min(cap, cluster_count));
    hipError_t e = clusters.allocate(std::max<size_t>(t->cluster_count, 1) * 64);
    if (e == hipSuccess) e = leaves.allocate(std::max<size_t>(t->leaf_ref_count, 1) * 4);
    if (e == hipSuccess && t->cluster_count)
      e = hipMemcpyAsync(clusters.ptr, t->clusters, t->cluster_count * 64, hipMemcpyHostToDevice, stream);
    if (e == hipSuccess && t->leaf_ref_count)
      e = hipMemcpyAsync(leaves.ptr, t->leaves, t->leaf_ref_count * 4, hipMemcpyHostToDevice, stream);
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    return e;
  }

  // A cut's state sized for this tree, cleared on `stream`.
  hipError_t allocate_cut(LodCut& c, hipStream_t stream) const {
    const size_t fc = frontier_capacity;
    const size_t blocks = (fc + 255) / 256 + 1;
    hipError_t e = c.state.allocate((size_t{rounds} + 1) * sizeof(skm::LodState));
    if (e == hipSuccess) e = c.ctrl.allocate(sizeof(skm::LodCtrl));
    if (e == hipSuccess) e = c.frontier[0].allocate(fc * 4);
    if (e == hipSuccess) e = c.frontier[1].allocate(fc * 4);
    if (e == hipSuccess) e = c.cost.allocate(fc * 8);
    if (e == hipSuccess) e = c.packets.allocate(fc * 16);
    if (e == hipSuccess) e = c.indices.allocate(std::max<size_t>(capacity, 1) * 4);
    for (auto& st : c.status)
      if (e == hipSuccess) e = st.allocate(blocks * 8);
    if (e == hipSuccess) e = hipMemsetAsync(c.ctrl.ptr, 0, sizeof(skm::LodCtrl), stream);
    for (auto& st : c.status)
      if (e == hipSuccess) e = hipMemsetAsync(st.ptr, 0, st.bytes, stream);
    return e;
  }

  // The leading rounds whose frontier cannot exceed head_frontier clusters (k_lod_head runs them).
  uint32_t head_rounds(uint32_t head_frontier) const {
    uint32_t r = 0;
    while (r < rounds && r < depth_clusters.size() && depth_clusters[r] <= head_frontier) ++r;
    return head_frontier == 0 ? 0u : r;
  }

  skm::LodArgs args(const Camera* cam, const LodCut& c, uint32_t limit, float pixel_limit, float color_weight,
                    int cull, uint32_t head_frontier) const {
    skm::LodArgs a{};
    a.cam = cam;
    a.clusters = clusters.as<skm::LodCluster>();
    a.leaves = leaves.as<uint32_t>();
    a.cluster_count = cluster_count;
    a.rounds = rounds;
    a.head_rounds = head_rounds(head_frontier);
    a.capacity = std::max(1u, std::min(limit == 0 ? capacity : limit, capacity));
    a.frontier_capacity = frontier_capacity;
    a.pixel_limit = pixel_limit;
    a.color_weight = color_weight;
    a.cull = cull;
    a.ctrl = c.ctrl.as<skm::LodCtrl>();
    a.state = c.state.as<skm::LodState>();
    a.frontier[0] = c.frontier[0].as<uint32_t>();
    a.frontier[1] = c.frontier[1].as<uint32_t>();
    a.cost = c.cost.as<uint32_t>();
    a.packets = c.packets.as<uint4>();
    a.indices = c.indices.as<uint32_t>();
    for (int k = 0; k < 3; ++k) a.status[k] = c.status[k].as<uint64_t>();
    return a;
  }
};

// The tree's indices must stay inside its arrays before any of it reaches the device: clusters
// breadth-first (children after their parent), leaf references and nodes in range.
const char* validate_tree(const skm_lod_tree_view* t, uint32_t* rounds) {
  if (t == nullptr || (t->layout == nullptr && t->nodes.count > 0)) return "null tree";
  if (t->nodes.count == 0 || t->cluster_count == 0 || t->clusters == nullptr) return "empty tree";
  if (t->leaf_ref_count > 0 && t->leaves == nullptr) return "missing leaf references";
  const uint64_t n = t->nodes.count;
  for (uint64_t k = 0; k < t->cluster_count; ++k) {
    const skm_lod_cluster& c = t->clusters[k];
    if (c.node >= n) return "cluster node out of range";
    if (c.child_count > 0 && (c.child_start <= k || uint64_t{c.child_start} + c.child_count > t->cluster_count))
      return "cluster children out of order or range";
    if (uint64_t{c.leaf_start} + c.leaf_count > t->leaf_ref_count) return "cluster leaves out of range";
  }
  for (uint64_t i = 0; i < t->leaf_ref_count; ++i)
    if (t->leaves[i] >= n) return "leaf reference out of range";
  *rounds = skm_lod_rounds(t);
  if (*rounds == 0 || *rounds > skm::kLodMaxRounds) return "tree deeper than the selection supports";
  return nullptr;
}

bool valid_cloud(const skm_cloud_view* c) {
  return c != nullptr && (c->count == 0 || (c->positions && c->covariances && c->colors && c->alphas)) &&
         c->sh_degree >= 0 && c->sh_degree <= 3;
}

}  // namespace

struct skm_renderer {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string error;
  std::string description;
  uint32_t tile = 16;
  uint32_t raster_px = 1;
  uint32_t tile_radix_bits = 8;  // digit width of the tile sort
  uint32_t pair_factor = 8;

  float render_scale = 1.0f;
  bool linear = false;
  bool count_tests = false;  // skm_count_raster_tests
  uint32_t lod_head_frontier = skm::kLodHeadFrontier;  // skm_set_lod_head
  bool vsync = true;
  uint32_t surface_w = 0, surface_h = 0;
  uint32_t generation = 0;
  Policy policy;

  WorldBuffers world;
  bool has_world = false;
  bool lod_world = false;
  LodBuffers lod;
  LodThreshold threshold;
  uint32_t last_selected = 0;
  uint32_t lod_stats[6] = {};

  // Sizes every slot's scratch reaches before its next frame: survivor lanes (whole cull tiles)
  // and tile pairs (grown after a frame overflows).
  uint32_t capacity = 0;
  uint64_t pair_target = 0;
  int sort_variant = skm::kSortReduceScan;
  DeviceBuffer staging;  // upload staging, on `stream`

  // Standalone sort scratch, on `stream`.
  Scratch sort;
  DeviceBuffer sort_keys, sort_vals, sort_n;
  hipEvent_t sort_marks[2] = {};
  uint32_t sort_capacity = 0;

  Slot slots[kMaxSlots];
  uint32_t slot_count = kDefaultSlots;
  struct Slots {
    Slot *b, *e;
    Slot* begin() const { return b; }
    Slot* end() const { return e; }
  };
  Slots active() { return {slots, slots + slot_count}; }
  uint64_t submitted = 0, completed = 0;
  int last_slot = -1;
  bool gpu_failed = false;
  bool completed_world_frame = false;
  bool grow_pairs = false;
  uint64_t pairs_needed = 0;
  uint64_t batch_pairs_needed = 0;  // the same for view batches (their pair buffers are the batch's)
  // A frame or batch whose tile pairs overflowed was dropped (its image is background only).
  // last_dropped: the newest completed frame was; drops: frames and batches dropped since the last
  // skm_finish, which reports them (however many frames completed after them).
  bool last_dropped = false;
  uint64_t drops = 0;
  double stage_ms[6] = {};
  uint32_t last_draw_count = 0;
  // The latest finished frame's cull lanes, survivors and tile pairs (valid once counts_known), from
  // which the next frames' grids are estimated.
  uint32_t last_lanes = 0, last_pairs = 0;
  uint64_t last_pixels = 0;  // that frame's extent
  bool counts_known = false;
  skm_frame_totals totals{};

  // View batches (skm_draw_views).
  std::unique_ptr<skm::Batch> batch;
  skm_batch_tuning batch_tuning{16, 8, 1, 0, 1, 0};
  bool batch_dropped = false;  // the newest completed batch was dropped
  bool batch_done = false;

  skm_status fail(skm_status s, const char* fmt, ...) {
    char buf[512];
    va_list args;
    va_start(args, fmt);
    std::vsnprintf(buf, sizeof(buf), fmt, args);
    va_end(args);
    error = buf;
    if (s == SKM_ERR_DEVICE) gpu_failed = true;
    return s;
  }
  skm_status check(hipError_t e, const char* what) {
    if (e == hipSuccess) return SKM_OK;
    return fail(e == hipErrorOutOfMemory ? SKM_ERR_MEMORY : SKM_ERR_DEVICE, "%s: %s", what,
                hipGetErrorString(e));
  }

  uint32_t draw_w() const {
    return surface_w == 0 ? 0 : std::max(1u, static_cast<uint32_t>(surface_w * render_scale + 0.5f));
  }
  uint32_t draw_h() const {
    return surface_h == 0 ? 0 : std::max(1u, static_cast<uint32_t>(surface_h * render_scale + 0.5f));
  }

  // Waits for a slot's frame and takes its counters and timings.
  skm_status harvest(Slot& s) {
    if (!s.in_flight) return SKM_OK;
    const hipError_t e = hipEventSynchronize(s.done);
    s.in_flight = false;
    if (e != hipSuccess) return check(e, "frame");
    const Ctrl* c = s.ctrl_host.as<Ctrl>();
    float ms[6] = {0, 0, 0, 0, 0, 0};
    for (int k = 0; k < 5; ++k) (void)hipEventElapsedTime(&ms[k], s.marks[k], s.marks[k + 1]);
    (void)hipEventElapsedTime(&ms[5], s.marks[0], s.marks[5]);
    for (int k = 0; k < 6; ++k) stage_ms[k] = ms[k];
    if (s.with_lod) {
      // k_prologue copied the cut's final state and denied refinements into ctrl->lod.
      skm::LodState cut;
      std::memcpy(&cut, c->lod, sizeof(cut));
      const skm::LodState* st = &cut;
      const uint32_t limited = c->lod[8];
      last_selected = st->count;
      lod_stats[0] = st->count;
      lod_stats[1] = st->active;
      lod_stats[2] = st->packets;
      lod_stats[3] = st->accepted;
      lod_stats[4] = limited;
      lod_stats[5] = st->evaluated;
      // adaptLodThreshold (MetalFrameCompute.mm:166-172).
      threshold.update(st->count, limited, s.lod_limit);
    }
    last_draw_count = c->visible;
    last_lanes = c->cull_lanes;
    last_pairs = c->pairs;
    last_pixels = uint64_t{s.width} * s.height;
    counts_known = true;
    last_dropped = c->overflow != 0;
    if (last_dropped) {
      // Counted apart: a dropped frame's work stopped after the binning, so its timings and pairs
      // are not a drawn frame's.
      totals.dropped += 1;
      ++drops;
      grow_pairs = true;
      pairs_needed = std::max<uint64_t>(pairs_needed, c->pairs);
    } else {
      totals.frames += 1;
      for (int k = 0; k < 6; ++k) totals.stage_ms[k] += ms[k];
      totals.visible += c->visible;
      totals.pairs += c->pairs;
      for (unsigned long long t : c->raster_tests) totals.raster_tests += t;
    }
    ++completed;
    if (s.with_world) completed_world_frame = true;
    return SKM_OK;
  }

  // The slot of the k-th oldest frame that may be in flight (frame f uses slot f % slot_count).
  Slot& slot_by_age(uint32_t k) { return slots[(submitted + k) % slot_count]; }

  // Takes the finished frames in submission order, stopping at the first still running (frames on
  // different slots may finish out of order; the last-frame statistics follow the newest in order).
  void poll() {
    for (uint32_t k = 0; k < slot_count; ++k) {
      Slot& s = slot_by_age(k);
      if (!s.in_flight) continue;
      if (hipEventQuery(s.done) != hipSuccess) break;
      (void)harvest(s);
    }
  }

  skm_status drain() {
    skm_status st = SKM_OK;
    for (uint32_t k = 0; k < slot_count; ++k) {
      const skm_status h = harvest(slot_by_age(k));
      if (h != SKM_OK) st = h;
    }
    if (batch) {
      bool dropped = false;
      const bool batch_waited = batch->in_flight();
      const skm_status k = check(batch->wait(&dropped), "batch");
      if (k != SKM_OK) st = k;
      if (k == SKM_OK && batch_waited) {
        batch_dropped = dropped;
        if (dropped) {
          totals.dropped += 1;
          ++drops;
          batch_pairs_needed = std::max<uint64_t>(batch_pairs_needed, batch->pairs());
        }
      }
    }
    return st;
  }

  skm_status ensure_status(Scratch& x, uint32_t items, hipStream_t st_) {
    const size_t blocks = (size_t{items} + skm::kSortTile - 1) / skm::kSortTile + 1;
    if (blocks <= x.sort_status_blocks) return SKM_OK;
    DeviceBuffer b, c;
    if (auto st = check(b.allocate(blocks * skm::kRadix * 8u), "sort status"); st != SKM_OK) return st;
    // The tile counts, then the binning's block totals and starts (kBinBlocksPerSortTile per sort tile
    // of survivors, and 2 spare).
    if (auto st = check(c.allocate((blocks * (skm::kRadix + skm::kBinBlocksPerSortTile) + 2u) * 4u), "sort counts"); st != SKM_OK) return st;
    if (auto st = check(hipMemsetAsync(b.ptr, 0, b.bytes, st_), "clear"); st != SKM_OK) return st;
    x.sort_status = std::move(b);
    x.sort_counts = std::move(c);
    x.sort_status_blocks = blocks;
    return SKM_OK;
  }

  // The control block and histograms every pipeline starts from.
  skm_status ensure_ctrl(Scratch& x, hipStream_t st_) {
    if (x.ctrl.ptr != nullptr) return SKM_OK;
    DeviceBuffer c, h;
    if (auto st = check(c.allocate(sizeof(Ctrl)), "control"); st != SKM_OK) return st;
    if (auto st = check(h.allocate(skm::kHistWords * 4), "histograms"); st != SKM_OK) return st;
    if (auto st = check(hipMemsetAsync(c.ptr, 0, sizeof(Ctrl), st_), "clear"); st != SKM_OK) return st;
    x.ctrl = std::move(c);
    x.hist = std::move(h);
    return SKM_OK;
  }

  // Survivor scratch for `cap` lanes in one slot, which has no frame in flight. Transactional:
  // nothing changes unless every buffer fits.
  skm_status ensure_capacity(Slot& s, uint32_t cap) {
    Scratch& x = s.x;
    if (cap <= x.capacity) return SKM_OK;
    DeviceBuffer ck, k0, k1, o0, o1, pr, rc, ss;
    const size_t n = size_t{cap};
    if (auto st = check(rc.allocate(n * 8), "rects"); st != SKM_OK) return st;
    if (auto st = check(ck.allocate(n * 4), "keys"); st != SKM_OK) return st;
    if (auto st = check(k0.allocate(n * 4), "keys"); st != SKM_OK) return st;
    if (auto st = check(k1.allocate(n * 4), "keys"); st != SKM_OK) return st;
    if (auto st = check(o0.allocate(n * 4), "order"); st != SKM_OK) return st;
    if (auto st = check(o1.allocate(n * 4), "order"); st != SKM_OK) return st;
    if (auto st = check(pr.allocate(n * 32), "projected"); st != SKM_OK) return st;
    if (auto st = check(ss.allocate(((n + 255) / 256 + 1) * 8), "scan status"); st != SKM_OK) return st;
    if (auto st = check(hipMemsetAsync(ss.ptr, 0, ss.bytes, s.stream), "clear"); st != SKM_OK) return st;
    if (auto st = ensure_status(x, cap, s.stream); st != SKM_OK) return st;
    x.cull_keys = std::move(ck);
    x.keys[0] = std::move(k0);
    x.keys[1] = std::move(k1);
    x.order[0] = std::move(o0);
    x.order[1] = std::move(o1);
    x.projected = std::move(pr);
    x.rects = std::move(rc);
    x.scan_status = std::move(ss);
    x.capacity = cap;
    return SKM_OK;
  }

  // Every slot's survivor scratch; the caller has drained the frames.
  skm_status ensure_capacity(uint32_t cap) {
    // Whole cull tiles: the cull writes survivors in a gapped layout of kCullTile lanes per tile.
    cap = static_cast<uint32_t>(align_up(std::max(cap, 1u), skm::kCullTile));
    for (Slot& s : active())
      if (auto st = ensure_capacity(s, cap); st != SKM_OK) return st;
    capacity = std::max(capacity, cap);
    return SKM_OK;
  }

  static uint64_t clamp_pairs(uint64_t want) {
    return std::min<uint64_t>(std::max<uint64_t>(want, uint64_t{1} << 22), 0xfffff000ull);
  }

  skm_status ensure_pairs(Slot& s, uint64_t want) {
    Scratch& x = s.x;
    want = clamp_pairs(want);
    if (want <= x.pair_capacity) return SKM_OK;
    DeviceBuffer k0, k1, v0, v1;
    const size_t n = static_cast<size_t>(want);
    if (auto st = check(k0.allocate(n * 4), "pair keys"); st != SKM_OK) return st;
    if (auto st = check(k1.allocate(n * 4), "pair keys"); st != SKM_OK) return st;
    if (auto st = check(v0.allocate(n * 4), "pair values"); st != SKM_OK) return st;
    if (auto st = check(v1.allocate(n * 4), "pair values"); st != SKM_OK) return st;
    if (auto st = ensure_status(x, static_cast<uint32_t>(want), s.stream); st != SKM_OK) return st;
    x.pair_keys[0] = std::move(k0);
    x.pair_keys[1] = std::move(k1);
    x.pair_vals[0] = std::move(v0);
    x.pair_vals[1] = std::move(v1);
    x.pair_capacity = static_cast<uint32_t>(want);
    return SKM_OK;
  }

  // Every slot's pair scratch; the caller has drained the frames.
  skm_status ensure_pairs(uint64_t want) {
    for (Slot& s : active())
      if (auto st = ensure_pairs(s, want); st != SKM_OK) return st;
    pair_target = std::max(pair_target, clamp_pairs(want));
    return SKM_OK;
  }

  skm_status ensure_tiles(Slot& s, uint32_t tiles) {
    Scratch& x = s.x;
    if (tiles <= x.tile_capacity) return SKM_OK;
    DeviceBuffer t;
    // Ranges (2 words per tile), then the raster's tile order (1 word per tile).
    if (auto st = check(t.allocate(size_t{tiles} * 12u), "tile ranges"); st != SKM_OK) return st;
    x.tile_ranges = std::move(t);
    x.tile_capacity = tiles;
    return SKM_OK;
  }

  // Clears look-back status words well before the 30-bit tags they carry could repeat. The
  // scratch has no pipeline in flight.
  skm_status maybe_clear_status(Scratch& x, hipStream_t st_) {
    if (++x.pipelines < (1ull << 24)) return SKM_OK;
    x.pipelines = 1;
    if (auto st = check(hipMemsetAsync(x.sort_status.ptr, 0, x.sort_status.bytes, st_), "clear"); st != SKM_OK)
      return st;
    if (x.scan_status.ptr == nullptr) return SKM_OK;
    return check(hipMemsetAsync(x.scan_status.ptr, 0, x.scan_status.bytes, st_), "clear");
  }

  // Copies a cloud's floats through the upload staging buffer (kept between uploads: freeing device
  // memory waits for the whole device, and tile uploads run while frames are in flight) and packs
  // them into records [offset, offset + count) on the renderer's stream. Blocking.
  skm_status upload_range(WorldBuffers& w, uint32_t offset, const skm_cloud_view* c, int src_degree) {
    const uint32_t n = c->count;
    const size_t shf = src_degree > 0 ? static_cast<size_t>((src_degree + 1) * (src_degree + 1) - 1) * 3u : 0;
    const uint32_t chunk = std::min(n, kUploadChunk);
    const size_t per = 12 + 24 + 12 + 4 + shf * 4;
    const size_t want = size_t{std::max(chunk, 1u)} * per + 1024;
    if (want > staging.bytes)
      if (auto st = check(staging.allocate(want), "upload staging"); st != SKM_OK) return st;
    float* d_pos = staging.as<float>(0);
    float* d_cov = d_pos + size_t{chunk} * 3;
    float* d_col = d_cov + size_t{chunk} * 6;
    float* d_alpha = d_col + size_t{chunk} * 3;
    float* d_sh = d_alpha + chunk;
    for (uint32_t base = 0; base < n; base += chunk) {
      const uint32_t m = std::min(chunk, n - base);
      hipError_t e = hipMemcpyAsync(d_pos, c->positions + size_t{base} * 3, size_t{m} * 12, hipMemcpyHostToDevice, stream);
      if (e == hipSuccess) e = hipMemcpyAsync(d_cov, c->covariances + size_t{base} * 6, size_t{m} * 24, hipMemcpyHostToDevice, stream);
      if (e == hipSuccess) e = hipMemcpyAsync(d_col, c->colors + size_t{base} * 3, size_t{m} * 12, hipMemcpyHostToDevice, stream);
      if (e == hipSuccess) e = hipMemcpyAsync(d_alpha, c->alphas + base, size_t{m} * 4, hipMemcpyHostToDevice, stream);
      if (e == hipSuccess && shf > 0)
        e = hipMemcpyAsync(d_sh, c->sh + size_t{base} * shf, size_t{m} * shf * 4, hipMemcpyHostToDevice, stream);
      if (e == hipSuccess)
        e = skm::launch_pack(w.out(), offset + base, m, d_pos, d_cov, d_col, d_alpha, shf > 0 ? d_sh : nullptr,
                             shf > 0 ? src_degree : 0, stream);
      if (e == hipSuccess) e = hipStreamSynchronize(stream);
      if (auto st = check(e, "upload"); st != SKM_OK) return st;
    }
    return SKM_OK;
  }
};

namespace {
}  // namespace

namespace skm {
Batch* renderer_batch(skm_renderer* r) {
  if (!r->batch) r->batch = std::make_unique<Batch>();
  return r->batch.get();
}
hipStream_t renderer_stream(skm_renderer* r) { return r->stream; }
uint64_t renderer_world_bytes(const skm_renderer* r) { return r->world.mem.bytes; }
int renderer_device(const skm_renderer* r) { return r->device; }
skm_status renderer_fail(skm_renderer* r, skm_status s, const char* message) { return r->fail(s, "%s", message); }
void static_scratch_setup(int device);  // noalloc.hip
}  // namespace skm

extern "C" {

skm_status skm_renderer_create(const skm_config* config, skm_renderer** out) {
  if (out == nullptr) return SKM_ERR_INVALID;
  *out = nullptr;
  int device = config ? config->device : 0;
  const uint32_t tile = config && config->tile_size ? config->tile_size : 16u;
  const uint32_t px = config && config->raster_px ? config->raster_px : (tile == 32u ? 2u : 1u);
  if (!skm::raster_shape_supported(tile, px)) return SKM_ERR_UNSUPPORTED;
  if (config && (config->sort_kernel > 1 || config->frames_in_flight > kMaxSlots)) return SKM_ERR_INVALID;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || device < 0 || device >= count) return SKM_ERR_DEVICE;
  auto* r = new skm_renderer();
  r->device = device;
  r->tile = tile;
  r->raster_px = px;
  if (config && config->pair_factor) r->pair_factor = config->pair_factor;
  if (config) r->sort_variant = static_cast<int>(config->sort_kernel);
  if (config && config->frames_in_flight) r->slot_count = config->frames_in_flight;
  hipDeviceProp_t prop;
  if (hipSetDevice(device) != hipSuccess || hipGetDeviceProperties(&prop, device) != hipSuccess ||
      hipStreamCreateWithFlags(&r->stream, hipStreamNonBlocking) != hipSuccess) {
    delete r;
    return SKM_ERR_DEVICE;
  }
  int runtime = 0;
  (void)hipRuntimeGetVersion(&runtime);
  char desc[320];
  std::snprintf(desc, sizeof(desc), "%s (%s, %d CUs), HIP %d.%d", prop.name[0] ? prop.name : "AMD GPU", prop.gcnArchName,
                prop.multiProcessorCount, runtime / 10000000, (runtime / 100000) % 100);
  r->description = desc;
  // The static sort's scratch on this device (its address, resolved once per device).
  skm::static_scratch_setup(device);
  bool ok = hipEventCreate(&r->sort_marks[0]) == hipSuccess && hipEventCreate(&r->sort_marks[1]) == hipSuccess;
  for (Slot& s : r->active()) {
    ok = ok && hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking) == hipSuccess &&
         r->ensure_ctrl(s.x, s.stream) == SKM_OK && r->ensure_tiles(s, 1) == SKM_OK &&
         s.inputs.allocate(kInputBytes) == hipSuccess && s.inputs_host.allocate(kInputBytes) == hipSuccess &&
         s.ctrl_host.allocate(sizeof(Ctrl)) == hipSuccess &&
         hipEventCreate(&s.done) == hipSuccess;
    for (auto& m : s.marks) ok = ok && hipEventCreate(&m) == hipSuccess;
  }
  ok = ok && r->ensure_capacity(1) == SKM_OK && r->ensure_pairs(0) == SKM_OK;
  for (Slot& s : r->active()) ok = ok && hipStreamSynchronize(s.stream) == hipSuccess;
  if (!ok) {
    skm_renderer_destroy(r);
    return SKM_ERR_DEVICE;
  }
  *out = r;
  return SKM_OK;
}

void skm_renderer_destroy(skm_renderer* r) {
  if (r == nullptr) return;
  (void)hipSetDevice(r->device);
  if (r->stream) (void)hipStreamSynchronize(r->stream);
  for (Slot& s : r->active())
    if (s.stream) (void)hipStreamSynchronize(s.stream);
  r->batch.reset();
  for (auto& m : r->sort_marks)
    if (m) (void)hipEventDestroy(m);
  for (Slot& s : r->active()) {
    if (s.stream) (void)hipStreamDestroy(s.stream);
    if (s.done) (void)hipEventDestroy(s.done);
    for (auto& m : s.marks)
      if (m) (void)hipEventDestroy(m);
  }
  if (r->stream) (void)hipStreamDestroy(r->stream);
  delete r;
}

const char* skm_last_error(const skm_renderer* r) { return r ? r->error.c_str() : "null renderer"; }

void skm_set_render_scale(skm_renderer* r, float scale) {
  if (r == nullptr || !std::isfinite(scale)) return;
  const float s = std::clamp(scale, 0.1f, 2.0f);
  if (s != r->render_scale) {
    r->render_scale = s;
    ++r->generation;
  }
}
float skm_render_scale(const skm_renderer* r) { return r ? r->render_scale : 1.0f; }
void skm_set_linear_blending(skm_renderer* r, int linear) {
  if (r) r->linear = linear != 0;
}
int skm_linear_blending(const skm_renderer* r) { return r && r->linear ? 1 : 0; }
void skm_count_raster_tests(skm_renderer* r, int enable) {
  if (r) r->count_tests = enable != 0;
}
void skm_set_lod_head(skm_renderer* r, uint32_t max_frontier) {
  if (r) r->lod_head_frontier = max_frontier;
}
void skm_set_vsync(skm_renderer* r, int vsync) {
  if (r) r->vsync = vsync != 0;
}

skm_status skm_set_draw_extent(skm_renderer* r, uint32_t width, uint32_t height) {
  if (r == nullptr) return SKM_ERR_INVALID;
  if (width == 0 || height == 0 || width > kMaxExtent || height > kMaxExtent)
    return r->fail(SKM_ERR_INVALID, "extent %ux%u outside 1..%u", width, height, kMaxExtent);
  if (width != r->surface_w || height != r->surface_h) {
    r->surface_w = width;
    r->surface_h = height;
    ++r->generation;
  }
  return SKM_OK;
}

// MetalSplatRenderer::deviceCapabilities (MetalSplatRenderer.mm:119-150).
skm_status skm_get_device_capabilities(const skm_renderer* r, skm_device_capabilities* out) {
  if (r == nullptr || out == nullptr) return SKM_ERR_INVALID;
  std::memset(out, 0, sizeof(*out));
  out->max_lod_capacity_splats = 1u << 27;
  out->min_residency_capacity_splats = 1;
  out->max_residency_capacity_splats = 0x7fffffffu;
  out->supports_hiz_occlusion = 0;
  out->supports_subgroups = 1;
  out->max_texture_dimension = kMaxExtent;
  skm_policy_support& p = out->policy;
  p.lod_error_pixels = 1;
  p.lod_splat_limit = 1;
  p.subpixel_threshold = 1;
  p.sort_depth = 1;
  p.background_color = 1;
  p.min_lod_error_pixels = 0.1f;   // LodThreshold::kMinPolicyPixels
  p.max_lod_error_pixels = 16.0f;  // LodThreshold::kMaxPolicyPixels
  p.min_subpixel_threshold = 0.0f;
  p.max_subpixel_threshold = 3.402823466e38f;
  p.fallback.lod_error_pixels = 1.0f;
  p.fallback.alpha_threshold = 2.0f / 255.0f;
  p.fallback.subpixel_threshold = 1.0f;
  p.fallback.enable_frustum_culling = 1;
  p.fallback.enable_early_termination = 1;
  p.fallback.sort_depth = r->policy.key16 ? 16u : 32u;
  p.fallback.background_color[0] = 0.05;
  p.fallback.background_color[1] = 0.05;
  p.fallback.background_color[2] = 0.08;
  return SKM_OK;
}

// validRenderPolicy (RenderPolicy.cpp:30-53) for the fields this backend applies; a rejected
// policy leaves the previous one in place (SplatRenderer.h:58-61).
skm_status skm_apply_render_policy(skm_renderer* r, const skm_render_policy* p, char* reason,
                                   size_t reason_size) {
  if (r == nullptr || p == nullptr) return SKM_ERR_INVALID;
  const char* why = nullptr;
  if (!std::isfinite(p->lod_error_pixels) || !(p->lod_error_pixels > 0.0f))
    why = "lodErrorPixels must be finite and positive";
  else if (!std::isfinite(p->alpha_threshold) || p->alpha_threshold < 0.0f || p->alpha_threshold > 1.0f)
    why = "alphaThreshold must be finite and in [0, 1]";
  else if (!std::isfinite(p->subpixel_threshold) || p->subpixel_threshold < 0.0f)
    why = "subpixelThreshold must be finite and non-negative";
  else if (p->sort_depth != 16 && p->sort_depth != 32)
    why = "sortDepth must be 16 or 32";
  else
    for (double c : p->background_color)
      if (!std::isfinite(c) || c < 0 || c > 1) why = "backgroundColor must be finite and in [0, 1]";
  if (why != nullptr) {
    if (reason && reason_size) std::snprintf(reason, reason_size, "%s", why);
    return r->fail(SKM_ERR_INVALID, "%s", why);
  }
  r->policy.lod_error_pixels = std::clamp(p->lod_error_pixels, 0.1f, 16.0f);
  r->policy.lod_splat_limit = p->lod_splat_limit;
  r->policy.subpixel = p->subpixel_threshold;
  r->policy.key16 = p->sort_depth == 16;
  for (int c = 0; c < 3; ++c) r->policy.background[c] = p->background_color[c];
  // A policy change restarts the LOD threshold at the policy's value (MetalFrameCompute.mm:71-103).
  r->threshold.reset(r->policy.lod_error_pixels);
  if (reason && reason_size) reason[0] = '\0';
  return SKM_OK;
}

int skm_available_memory(const skm_renderer* r, uint64_t* bytes) {
  if (r == nullptr || bytes == nullptr) return 0;
  size_t free_b = 0, total_b = 0;
  if (hipSetDevice(r->device) != hipSuccess || hipMemGetInfo(&free_b, &total_b) != hipSuccess) return 0;
  *bytes = free_b;
  return 1;
}

int skm_ready(const skm_renderer* r) { return r && r->surface_w > 0 && !r->gpu_failed ? 1 : 0; }
int skm_gpu_failed(skm_renderer* r) { return r && r->gpu_failed ? 1 : 0; }
int skm_has_completed_world_frame(skm_renderer* r) {
  if (r == nullptr) return 0;
  r->poll();
  return r->completed_world_frame ? 1 : 0;
}
void skm_draw_extent(const skm_renderer* r, uint32_t* width, uint32_t* height) {
  if (width) *width = r ? r->draw_w() : 0;
  if (height) *height = r ? r->draw_h() : 0;
}
uint32_t skm_generation(const skm_renderer* r) { return r ? r->generation : 0; }
int skm_surface_rebuild_pending(const skm_renderer*) { return 0; }

// MetalSplatRenderer::uploadWorld (:166-178) and MetalWorld::upload (MetalWorld.mm:24-57): the
// drawn degree is min(cloud, requested) when the cloud carries it, else 0.
skm_status skm_upload_world(skm_renderer* r, const skm_cloud_view* cloud, int max_sh_degree) {
  if (r == nullptr) return SKM_ERR_INVALID;
  if (!valid_cloud(cloud)) return r->fail(SKM_ERR_INVALID, "incomplete cloud");
  if (r->gpu_failed) return r->fail(SKM_ERR_DEVICE, "GPU failed earlier");
  if (auto st = r->check(hipSetDevice(r->device), "device"); st != SKM_OK) return st;
  if (auto st = r->drain(); st != SKM_OK) return st;
  int degree = std::clamp(std::min(cloud->sh_degree, max_sh_degree), 0, 3);
  if (degree > 0 && cloud->sh == nullptr) degree = 0;
  WorldBuffers w;
  if (auto st = r->check(w.allocate(std::max(cloud->count, 1u), degree), "world"); st != SKM_OK) return st;
  w.count = cloud->count;
  if (auto st = r->upload_range(w, 0, cloud, degree > 0 ? cloud->sh_degree : 0); st != SKM_OK) return st;
  if (auto st = r->ensure_capacity(cloud->count); st != SKM_OK) return st;
  if (auto st = r->ensure_pairs(uint64_t{cloud->count} * r->pair_factor); st != SKM_OK) return st;
  r->world = std::move(w);
  r->has_world = true;
  r->lod_world = false;
  r->lod = LodBuffers{};
  for (Slot& sl : r->active()) sl.cut = LodCut{};  // a flat world keeps no cut state
  r->completed_world_frame = false;
  r->counts_known = false;
  return r->check(hipStreamSynchronize(r->stream), "upload");
}

int skm_selects_lod_on_gpu(const skm_renderer*) { return 1; }

// MetalSplatRenderer::uploadLodWorld (MetalSplatRenderer.mm:180-194) with
// MetalLodSelection::upload (MetalLodSelection.mm:60-108).
skm_status skm_upload_lod_world(skm_renderer* r, const skm_lod_tree_view* tree, int max_sh_degree, uint32_t budget) {
  if (r == nullptr) return SKM_ERR_INVALID;
  uint32_t rounds = 0;
  if (const char* why = validate_tree(tree, &rounds)) return r->fail(SKM_ERR_INVALID, "invalid LOD tree: %s", why);
  if (!valid_cloud(&tree->nodes)) return r->fail(SKM_ERR_INVALID, "incomplete node cloud");
  if (r->gpu_failed) return r->fail(SKM_ERR_DEVICE, "GPU failed earlier");
  if (auto st = r->check(hipSetDevice(r->device), "device"); st != SKM_OK) return st;
  if (auto st = r->drain(); st != SKM_OK) return st;
  const uint64_t leaves = std::max<uint64_t>(tree->leaf_count, 1);
  uint32_t cap = static_cast<uint32_t>(std::min<uint64_t>({budget == 0 ? leaves : budget, leaves, 1ull << 27}));
  cap = std::max(cap, 1u);
  const skm_cloud_view* cloud = &tree->nodes;
  int degree = std::clamp(std::min(cloud->sh_degree, max_sh_degree), 0, 3);
  if (degree > 0 && cloud->sh == nullptr) degree = 0;
  WorldBuffers w;
  if (auto st = r->check(w.allocate(cloud->count, degree), "world"); st != SKM_OK) return st;
  w.count = cloud->count;
  if (auto st = r->upload_range(w, 0, cloud, degree > 0 ? cloud->sh_degree : 0); st != SKM_OK) return st;
  LodBuffers lod;
  if (auto st = r->check(lod.upload(tree, cap, rounds, r->stream), "LOD upload"); st != SKM_OK) return st;
  // One cut state per frame slot, each cleared on its slot's stream.
  LodCut cuts[kMaxSlots];
  for (uint32_t k = 0; k < r->slot_count; ++k)
    if (auto st = r->check(lod.allocate_cut(cuts[k], r->slots[k].stream), "LOD cut"); st != SKM_OK) return st;
  if (auto st = r->ensure_capacity(cap); st != SKM_OK) return st;
  if (auto st = r->ensure_pairs(uint64_t{cap} * r->pair_factor); st != SKM_OK) return st;
  r->world = std::move(w);
  r->lod = std::move(lod);
  for (uint32_t k = 0; k < r->slot_count; ++k) r->slots[k].cut = std::move(cuts[k]);
  r->has_world = true;
  r->lod_world = true;
  r->threshold.reset(r->policy.lod_error_pixels);
  r->completed_world_frame = false;
  r->counts_known = false;
  return r->check(hipStreamSynchronize(r->stream), "upload");
}

// MetalWorld::slab (MetalWorld.mm:59-79).
skm_status skm_create_slab(skm_renderer* r, uint32_t capacity, int sh_degree) {
  if (r == nullptr) return SKM_ERR_INVALID;
  if (capacity == 0) return r->fail(SKM_ERR_INVALID, "empty slab");
  if (auto st = r->drain(); st != SKM_OK) return st;
  WorldBuffers w;
  if (auto st = r->check(w.allocate(capacity, std::clamp(sh_degree, 0, 3)), "slab"); st != SKM_OK) return st;
  if (auto st = r->check(hipMemsetAsync(w.mem.ptr, 0, w.mem.bytes, r->stream), "slab"); st != SKM_OK) return st;
  w.count = capacity;
  w.slab = true;
  if (auto st = r->ensure_capacity(capacity); st != SKM_OK) return st;
  if (auto st = r->ensure_pairs(uint64_t{capacity} * r->pair_factor); st != SKM_OK) return st;
  r->world = std::move(w);
  r->has_world = true;
  r->lod_world = false;
  r->lod = LodBuffers{};
  for (Slot& sl : r->active()) sl.cut = LodCut{};  // a flat world keeps no cut state
  r->completed_world_frame = false;
  r->counts_known = false;
  return r->check(hipStreamSynchronize(r->stream), "slab");
}

// MetalWorld::uploadTile (MetalWorld.mm:81-100). Blocking, and it does not wait for the frames in
// flight: the caller rewrites a range only after every frame that drew it completed
// (SplatRenderer.h:94-96), so frames drawing the slab's other ranges keep running.
skm_status skm_upload_tile(skm_renderer* r, uint32_t offset, const skm_cloud_view* cloud) {
  if (r == nullptr) return SKM_ERR_INVALID;
  if (!r->has_world || !r->world.slab) return r->fail(SKM_ERR_STATE, "no slab");
  if (!valid_cloud(cloud)) return r->fail(SKM_ERR_INVALID, "incomplete cloud");
  if (offset > r->world.count || cloud->count > r->world.count - offset)
    return r->fail(SKM_ERR_INVALID, "tile [%u, +%u) outside the slab of %u", offset, cloud->count, r->world.count);
  if (cloud->count == 0) return SKM_OK;
  if (r->gpu_failed) return r->fail(SKM_ERR_DEVICE, "GPU failed earlier");
  if (auto st = r->check(hipSetDevice(r->device), "device"); st != SKM_OK) return st;
  const int deg = r->world.sh_degree;
  const bool carries = deg > 0 && cloud->sh_degree >= deg && cloud->sh != nullptr;
  skm_cloud_view c = *cloud;
  if (!carries) c.sh = nullptr;
  return r->upload_range(r->world, offset, &c, carries ? cloud->sh_degree : 0);
}

int skm_world(const skm_renderer* r, skm_world_info* out) {
  if (r == nullptr || !r->has_world) return 0;
  if (out) {
    out->count = r->world.count;
    out->sh_degree = r->world.sh_degree;
  }
  return 1;
}

int skm_sorts_on_gpu(const skm_renderer*) { return 1; }

skm_status skm_draw(skm_renderer* r, const skm_frame* f) {
  if (r == nullptr || f == nullptr) return SKM_ERR_INVALID;
  if (r->gpu_failed) return r->fail(SKM_ERR_DEVICE, "GPU failed earlier");
  if (r->surface_w == 0) return r->fail(SKM_ERR_STATE, "no draw extent");
  if (f->order_source != SKM_ORDER_GPU || f->order != nullptr)
    return r->fail(SKM_ERR_UNSUPPORTED, "this renderer orders on the GPU only");
  if (f->label_styles != nullptr) return r->fail(SKM_ERR_UNSUPPORTED, "part labels are not supported");
  if (f->range_count > skm::kMaxRanges || (f->range_count > 0 && f->ranges == nullptr))
    return r->fail(SKM_ERR_INVALID, "invalid ranges");
  // Validate before writing any slot input (MetalVisibilityPass.mm:155-162). A tree world draws its
  // cut and ignores the ranges (MetalFrameCompute.mm:213-215).
  const uint32_t source = r->has_world ? r->world.count : 0;
  uint64_t total = 0;
  for (uint32_t i = 0; i < (r->lod_world ? 0u : f->range_count); ++i) {
    const skm_range& g = f->ranges[i];
    if (g.offset > source || g.count > source - g.offset || total + g.count > r->capacity)
      return r->fail(SKM_ERR_INVALID, "range %u [%u, +%u) invalid for %u splats", i, g.offset, g.count, source);
    total += g.count;
  }
  if (auto st = r->check(hipSetDevice(r->device), "device"); st != SKM_OK) return st;
  const uint32_t W = r->draw_w(), H = r->draw_h();
  const uint32_t tiles_x = (W + r->tile - 1) / r->tile, tiles_y = (H + r->tile - 1) / r->tile;
  const uint32_t slot_id = static_cast<uint32_t>(r->submitted % r->slot_count);
  Slot& s = r->slots[slot_id];
  // The slot's previous frame ends here; its scratch is free from then on.
  if (auto st = r->harvest(s); st != SKM_OK) return st;
  if (r->grow_pairs) {
    r->pair_target = skm_renderer::clamp_pairs(std::max<uint64_t>(r->pairs_needed + r->pairs_needed / 4, r->pair_target * 2));
    r->grow_pairs = false;
  }
  if (auto st = r->ensure_pairs(s, r->pair_target); st != SKM_OK) return st;
  if (auto st = r->ensure_tiles(s, tiles_x * tiles_y); st != SKM_OK) return st;
  // The slot's image grows to the largest extent drawn (draws that alternate extents reallocate
  // nothing).
  if (size_t{W} * H * 16u > s.image.bytes) {
    DeviceBuffer img;
    if (auto st = r->check(img.allocate(size_t{W} * H * 16u), "image"); st != SKM_OK) return st;
    s.image = std::move(img);
  }
  s.width = W;
  s.height = H;
  Scratch& x = s.x;
  if (auto st = r->maybe_clear_status(x, s.stream); st != SKM_OK) return st;

  char* inputs = s.inputs_host.as<char>();
  Camera& cam = *reinterpret_cast<Camera*>(inputs);
  fill_camera(cam, f->view, f->proj, W, H, r->linear, r->policy.subpixel, f->camera_position, &f->reveal);

  const size_t starts_at = kInputRanges + align_up(size_t{f->range_count} * 8, 16);
  uint32_t* rh = reinterpret_cast<uint32_t*>(inputs + kInputRanges);
  uint32_t* sh = reinterpret_cast<uint32_t*>(inputs + starts_at);
  sh[0] = 0;
  for (uint32_t i = 0; i < f->range_count; ++i) {
    rh[2 * i] = f->ranges[i].offset;
    rh[2 * i + 1] = f->ranges[i].count;
    sh[i + 1] = sh[i] + f->ranges[i].count;
  }
  const size_t input_bytes = f->range_count > 0 ? starts_at + size_t{f->range_count + 1} * 4 : sizeof(Camera);
  hipError_t e = hipMemcpyAsync(s.inputs.ptr, inputs, input_bytes, hipMemcpyHostToDevice, s.stream);
  if (auto st = r->check(e, "frame inputs"); st != SKM_OK) return st;
  const Camera* cam_dev = s.inputs.as<Camera>();

  const bool world = r->has_world;
  const bool lod_frame = world && r->lod_world;
  skm::LodArgs lod{};
  if (lod_frame) {
    // The slot's own cut state: the cut overlaps the other slots' frames.
    s.lod_limit = r->policy.lod_splat_limit == 0 ? r->lod.capacity : std::min(r->policy.lod_splat_limit, r->lod.capacity);
    lod = r->lod.args(cam_dev, s.cut, s.lod_limit, r->threshold.pixels(), 4.0f, 1, r->lod_head_frontier);
  }
  skm::FrameArgs a{};
  a.cam = cam_dev;
  a.world = world ? r->world.view() : skm::World{};
  a.ranges = reinterpret_cast<const uint32_t*>(s.inputs.as<char>() + kInputRanges);
  a.range_starts = reinterpret_cast<const uint32_t*>(s.inputs.as<char>() + starts_at);
  a.range_count = world && !lod_frame ? f->range_count : 0;
  a.total = world && !lod_frame ? static_cast<uint32_t>(total) : 0;
  a.active = lod_frame ? lod.indices : nullptr;
  a.active_count = lod_frame ? &lod.state[lod.rounds].count : nullptr;
  a.key16 = r->policy.key16 ? 1 : 0;
  a.width = W;
  a.height = H;
  a.tile = r->tile;
  a.count_tests = r->count_tests ? 1u : 0u;
  a.split_cap = r->slot_count == 1 ? skm::kSplitTiles : 0u;
  a.tiles_x = tiles_x;
  a.tiles_y = tiles_y;
  for (int c = 0; c < 3; ++c) a.background[c] = static_cast<float>(r->policy.background[c]);
  a.ctrl = x.ctrl.as<Ctrl>();
  a.cull_keys = x.cull_keys.as<uint32_t>();
  for (int k = 0; k < 2; ++k) {
    a.keys[k] = x.keys[k].as<uint32_t>();
    a.order[k] = x.order[k].as<uint32_t>();
    a.pair_keys[k] = x.pair_keys[k].as<uint32_t>();
    a.pair_vals[k] = x.pair_vals[k].as<uint32_t>();
  }
  a.projected = x.projected.as<uint32_t>();
  a.rects = x.rects.as<uint64_t>();
  a.scan_status = x.scan_status.as<uint64_t>();
  a.cull_counts = reinterpret_cast<uint32_t*>(x.scan_status.ptr);
  a.sort_status = x.sort_status.as<uint64_t>();
  a.hist = x.hist.as<uint32_t>();
  a.pair_capacity = x.pair_capacity;
  a.tile_ranges = x.tile_ranges.as<uint32_t>();
  a.tile_order = x.tile_ranges.as<uint32_t>() + size_t{x.tile_capacity} * 2;
  a.image = s.image.as<float>();
  a.capacity = x.capacity;
  a.view_count = 1;
  a.total_tiles = tiles_x * tiles_y;
  a.sort_variant = r->sort_variant;
  if (r->counts_known) {
    // Half again the latest finished frame's counts, and a few blocks more: frames of one camera path
    // change gradually (a frame beyond the estimate is still drawn whole, its blocks loop).
    auto ahead = [](uint64_t v, uint32_t slack) {
      return static_cast<uint32_t>(std::min<uint64_t>(v + v / 2 + slack, 0xffffffffull));
    };
    a.lane_estimate = ahead(r->last_lanes, 8u * skm::kCullTile);
    a.survivor_estimate = ahead(r->last_draw_count, 8u * skm::kSortTile);
    // Pairs per pixel: a frame at another extent scales with its pixels.
    const uint64_t px = uint64_t{W} * H;
    a.pair_estimate = ahead(r->last_pixels != 0 ? r->last_pairs * px / r->last_pixels : r->last_pairs,
                            8u * skm::kSortTile);
  }
  a.sort_counts = x.sort_counts.as<uint32_t>();
  a.bin_starts = a.sort_counts + x.sort_status_blocks * skm::kRadix;
  a.raster_px = r->raster_px;
  s.tile = r->tile;
  s.depth_plan = skm::make_digit_plan(r->policy.key16 ? 16u : 32u, 8u);
  s.tile_plan = skm::make_digit_plan(skm::key_bits_for(tiles_x * tiles_y - 1), r->tile_radix_bits);
  s.depth_passes = s.depth_plan.passes;
  s.tile_passes = s.tile_plan.passes;
  e = skm::launch_frame(a, lod_frame ? &lod : nullptr, s.depth_plan, s.tile_plan, s.marks, s.stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(s.ctrl_host.ptr, x.ctrl.ptr, sizeof(Ctrl), hipMemcpyDeviceToHost, s.stream);
  if (e == hipSuccess) e = hipEventRecord(s.done, s.stream);
  if (auto st = r->check(e, "frame launch"); st != SKM_OK) return st;
  s.in_flight = true;
  s.with_world = world;
  s.with_lod = lod_frame;
  r->last_slot = static_cast<int>(slot_id);
  ++r->submitted;
  return SKM_OK;
}

skm_status skm_finish(skm_renderer* r) {
  if (r == nullptr) return SKM_ERR_INVALID;
  if (auto st = r->drain(); st != SKM_OK) return st;
  if (r->drops > 0) {
    const unsigned long long n = r->drops;
    r->drops = 0;
    return r->fail(SKM_ERR_CAPACITY,
                   "%llu frame(s) or batch(es) since the last finish were dropped: tile pairs overflowed (up to %llu "
                   "needed); capacity grows on the next draw",
                   n, static_cast<unsigned long long>(std::max(r->pairs_needed, r->batch_pairs_needed)));
  }
  return SKM_OK;
}

uint64_t skm_submitted_frames(const skm_renderer* r) { return r ? r->submitted : 0; }
uint64_t skm_completed_frames(skm_renderer* r) {
  if (r == nullptr) return 0;
  if (r->gpu_failed) return 0;
  r->poll();
  return r->completed;
}
double skm_last_gpu_millis(const skm_renderer* r) { return r ? r->stage_ms[5] : 0; }
double skm_last_sort_millis(const skm_renderer* r) { return r ? r->stage_ms[2] + r->stage_ms[3] : 0; }
double skm_last_cull_millis(const skm_renderer* r) { return r ? r->stage_ms[1] : 0; }
uint32_t skm_last_draw_count(const skm_renderer* r) { return r ? r->last_draw_count : 0; }
uint32_t skm_last_selected_count(const skm_renderer* r) { return r ? r->last_selected : 0; }
double skm_last_select_millis(const skm_renderer* r) { return r ? r->stage_ms[0] : 0; }
const char* skm_device_description(const skm_renderer* r) { return r ? r->description.c_str() : ""; }

skm_status skm_read_stage_millis(skm_renderer* r, double* ms6) {
  if (r == nullptr || ms6 == nullptr) return SKM_ERR_INVALID;
  if (auto st = r->drain(); st != SKM_OK) return st;
  for (int k = 0; k < 6; ++k) ms6[k] = r->stage_ms[k];
  return SKM_OK;
}

skm_status skm_read_lod_cut(skm_renderer* r, uint32_t* count, uint32_t* indices, uint32_t capacity, uint32_t* stats6) {
  if (r == nullptr || count == nullptr) return SKM_ERR_INVALID;
  if (auto st = r->drain(); st != SKM_OK) return st;
  if (!r->lod_world) return r->fail(SKM_ERR_STATE, "no LOD world");
  *count = r->lod_stats[0];
  if (stats6) std::memcpy(stats6, r->lod_stats, sizeof(r->lod_stats));
  const uint32_t n = std::min(r->lod_stats[0], capacity);
  if (r->last_slot < 0) return SKM_OK;
  const LodCut& cut = r->slots[r->last_slot].cut;
  if (indices && n) return r->check(hipMemcpy(indices, cut.indices.ptr, size_t{n} * 4, hipMemcpyDeviceToHost), "read");
  return SKM_OK;
}

// LodThreshold::reset and ::update (LodThreshold.h:22-36): x1.1 while refinements are denied, up to
// max(kCeilingPixels = 64, policy); /1.03 while the cut stays under limit / 10 * 9, down to the policy.
void skm_lod_threshold_reset(skm_lod_threshold* t, float policy_pixels) {
  if (t == nullptr) return;
  t->policy_pixels = policy_pixels;
  t->pixels = policy_pixels;
}

float skm_lod_threshold_update(skm_lod_threshold* t, uint32_t selected, uint32_t limited, uint32_t limit) {
  if (t == nullptr) return 0.0f;
  constexpr float kRaise = 1.1f, kRelax = 1.03f, kCeiling = 64.0f;
  if (limited > 0) {
    t->pixels = std::min(t->pixels * kRaise, std::max(kCeiling, t->policy_pixels));
  } else if (selected < limit / 10 * 9) {
    t->pixels = std::max(t->pixels / kRelax, t->policy_pixels);
  }
  return t->pixels;
}

float skm_lod_threshold_pixels(skm_renderer* r) {
  if (r == nullptr) return 0.0f;
  r->poll();
  return r->threshold.pixels();
}

skm_status skm_lod_select(skm_renderer* r, const skm_lod_tree_view* tree, const float* view16, const float* proj16,
                          uint32_t width, uint32_t height, uint32_t capacity, float pixel_limit, float color_weight,
                          int cull, uint32_t* indices, uint32_t* stats6) {
  if (r == nullptr || view16 == nullptr || proj16 == nullptr || width == 0 || height == 0) return SKM_ERR_INVALID;
  uint32_t rounds = 0;
  if (const char* why = validate_tree(tree, &rounds)) return r->fail(SKM_ERR_INVALID, "invalid LOD tree: %s", why);
  if (!std::isfinite(pixel_limit) || pixel_limit < 0 || !std::isfinite(color_weight) || color_weight < 0)
    return r->fail(SKM_ERR_INVALID, "invalid selection limits");
  if (auto st = r->check(hipSetDevice(r->device), "device"); st != SKM_OK) return st;
  if (auto st = r->drain(); st != SKM_OK) return st;
  const uint64_t leaves = std::max<uint64_t>(tree->leaf_count, 1);
  const uint32_t cap = std::max(1u, static_cast<uint32_t>(std::min<uint64_t>({capacity == 0 ? leaves : capacity, leaves, 1ull << 27})));
  LodBuffers lod;
  LodCut cut;
  if (auto st = r->check(lod.upload(tree, cap, rounds, r->stream), "LOD upload"); st != SKM_OK) return st;
  if (auto st = r->check(lod.allocate_cut(cut, r->stream), "LOD cut"); st != SKM_OK) return st;
  DeviceBuffer cam_dev;
  Camera cam;
  fill_camera(cam, view16, proj16, width, height, false, 0.0f, nullptr, nullptr);
  if (auto st = r->check(cam_dev.allocate(sizeof(Camera)), "camera"); st != SKM_OK) return st;
  const skm::LodArgs a = lod.args(cam_dev.as<Camera>(), cut, cap, pixel_limit, color_weight, cull, r->lod_head_frontier);
  skm::LodState final_state{};
  uint32_t limited = 0;
  hipError_t e = hipMemcpyAsync(cam_dev.ptr, &cam, sizeof(Camera), hipMemcpyHostToDevice, r->stream);
  if (e == hipSuccess) e = skm::launch_lod(a, r->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(&final_state, a.state + rounds, sizeof(final_state), hipMemcpyDeviceToHost, r->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(&limited, &a.ctrl->final_limited, 4, hipMemcpyDeviceToHost, r->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(r->stream);
  if (auto st = r->check(e, "LOD select"); st != SKM_OK) return st;
  if (stats6) {
    const uint32_t v[6] = {final_state.count, final_state.active, final_state.packets, final_state.accepted, limited,
                           final_state.evaluated};
    std::memcpy(stats6, v, sizeof(v));
  }
  const uint32_t n = std::min(final_state.count, cap);
  if (indices && n) return r->check(hipMemcpy(indices, cut.indices.ptr, size_t{n} * 4, hipMemcpyDeviceToHost), "read");
  return SKM_OK;
}

skm_status skm_read_frame_totals(skm_renderer* r, skm_frame_totals* out, int reset) {
  if (r == nullptr || out == nullptr) return SKM_ERR_INVALID;
  if (auto st = r->drain(); st != SKM_OK) return st;
  *out = r->totals;
  if (reset) r->totals = skm_frame_totals{};
  return SKM_OK;
}

skm_status skm_read_frame(skm_renderer* r, int format, void* dst, size_t dst_bytes) {
  if (r == nullptr || dst == nullptr) return SKM_ERR_INVALID;
  if (auto st = r->drain(); st != SKM_OK) return st;
  if (r->last_slot < 0) return r->fail(SKM_ERR_STATE, "no frame drawn");
  if (r->last_dropped) return r->fail(SKM_ERR_CAPACITY, "the last frame was dropped (tile pairs overflowed)");
  Slot& s = r->slots[r->last_slot];
  const size_t px = size_t{s.width} * s.height;
  if (format == SKM_PIXELS_RGBA32F) {
    if (dst_bytes < px * 16) return r->fail(SKM_ERR_INVALID, "destination too small");
    return r->check(hipMemcpy(dst, s.image.ptr, px * 16, hipMemcpyDeviceToHost), "read frame");
  }
  if (format != SKM_PIXELS_BGRA8) return r->fail(SKM_ERR_INVALID, "unknown pixel format");
  if (dst_bytes < px * 4) return r->fail(SKM_ERR_INVALID, "destination too small");
  std::vector<float> rgba(px * 4);
  if (auto st = r->check(hipMemcpy(rgba.data(), s.image.ptr, px * 16, hipMemcpyDeviceToHost), "read frame");
      st != SKM_OK)
    return st;
  auto* out = static_cast<uint8_t*>(dst);
  auto q = [](float v) { return static_cast<uint8_t>(std::lround(std::clamp(v, 0.0f, 1.0f) * 255.0f)); };
  for (size_t i = 0; i < px; ++i) {
    out[4 * i + 0] = q(rgba[4 * i + 2]);
    out[4 * i + 1] = q(rgba[4 * i + 1]);
    out[4 * i + 2] = q(rgba[4 * i + 0]);
    out[4 * i + 3] = q(rgba[4 * i + 3]);
  }
  return SKM_OK;
}

// The last frame's gapped survivor layout: the packed index of each cull tile's first survivor
// (bases[b], with bases[tiles] the survivor count) from the per-tile counts.
static skm_status survivor_bases(skm_renderer* r, const Scratch& x, const Ctrl& c, std::vector<uint32_t>* counts,
                                 std::vector<uint32_t>* bases) {
  const uint32_t tiles = c.cull_lanes / skm::kCullTile;
  counts->assign(tiles, 0u);
  bases->assign(size_t{tiles} + 1, 0u);
  if (tiles)
    if (auto st = r->check(hipMemcpy(counts->data(), x.scan_status.ptr, size_t{tiles} * 4, hipMemcpyDeviceToHost),
                           "read");
        st != SKM_OK)
      return st;
  for (uint32_t b = 0; b < tiles; ++b) (*bases)[b + 1] = (*bases)[b] + (*counts)[b];
  return SKM_OK;
}

// A survivor's packed index (compaction order) from its position in the gapped layout.
static uint32_t packed_index(const std::vector<uint32_t>& bases, uint32_t pos) {
  return bases[pos / skm::kCullTile] + pos % skm::kCullTile;
}

skm_status skm_read_visible(skm_renderer* r, uint32_t* count, uint32_t* keys, uint32_t* projected,
                            uint32_t capacity) {
  if (r == nullptr || count == nullptr) return SKM_ERR_INVALID;
  if (auto st = r->drain(); st != SKM_OK) return st;
  *count = 0;
  if (r->last_slot < 0) return SKM_OK;
  const Scratch& x = r->slots[r->last_slot].x;
  Ctrl c;
  if (auto st = r->check(hipMemcpy(&c, x.ctrl.ptr, sizeof(Ctrl), hipMemcpyDeviceToHost), "read"); st != SKM_OK)
    return st;
  *count = c.visible;
  const uint32_t n = std::min(c.visible, capacity);
  if (n == 0 || (keys == nullptr && projected == nullptr)) return SKM_OK;
  std::vector<uint32_t> counts, bases;
  if (auto st = survivor_bases(r, x, c, &counts, &bases); st != SKM_OK) return st;
  const size_t lanes = c.cull_lanes;
  std::vector<uint32_t> gk(keys ? lanes : 0), gp(projected ? lanes * 8 : 0);
  if (keys)
    if (auto st = r->check(hipMemcpy(gk.data(), x.cull_keys.ptr, lanes * 4, hipMemcpyDeviceToHost), "read");
        st != SKM_OK)
      return st;
  if (projected)
    if (auto st = r->check(hipMemcpy(gp.data(), x.projected.ptr, lanes * 32, hipMemcpyDeviceToHost), "read");
        st != SKM_OK)
      return st;
  for (uint32_t b = 0; b < counts.size(); ++b)
    for (uint32_t j = 0; j < counts[b] && bases[b] + j < n; ++j) {
      const size_t g = size_t{b} * skm::kCullTile + j, k = bases[b] + j;
      if (keys) keys[k] = gk[g];
      if (projected) std::memcpy(projected + k * 8, gp.data() + g * 8, 32);
    }
  return SKM_OK;
}

skm_status skm_read_order(skm_renderer* r, uint32_t* order, uint32_t capacity) {
  if (r == nullptr || order == nullptr) return SKM_ERR_INVALID;
  if (auto st = r->drain(); st != SKM_OK) return st;
  if (r->last_slot < 0) return r->fail(SKM_ERR_STATE, "no frame drawn");
  const Scratch& x = r->slots[r->last_slot].x;
  Ctrl c;
  if (auto st = r->check(hipMemcpy(&c, x.ctrl.ptr, sizeof(Ctrl), hipMemcpyDeviceToHost), "read"); st != SKM_OK)
    return st;
  const int passes = r->slots[r->last_slot].depth_passes;
  const uint32_t n = std::min(c.visible, capacity);
  if (n == 0) return SKM_OK;
  std::vector<uint32_t> counts, bases;
  if (auto st = survivor_bases(r, x, c, &counts, &bases); st != SKM_OK) return st;
  if (auto st = r->check(hipMemcpy(order, x.order[(passes - 1) & 1].ptr, size_t{n} * 4, hipMemcpyDeviceToHost),
                         "read");
      st != SKM_OK)
    return st;
  for (uint32_t i = 0; i < n; ++i) order[i] = packed_index(bases, order[i]);
  return SKM_OK;
}

skm_status skm_read_tiles(skm_renderer* r, uint64_t* pairs, uint32_t* pair_tiles, uint32_t* pair_values,
                          uint64_t capacity, uint32_t* ranges, uint32_t tile_capacity) {
  if (r == nullptr || pairs == nullptr) return SKM_ERR_INVALID;
  if (auto st = r->drain(); st != SKM_OK) return st;
  if (r->last_slot < 0) return r->fail(SKM_ERR_STATE, "no frame drawn");
  const Scratch& x = r->slots[r->last_slot].x;
  Ctrl c;
  if (auto st = r->check(hipMemcpy(&c, x.ctrl.ptr, sizeof(Ctrl), hipMemcpyDeviceToHost), "read"); st != SKM_OK)
    return st;
  *pairs = c.pairs;
  if (c.overflow != 0 && (pair_tiles != nullptr || pair_values != nullptr || ranges != nullptr))
    return r->fail(SKM_ERR_CAPACITY, "the last frame was dropped: %u tile pairs exceeded its capacity", c.pairs);
  const Slot& s = r->slots[r->last_slot];
  const int fin = s.tile_passes & 1;
  const uint64_t n = std::min<uint64_t>(c.sort_pairs, capacity);
  if (n > 0 && pair_tiles)
    if (auto st = r->check(hipMemcpy(pair_tiles, x.pair_keys[fin].ptr, n * 4, hipMemcpyDeviceToHost), "read");
        st != SKM_OK)
      return st;
  if (n > 0 && pair_values) {
    std::vector<uint32_t> counts, bases;
    if (auto st = survivor_bases(r, x, c, &counts, &bases); st != SKM_OK) return st;
    if (auto st = r->check(hipMemcpy(pair_values, x.pair_vals[fin].ptr, n * 4, hipMemcpyDeviceToHost), "read");
        st != SKM_OK)
      return st;
    for (uint64_t i = 0; i < n; ++i) pair_values[i] = packed_index(bases, pair_values[i]);
  }
  const uint32_t tiles = ((s.width + s.tile - 1) / s.tile) * ((s.height + s.tile - 1) / s.tile);
  if (ranges && tile_capacity)
    return r->check(hipMemcpy(ranges, x.tile_ranges.ptr, size_t{std::min(tiles, tile_capacity)} * 8,
                              hipMemcpyDeviceToHost),
                    "read");
  return SKM_OK;
}

skm_status skm_argsort_u32_device(skm_renderer* r, void* keys_dev, uint32_t n, int key_bits,
                                  void* order_dev, double* ms) {
  if (r == nullptr || (n > 0 && (keys_dev == nullptr || order_dev == nullptr))) return SKM_ERR_INVALID;
  if (key_bits < 1 || key_bits > 32) return r->fail(SKM_ERR_INVALID, "key_bits must be 1 to 32");
  if (auto st = r->check(hipSetDevice(r->device), "device"); st != SKM_OK) return st;
  // The sort has its own scratch on the renderer's stream, apart from the frames in flight; only its
  // own previous sort (and any batch) is ordered before it.
  if (auto st = r->check(hipStreamSynchronize(r->stream), "sort"); st != SKM_OK) return st;
  Scratch& x = r->sort;
  if (auto st = r->ensure_ctrl(x, r->stream); st != SKM_OK) return st;
  if (n > r->sort_capacity || r->sort_n.ptr == nullptr) {
    DeviceBuffer k, v, c;
    const size_t cap = std::max<size_t>(n, 1);
    if (auto st = r->check(k.allocate(cap * 4), "sort keys"); st != SKM_OK) return st;
    if (auto st = r->check(v.allocate(cap * 4), "sort values"); st != SKM_OK) return st;
    if (auto st = r->check(c.allocate(256), "sort count"); st != SKM_OK) return st;
    if (auto st = r->ensure_status(x, static_cast<uint32_t>(cap), r->stream); st != SKM_OK) return st;
    r->sort_keys = std::move(k);
    r->sort_vals = std::move(v);
    r->sort_n = std::move(c);
    r->sort_capacity = static_cast<uint32_t>(cap);
  }
  if (auto st = r->maybe_clear_status(x, r->stream); st != SKM_OK) return st;
  uint32_t* keys[2] = {static_cast<uint32_t*>(keys_dev), r->sort_keys.as<uint32_t>()};
  // The order ends in vals[(passes - 1) & 1]: that slot is the caller's buffer.
  const skm::DigitPlan plan = skm::make_digit_plan(static_cast<uint32_t>(key_bits), 8u);
  uint32_t* vals[2] = {r->sort_vals.as<uint32_t>(), r->sort_vals.as<uint32_t>()};
  vals[(plan.passes - 1) & 1] = static_cast<uint32_t*>(order_dev);
  hipEvent_t t0 = r->sort_marks[0], t1 = r->sort_marks[1];
  hipError_t e = hipMemcpyAsync(r->sort_n.ptr, &n, 4, hipMemcpyHostToDevice, r->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(r->stream);
  if (e == hipSuccess) e = hipEventRecord(t0, r->stream);
  if (e == hipSuccess)
    e = skm::launch_argsort(x.ctrl.as<Ctrl>(), keys, vals, r->sort_n.as<uint32_t>(), std::max(n, 1u),
                            plan, x.hist.as<uint32_t>(),
                            x.sort_status.as<uint64_t>(), x.sort_counts.as<uint32_t>(), r->sort_variant,
                            r->stream);
  if (e == hipSuccess) e = hipEventRecord(t1, r->stream);
  if (e == hipSuccess) e = hipEventSynchronize(t1);
  if (auto st = r->check(e, "sort"); st != SKM_OK) return st;
  float elapsed = 0;
  (void)hipEventElapsedTime(&elapsed, t0, t1);
  if (ms) *ms = elapsed;
  return SKM_OK;
}

uint32_t skm_morton_code(const float position[3], const float lo[3], const float hi[3]) {
  if (position == nullptr || lo == nullptr || hi == nullptr) return 0;
  return skm::morton_code(position, lo, hi);
}

// reorderSpatially (SpatialOrder.cpp:48-92): codes on the GPU, the stable argsort of their 30 bits
// (equal codes keep decode order, as the reference's three stable counting passes do), then each
// attribute array gathered in that order and copied back over the host array.
skm_status skm_reorder_spatially(skm_renderer* r, float* positions, float* covariances, float* colors,
                                 float* alphas, float* sh, uint32_t sh_floats, uint32_t count, const float lo[3],
                                 const float hi[3], uint32_t* order) {
  if (r == nullptr || lo == nullptr || hi == nullptr) return SKM_ERR_INVALID;
  if (count > 0 && (positions == nullptr || covariances == nullptr || colors == nullptr || alphas == nullptr ||
                    (sh_floats > 0 && sh == nullptr)))
    return r->fail(SKM_ERR_INVALID, "incomplete cloud");
  if (count < 2) {
    if (order && count == 1) order[0] = 0;
    return SKM_OK;
  }
  if (auto st = r->check(hipSetDevice(r->device), "device"); st != SKM_OK) return st;
  DeviceBuffer codes, idx, in, out;
  const size_t n = count;
  const size_t widest = std::max<size_t>(6, sh_floats);
  if (auto st = r->check(codes.allocate(n * 4), "codes"); st != SKM_OK) return st;
  if (auto st = r->check(idx.allocate(n * 4), "order"); st != SKM_OK) return st;
  if (auto st = r->check(in.allocate(n * widest * 4), "reorder"); st != SKM_OK) return st;
  if (auto st = r->check(out.allocate(n * widest * 4), "reorder"); st != SKM_OK) return st;
  hipError_t e = hipMemcpy(in.ptr, positions, n * 12, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = skm::launch_morton(in.as<float>(), count, lo, hi, codes.as<uint32_t>(), r->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(r->stream);
  if (auto st = r->check(e, "morton codes"); st != SKM_OK) return st;
  if (auto st = skm_argsort_u32_device(r, codes.ptr, count, 30, idx.ptr, nullptr); st != SKM_OK) return st;
  struct Rows {
    float* host;
    size_t stride;
  };
  const Rows rows[] = {{positions, 3}, {covariances, 6}, {colors, 3}, {alphas, 1}, {sh, sh_floats}};
  for (const Rows& a : rows) {
    if (a.stride == 0 || a.host == nullptr) continue;
    e = hipMemcpy(in.ptr, a.host, n * a.stride * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess)
      e = skm::launch_gather_rows(in.as<float>(), out.as<float>(), idx.as<uint32_t>(), count,
                                  static_cast<uint32_t>(a.stride), r->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(r->stream);
    if (e == hipSuccess) e = hipMemcpy(a.host, out.ptr, n * a.stride * 4, hipMemcpyDeviceToHost);
    if (auto st = r->check(e, "reorder"); st != SKM_OK) return st;
  }
  if (order) return r->check(hipMemcpy(order, idx.ptr, n * 4, hipMemcpyDeviceToHost), "order");
  return SKM_OK;
}

skm_status skm_argsort_u32(skm_renderer* r, const uint32_t* keys, uint32_t n, int key_bits, uint32_t* order) {
  if (r == nullptr || (n > 0 && (keys == nullptr || order == nullptr))) return SKM_ERR_INVALID;
  if (n == 0) return SKM_OK;
  if (auto st = r->check(hipSetDevice(r->device), "device"); st != SKM_OK) return st;
  DeviceBuffer k, o;
  if (auto st = r->check(k.allocate(size_t{n} * 4), "keys"); st != SKM_OK) return st;
  if (auto st = r->check(o.allocate(size_t{n} * 4), "order"); st != SKM_OK) return st;
  if (auto st = r->check(hipMemcpy(k.ptr, keys, size_t{n} * 4, hipMemcpyHostToDevice), "keys"); st != SKM_OK)
    return st;
  if (auto st = skm_argsort_u32_device(r, k.ptr, n, key_bits, o.ptr, nullptr); st != SKM_OK) return st;
  return r->check(hipMemcpy(order, o.ptr, size_t{n} * 4, hipMemcpyDeviceToHost), "order");
}


skm_status skm_draw_views(skm_renderer* r, const skm_view* views, uint32_t count, const skm_range* ranges,
                          uint32_t range_count, const skm_batch_tuning* tuning) {
  if (r == nullptr || views == nullptr || count == 0) return SKM_ERR_INVALID;
  if (r->gpu_failed) return r->fail(SKM_ERR_DEVICE, "GPU failed earlier");
  if (!r->has_world) return r->fail(SKM_ERR_STATE, "no world");
  if (r->lod_world) return r->fail(SKM_ERR_UNSUPPORTED, "view batches draw flat worlds; a tree world draws per view");
  if (range_count > skm::kMaxRanges || (range_count > 0 && ranges == nullptr))
    return r->fail(SKM_ERR_INVALID, "invalid ranges");
  const uint32_t source = r->world.count;
  uint64_t total = 0;
  for (uint32_t i = 0; i < range_count; ++i) {
    const skm_range& g = ranges[i];
    if (g.offset > source || g.count > source - g.offset || total + g.count > source)
      return r->fail(SKM_ERR_INVALID, "range %u [%u, +%u) invalid for %u splats", i, g.offset, g.count, source);
    total += g.count;
  }
  const skm_range whole{0, source};
  if (ranges == nullptr) {
    ranges = &whole;
    range_count = 1;
  }
  if (auto st = r->check(hipSetDevice(r->device), "device"); st != SKM_OK) return st;
  if (auto st = r->drain(); st != SKM_OK) return st;
  r->batch_done = false;
  if (!r->batch) r->batch = std::make_unique<skm::Batch>();
  skm::BatchPolicy p;
  p.key16 = r->policy.key16;
  p.linear = r->linear;
  p.subpixel = r->policy.subpixel;
  for (int c = 0; c < 3; ++c) p.background[c] = static_cast<float>(r->policy.background[c]);
  skm_batch_tuning t = tuning ? *tuning : r->batch_tuning;
  std::string why;
  const hipError_t e = r->batch->launch(r->world.view(), r->world.count, r->world.sh_degree, views, count, ranges,
                                        range_count, p, t, r->stream, &why);
  if (e == hipErrorInvalidValue) return r->fail(SKM_ERR_INVALID, "%s", why.c_str());
  if (e != hipSuccess) return r->check(e, why.c_str());
  r->batch_done = true;
  return SKM_OK;
}

skm_status skm_read_batch_stats(skm_renderer* r, skm_batch_stats* out) {
  if (r == nullptr || out == nullptr) return SKM_ERR_INVALID;
  if (auto st = r->drain(); st != SKM_OK) return st;
  if (!r->batch_done) return r->fail(SKM_ERR_STATE, "no batch drawn");
  out->views = r->batch->view_count();
  out->graph = r->batch->graph_used() ? 1u : 0u;
  out->gpu_ms = r->batch->gpu_ms();
  out->visible = r->batch->visible();
  out->pairs = r->batch->pairs();
  out->tile_passes = static_cast<uint32_t>(r->batch->tile_passes());
  out->view_sort_passes = static_cast<uint32_t>(r->batch->view_sort_passes());
  return SKM_OK;
}

skm_status skm_autotune_views(skm_renderer* r, const skm_view* views, uint32_t count, const skm_range* ranges,
                              uint32_t range_count, uint32_t repeats, skm_tuning_result* results,
                              uint32_t capacity, uint32_t* result_count, skm_batch_tuning* best) {
  if (r == nullptr || views == nullptr || count == 0 || (capacity > 0 && results == nullptr)) return SKM_ERR_INVALID;
  static const uint32_t kShapes[][2] = {{16, 1}, {16, 2}, {32, 2}, {8, 1}};
  repeats = std::max(1u, repeats);
  skm_batch_tuning chosen{};
  double chosen_ms = 0.0;
  uint32_t n = 0;
  for (const auto& shape : kShapes) {
    for (uint32_t sort_kernel = 0; sort_kernel < 2; ++sort_kernel) {
     for (uint32_t radix_bits = 8; radix_bits >= 6; --radix_bits) {
      const skm_batch_tuning t{shape[0], radix_bits, 1, sort_kernel, shape[1], 0};
      double ms = 0.0;
      for (uint32_t k = 0; k <= repeats; ++k) {
        skm_status st = skm_draw_views(r, views, count, ranges, range_count, &t);
        if (st == SKM_OK) st = skm_finish(r);
        if (st == SKM_ERR_CAPACITY) {  // the pairs outgrew the scratch: it has grown, run again
          st = skm_draw_views(r, views, count, ranges, range_count, &t);
          if (st == SKM_OK) st = skm_finish(r);
        }
        if (st != SKM_OK) return st;
        const double g = r->batch->gpu_ms();
        if (k == 1 || (k > 1 && g < ms)) ms = g;  // launch 0 is untimed
      }
      if (n < capacity) results[n] = skm_tuning_result{t, ms};
      ++n;
      if (chosen_ms == 0.0 || ms < chosen_ms) {
        chosen = t;
        chosen_ms = ms;
      }
     }
    }
  }
  r->batch_tuning = chosen;
  if (result_count) *result_count = n;
  if (best) *best = chosen;
  return SKM_OK;
}

skm_status skm_set_tuning(skm_renderer* r, const skm_batch_tuning* t) {
  if (r == nullptr || t == nullptr) return SKM_ERR_INVALID;
  const uint32_t tile = t->tile_size ? t->tile_size : 16u;
  const uint32_t px = t->raster_px ? t->raster_px : (tile == 32u ? 2u : 1u);
  const uint32_t rb = t->tile_radix_bits ? t->tile_radix_bits : 8u;
  if (!skm::raster_shape_supported(tile, px) || t->sort_kernel > 1 || rb < 6 || rb > 8)
    return r->fail(SKM_ERR_INVALID, "unsupported tuning");
  if (auto st = r->drain(); st != SKM_OK) return st;
  r->tile = tile;
  r->raster_px = px;
  r->tile_radix_bits = rb;
  r->sort_variant = static_cast<int>(t->sort_kernel);
  r->batch_tuning = *t;
  r->batch_tuning.tile_size = tile;
  r->batch_tuning.raster_px = px;
  r->batch_tuning.tile_radix_bits = rb;
  return SKM_OK;
}

static skm_status batch_view(skm_renderer* r, uint32_t view) {
  if (auto st = r->drain(); st != SKM_OK) return st;
  if (!r->batch_done) return r->fail(SKM_ERR_STATE, "no batch drawn");
  if (r->batch_dropped) return r->fail(SKM_ERR_CAPACITY, "the last batch was dropped");
  if (view >= r->batch->view_count()) return r->fail(SKM_ERR_INVALID, "view %u out of range", view);
  return SKM_OK;
}

skm_status skm_read_view_frame(skm_renderer* r, uint32_t view, int format, void* dst, size_t dst_bytes) {
  if (r == nullptr || dst == nullptr) return SKM_ERR_INVALID;
  if (auto st = batch_view(r, view); st != SKM_OK) return st;
  const skm::ViewInfo& vi = r->batch->info(view);
  const size_t px = size_t{vi.width} * vi.height;
  if (format == SKM_PIXELS_RGBA32F) {
    if (dst_bytes < px * 16) return r->fail(SKM_ERR_INVALID, "destination too small");
    return r->check(r->batch->read_image(view, static_cast<float*>(dst)), "read view");
  }
  if (format != SKM_PIXELS_BGRA8) return r->fail(SKM_ERR_INVALID, "unknown pixel format");
  if (dst_bytes < px * 4) return r->fail(SKM_ERR_INVALID, "destination too small");
  std::vector<float> rgba(px * 4);
  if (auto st = r->check(r->batch->read_image(view, rgba.data()), "read view"); st != SKM_OK) return st;
  skm::to_bgra8(rgba.data(), px, static_cast<uint8_t*>(dst));
  return SKM_OK;
}

skm_status skm_read_view_visible(skm_renderer* r, uint32_t view, uint32_t* count, uint32_t* keys,
                                 uint32_t* projected, uint32_t capacity) {
  if (r == nullptr || count == nullptr) return SKM_ERR_INVALID;
  if (auto st = batch_view(r, view); st != SKM_OK) return st;
  *count = r->batch->view_visible(view);
  return r->check(r->batch->read_visible(view, keys, projected, capacity), "read view");
}

skm_status skm_read_view_order(skm_renderer* r, uint32_t view, uint32_t* order, uint32_t capacity) {
  if (r == nullptr || (order == nullptr && capacity > 0)) return SKM_ERR_INVALID;
  if (auto st = batch_view(r, view); st != SKM_OK) return st;
  return r->check(r->batch->read_order(view, order, capacity), "read view");
}

skm_status skm_read_view_tiles(skm_renderer* r, uint32_t view, uint64_t* pairs, uint32_t* pair_tiles,
                               uint32_t* pair_values, uint64_t capacity, uint32_t* ranges,
                               uint32_t tile_capacity) {
  if (r == nullptr || pairs == nullptr) return SKM_ERR_INVALID;
  if (auto st = batch_view(r, view); st != SKM_OK) return st;
  return r->check(r->batch->read_tiles(view, pairs, pair_tiles, pair_values, capacity, ranges, tile_capacity),
                  "read view");
}

}  // extern "C"
