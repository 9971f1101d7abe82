// HipSplatRenderer: the reference engine's SplatRenderer
// (packages/splatkit-engine/include/splatkit/rendering/SplatRenderer.h:36-160) over the skm C ABI
// (include/skm.h), so that SplatEngine(std::unique_ptr<SplatRenderer>) (SplatEngine.cpp:29-34) draws
// with the MI355X path unchanged. Every virtual forwards to its skm_ function; the engine's types
// cross as the ABI's POD mirrors, whose layouts engine_harness.cpp checks with static_assert against
// the engine's own headers.
//
// Compiled on the engine side (it includes the engine's headers) and linked with libskm.so.
#pragma once

#include <cstring>
#include <optional>
#include <string>
#include <vector>

#include "skm.h"
#include "skm_lod.h"
#include "splatkit/rendering/SplatRenderer.h"

namespace splatkit {

class HipSplatRenderer final : public SplatRenderer {
 public:
  // 16-pixel tiles, 8 tile pairs per splat to start with, the reduce-then-scan radix passes, one
  // pixel per lane and `frames_in_flight` frame slots (the reference keeps two, MetalFrameSlots.h:8).
  explicit HipSplatRenderer(int device = 0, uint32_t frames_in_flight = 2) {
    const skm_config c{device, 16, 8, 0, 1, frames_in_flight};
    if (skm_renderer_create(&c, &r_) != SKM_OK) r_ = nullptr;  // ready() and draw() then report false
    if (r_ != nullptr) description_ = skm_device_description(r_);
  }
  ~HipSplatRenderer() override { skm_renderer_destroy(r_); }
  HipSplatRenderer(const HipSplatRenderer&) = delete;
  HipSplatRenderer& operator=(const HipSplatRenderer&) = delete;

  // ---- presentation controls (SplatRenderer.h:42-52)
  void setRenderScale(float scale) override { skm_set_render_scale(r_, scale); }
  float renderScale() const override { return skm_render_scale(r_); }
  void setLinearBlending(bool linear) override { skm_set_linear_blending(r_, linear ? 1 : 0); }
  bool linearBlending() const override { return skm_linear_blending(r_) != 0; }
  void setVsync(bool vsync) override { skm_set_vsync(r_, vsync ? 1 : 0); }

  // ---- capabilities and policy (SplatRenderer.h:56-70)
  DeviceCapabilities deviceCapabilities() const override {
    DeviceCapabilities caps;
    skm_device_capabilities c{};
    if (skm_get_device_capabilities(r_, &c) != SKM_OK) return caps;
    caps.limits.maxLodCapacitySplats = c.max_lod_capacity_splats;
    caps.limits.minResidencyCapacitySplats = c.min_residency_capacity_splats;
    caps.limits.maxResidencyCapacitySplats = c.max_residency_capacity_splats;
    caps.supportsHiZOcclusion = c.supports_hiz_occlusion != 0;
    caps.supportsSubgroups = c.supports_subgroups != 0;
    caps.maxTextureDimension = c.max_texture_dimension;
    const skm_policy_support& s = c.policy;
    RenderPolicySupport& p = caps.policy;
    p.lodErrorPixels = s.lod_error_pixels != 0;
    p.lodSplatLimit = s.lod_splat_limit != 0;
    p.alphaThreshold = s.alpha_threshold != 0;
    p.subpixelThreshold = s.subpixel_threshold != 0;
    p.enableFrustumCulling = s.enable_frustum_culling != 0;
    p.enableHiZOcclusion = s.enable_hiz_occlusion != 0;
    p.enableEarlyTermination = s.enable_early_termination != 0;
    p.sortDepth = s.sort_depth != 0;
    p.backgroundColor = s.background_color != 0;
    p.minLodErrorPixels = s.min_lod_error_pixels;
    p.maxLodErrorPixels = s.max_lod_error_pixels;
    p.minSubpixelThreshold = s.min_subpixel_threshold;
    p.maxSubpixelThreshold = s.max_subpixel_threshold;
    p.fallback = fromSkm(s.fallback);
    return caps;
  }
  bool applyRenderPolicy(const RenderPolicy& policy, std::string* reason) override {
    const skm_render_policy q = toSkm(policy);
    char why[256] = {};
    const bool ok = skm_apply_render_policy(r_, &q, why, sizeof why) == SKM_OK;
    if (reason != nullptr) *reason = ok ? std::string() : std::string(r_ ? why : "no HIP device");
    return ok;
  }
  std::optional<std::uint64_t> availableMemory() const override {
    uint64_t bytes = 0;
    return skm_available_memory(r_, &bytes) ? std::optional<std::uint64_t>(bytes) : std::nullopt;
  }

  // ---- state (SplatRenderer.h:73-83)
  bool ready() const override { return r_ != nullptr && skm_ready(r_) != 0; }
  bool gpuFailed() override { return r_ == nullptr || skm_gpu_failed(r_) != 0; }
  bool hasCompletedWorldFrame() override { return skm_has_completed_world_frame(r_) != 0; }
  Extent drawExtent() const override {
    Extent e;
    skm_draw_extent(r_, &e.width, &e.height);
    return e;
  }
  uint32_t generation() const override { return skm_generation(r_); }
  bool surfaceRebuildPending() const override { return skm_surface_rebuild_pending(r_) != 0; }

  // ---- worlds (SplatRenderer.h:87-107)
  bool uploadWorld(const splat::SplatCloud& cloud, int maxShDegree) override {
    const skm_cloud_view v = view(cloud);
    return skm_upload_world(r_, &v, maxShDegree) == SKM_OK;
  }
  bool selectsLodOnGpu() const override { return skm_selects_lod_on_gpu(r_) != 0; }
  bool uploadLodWorld(const splat::LodTree& tree, int maxShDegree, uint32_t budget) override {
    skm_lod_tree_view v{};
    v.nodes = view(tree.nodes);
    v.layout = reinterpret_cast<const skm_lod_node*>(tree.layout.data());
    v.leaf_count = tree.leafCount;
    v.clusters = reinterpret_cast<const skm_lod_cluster*>(tree.selection.clusters.data());
    v.cluster_count = tree.selection.clusters.size();
    v.leaves = tree.selection.leaves.data();
    v.leaf_ref_count = tree.selection.leaves.size();
    return skm_upload_lod_world(r_, &v, maxShDegree, budget) == SKM_OK;
  }
  bool createSlab(uint32_t capacity, int shDegree) override {
    return skm_create_slab(r_, capacity, shDegree) == SKM_OK;
  }
  bool uploadTile(uint32_t offset, const splat::SplatCloud& cloud) override {
    const skm_cloud_view v = view(cloud);
    return skm_upload_tile(r_, offset, &v) == SKM_OK;
  }
  std::optional<GpuWorldInfo> world() const override {
    skm_world_info w{};
    if (!skm_world(r_, &w)) return std::nullopt;
    return GpuWorldInfo{w.count, w.sh_degree};
  }
  bool sortsOnGpu() const override { return skm_sorts_on_gpu(r_) != 0; }

  // ---- frames (SplatRenderer.h:111-159)
  bool draw(const Frame& frame) override {
    const skm_frame f = toSkm(frame);
    return skm_draw(r_, &f) == SKM_OK;
  }
  uint64_t submittedFrames() const override { return skm_submitted_frames(r_); }
  uint64_t completedFrames() override { return skm_completed_frames(r_); }
  double lastGpuMillis() const override { return skm_last_gpu_millis(r_); }
  double lastSortMillis() const override { return skm_last_sort_millis(r_); }
  double lastCullMillis() const override { return skm_last_cull_millis(r_); }
  uint32_t lastDrawCount() const override { return skm_last_draw_count(r_); }
  uint32_t lastSelectedCount() const override { return skm_last_selected_count(r_); }
  double lastSelectMillis() const override { return skm_last_select_millis(r_); }
  const std::string& deviceDescription() const override { return description_; }

  // ---- what the platform's view code calls instead of attaching a surface
  bool setDrawExtent(uint32_t width, uint32_t height) { return skm_set_draw_extent(r_, width, height) == SKM_OK; }
  // captureNextFrame (MetalRenderer.h:29): the last completed frame as BGRA8, top-down rows.
  bool capture(std::vector<uint8_t>* bgra) {
    uint32_t w = 0, h = 0;
    skm_draw_extent(r_, &w, &h);
    bgra->resize(size_t{w} * h * 4);
    return skm_read_frame(r_, SKM_PIXELS_BGRA8, bgra->data(), bgra->size()) == SKM_OK;
  }
  const char* lastError() const { return skm_last_error(r_); }
  skm_renderer* handle() const { return r_; }

  // The engine's types as the ABI's: field for field, nothing converted but bools to bytes.
  static skm_frame toSkm(const Frame& frame) {
    skm_frame g{};
    g.order_source = frame.orderSource == OrderSource::gpu ? SKM_ORDER_GPU : SKM_ORDER_CPU;
    g.order = frame.order;
    g.order_count = frame.orderCount;
    g.draw_count = frame.drawCount;
    g.ranges = reinterpret_cast<const skm_range*>(frame.ranges);
    g.range_count = frame.rangeCount;
    g.sh_degree = frame.shDegree;
    std::memcpy(g.view, frame.view.m.data(), sizeof g.view);  // column-major (Mat4.h:10-14)
    std::memcpy(g.proj, frame.proj.m.data(), sizeof g.proj);
    g.camera_position[0] = frame.cameraPosition.x;
    g.camera_position[1] = frame.cameraPosition.y;
    g.camera_position[2] = frame.cameraPosition.z;
    g.label_styles = frame.labelStyles;  // part labels are refused by skm_draw
    std::memcpy(&g.reveal, &frame.reveal, sizeof g.reveal);
    return g;
  }
  static skm_render_policy toSkm(const RenderPolicy& p) {
    skm_render_policy q{};
    q.lod_error_pixels = p.lodErrorPixels;
    q.lod_splat_limit = p.lodSplatLimit;
    q.alpha_threshold = p.alphaThreshold;
    q.subpixel_threshold = p.subpixelThreshold;
    q.enable_frustum_culling = p.enableFrustumCulling ? 1 : 0;
    q.enable_hiz_occlusion = p.enableHiZOcclusion ? 1 : 0;
    q.enable_early_termination = p.enableEarlyTermination ? 1 : 0;
    q.sort_depth = static_cast<uint32_t>(p.sortDepth);
    for (int c = 0; c < 3; ++c) q.background_color[c] = p.backgroundColor[c];
    return q;
  }
  static RenderPolicy fromSkm(const skm_render_policy& q) {
    RenderPolicy p;
    p.lodErrorPixels = q.lod_error_pixels;
    p.lodSplatLimit = q.lod_splat_limit;
    p.alphaThreshold = q.alpha_threshold;
    p.subpixelThreshold = q.subpixel_threshold;
    p.enableFrustumCulling = q.enable_frustum_culling != 0;
    p.enableHiZOcclusion = q.enable_hiz_occlusion != 0;
    p.enableEarlyTermination = q.enable_early_termination != 0;
    p.sortDepth = q.sort_depth == 32 ? SortKeyBits::full32 : SortKeyBits::low16;
    for (int c = 0; c < 3; ++c) p.backgroundColor[c] = q.background_color[c];
    return p;
  }
  static skm_cloud_view view(const splat::SplatCloud& c) {
    return {static_cast<uint32_t>(c.count()), c.shDegree, c.positions.data(), c.covariances.data(),
            c.colors.data(), c.alphas.data(), c.sh.empty() ? nullptr : c.sh.data()};
  }

 private:
  skm_renderer* r_ = nullptr;
  std::string description_;
};

}  // namespace splatkit
