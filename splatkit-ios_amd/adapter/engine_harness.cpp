// The adapter (HipSplatRenderer.h) compiled against the reference engine's own headers, with
//   * compile-time checks that every POD the ABI mirrors has the engine type's size and field
//     offsets (the adapter reinterprets ranges, reveal uniforms, LOD nodes and clusters in place);
//   * the renderer-facing steps of SplatEngine restated over the adapter: the policy
//     (SplatEngine.cpp:29-56, with the reference's own resolveRenderPolicy, RenderPolicy.cpp, built
//     from its source beside this file), world uploads (uploadFinishedLoads, :226-267), tile uploads
//     (streamTiles, :429-449) and the frame (render and configureWorldFrame, :559-638). SplatEngine
//     itself links the SPZ decoder, a dependency absent here, so the engine proper is not built;
//   * a C entry layer (hsr_*) for the Python tests.
#include <chrono>
#include <cstddef>
#include <cstring>
#include <memory>
#include <optional>
#include <string>
#include <thread>
#include <vector>

#include "HipSplatRenderer.h"
#include "splatkit/rendering/RenderPolicy.h"

namespace {

using splatkit::DeviceCapabilities;
using splatkit::GpuWorldInfo;
using splatkit::HipSplatRenderer;
using splatkit::RenderPolicy;
using splatkit::RenderPolicySupport;
using splatkit::RevealUniform;
using splatkit::SplatRenderer;

// ---- the ABI's POD mirrors against the engine's types (SplatRenderer.h:25-28,100-103;
// RenderPolicy.h:22-77; Reveal.h:25-35; LodTree.h:13-35)
#define SAME_FIELD(A, fa, B, fb) static_assert(offsetof(A, fa) == offsetof(B, fb), #A "." #fa " vs " #B "." #fb)
static_assert(sizeof(skm_range) == sizeof(SplatRenderer::Range));
SAME_FIELD(skm_range, offset, SplatRenderer::Range, offset);
SAME_FIELD(skm_range, count, SplatRenderer::Range, count);
static_assert(sizeof(skm_world_info) == sizeof(GpuWorldInfo));
SAME_FIELD(skm_world_info, count, GpuWorldInfo, count);
SAME_FIELD(skm_world_info, sh_degree, GpuWorldInfo, shDegree);
static_assert(sizeof(skm_reveal) == sizeof(RevealUniform));
SAME_FIELD(skm_reveal, level, RevealUniform, level);
SAME_FIELD(skm_reveal, band, RevealUniform, band);
SAME_FIELD(skm_reveal, tint_amount, RevealUniform, tintAmount);
SAME_FIELD(skm_reveal, brightness_boost, RevealUniform, brightnessBoost);
SAME_FIELD(skm_reveal, tint, RevealUniform, tint);
static_assert(sizeof(skm_render_policy) == sizeof(RenderPolicy));
SAME_FIELD(skm_render_policy, lod_error_pixels, RenderPolicy, lodErrorPixels);
SAME_FIELD(skm_render_policy, lod_splat_limit, RenderPolicy, lodSplatLimit);
SAME_FIELD(skm_render_policy, alpha_threshold, RenderPolicy, alphaThreshold);
SAME_FIELD(skm_render_policy, subpixel_threshold, RenderPolicy, subpixelThreshold);
SAME_FIELD(skm_render_policy, enable_frustum_culling, RenderPolicy, enableFrustumCulling);
SAME_FIELD(skm_render_policy, enable_hiz_occlusion, RenderPolicy, enableHiZOcclusion);
SAME_FIELD(skm_render_policy, enable_early_termination, RenderPolicy, enableEarlyTermination);
SAME_FIELD(skm_render_policy, sort_depth, RenderPolicy, sortDepth);
SAME_FIELD(skm_render_policy, background_color, RenderPolicy, backgroundColor);
static_assert(sizeof(skm_policy_support) == sizeof(RenderPolicySupport));
SAME_FIELD(skm_policy_support, lod_error_pixels, RenderPolicySupport, lodErrorPixels);
SAME_FIELD(skm_policy_support, background_color, RenderPolicySupport, backgroundColor);
SAME_FIELD(skm_policy_support, min_lod_error_pixels, RenderPolicySupport, minLodErrorPixels);
SAME_FIELD(skm_policy_support, max_subpixel_threshold, RenderPolicySupport, maxSubpixelThreshold);
SAME_FIELD(skm_policy_support, fallback, RenderPolicySupport, fallback);
static_assert(sizeof(skm_device_capabilities) == sizeof(DeviceCapabilities));
SAME_FIELD(skm_device_capabilities, max_lod_capacity_splats, DeviceCapabilities, limits.maxLodCapacitySplats);
SAME_FIELD(skm_device_capabilities, max_residency_capacity_splats, DeviceCapabilities,
           limits.maxResidencyCapacitySplats);
SAME_FIELD(skm_device_capabilities, supports_hiz_occlusion, DeviceCapabilities, supportsHiZOcclusion);
SAME_FIELD(skm_device_capabilities, supports_subgroups, DeviceCapabilities, supportsSubgroups);
SAME_FIELD(skm_device_capabilities, max_texture_dimension, DeviceCapabilities, maxTextureDimension);
SAME_FIELD(skm_device_capabilities, policy, DeviceCapabilities, policy);
static_assert(sizeof(skm_lod_node) == sizeof(splat::LodNode));
SAME_FIELD(skm_lod_node, position, splat::LodNode, position);
SAME_FIELD(skm_lod_node, size, splat::LodNode, size);
SAME_FIELD(skm_lod_node, child_start, splat::LodNode, childStart);
SAME_FIELD(skm_lod_node, child_count, splat::LodNode, childCount);
static_assert(sizeof(skm_lod_cluster) == sizeof(splat::LodCluster));
SAME_FIELD(skm_lod_cluster, center, splat::LodCluster, center);
SAME_FIELD(skm_lod_cluster, radius, splat::LodCluster, radius);
SAME_FIELD(skm_lod_cluster, extent, splat::LodCluster, extent);
SAME_FIELD(skm_lod_cluster, error, splat::LodCluster, error);
SAME_FIELD(skm_lod_cluster, color_variance, splat::LodCluster, colorVariance);
SAME_FIELD(skm_lod_cluster, opacity, splat::LodCluster, opacity);
SAME_FIELD(skm_lod_cluster, node, splat::LodCluster, node);
SAME_FIELD(skm_lod_cluster, child_start, splat::LodCluster, childStart);
SAME_FIELD(skm_lod_cluster, child_count, splat::LodCluster, childCount);
SAME_FIELD(skm_lod_cluster, leaf_start, splat::LodCluster, leafStart);
SAME_FIELD(skm_lod_cluster, leaf_count, splat::LodCluster, leafCount);
SAME_FIELD(skm_lod_cluster, subtree_leaves, splat::LodCluster, subtreeLeaves);
// The frame's matrices and eye are copied, not reinterpreted; their sizes still have to agree.
static_assert(sizeof(splat::Mat4) == sizeof(float) * 16 && sizeof(splat::Vec3) == sizeof(float) * 3);
#undef SAME_FIELD
constexpr int kCheckedLayouts = 9;

splat::SplatCloud cloud_of(const skm_cloud_view& v) {
  splat::SplatCloud c;
  const size_t n = v.count;
  c.positions.assign(v.positions, v.positions + n * 3);
  c.covariances.assign(v.covariances, v.covariances + n * 6);
  c.colors.assign(v.colors, v.colors + n * 3);
  c.alphas.assign(v.alphas, v.alphas + n);
  c.shDegree = v.sh != nullptr ? v.sh_degree : 0;
  if (c.shDegree > 0) c.sh.assign(v.sh, v.sh + n * size_t((c.shDegree + 1) * (c.shDegree + 1) - 1) * 3);
  return c;
}

// The renderer-facing part of SplatEngine over one renderer (see the file comment).
struct EngineLoop {
  std::unique_ptr<SplatRenderer> renderer;
  RenderPolicy policy;
  std::string failure;   // the engine's worldFailed message of the last load
  std::vector<SplatRenderer::Range> ranges;  // what the last frame drew
  std::vector<SplatRenderer::Range> tile_ranges;  // a tiled world's resident tiles
  bool tiled = false;

  explicit EngineLoop(std::unique_ptr<SplatRenderer> r) : renderer(std::move(r)) {
    // The backend's fallback is the policy in effect until the host sets one (SplatEngine.cpp:29-34).
    policy = renderer->deviceCapabilities().policy.fallback;
  }

  // setRenderPolicy (SplatEngine.cpp:36-56).
  bool set_policy(const RenderPolicy& requested, std::string* error) {
    splatkit::RenderPolicyResolution res =
        splatkit::resolveRenderPolicy(requested, renderer->deviceCapabilities().policy);
    if (!res.accepted) {
      *error = res.error;
      return false;
    }
    std::string reason;
    if (!renderer->applyRenderPolicy(res.effective, &reason)) {
      *error = reason.empty() ? "renderer refused the policy" : reason;
      return false;
    }
    policy = res.effective;
    return true;
  }

  // uploadFinishedLoads (SplatEngine.cpp:226-267) for a cloud, a tree or a slab.
  bool load(const splat::SplatCloud* cloud, const splat::LodTree* tree, uint32_t budget, uint32_t residency,
            int slab_degree, int max_sh) {
    failure.clear();
    if (!renderer->sortsOnGpu()) {
      failure = "renderer lacks GPU sort support";
      return false;
    }
    if (residency > 0) {
      if (!renderer->createSlab(residency, std::min(slab_degree, max_sh))) {
        failure = "GPU upload failed";
        return false;
      }
      tiled = true;
      tile_ranges.clear();
      return true;
    }
    if (tree != nullptr && !renderer->selectsLodOnGpu()) {
      failure = "renderer lacks GPU level of detail selection";
      return false;
    }
    const bool uploaded = tree != nullptr ? renderer->uploadLodWorld(*tree, max_sh, budget)
                                          : renderer->uploadWorld(*cloud, max_sh);
    if (!uploaded) {
      failure = "GPU upload failed";
      return false;
    }
    tiled = false;
    return true;
  }

  // streamTiles (SplatEngine.cpp:429-449): an arrived tile goes into its slab range and is drawn
  // from the next frame on.
  bool arrive(uint32_t offset, const splat::SplatCloud& tile) {
    if (!renderer->uploadTile(offset, tile)) return false;
    tile_ranges.push_back({offset, static_cast<uint32_t>(tile.count())});
    return true;
  }

  // render + configureWorldFrame (SplatEngine.cpp:559-638) for a world frame.
  bool render(const float* view, const float* proj, const float* eye, int sh_degree) {
    const std::optional<GpuWorldInfo> world = renderer->world();
    SplatRenderer::Frame frame;
    if (world) {
      frame.orderSource = SplatRenderer::OrderSource::gpu;
      if (tiled)
        ranges = tile_ranges;
      else
        ranges.assign(1, {0, world->count});
      frame.ranges = ranges.data();
      frame.rangeCount = static_cast<uint32_t>(ranges.size());
      frame.drawCount = renderer->lastDrawCount();
      frame.shDegree = sh_degree;
      std::memcpy(frame.view.m.data(), view, 64);
      std::memcpy(frame.proj.m.data(), proj, 64);
      frame.cameraPosition = {eye[0], eye[1], eye[2]};
    }
    return renderer->draw(frame);
  }

  // pollGpuEvents waits no frame; a test does, for the last one, by the completion serials.
  bool wait(double timeout_s) {
    const auto t0 = std::chrono::steady_clock::now();
    while (renderer->completedFrames() < renderer->submittedFrames()) {
      if (renderer->gpuFailed()) return false;
      if (std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count() > timeout_s) return false;
      std::this_thread::sleep_for(std::chrono::microseconds(50));
    }
    return true;
  }
};

// A renderer that sorts or selects on the CPU, for the engine's refusal paths (FrameOrderTest.cpp:
// 112-145): everything else is HipSplatRenderer's.
class WithoutGpu final : public SplatRenderer {
 public:
  WithoutGpu(std::unique_ptr<HipSplatRenderer> inner, bool gpu_sort, bool gpu_lod)
      : inner_(std::move(inner)), gpu_sort_(gpu_sort), gpu_lod_(gpu_lod) {}
  void setRenderScale(float s) override { inner_->setRenderScale(s); }
  float renderScale() const override { return inner_->renderScale(); }
  void setLinearBlending(bool l) override { inner_->setLinearBlending(l); }
  bool linearBlending() const override { return inner_->linearBlending(); }
  void setVsync(bool v) override { inner_->setVsync(v); }
  DeviceCapabilities deviceCapabilities() const override { return inner_->deviceCapabilities(); }
  bool ready() const override { return inner_->ready(); }
  splatkit::Extent drawExtent() const override { return inner_->drawExtent(); }
  uint32_t generation() const override { return inner_->generation(); }
  bool uploadWorld(const splat::SplatCloud& c, int d) override { return inner_->uploadWorld(c, d); }
  bool selectsLodOnGpu() const override { return gpu_lod_ && inner_->selectsLodOnGpu(); }
  bool uploadLodWorld(const splat::LodTree& t, int d, uint32_t b) override { return inner_->uploadLodWorld(t, d, b); }
  bool createSlab(uint32_t c, int d) override { return inner_->createSlab(c, d); }
  bool uploadTile(uint32_t o, const splat::SplatCloud& c) override { return inner_->uploadTile(o, c); }
  std::optional<GpuWorldInfo> world() const override { return inner_->world(); }
  bool sortsOnGpu() const override { return gpu_sort_ && inner_->sortsOnGpu(); }
  bool draw(const Frame& f) override { return inner_->draw(f); }
  uint64_t submittedFrames() const override { return inner_->submittedFrames(); }
  uint64_t completedFrames() override { return inner_->completedFrames(); }
  double lastGpuMillis() const override { return inner_->lastGpuMillis(); }
  const std::string& deviceDescription() const override { return inner_->deviceDescription(); }

 private:
  std::unique_ptr<HipSplatRenderer> inner_;
  bool gpu_sort_, gpu_lod_;
};

struct Session {
  HipSplatRenderer* hip = nullptr;  // owned by loop.renderer (directly or inside WithoutGpu)
  std::unique_ptr<EngineLoop> loop;
  std::string error;
};

}  // namespace

extern "C" {

// The number of mirrored types whose layouts were checked (each check is a static_assert above).
int hsr_checked_layouts(void) { return kCheckedLayouts; }

// The adapter's translation of a world frame as configureWorldFrame fills it (SplatEngine.cpp:620-638).
int hsr_translate(const skm_range* ranges, uint32_t range_count, const float* view, const float* proj,
                  const float* eye, const skm_reveal* reveal, int sh_degree, skm_frame* out) {
  SplatRenderer::Frame f;
  f.orderSource = SplatRenderer::OrderSource::gpu;
  f.ranges = reinterpret_cast<const SplatRenderer::Range*>(ranges);
  f.rangeCount = range_count;
  f.shDegree = sh_degree;
  std::memcpy(f.view.m.data(), view, 64);
  std::memcpy(f.proj.m.data(), proj, 64);
  f.cameraPosition = {eye[0], eye[1], eye[2]};
  if (reveal != nullptr) std::memcpy(static_cast<void*>(&f.reveal), reveal, sizeof(RevealUniform));
  *out = HipSplatRenderer::toSkm(f);
  return 1;
}

// A session: the engine loop over HipSplatRenderer (gpu_sort and gpu_lod 0 wrap it in a renderer
// that reports no GPU sort / no GPU level of detail).
void* hsr_session_create(int device, uint32_t frames_in_flight, int gpu_sort, int gpu_lod) {
  auto s = std::make_unique<Session>();
  auto hip = std::make_unique<HipSplatRenderer>(device, frames_in_flight);
  s->hip = hip.get();
  std::unique_ptr<SplatRenderer> r;
  if (gpu_sort && gpu_lod)
    r = std::move(hip);
  else
    r = std::make_unique<WithoutGpu>(std::move(hip), gpu_sort != 0, gpu_lod != 0);
  s->loop = std::make_unique<EngineLoop>(std::move(r));
  return s.release();
}

void hsr_session_destroy(void* h) { delete static_cast<Session*>(h); }

const char* hsr_session_error(void* h) {
  auto* s = static_cast<Session*>(h);
  if (!s->loop->failure.empty()) return s->loop->failure.c_str();
  if (!s->error.empty()) return s->error.c_str();
  return s->hip->lastError();
}

// flags[0..7]: sortsOnGpu, selectsLodOnGpu, ready, gpuFailed, world present, fallback policy's sort
// depth, max texture dimension, description non-empty.
void hsr_session_probe(void* h, int32_t* flags) {
  auto* s = static_cast<Session*>(h);
  SplatRenderer& r = *s->loop->renderer;
  const DeviceCapabilities caps = r.deviceCapabilities();
  flags[0] = r.sortsOnGpu();
  flags[1] = r.selectsLodOnGpu();
  flags[2] = r.ready();
  flags[3] = r.gpuFailed();
  flags[4] = r.world().has_value();
  flags[5] = static_cast<int32_t>(s->loop->policy.sortDepth);
  flags[6] = static_cast<int32_t>(caps.maxTextureDimension);
  flags[7] = !r.deviceDescription().empty();
}

int hsr_session_policy(void* h, uint32_t sort_depth, float lod_error_pixels, uint32_t lod_splat_limit,
                       float subpixel) {
  auto* s = static_cast<Session*>(h);
  RenderPolicy p = s->loop->policy;
  p.sortDepth = sort_depth == 32 ? splatkit::SortKeyBits::full32 : splatkit::SortKeyBits::low16;
  p.lodErrorPixels = lod_error_pixels;
  p.lodSplatLimit = lod_splat_limit;
  p.subpixelThreshold = subpixel;
  return s->loop->set_policy(p, &s->error) ? 1 : 0;
}

int hsr_session_extent(void* h, uint32_t width, uint32_t height) {
  return static_cast<Session*>(h)->hip->setDrawExtent(width, height) ? 1 : 0;
}

int hsr_session_load_cloud(void* h, const skm_cloud_view* cloud, int max_sh) {
  auto* s = static_cast<Session*>(h);
  const splat::SplatCloud c = cloud_of(*cloud);
  return s->loop->load(&c, nullptr, 0, 0, 0, max_sh) ? 1 : 0;
}

// The tree as the engine's LodTree (nodes, layout, leaf count, selection clusters and leaves).
int hsr_session_load_tree(void* h, const skm_lod_tree_view* tree, int max_sh, uint32_t budget) {
  auto* s = static_cast<Session*>(h);
  splat::LodTree t;
  t.nodes = cloud_of(tree->nodes);
  t.layout.resize(tree->nodes.count);
  std::memcpy(static_cast<void*>(t.layout.data()), tree->layout, t.layout.size() * sizeof(splat::LodNode));
  t.leafCount = tree->leaf_count;
  t.selection.clusters.resize(tree->cluster_count);
  std::memcpy(static_cast<void*>(t.selection.clusters.data()), tree->clusters,
              tree->cluster_count * sizeof(splat::LodCluster));
  t.selection.leaves.assign(tree->leaves, tree->leaves + tree->leaf_ref_count);
  return s->loop->load(nullptr, &t, budget, 0, 0, max_sh) ? 1 : 0;
}

int hsr_session_load_slab(void* h, uint32_t residency, int sh_degree, int max_sh) {
  return static_cast<Session*>(h)->loop->load(nullptr, nullptr, 0, residency, sh_degree, max_sh) ? 1 : 0;
}

int hsr_session_arrive(void* h, uint32_t offset, const skm_cloud_view* tile) {
  const splat::SplatCloud c = cloud_of(*tile);
  return static_cast<Session*>(h)->loop->arrive(offset, c) ? 1 : 0;
}

// One engine frame; ranges_out (capacity entries) receives the ranges it handed the renderer.
int hsr_session_render(void* h, const float* view, const float* proj, const float* eye, int sh_degree,
                       skm_range* ranges_out, uint32_t capacity, uint32_t* range_count) {
  auto* s = static_cast<Session*>(h);
  const bool ok = s->loop->render(view, proj, eye, sh_degree);
  const auto& rg = s->loop->ranges;
  *range_count = static_cast<uint32_t>(rg.size());
  for (uint32_t i = 0; i < std::min<uint32_t>(capacity, *range_count); ++i) ranges_out[i] = {rg[i].offset, rg[i].count};
  return ok ? 1 : 0;
}

int hsr_session_wait(void* h, double timeout_s) { return static_cast<Session*>(h)->loop->wait(timeout_s) ? 1 : 0; }

// Serials and the last frame's counters: submitted, completed, lastDrawCount, lastSelectedCount.
void hsr_session_counters(void* h, uint64_t* out4) {
  SplatRenderer& r = *static_cast<Session*>(h)->loop->renderer;
  out4[0] = r.submittedFrames();
  out4[1] = r.completedFrames();
  out4[2] = r.lastDrawCount();
  out4[3] = r.lastSelectedCount();
}

int hsr_session_capture(void* h, uint8_t* bgra, size_t bytes) {
  auto* s = static_cast<Session*>(h);
  std::vector<uint8_t> px;
  if (!s->hip->capture(&px) || px.size() > bytes) return 0;
  std::memcpy(bgra, px.data(), px.size());
  return 1;
}

// The renderer underneath, for read-backs through the C ABI (skm_read_lod_cut and the like).
void* hsr_session_renderer(void* h) { return static_cast<Session*>(h)->hip->handle(); }

}  // extern "C"
