// C entry points over the reference's own CPU sources, compiled where they lie under
// /root/reference into oracle/_ref/libref.so by oracle/Makefile. Test infrastructure only: the
// tests use it to pin the oracle's restatement (oracle.cpp) against the real reference code.
// Nothing here re-implements the reference; each function only adapts types and calls it.
#include <algorithm>
#include <cstdint>
#include <cstring>
#include <vector>

#include <cstdio>
#include <string>

#include "splat/lod/LodFile.h"
#include "splat/lod/LodTree.h"
#include "splat/math/Covariance.h"
#include "splat/math/Half.h"
#include "splat/math/Mat4.h"
#include "splat/sorting/SpatialOrder.h"
#include "splatkit/rendering/GpuLayout.h"
#include "splatkit/rendering/LodThreshold.h"

namespace {
splat::SplatCloud make_cloud(uint32_t n, int sh_degree, const float* positions,
                             const float* covariances, const float* colors, const float* alphas,
                             const float* sh) {
  splat::SplatCloud cloud;
  cloud.positions.assign(positions, positions + size_t{n} * 3);
  cloud.covariances.assign(covariances, covariances + size_t{n} * 6);
  cloud.colors.assign(colors, colors + size_t{n} * 3);
  cloud.alphas.assign(alphas, alphas + n);
  cloud.shDegree = sh_degree;
  if (sh_degree > 0 && sh != nullptr) {
    const size_t k = static_cast<size_t>((sh_degree + 1) * (sh_degree + 1) - 1) * 3;
    cloud.sh.assign(sh, sh + size_t{n} * k);
  }
  return cloud;
}
}  // namespace

extern "C" {

uint32_t ref_morton_code(const float* p3, const float* lo3, const float* hi3) {
  splat::Bounds b;
  for (int k = 0; k < 3; ++k) {
    b.min[k] = lo3[k];
    b.max[k] = hi3[k];
  }
  return splat::mortonCode(p3, b);
}

// reorderSpatially on a cloud whose alphas are the source indices: the permutation it applied.
void ref_spatial_order(const float* positions, uint32_t n, const float* lo3, const float* hi3, uint32_t* order) {
  splat::SplatCloud c;
  c.positions.assign(positions, positions + size_t{n} * 3);
  c.covariances.assign(size_t{n} * 6, 0.0f);
  c.colors.assign(size_t{n} * 3, 0.0f);
  for (uint32_t i = 0; i < n; ++i) c.alphas.push_back(static_cast<float>(i));
  for (int k = 0; k < 3; ++k) {
    c.bounds.min[k] = lo3[k];
    c.bounds.max[k] = hi3[k];
  }
  splat::reorderSpatially(c);
  for (uint32_t i = 0; i < n; ++i) order[i] = static_cast<uint32_t>(c.alphas[i]);
}

uint16_t ref_to_half(float v) { return splat::toHalf(v); }
float ref_from_half(uint16_t h) { return splat::fromHalf(h); }

// packSplatRange (GpuLayout.cpp:122-137): 8 words per splat.
void ref_pack_splats(uint32_t n, const float* positions, const float* covariances,
                     const float* colors, const float* alphas, uint32_t* out) {
  const auto cloud = make_cloud(n, 0, positions, covariances, colors, alphas, nullptr);
  std::vector<splatkit::GpuSplat> packed(n);
  splatkit::packSplatRange(cloud, 0, n, packed.data());
  std::memcpy(out, packed.data(), size_t{n} * sizeof(splatkit::GpuSplat));
}

// packShRange, float16 storage (GpuLayout.cpp:92-113).
uint32_t ref_sh_stride(int degree) { return static_cast<uint32_t>(splatkit::shStride(degree)); }
void ref_pack_sh(uint32_t n, int src_degree, const float* positions, const float* sh, int degree,
                 uint32_t* out) {
  std::vector<float> zeros(size_t{n} * 6, 0.0f), ones(n, 1.0f);
  const auto cloud = make_cloud(n, src_degree, positions, zeros.data(), zeros.data(), ones.data(), sh);
  splatkit::packShRange(cloud, degree, 0, n, out);
}

// covarianceFromQuaternionScale (Covariance.cpp:7-17).
void ref_covariance(const float* quaternion, const float* scale, float* out6) {
  const auto c = splat::covarianceFromQuaternionScale(quaternion, scale);
  std::memcpy(out6, c.data(), 6 * sizeof(float));
}

// Mat4::perspective (Mat4.h:75-86), column-major.
void ref_perspective(float fovy, float aspect, float near_plane, float far_plane, float* out16) {
  const auto m = splat::Mat4::perspective(fovy, aspect, near_plane, far_plane);
  std::memcpy(out16, m.m.data(), 64);
}

// LodThreshold::update (LodThreshold.h:27-36) over a sequence of readbacks.
void ref_lod_threshold(float policy, uint32_t steps, const uint32_t* selected,
                       const uint32_t* limited, const uint32_t* limit, float* out) {
  splatkit::LodThreshold t(policy);
  for (uint32_t i = 0; i < steps; ++i) out[i] = t.update(selected[i], limited[i], limit[i]);
}

// buildLodTree (LodTree.cpp:323-342) + buildLodSelectionData (LodSelection.cpp:47-133).
// Returns an opaque tree; read it back with the accessors below, then free it.
void* ref_lod_build(uint32_t n, int sh_degree, const float* positions, const float* covariances,
                    const float* colors, const float* alphas, const float* sh, uint32_t depth,
                    uint32_t cluster_leaves) {
  auto cloud = make_cloud(n, sh_degree, positions, covariances, colors, alphas, sh);
  // The decoder fills bounds with the positions' box (SpzDecoder); buildLodTree reads them.
  for (int c = 0; c < 3; ++c) {
    cloud.bounds.min[c] = n ? positions[c] : 0.0f;
    cloud.bounds.max[c] = n ? positions[c] : 0.0f;
  }
  for (uint32_t i = 0; i < n; ++i)
    for (int c = 0; c < 3; ++c) {
      cloud.bounds.min[c] = std::min(cloud.bounds.min[c], positions[size_t{i} * 3 + c]);
      cloud.bounds.max[c] = std::max(cloud.bounds.max[c], positions[size_t{i} * 3 + c]);
    }
  splat::LodBuildOptions options;
  options.octreeDepth = depth;
  options.clusterLeaves = cluster_leaves;
  auto* tree = new splat::LodTree(splat::buildLodTree(std::move(cloud), options));
  tree->selection = splat::buildLodSelectionData(*tree);
  return tree;
}
void ref_lod_free(void* t) { delete static_cast<splat::LodTree*>(t); }
// counts[0..3] = nodes, leaves, clusters, leaf refs.
void ref_lod_counts(const void* t, uint64_t* counts) {
  const auto* tree = static_cast<const splat::LodTree*>(t);
  counts[0] = tree->nodeCount();
  counts[1] = tree->leafCount;
  counts[2] = tree->selection.clusters.size();
  counts[3] = tree->selection.leaves.size();
}
// Node attributes (SplatCloud SoA), layout (6 words per node), clusters (16 words), leaves.
void ref_lod_read(const void* t, float* positions, float* covariances, float* colors,
                  float* alphas, float* sh, uint32_t* layout, uint32_t* clusters,
                  uint32_t* leaves) {
  const auto* tree = static_cast<const splat::LodTree*>(t);
  const auto& c = tree->nodes;
  if (positions) std::memcpy(positions, c.positions.data(), c.positions.size() * 4);
  if (covariances) std::memcpy(covariances, c.covariances.data(), c.covariances.size() * 4);
  if (colors) std::memcpy(colors, c.colors.data(), c.colors.size() * 4);
  if (alphas) std::memcpy(alphas, c.alphas.data(), c.alphas.size() * 4);
  if (sh && !c.sh.empty()) std::memcpy(sh, c.sh.data(), c.sh.size() * 4);
  if (layout) std::memcpy(layout, tree->layout.data(), tree->layout.size() * sizeof(splat::LodNode));
  if (clusters)
    std::memcpy(clusters, tree->selection.clusters.data(),
                tree->selection.clusters.size() * sizeof(splat::LodCluster));
  if (leaves) std::memcpy(leaves, tree->selection.leaves.data(), tree->selection.leaves.size() * 4);
}

// writeLodSplat (LodFile.cpp:430-460): 1 when written; the reason otherwise.
int ref_lod_write(const void* t, const char* path, char* error, size_t error_size) {
  const auto r = splat::writeLodSplat(*static_cast<const splat::LodTree*>(t), path);
  if (!r) std::snprintf(error, error_size, "%s", r.error().message.c_str());
  return r ? 1 : 0;
}
// decodeLodSplat (LodFile.cpp:382-427): a tree to read back with the accessors above, or null and
// the reason.
void* ref_lod_decode(const uint8_t* data, size_t size, int max_sh_degree, char* error, size_t error_size) {
  auto r = splat::decodeLodSplat(data, size, max_sh_degree);
  if (!r) {
    std::snprintf(error, error_size, "%s", r.error().message.c_str());
    return nullptr;
  }
  return new splat::LodTree(std::move(r.value()));
}
// readLodSplatHeader (LodFile.cpp:355-380): 1 and out[0..6] = version, nodes, leaves, SH degree,
// depth, clusters, leaf references; 0 and the reason otherwise.
int ref_lod_header(const uint8_t* data, size_t size, uint64_t* out, char* error, size_t error_size) {
  const auto r = splat::readLodSplatHeader(data, size);
  if (!r) {
    std::snprintf(error, error_size, "%s", r.error().message.c_str());
    return 0;
  }
  const auto& h = r.value();
  const uint64_t v[7] = {h.version, h.nodeCount, h.leafCount, static_cast<uint64_t>(h.shDegree), h.depth,
                         h.clusterCount, h.leafRefCount};
  std::memcpy(out, v, sizeof v);
  return 1;
}
// validateLodTree (LodFile.cpp:336-344): 1 and the depth, or 0 and the reason.
int ref_lod_validate(const void* t, uint32_t* depth, char* error, size_t error_size) {
  const auto r = splat::validateLodTree(*static_cast<const splat::LodTree*>(t));
  if (!r) {
    std::snprintf(error, error_size, "%s", r.error().message.c_str());
    return 0;
  }
  *depth = r.value();
  return 1;
}
// The tree's bounds: min xyz, max xyz.
void ref_lod_bounds(const void* t, float* out6) {
  const auto& b = static_cast<const splat::LodTree*>(t)->nodes.bounds;
  for (int j = 0; j < 3; ++j) {
    out6[j] = b.min[j];
    out6[3 + j] = b.max[j];
  }
}

}  // extern "C"
