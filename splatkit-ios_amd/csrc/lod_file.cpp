// The LODSPLAT file (packages/splat-core/include/splat/lod/LodFile.h:12-16): a 64-byte header, one
// record per node (LodNode, covariance[6], RGB[3], alpha, then the higher SH coefficients, all
// little-endian 32-bit words) and, in version 2, the LodCluster records and the leaf references.
// The decoder and the checks it runs restate LodFile.cpp (readLodSplatHeader :355-380,
// decodeLodSplat :382-427, validateLodTree :336-344) with one change: the reference refuses files
// of more than 20,000,000 nodes (kMaxNodes, :24), which the 20M-leaf world's tree exceeds; here the
// limit is what the header's 32-bit counts and the renderer's 32-bit indices hold (2^31 nodes). The
// writer produces the reference writer's bytes (writeLodSplat :430-460).
#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "skm_lod.h"
#include "skm_lod_host.h"

namespace {

constexpr uint8_t kMagic[8] = {'L', 'O', 'D', 'S', 'P', 'L', 'A', 'T'};
constexpr size_t kHeader = 64;
constexpr size_t kNodeFixedBytes = 64;  // LodNode (24) + covariance (24) + RGB (12) + alpha (4)
constexpr size_t kClusterBytes = 64;
constexpr uint64_t kMaxNodes = uint64_t{1} << 31;
constexpr uint32_t kMaxDepth = 64;  // LodFile.cpp:30: bounds a corrupt file's depth

uint32_t rd32(const uint8_t* p) {
  return uint32_t{p[0]} | (uint32_t{p[1]} << 8) | (uint32_t{p[2]} << 16) | (uint32_t{p[3]} << 24);
}
float rdf(const uint8_t* p) {
  const uint32_t b = rd32(p);
  float f;
  std::memcpy(&f, &b, 4);
  return f;
}
void wr32(uint8_t* p, uint32_t v) {
  for (int j = 0; j < 4; ++j) p[j] = static_cast<uint8_t>(v >> (8 * j));
}
void wrf(uint8_t* p, float f) {
  uint32_t b;
  std::memcpy(&b, &f, 4);
  wr32(p, b);
}
size_t sh_floats(int degree) { return static_cast<size_t>((degree + 1) * (degree + 1) - 1) * 3; }

bool finite_nonneg(float v) { return std::isfinite(v) && v >= 0.0f; }

// validateLodTree (LodFile.cpp:336-344) over a tree in this library's layout: the first failure's
// message, or nullptr with *depth the refinement rounds (the deepest interior level + 1).
const char* validate(const skm_lod_tree& t, uint32_t* depth) {
  const size_t n = t.node_count();
  // checkDimensions (:62-70).
  if (n == 0 || n > kMaxNodes || t.sh_degree < 0 || t.sh_degree > 3 || t.leaf_count > n ||
      t.layout.size() != n || t.positions.size() != n * 3 || t.covariances.size() != n * 6 ||
      t.colors.size() != n * 3 || t.sh.size() != n * sh_floats(t.sh_degree))
    return "LOD attribute dimensions or node count invalid";
  // checkBounds (:72-79).
  for (int j = 0; j < 3; ++j)
    if (!std::isfinite(t.bounds_min[j]) || !std::isfinite(t.bounds_max[j]) || t.bounds_min[j] > t.bounds_max[j])
      return "LOD bounds invalid";
  // scanNodes (:103-141): node attributes, covariance, and children as the next unclaimed run
  // (breadth-first), depths bounded.
  std::vector<uint8_t> level(n, 0);
  size_t next = 1, leaves = 0;
  uint32_t deepest = 0;
  for (size_t i = 0; i < n; ++i) {
    if (i >= next) return "LOD has unreachable nodes";
    const skm_lod_node& node = t.layout[i];
    const float a = t.alphas[i];
    if (!std::isfinite(node.size) || node.size < 0 || !std::isfinite(a) || a < 0 || a > 1000)
      return "LOD size or opacity invalid";
    for (int j = 0; j < 3; ++j) {
      const float c = t.colors[i * 3 + j];
      if (!std::isfinite(node.position[j]) || node.position[j] != t.positions[i * 3 + j] || !std::isfinite(c) ||
          c < 0 || c > 1)
        return "LOD position or colour invalid";
    }
    for (int j = 0; j < 6; ++j)
      if (!std::isfinite(t.covariances[i * 6 + j])) return "LOD covariance is non-finite";
    const auto e = skm::symmetric_eigenvalues(&t.covariances[i * 6]);
    if (e[2] < -1e-5f * std::max(e[0], 1e-20f)) return "LOD covariance is not positive semidefinite";
    if (node.child_count == 0) {
      ++leaves;
      continue;
    }
    if (node.child_start != next || node.child_count > n - next || level[i] >= kMaxDepth)
      return "LOD child ranges are not a bounded BFS tree";
    for (size_t k = next; k < next + node.child_count; ++k) level[k] = static_cast<uint8_t>(level[i] + 1);
    deepest = std::max(deepest, uint32_t{level[i]} + 1u);
    next += node.child_count;
  }
  if (next != n || leaves != t.leaf_count) return "LOD topology or leaf count invalid";
  for (float v : t.sh)
    if (!std::isfinite(v)) return "LOD SH is non-finite";
  // checkSelection (:237-250).
  const auto& cl = t.clusters;
  if (cl.empty()) {
    if (!t.leaves.empty()) return "LOD selection leaves without clusters";
    *depth = deepest;
    return nullptr;
  }
  if (cl.size() != std::max<size_t>(1, n - leaves) || t.leaves.size() != leaves || cl[0].node != 0)
    return "LOD selection dimensions invalid";
  size_t cursor_cluster = 1, cursor_leaf = 0;
  for (size_t k = 0; k < cl.size(); ++k) {
    const skm_lod_cluster& c = cl[k];
    // checkClusterRanges (:150-158) and checkClusterMetadata (:160-167).
    if (k >= cursor_cluster || c.node >= n || c.child_start != cursor_cluster ||
        c.child_count > cl.size() - cursor_cluster || c.leaf_start != cursor_leaf ||
        c.leaf_count > leaves - cursor_leaf)
      return "LOD selection ranges or error metadata invalid";
    if (!finite_nonneg(c.error) || !finite_nonneg(c.color_variance) || !finite_nonneg(c.opacity) ||
        c.opacity > 1 || !finite_nonneg(c.radius))
      return "LOD selection ranges or error metadata invalid";
    // checkClusterSphere (:170-180).
    float radius2 = 0;
    for (int j = 0; j < 3; ++j) {
      if (!std::isfinite(c.center[j]) || !finite_nonneg(c.extent[j])) return "LOD selection bounds invalid";
      radius2 += c.extent[j] * c.extent[j];
    }
    if (c.radius + 1e-5f * std::max(c.radius, 1.0f) < std::sqrt(radius2)) return "LOD sphere does not enclose its bounds";
    // checkClusterMembers (:218-228) with checkClusterMember (:202-216) and checkMemberSupport
    // (:184-198): the node's children (or the node itself) map onto the cluster's interior children and
    // leaf references in order, and each member's box sits inside the cluster's.
    const skm_lod_node& original = t.layout[c.node];
    uint32_t interior = 0, member_leaves = 0, subtree = 0;
    for (uint32_t j = 0; j < std::max(original.child_count, 1u); ++j) {
      const uint32_t index = original.child_count ? original.child_start + j : c.node;
      const skm_lod_cluster* child = nullptr;
      if (t.layout[index].child_count > 0) {
        if (interior >= c.child_count) return "LOD missing interior child";
        child = &cl[c.child_start + interior++];
        if (child->node != index) return "LOD selection interior mapping invalid";
        subtree += child->subtree_leaves;
      } else {
        if (member_leaves >= c.leaf_count || t.leaves[c.leaf_start + member_leaves] != index)
          return "LOD selection leaf mapping invalid";
        ++member_leaves;
        ++subtree;
      }
      const float reach = std::sqrt(2 * std::log(std::max(255.0f * t.alphas[index], 1.0f)));
      static constexpr size_t kDiagonal[3] = {0, 3, 5};
      for (int axis = 0; axis < 3; ++axis) {
        const float center = child ? child->center[axis] : t.positions[size_t{index} * 3 + axis];
        const float variance = std::max(t.covariances[size_t{index} * 6 + kDiagonal[axis]], 0.0f);
        const float extent = child ? child->extent[axis] : reach * std::sqrt(variance);
        const float tolerance = 1e-4f * std::max({1.0f, std::abs(center), extent});
        if (std::abs(center - c.center[axis]) + extent > c.extent[axis] + tolerance)
          return "LOD bounds do not enclose descendant support";
      }
    }
    if (interior != c.child_count || member_leaves != c.leaf_count || subtree != c.subtree_leaves)
      return "LOD selection subtree mismatch";
    cursor_cluster += interior;
    cursor_leaf += member_leaves;
  }
  if (cursor_cluster != cl.size() || cursor_leaf != leaves) return "LOD selection has unreachable records";
  *depth = deepest;
  return nullptr;
}

void set_error(char* error, size_t size, const char* message) {
  if (error != nullptr && size > 0) std::snprintf(error, size, "%s", message);
}

// readLodSplatHeader (LodFile.cpp:355-380), node cap lifted.
const char* read_header(const uint8_t* data, size_t size, skm_lod_file_header* h) {
  if (data == nullptr || size < sizeof kMagic || std::memcmp(data, kMagic, sizeof kMagic) != 0)
    return "not a LODSPLAT file";
  if (size < kHeader) return "truncated LODSPLAT header";
  std::memset(h, 0, sizeof *h);
  h->version = rd32(data + 8);
  if (h->version != 1 && h->version != 2) return "unsupported LODSPLAT version";
  h->node_count = rd32(data + 12);
  h->leaf_count = rd32(data + 16);
  const uint32_t degree = rd32(data + 20);
  h->depth = rd32(data + 24);
  h->cluster_count = rd32(data + 28);
  for (int j = 0; j < 3; ++j) {
    h->bounds_min[j] = rdf(data + 32 + j * 4);
    h->bounds_max[j] = rdf(data + 44 + j * 4);
  }
  h->leaf_ref_count = rd32(data + 56);
  const uint32_t cluster_bytes = rd32(data + 60);
  const uint64_t n = h->node_count, clusters = h->cluster_count, refs = h->leaf_ref_count;
  if (n == 0 || n > kMaxNodes || degree > 3 || h->depth > kMaxDepth || clusters > n || refs > n ||
      (h->version == 1 && (clusters || refs || cluster_bytes)) ||
      (h->version == 2 && (!clusters || refs != h->leaf_count || cluster_bytes != kClusterBytes)))
    return "invalid LODSPLAT header";
  h->sh_degree = static_cast<int32_t>(degree);
  const uint64_t stride = kNodeFixedBytes + sh_floats(h->sh_degree) * 4;
  if (size != kHeader + n * stride + clusters * kClusterBytes + refs * 4)
    return "LODSPLAT length does not match node records";
  return nullptr;
}

// decodeLodSplat (LodFile.cpp:382-427): the records, SH kept up to max_sh_degree, then the checks.
skm_lod_tree* decode(const uint8_t* data, size_t size, int max_sh_degree, char* error, size_t error_size) {
  skm_lod_file_header h;
  if (const char* why = read_header(data, size, &h)) {
    set_error(error, error_size, why);
    return nullptr;
  }
  const size_t n = h.node_count;
  const size_t stride = kNodeFixedBytes + sh_floats(h.sh_degree) * 4;
  auto t = std::make_unique<skm_lod_tree>();
  t->leaf_count = h.leaf_count;
  t->sh_degree = std::min(h.sh_degree, std::clamp(max_sh_degree, 0, 3));
  t->sh_stride = t->sh_degree > 0 ? sh_floats(t->sh_degree) : 0;
  const size_t sh = t->sh_stride;
  t->layout.resize(n);
  t->positions.resize(n * 3);
  t->covariances.resize(n * 6);
  t->colors.resize(n * 3);
  t->alphas.resize(n);
  t->sh.resize(n * sh);
  for (int j = 0; j < 3; ++j) {
    t->bounds_min[j] = h.bounds_min[j];
    t->bounds_max[j] = h.bounds_max[j];
  }
  for (size_t i = 0; i < n; ++i) {
    const uint8_t* p = data + kHeader + i * stride;
    skm_lod_node& node = t->layout[i];
    for (int j = 0; j < 3; ++j) node.position[j] = t->positions[i * 3 + j] = rdf(p + j * 4);
    node.size = rdf(p + 12);
    node.child_start = rd32(p + 16);
    node.child_count = rd32(p + 20);
    for (int j = 0; j < 6; ++j) t->covariances[i * 6 + j] = rdf(p + 24 + j * 4);
    for (int j = 0; j < 3; ++j) t->colors[i * 3 + j] = rdf(p + 48 + j * 4);
    t->alphas[i] = rdf(p + 60);
    for (size_t j = 0; j < sh; ++j) t->sh[i * sh + j] = rdf(p + kNodeFixedBytes + j * 4);
  }
  if (h.version == 2) {
    t->clusters.resize(h.cluster_count);
    t->leaves.resize(h.leaf_ref_count);
    const uint8_t* p = data + kHeader + n * stride;
    for (size_t k = 0; k < h.cluster_count; ++k) {
      uint32_t words[16];
      for (int j = 0; j < 16; ++j) words[j] = rd32(p + k * kClusterBytes + j * 4);
      std::memcpy(&t->clusters[k], words, kClusterBytes);
    }
    p += h.cluster_count * kClusterBytes;
    for (size_t k = 0; k < h.leaf_ref_count; ++k) t->leaves[k] = rd32(p + k * 4);
  }
  uint32_t depth = 0;
  if (const char* why = validate(*t, &depth)) {
    set_error(error, error_size, why);
    return nullptr;
  }
  if (depth != h.depth) {
    set_error(error, error_size, "LODSPLAT declared depth differs from topology");
    return nullptr;
  }
  return t.release();
}

}  // namespace

extern "C" {

skm_status skm_lod_read_header(const uint8_t* data, size_t size, skm_lod_file_header* out, char* error,
                               size_t error_size) {
  if (out == nullptr) return SKM_ERR_INVALID;
  if (const char* why = read_header(data, size, out)) {
    set_error(error, error_size, why);
    return SKM_ERR_INVALID;
  }
  return SKM_OK;
}

skm_lod_tree* skm_lod_decode(const uint8_t* data, size_t size, int max_sh_degree, char* error, size_t error_size) {
  return decode(data, size, max_sh_degree, error, error_size);
}

skm_lod_tree* skm_lod_read(const char* path, int max_sh_degree, char* error, size_t error_size) {
  if (path == nullptr) {
    set_error(error, error_size, "no path");
    return nullptr;
  }
  const int fd = ::open(path, O_RDONLY);
  if (fd < 0) {
    set_error(error, error_size, "cannot open LODSPLAT file");
    return nullptr;
  }
  struct stat st{};
  if (::fstat(fd, &st) != 0 || st.st_size <= 0) {
    ::close(fd);
    set_error(error, error_size, "cannot read LODSPLAT file");
    return nullptr;
  }
  const size_t size = static_cast<size_t>(st.st_size);
  void* map = ::mmap(nullptr, size, PROT_READ, MAP_PRIVATE, fd, 0);
  ::close(fd);
  if (map == MAP_FAILED) {
    set_error(error, error_size, "cannot map LODSPLAT file");
    return nullptr;
  }
  (void)::madvise(map, size, MADV_SEQUENTIAL);
  skm_lod_tree* t = decode(static_cast<const uint8_t*>(map), size, max_sh_degree, error, error_size);
  ::munmap(map, size);
  return t;
}

skm_status skm_lod_validate(const skm_lod_tree* tree, uint32_t* depth, char* error, size_t error_size) {
  if (tree == nullptr) return SKM_ERR_INVALID;
  uint32_t d = 0;
  if (const char* why = validate(*tree, &d)) {
    set_error(error, error_size, why);
    return SKM_ERR_INVALID;
  }
  if (depth) *depth = d;
  return SKM_OK;
}

void skm_lod_bounds(const skm_lod_tree* tree, float* bounds6) {
  if (tree == nullptr || bounds6 == nullptr) return;
  for (int j = 0; j < 3; ++j) {
    bounds6[j] = tree->bounds_min[j];
    bounds6[3 + j] = tree->bounds_max[j];
  }
}

// writeLodSplat (LodFile.cpp:430-460): validates first, never replaces an existing file.
skm_status skm_lod_write(const skm_lod_tree* tree, const char* path, char* error, size_t error_size) {
  if (tree == nullptr || path == nullptr) return SKM_ERR_INVALID;
  uint32_t depth = 0;
  if (const char* why = validate(*tree, &depth)) {
    set_error(error, error_size, why);
    return SKM_ERR_INVALID;
  }
  const skm_lod_tree& t = *tree;
  const size_t n = t.node_count();
  if (n > 0xffffffffull) {
    set_error(error, error_size, "too many nodes for the header's 32-bit count");
    return SKM_ERR_INVALID;
  }
  const auto closer = [](FILE* f) { std::fclose(f); };
  std::unique_ptr<FILE, decltype(closer)> file(std::fopen(path, "wbx"), closer);
  if (!file) {
    set_error(error, error_size, "cannot create LODSPLAT output (exists or unwritable)");
    return SKM_ERR_STATE;
  }
  uint8_t header[kHeader] = {};
  std::memcpy(header, kMagic, sizeof kMagic);
  wr32(header + 8, t.clusters.empty() ? 1u : 2u);
  wr32(header + 12, static_cast<uint32_t>(n));
  wr32(header + 16, static_cast<uint32_t>(t.leaf_count));
  wr32(header + 20, static_cast<uint32_t>(t.sh_degree));
  wr32(header + 24, depth);
  wr32(header + 28, static_cast<uint32_t>(t.clusters.size()));
  for (int j = 0; j < 3; ++j) {
    wrf(header + 32 + j * 4, t.bounds_min[j]);
    wrf(header + 44 + j * 4, t.bounds_max[j]);
  }
  wr32(header + 56, static_cast<uint32_t>(t.leaves.size()));
  wr32(header + 60, t.clusters.empty() ? 0u : static_cast<uint32_t>(kClusterBytes));
  bool ok = std::fwrite(header, 1, kHeader, file.get()) == kHeader;
  const size_t sh = sh_floats(t.sh_degree) * (t.sh_degree > 0 ? 1 : 0);
  const size_t stride = kNodeFixedBytes + sh * 4;
  constexpr size_t kBatch = 4096;
  std::vector<uint8_t> block(stride * kBatch);
  for (size_t start = 0; ok && start < n; start += kBatch) {
    const size_t count = std::min(kBatch, n - start);
    for (size_t k = 0; k < count; ++k) {
      const size_t i = start + k;
      uint8_t* p = block.data() + k * stride;
      const skm_lod_node& node = t.layout[i];
      for (int j = 0; j < 3; ++j) wrf(p + j * 4, node.position[j]);
      wrf(p + 12, node.size);
      wr32(p + 16, node.child_start);
      wr32(p + 20, node.child_count);
      for (int j = 0; j < 6; ++j) wrf(p + 24 + j * 4, t.covariances[i * 6 + j]);
      for (int j = 0; j < 3; ++j) wrf(p + 48 + j * 4, t.colors[i * 3 + j]);
      wrf(p + 60, t.alphas[i]);
      for (size_t j = 0; j < sh; ++j) wrf(p + kNodeFixedBytes + j * 4, t.sh[i * sh + j]);
    }
    ok = std::fwrite(block.data(), stride, count, file.get()) == count;
  }
  for (size_t start = 0; ok && start < t.clusters.size(); start += kBatch) {
    const size_t count = std::min(kBatch, t.clusters.size() - start);
    for (size_t k = 0; k < count; ++k) {
      uint32_t words[16];
      std::memcpy(words, &t.clusters[start + k], kClusterBytes);
      for (int j = 0; j < 16; ++j) wr32(block.data() + k * kClusterBytes + j * 4, words[j]);
    }
    ok = std::fwrite(block.data(), kClusterBytes, count, file.get()) == count;
  }
  for (size_t start = 0; ok && start < t.leaves.size(); start += kBatch) {
    const size_t count = std::min(kBatch, t.leaves.size() - start);
    for (size_t k = 0; k < count; ++k) wr32(block.data() + k * 4, t.leaves[start + k]);
    ok = std::fwrite(block.data(), 4, count, file.get()) == count;
  }
  ok = std::fclose(file.release()) == 0 && ok;
  if (!ok) {
    set_error(error, error_size, "LODSPLAT write failed; output is incomplete");
    return SKM_ERR_STATE;
  }
  return SKM_OK;
}

}  // extern "C"
