// Level-of-detail trees (include/skm_lod.h): the octree with moment-matched merges of
// packages/splat-core/src/lod/LodTree.cpp and the selection metadata of LodSelection.cpp, restated
// for a large host: the octree's shape is found first (cheap, sequential), then every merge of one
// height runs in parallel, bottom-up; the selection metadata is built level by level the same way.
// Merges see their members in the same order and with the same float operations as the reference,
// so the trees are identical to its builder's.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <limits>
#include <memory>
#include <thread>
#include <vector>

#include "skm_lod.h"
#include "skm_lod_host.h"

namespace {

constexpr uint32_t kMaxOctreeDepth = 21;  // LodTree.h:57

template <typename F>
void parallel_for(size_t n, uint32_t threads, F&& body) {
  threads = std::max(1u, std::min<uint32_t>(threads, static_cast<uint32_t>(std::min<size_t>(n, 256))));
  if (threads <= 1 || n < 64) {
    for (size_t i = 0; i < n; ++i) body(i);
    return;
  }
  std::vector<std::thread> pool;
  for (uint32_t t = 0; t < threads; ++t)
    pool.emplace_back([&, t] {
      for (size_t i = t; i < n; i += threads) body(i);
    });
  for (auto& th : pool) th.join();
}

}  // namespace

namespace skm {
// Eigenvalues of a symmetric 3x3 matrix (upper triangle xx xy xz yy yz zz), descending, by the
// trigonometric closed form (Smith 1961), in double (SymmetricEigen.cpp:8-40).
std::array<float, 3> symmetric_eigenvalues(const float* u) {
  const double xx = u[0], xy = u[1], xz = u[2], yy = u[3], yz = u[4], zz = u[5];
  const double p1 = xy * xy + xz * xz + yz * yz;
  std::array<double, 3> e;
  if (p1 < 1e-30) {
    e = {xx, yy, zz};
  } else {
    const double q = (xx + yy + zz) / 3;
    const double p2 = (xx - q) * (xx - q) + (yy - q) * (yy - q) + (zz - q) * (zz - q) + 2 * p1;
    const double p = std::sqrt(p2 / 6);
    const double bxx = (xx - q) / p, byy = (yy - q) / p, bzz = (zz - q) / p;
    const double bxy = xy / p, bxz = xz / p, byz = yz / p;
    const double det = bxx * (byy * bzz - byz * byz) - bxy * (bxy * bzz - byz * bxz) +
                       bxz * (bxy * byz - byy * bxz);
    const double r = std::clamp(det / 2, -1.0, 1.0);
    const double phi = std::acos(r) / 3;
    const double e0 = q + 2 * p * std::cos(phi);
    const double e2 = q + 2 * p * std::cos(phi + 2 * M_PI / 3);
    e = {e0, 3 * q - e0 - e2, e2};
  }
  std::sort(e.begin(), e.end(), [](double a, double b) { return a > b; });
  return {static_cast<float>(e[0]), static_cast<float>(e[1]), static_cast<float>(e[2])};
}
}  // namespace skm

namespace {
using skm::symmetric_eigenvalues;

std::array<float, 3> semi_axes(const float* cov) {
  const auto e = symmetric_eigenvalues(cov);
  return {std::sqrt(std::max(e[0], 0.0f)), std::sqrt(std::max(e[1], 0.0f)), std::sqrt(std::max(e[2], 0.0f))};
}

// Surface of an ellipsoid by Knud Thomsen's approximation (p = 1.6075): a splat's merge weight
// is its area times its opacity (LodTree.cpp:14-22).
float ellipsoid_area(const std::array<float, 3>& s) {
  constexpr float kP = 1.6075f;
  const float sum = std::pow(s[0] * s[1], kP) + std::pow(s[0] * s[2], kP) + std::pow(s[1] * s[2], kP);
  return 4.0f * static_cast<float>(M_PI) * std::pow(sum / 3.0f, 1.0f / kP);
}

// How far an opacity above one stays solid, in sigmas (LodTree.cpp:24-29).
float opacity_reach(float alpha) {
  return alpha > 1.0f ? std::sqrt(1.0f + static_cast<float>(M_E) * std::log(alpha)) : 1.0f;
}

// Node attributes during the build: leaves first, merged nodes appended in post-order.
struct Arena {
  std::vector<float> pos, cov, col, alpha, sh, size;
  size_t sh_stride = 0;
  // Children of merged node j (index leaves + j): pool[first[j] .. first[j] + count[j]).
  std::vector<uint32_t> first, count, pool;
  uint32_t leaves = 0;

  const uint32_t* kids(uint32_t node, uint32_t* n) const {
    if (node < leaves) {
      *n = 0;
      return nullptr;
    }
    *n = count[node - leaves];
    return pool.data() + first[node - leaves];
  }

  // The moment-matched stand-in of members (LodTree.cpp:46-102), written into slot `index`.
  void merge(uint32_t index, const uint32_t* members, uint32_t n) {
    std::vector<float> weights(n);
    float total = 0.0f;
    for (uint32_t k = 0; k < n; ++k) {
      const uint32_t i = members[k];
      weights[k] = ellipsoid_area(semi_axes(&cov[size_t{i} * 6])) * alpha[i];
      total += weights[k];
    }
    for (float& w : weights) w = total > 1e-30f ? w / total : 1.0f / static_cast<float>(n);
    float center[3] = {0, 0, 0}, rgb[3] = {0, 0, 0};
    std::vector<float> sh_sum(sh_stride, 0.0f);
    for (uint32_t k = 0; k < n; ++k) {
      const uint32_t i = members[k];
      for (int c = 0; c < 3; ++c) {
        center[c] += weights[k] * pos[size_t{i} * 3 + c];
        rgb[c] += weights[k] * col[size_t{i} * 3 + c];
      }
      for (size_t c = 0; c < sh_stride; ++c) sh_sum[c] += weights[k] * sh[i * sh_stride + c];
    }
    for (float& channel : rgb) channel = std::clamp(channel, 0.0f, 1.0f);
    float m[6] = {0, 0, 0, 0, 0, 0};
    for (uint32_t k = 0; k < n; ++k) {
      const uint32_t i = members[k];
      const float dx = pos[size_t{i} * 3] - center[0];
      const float dy = pos[size_t{i} * 3 + 1] - center[1];
      const float dz = pos[size_t{i} * 3 + 2] - center[2];
      const float* c = &cov[size_t{i} * 6];
      const float w = weights[k];
      m[0] += w * (dx * dx + c[0]);
      m[1] += w * (dx * dy + c[1]);
      m[2] += w * (dx * dz + c[2]);
      m[3] += w * (dy * dy + c[3]);
      m[4] += w * (dy * dz + c[4]);
      m[5] += w * (dz * dz + c[5]);
    }
    const auto axes = semi_axes(m);
    const float a = std::clamp(total / std::max(ellipsoid_area(axes), 1e-30f), 0.0f, 1000.0f);
    std::memcpy(&pos[size_t{index} * 3], center, 12);
    std::memcpy(&cov[size_t{index} * 6], m, 24);
    std::memcpy(&col[size_t{index} * 3], rgb, 12);
    alpha[index] = a;
    if (sh_stride) std::memcpy(&sh[index * sh_stride], sh_sum.data(), sh_stride * 4);
    size[index] = 2.0f * axes[0] * opacity_reach(a);
  }
};

struct Cell {
  uint64_t key;
  uint32_t node;
};

// Octree shape over Morton-sorted leaves (LodTree.cpp:158-196): a cell of one leaf is that leaf;
// a cell of at most cluster_leaves (or at full depth) merges its leaves; a cell with one occupied
// octant is that octant; otherwise it merges its octants' nodes. Merged nodes are numbered in
// post-order after the leaves, as the reference appends them.
struct Shape {
  const std::vector<Cell>& sorted;
  size_t cluster_leaves;
  uint32_t depth;
  Arena& arena;
  std::vector<uint32_t> height;  // per merged node

  uint32_t add(const std::vector<uint32_t>& members) {
    const uint32_t j = static_cast<uint32_t>(arena.first.size());
    arena.first.push_back(static_cast<uint32_t>(arena.pool.size()));
    arena.count.push_back(static_cast<uint32_t>(members.size()));
    arena.pool.insert(arena.pool.end(), members.begin(), members.end());
    uint32_t h = 0;
    for (uint32_t m : members) h = std::max(h, m < arena.leaves ? 0u : height[m - arena.leaves]);
    height.push_back(h + 1);
    return arena.leaves + j;
  }

  uint32_t build(size_t begin, size_t end, uint32_t level) {
    if (end - begin == 1) return sorted[begin].node;
    if (end - begin <= cluster_leaves || level == depth) {
      std::vector<uint32_t> members;
      members.reserve(end - begin);
      for (size_t j = begin; j < end; ++j) members.push_back(sorted[j].node);
      return add(members);
    }
    const uint32_t shift = 3 * (depth - level - 1);
    if ((sorted[begin].key >> shift) == (sorted[end - 1].key >> shift)) return build(begin, end, level + 1);
    std::vector<uint32_t> members;
    for (size_t first = begin; first < end;) {
      size_t last = first + 1;
      while (last < end && (sorted[last].key >> shift) == (sorted[first].key >> shift)) ++last;
      members.push_back(build(first, last, level + 1));
      first = last;
    }
    return add(members);
  }
};

float length3(const float* a, const float* b) {
  float s = 0;
  for (int j = 0; j < 3; ++j) s += (a[j] - b[j]) * (a[j] - b[j]);
  return std::sqrt(s);
}

// Frobenius covariance discrepancy, square-rooted twice to world length (LodSelection.cpp:15-24).
float covariance_error(const float* a, const float* b) {
  float s = 0;
  for (int j = 0; j < 6; ++j) {
    const float d = a[j] - b[j];
    s += d * d * ((j == 1 || j == 2 || j == 4) ? 2 : 1);
  }
  return std::sqrt(std::sqrt(s));
}

// Largest per-channel colour difference plus the SH bands' directional amplitude bound from the
// addition theorem (LodSelection.cpp:25-44).
float appearance_difference(const skm_lod_tree& t, uint32_t a, uint32_t b) {
  float difference = 0;
  const size_t stride = t.sh_stride;
  for (int c = 0; c < 3; ++c) {
    float d = std::abs(t.colors[size_t{a} * 3 + c] - t.colors[size_t{b} * 3 + c]);
    for (int band = 1, start = 0; band <= t.sh_degree; ++band) {
      float norm = 0;
      for (int k = 0; k < 2 * band + 1; ++k) {
        const float v = t.sh[a * stride + (start + k) * 3 + c] - t.sh[b * stride + (start + k) * 3 + c];
        norm += v * v;
      }
      d += std::sqrt(norm * (2 * band + 1) / (4.0f * static_cast<float>(M_PI)));
      start += 2 * band + 1;
    }
    difference = std::max(difference, d);
  }
  return difference;
}

// Interior clusters breadth-first with their direct leaves, then bounds, error, colour variance
// and opacity bottom-up (LodSelection.cpp:47-133); clusters of one depth are independent.
void build_selection(skm_lod_tree& t, uint32_t threads) {
  t.clusters.clear();
  t.leaves.clear();
  if (t.layout.empty()) return;
  std::vector<uint32_t> order{0}, depth_of{0};
  t.clusters.resize(1);
  t.leaves.reserve(t.leaf_count);
  for (size_t head = 0; head < order.size(); ++head) {
    const uint32_t index = order[head];
    const skm_lod_node& node = t.layout[index];
    skm_lod_cluster cl{};
    cl.node = index;
    cl.child_start = static_cast<uint32_t>(order.size());
    cl.leaf_start = static_cast<uint32_t>(t.leaves.size());
    if (node.child_count == 0) t.leaves.push_back(index);
    for (uint32_t j = 0; j < node.child_count; ++j) {
      const uint32_t child = node.child_start + j;
      if (t.layout[child].child_count) {
        order.push_back(child);
        depth_of.push_back(depth_of[head] + 1);
      } else {
        t.leaves.push_back(child);
      }
    }
    cl.child_count = static_cast<uint32_t>(order.size()) - cl.child_start;
    cl.leaf_count = static_cast<uint32_t>(t.leaves.size()) - cl.leaf_start;
    t.clusters.resize(order.size());
    t.clusters[head] = cl;
  }
  const uint32_t max_depth = depth_of.back();
  std::vector<std::vector<uint32_t>> by_depth(max_depth + 1);
  for (uint32_t k = 0; k < t.clusters.size(); ++k) by_depth[depth_of[k]].push_back(k);
  for (uint32_t d = max_depth + 1; d-- > 0;) {
    const auto& level = by_depth[d];
    parallel_for(level.size(), threads, [&](size_t q) {
      skm_lod_cluster& cl = t.clusters[level[q]];
      const float* parent = &t.positions[size_t{cl.node} * 3];
      const float* parent_cov = &t.covariances[size_t{cl.node} * 6];
      std::array<float, 3> lo, hi;
      lo.fill(std::numeric_limits<float>::max());
      hi.fill(std::numeric_limits<float>::lowest());
      float weighted_variance = 0, weight_sum = 0, coincident = 1;
      cl.opacity = std::clamp(t.alphas[cl.node], 0.0f, 1.0f);
      auto include = [&](uint32_t index, const float* center, const float* extent, float error, float variance,
                         uint32_t leaves) {
        for (int j = 0; j < 3; ++j) {
          lo[j] = std::min(lo[j], center[j] - extent[j]);
          hi[j] = std::max(hi[j], center[j] + extent[j]);
        }
        const float displacement = length3(&t.positions[size_t{index} * 3], parent);
        const float shape = covariance_error(&t.covariances[size_t{index} * 6], parent_cov);
        cl.error = std::max(cl.error, error + displacement + shape);
        const float difference = appearance_difference(t, cl.node, index);
        const float weight = std::max(t.alphas[index], 0.0f) * leaves;
        weighted_variance += weight * (variance + difference * difference);
        weight_sum += weight;
        cl.subtree_leaves += leaves;
        coincident *= 1.0f - std::clamp(t.alphas[index], 0.0f, 1.0f);
        cl.opacity = std::max(cl.opacity, std::min(t.alphas[index], 1.0f));
      };
      for (uint32_t j = 0; j < cl.child_count; ++j) {
        const skm_lod_cluster& child = t.clusters[cl.child_start + j];
        include(child.node, child.center, child.extent, child.error, child.color_variance, child.subtree_leaves);
      }
      for (uint32_t j = 0; j < cl.leaf_count; ++j) {
        const uint32_t index = t.leaves[cl.leaf_start + j];
        const float* cv = &t.covariances[size_t{index} * 6];
        const float reach = std::sqrt(2 * std::log(std::max(255.0f * t.alphas[index], 1.0f)));
        const float extent[3] = {reach * std::sqrt(std::max(cv[0], 0.0f)), reach * std::sqrt(std::max(cv[3], 0.0f)),
                                 reach * std::sqrt(std::max(cv[5], 0.0f))};
        include(index, &t.positions[size_t{index} * 3], extent, 0, 0, 1);
      }
      for (int j = 0; j < 3; ++j) {
        cl.center[j] = (lo[j] + hi[j]) * 0.5f;
        cl.extent[j] = (hi[j] - lo[j]) * 0.5f;
      }
      cl.radius = length3(hi.data(), cl.center);
      const float peak = std::abs(std::min(t.alphas[cl.node], 1.0f) - (1.0f - coincident));
      cl.error = std::max(cl.error, cl.radius * peak);
      cl.color_variance = weight_sum > 0 ? weighted_variance / weight_sum : 0;
    });
  }
}

bool valid_cloud(const skm_cloud_view* c) {
  return c != nullptr && c->sh_degree >= 0 && c->sh_degree <= 3 &&
         (c->count == 0 || (c->positions && c->covariances && c->colors && c->alphas)) &&
         (c->sh_degree == 0 || c->count == 0 || c->sh != nullptr);
}

void copy_cloud(skm_lod_tree& t, const skm_cloud_view* c) {
  const size_t n = c->count;
  t.sh_degree = c->sh_degree;
  t.sh_stride = c->sh_degree > 0 ? static_cast<size_t>((c->sh_degree + 1) * (c->sh_degree + 1) - 1) * 3 : 0;
  t.positions.assign(c->positions, c->positions + n * 3);
  t.covariances.assign(c->covariances, c->covariances + n * 6);
  t.colors.assign(c->colors, c->colors + n * 3);
  t.alphas.assign(c->alphas, c->alphas + n);
  if (t.sh_stride) t.sh.assign(c->sh, c->sh + n * t.sh_stride);
}

}  // namespace

extern "C" {

skm_lod_tree* skm_lod_build(const skm_cloud_view* cloud, uint32_t octree_depth, uint32_t cluster_leaves,
                            uint32_t threads) {
  if (!valid_cloud(cloud)) return nullptr;
  auto tree = std::make_unique<skm_lod_tree>();
  const uint32_t n = cloud->count;
  tree->leaf_count = n;
  tree->sh_degree = cloud->sh_degree;
  if (n == 0) return tree.release();
  threads = std::max(1u, threads);

  Arena arena;
  arena.leaves = n;
  arena.sh_stride = cloud->sh_degree > 0 ? static_cast<size_t>((cloud->sh_degree + 1) * (cloud->sh_degree + 1) - 1) * 3 : 0;
  arena.pos.assign(cloud->positions, cloud->positions + size_t{n} * 3);
  arena.cov.assign(cloud->covariances, cloud->covariances + size_t{n} * 6);
  arena.col.assign(cloud->colors, cloud->colors + size_t{n} * 3);
  arena.alpha.assign(cloud->alphas, cloud->alphas + n);
  if (arena.sh_stride) arena.sh.assign(cloud->sh, cloud->sh + size_t{n} * arena.sh_stride);
  arena.size.resize(n);
  parallel_for(n, threads, [&](size_t i) { arena.size[i] = 2.0f * semi_axes(&arena.cov[i * 6])[0]; });

  // Morton keys over the positions' bounding box (LodTree.cpp:126-156).
  float lo[3], hi[3];
  for (int c = 0; c < 3; ++c) lo[c] = hi[c] = arena.pos[c];
  for (size_t i = 0; i < n; ++i)
    for (int c = 0; c < 3; ++c) {
      lo[c] = std::min(lo[c], arena.pos[i * 3 + c]);
      hi[c] = std::max(hi[c], arena.pos[i * 3 + c]);
    }
  for (int c = 0; c < 3; ++c) {
    tree->bounds_min[c] = lo[c];
    tree->bounds_max[c] = hi[c];
  }
  const uint32_t depth = std::clamp(octree_depth, 1u, kMaxOctreeDepth);
  float extent = 1e-6f;
  for (int c = 0; c < 3; ++c) extent = std::max(extent, hi[c] - lo[c]);
  std::vector<Cell> sorted(n);
  const uint32_t resolution = 1u << depth;
  parallel_for(n, threads, [&](size_t i) {
    uint64_t key = 0;
    for (uint32_t c = 0; c < 3; ++c) {
      const float unit = std::clamp((arena.pos[i * 3 + c] - lo[c]) / extent, 0.0f, 1.0f);
      const uint32_t grid = std::min(static_cast<uint32_t>(unit * resolution), resolution - 1);
      for (uint32_t bit = 0; bit < depth; ++bit) key |= uint64_t{(grid >> bit) & 1u} << (3 * bit + c);
    }
    sorted[i] = {key, static_cast<uint32_t>(i)};
  });
  std::sort(sorted.begin(), sorted.end(),
            [](const Cell& a, const Cell& b) { return a.key == b.key ? a.node < b.node : a.key < b.key; });

  Shape shape{sorted, std::max<size_t>(cluster_leaves, 2), depth, arena, {}};
  const uint32_t root = shape.build(0, n, 0);
  const size_t merged = arena.first.size();
  const size_t total = size_t{n} + merged;
  arena.pos.resize(total * 3);
  arena.cov.resize(total * 6);
  arena.col.resize(total * 3);
  arena.alpha.resize(total);
  arena.sh.resize(total * arena.sh_stride);
  arena.size.resize(total);
  // Merges bottom-up, one height at a time.
  uint32_t max_h = 0;
  for (uint32_t h : shape.height) max_h = std::max(max_h, h);
  std::vector<std::vector<uint32_t>> by_height(max_h + 1);
  for (uint32_t j = 0; j < merged; ++j) by_height[shape.height[j]].push_back(j);
  for (uint32_t h = 1; h <= max_h; ++h) {
    const auto& level = by_height[h];
    parallel_for(level.size(), threads, [&](size_t q) {
      const uint32_t j = level[q];
      arena.merge(n + j, arena.pool.data() + arena.first[j], arena.count[j]);
    });
  }

  // Breadth-first layout from the root (LodTree.cpp:206-243).
  std::vector<uint32_t> order;
  order.reserve(total);
  tree->layout.assign(total, skm_lod_node{});
  order.push_back(root);
  for (size_t head = 0; head < order.size(); ++head) {
    uint32_t count = 0;
    const uint32_t* kids = arena.kids(order[head], &count);
    tree->layout[head].child_start = static_cast<uint32_t>(order.size());
    tree->layout[head].child_count = count;
    for (uint32_t k = 0; k < count; ++k) order.push_back(kids[k]);
  }
  skm_lod_tree& t = *tree;
  t.sh_stride = arena.sh_stride;
  t.positions.resize(total * 3);
  t.covariances.resize(total * 6);
  t.colors.resize(total * 3);
  t.alphas.resize(total);
  t.sh.resize(total * arena.sh_stride);
  parallel_for(total, threads, [&](size_t i) {
    const uint32_t old = order[i];
    std::copy_n(&arena.pos[size_t{old} * 3], 3, &t.positions[i * 3]);
    std::copy_n(&arena.pos[size_t{old} * 3], 3, t.layout[i].position);
    t.layout[i].size = arena.size[old];
    std::copy_n(&arena.cov[size_t{old} * 6], 6, &t.covariances[i * 6]);
    std::copy_n(&arena.col[size_t{old} * 3], 3, &t.colors[i * 3]);
    t.alphas[i] = arena.alpha[old];
    if (arena.sh_stride) std::copy_n(&arena.sh[old * arena.sh_stride], arena.sh_stride, &t.sh[i * arena.sh_stride]);
  });
  build_selection(t, threads);
  return tree.release();
}

skm_lod_tree* skm_lod_from_layout(const skm_cloud_view* nodes, const skm_lod_node* layout, uint64_t leaf_count) {
  if (!valid_cloud(nodes) || (nodes->count > 0 && layout == nullptr)) return nullptr;
  for (uint64_t i = 0; i < nodes->count; ++i)
    if (layout[i].child_count > 0 && (layout[i].child_start <= i ||
                                      uint64_t{layout[i].child_start} + layout[i].child_count > nodes->count))
      return nullptr;
  auto tree = std::make_unique<skm_lod_tree>();
  copy_cloud(*tree, nodes);
  tree->set_bounds_from_positions();
  tree->layout.assign(layout, layout + nodes->count);
  tree->leaf_count = leaf_count;
  build_selection(*tree, 1);
  return tree.release();
}

void skm_lod_free(skm_lod_tree* tree) { delete tree; }

void skm_lod_view(const skm_lod_tree* t, skm_lod_tree_view* out) {
  if (t == nullptr || out == nullptr) return;
  out->nodes.count = static_cast<uint32_t>(t->alphas.size());
  out->nodes.sh_degree = t->sh_degree;
  out->nodes.positions = t->positions.data();
  out->nodes.covariances = t->covariances.data();
  out->nodes.colors = t->colors.data();
  out->nodes.alphas = t->alphas.data();
  out->nodes.sh = t->sh.empty() ? nullptr : t->sh.data();
  out->layout = t->layout.data();
  out->leaf_count = t->leaf_count;
  out->clusters = t->clusters.data();
  out->cluster_count = t->clusters.size();
  out->leaves = t->leaves.data();
  out->leaf_ref_count = t->leaves.size();
}

uint32_t skm_lod_rounds(const skm_lod_tree_view* t) {
  if (t == nullptr || t->cluster_count == 0) return 0;
  // Clusters are breadth-first: a child's index exceeds its parent's.
  std::vector<uint32_t> depth(t->cluster_count, 0);
  uint32_t rounds = 1;
  for (uint64_t k = 0; k < t->cluster_count; ++k) {
    const skm_lod_cluster& c = t->clusters[k];
    for (uint32_t j = 0; j < c.child_count; ++j) {
      depth[c.child_start + j] = depth[k] + 1;
      rounds = std::max(rounds, depth[k] + 2);
    }
  }
  return rounds;
}

}  // extern "C"
