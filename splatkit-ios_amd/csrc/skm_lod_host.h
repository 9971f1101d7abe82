// The owning level-of-detail tree behind skm_lod_tree (include/skm_lod.h), shared by the builder
// (lod_build.cpp) and the LODSPLAT file codec (lod_file.cpp). Host code only.
#pragma once
#include <array>
#include <cstddef>
#include <cstdint>
#include <vector>

#include "skm_lod.h"

struct skm_lod_tree {
  std::vector<float> positions, covariances, colors, alphas, sh;
  int sh_degree = 0;
  size_t sh_stride = 0;  // SH floats per node
  std::vector<skm_lod_node> layout;
  uint64_t leaf_count = 0;
  std::vector<skm_lod_cluster> clusters;
  std::vector<uint32_t> leaves;
  // The cloud's bounds (SplatCloud::bounds, SplatCloud.h:10-13): the box the builder's octree spans,
  // kept in the file header.
  float bounds_min[3] = {0, 0, 0};
  float bounds_max[3] = {0, 0, 0};

  size_t node_count() const { return alphas.size(); }
  void set_bounds_from_positions() {
    const size_t n = node_count();
    for (int c = 0; c < 3; ++c) bounds_min[c] = bounds_max[c] = n ? positions[c] : 0.0f;
    for (size_t i = 0; i < n; ++i)
      for (int c = 0; c < 3; ++c) {
        const float v = positions[i * 3 + c];
        bounds_min[c] = v < bounds_min[c] ? v : bounds_min[c];
        bounds_max[c] = v > bounds_max[c] ? v : bounds_max[c];
      }
  }
};

namespace skm {
// symmetricEigenvalues (SymmetricEigen.cpp:8-40): descending eigenvalues of a symmetric 3x3 matrix
// given as its upper triangle xx xy xz yy yz zz (lod_build.cpp).
std::array<float, 3> symmetric_eigenvalues(const float* upper);
}  // namespace skm
