/*
 * Level-of-detail trees for the GPU tree cut: the octree builder with moment-matched merges
 * (packages/splat-core/src/lod/LodTree.cpp:323-342) and the interior-only selection metadata
 * (packages/splat-core/src/lod/LodSelection.cpp:47-133), in host C++ with the reference's record
 * layouts (packages/splat-core/include/splat/lod/LodTree.h:13-76).
 *
 * A tree's nodes are a SplatCloud, root first, level by level (breadth-first), children of a node
 * contiguous; the leaves are the original splats. The selection walks interior clusters.
 */
#pragma once
#include <stddef.h>
#include <stdint.h>

#include "skm.h"

#ifdef __cplusplus
extern "C" {
#endif

/* LodNode (LodTree.h:13-22), 24 bytes. */
typedef struct skm_lod_node {
  float position[3];
  float size;
  uint32_t child_start;
  uint32_t child_count;
} skm_lod_node;

/* LodCluster (LodTree.h:28-35), 64 bytes. */
typedef struct skm_lod_cluster {
  float center[3];
  float radius;
  float extent[3];
  float error;
  float color_variance;
  float opacity;
  uint32_t node;
  uint32_t child_start;
  uint32_t child_count;
  uint32_t leaf_start;
  uint32_t leaf_count;
  uint32_t subtree_leaves;
} skm_lod_cluster;

/* A tree as borrowed arrays (LodTree, LodTree.h:48-56). */
typedef struct skm_lod_tree_view {
  skm_cloud_view nodes;            /* every node's splat, layout order */
  const skm_lod_node* layout;      /* nodes.count entries */
  uint64_t leaf_count;
  const skm_lod_cluster* clusters; /* interior clusters, breadth-first */
  uint64_t cluster_count;
  const uint32_t* leaves;          /* leaf node indices referenced by clusters */
  uint64_t leaf_ref_count;
} skm_lod_tree_view;

typedef struct skm_lod_tree skm_lod_tree; /* owning */

/* buildLodTree + buildLodSelectionData: octree_depth in 1..21, cluster_leaves >= 2 (LodTree.h:60-70).
 * The cloud's bounds are its positions' bounding box. Returns NULL on invalid input. */
skm_lod_tree* skm_lod_build(const skm_cloud_view* cloud, uint32_t octree_depth, uint32_t cluster_leaves,
                            uint32_t threads);
/* A tree from a given layout (nodes in layout order), with selection metadata built for it. */
skm_lod_tree* skm_lod_from_layout(const skm_cloud_view* nodes, const skm_lod_node* layout,
                                  uint64_t leaf_count);
void skm_lod_free(skm_lod_tree* tree);
/* Borrowed view of an owning tree; valid until skm_lod_free. */
void skm_lod_view(const skm_lod_tree* tree, skm_lod_tree_view* out);
/* Interior levels below the root cluster: the rounds a selection needs. */
uint32_t skm_lod_rounds(const skm_lod_tree_view* tree);

#ifdef __cplusplus
}
#endif
