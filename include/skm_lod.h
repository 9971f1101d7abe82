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
/* The cloud's bounds the tree was built over: min xyz, then max xyz. */
void skm_lod_bounds(const skm_lod_tree* tree, float* bounds6);

/* ---- the LODSPLAT file (packages/splat-core/include/splat/lod/LodFile.h:12-16) ----
 * 64-byte little-endian header, one record per node (LodNode, covariance[6], RGB[3], alpha, SH), then in
 * version 2 the LodCluster records and the leaf references. Same format and checks as the reference's
 * decodeLodSplat / validateLodTree / writeLodSplat (LodFile.cpp:336-460), except the node limit: the
 * reference refuses more than 20,000,000 nodes (LodFile.cpp:24); here up to 2^31 are read. */
typedef struct skm_lod_file_header {
  uint32_t version;        /* 1, or 2 with selection data */
  uint32_t node_count;
  uint32_t leaf_count;
  int32_t sh_degree;
  uint32_t depth;          /* refinement rounds, checked against the topology */
  uint32_t cluster_count;
  uint32_t leaf_ref_count;
  uint32_t reserved;
  float bounds_min[3];
  float bounds_max[3];
} skm_lod_file_header;

/* readLodSplatHeader (LodFile.cpp:355-380): the header, checked against `size`, the file's length. */
skm_status skm_lod_read_header(const uint8_t* data, size_t size, skm_lod_file_header* out, char* error,
                               size_t error_size);
/* decodeLodSplat (LodFile.cpp:382-427): the tree, SH kept up to max_sh_degree, validated; NULL with
 * the reason in `error` when the bytes are not a valid tree. */
skm_lod_tree* skm_lod_decode(const uint8_t* data, size_t size, int max_sh_degree, char* error,
                             size_t error_size);
/* The same over a file, memory-mapped. */
skm_lod_tree* skm_lod_read(const char* path, int max_sh_degree, char* error, size_t error_size);
/* validateLodTree (LodFile.cpp:336-344): SKM_OK and the refinement rounds, or the first failure. */
skm_status skm_lod_validate(const skm_lod_tree* tree, uint32_t* depth, char* error, size_t error_size);
/* writeLodSplat (LodFile.cpp:430-460): validates, then writes a new file (never replaces one). */
skm_status skm_lod_write(const skm_lod_tree* tree, const char* path, char* error, size_t error_size);

#ifdef __cplusplus
}
#endif
