// Layouts of the GPU tree cut (lod.hip), shared with the host renderer.
#pragma once
#include "skm_internal.h"

namespace skm {

// LodCluster (packages/splat-core/include/splat/lod/LodTree.h:28-35), 64 bytes.
struct LodCluster {
  float center[3];
  float radius;
  float extent[3];
  float error;
  float color_variance;
  float opacity;
  uint32_t node, child_start, child_count, leaf_start, leaf_count, subtree_leaves;
};
static_assert(sizeof(LodCluster) == 64, "cluster layout");

// The selection state of one round (LodState, SplatLOD.metal:12-20, without the indirect
// dispatch words HIP does not use).
struct LodState {
  uint32_t count;        // indices emitted before this round (including pending packets)
  uint32_t active;       // frontier clusters this round evaluates
  uint32_t packets;      // leaf packets before this round
  uint32_t accepted;     // extra cost the budget accepted this round
  uint32_t evaluated;    // clusters evaluated before this round
  uint32_t total_extra;  // this round's requested extra cost
  uint32_t drops;        // this round's culled clusters
  uint32_t reserved;
};

constexpr uint32_t kLodMaxRounds = 64;
// The single-block head of the cut (k_lod_head) takes the leading rounds whose frontier cannot exceed
// this many clusters.
constexpr uint32_t kLodHeadFrontier = 4096;
constexpr uint32_t kLodHeadThreads = 1024;

struct LodCtrl {
  uint32_t epoch;
  uint32_t limited;        // denied refinements, summed over the rounds
  uint32_t final_limited;  // copy of `limited` after the last round
  uint32_t reserved;
  uint32_t ticket[2 * kLodMaxRounds];
};

struct LodArgs {
  const Camera* cam;
  const LodCluster* clusters;
  const uint32_t* leaves;
  uint32_t cluster_count;
  uint32_t rounds;
  uint32_t head_rounds;        // leading rounds the single-block head runs (0: none)
  uint32_t capacity;           // most indices one cut emits
  uint32_t frontier_capacity;  // min(capacity, clusters)
  float pixel_limit;
  float color_weight;
  int cull;
  LodCtrl* ctrl;
  LodState* state;  // rounds + 1 records
  uint32_t* frontier[2];
  uint32_t* cost;   // 2 words per frontier slot: flags | extra, inclusive prefix end
  uint4* packets;
  uint32_t* indices;
  uint64_t* status[3];
};

hipError_t launch_lod(const LodArgs& a, hipStream_t stream);

}  // namespace skm
