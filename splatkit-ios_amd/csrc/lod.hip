// GPU level-of-detail tree cut (SplatLOD.metal:26-220 with the round structure of
// MetalLodSelection.mm:109-192), for gfx950.
//
// The reference runs seven dispatches per round with indirect arguments (evaluate, scan, budget,
// compact, scan, commit, scatter). HIP has no indirect dispatch, and a decoupled look-back turns
// each scan into part of the kernel that produces its inputs, so a round here is two kernels:
//   k_lod_eval     per frontier cluster: drop / split / keep and the extra cost of a split, with the
//                  frontier-order prefix of the extra cost (look-back);
//   k_lod_compact  the budget (the largest prefix of refinements that fits the capacity, a search
//                  over the prefix ends, :116-132, in every block), denials, then the next frontier,
//                  leaf emission and packet offsets by look-back scans, scattered in the same kernel,
//                  and the next round's state.
// k_lod_emit copies the leaf packets (> 32 leaves) cooperatively after the last round. Grids are
// sized for the frontier capacity; blocks past the round's frontier exit at once. Per-round state
// lives in its own record (state[r] read, state[r + 1] written), so no kernel overwrites what a
// concurrent block still reads.
#include "skm_lod_internal.h"
#include "skm_math.h"
#include "skm_wave.h"

namespace skm {
namespace {

constexpr uint32_t kSplit = 0x80000000u;
constexpr uint32_t kDrop = 0x40000000u;
constexpr uint32_t kCost = 0x3fffffffu;
constexpr uint32_t kFlagAggregate = 1;
constexpr uint32_t kFlagPrefix = 2;

__device__ __forceinline__ uint32_t make_tag(uint32_t epoch, uint32_t launch) {
  return (epoch * 128u + launch) & 0x3fffffffu;
}
__device__ __forceinline__ uint64_t status_word(uint32_t tag, uint32_t flag, uint32_t value) {
  return (static_cast<uint64_t>(tag) << 34) | (static_cast<uint64_t>(flag) << 32) | value;
}
__device__ __forceinline__ void status_store(uint64_t* p, uint64_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint64_t status_load(uint64_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t sat_add(uint32_t a, uint32_t b) {
  const uint32_t s = a + b;
  return s < a ? 0xffffffffu : s;
}

__device__ __forceinline__ uint64_t wave_sum64(uint64_t v) {
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
  return v;
}

// Decoupled look-back by one wave (see kernels.hip): the exclusive prefix of block `block`.
__device__ uint32_t wave_lookback(uint64_t* status, uint32_t block, uint32_t tag) {
  const uint32_t lane = lane_id();
  uint64_t excl = 0;
  int32_t k = static_cast<int32_t>(block) - 1;
  while (k >= 0) {
    const int32_t idx = k - static_cast<int32_t>(lane);
    uint32_t flag = kFlagPrefix, value = 0;
    bool ready = true;
    if (idx >= 0) {
      const uint64_t s = status_load(status + idx);
      flag = static_cast<uint32_t>(s >> 32) & 3u;
      value = static_cast<uint32_t>(s);
      ready = (static_cast<uint32_t>(s >> 34) == tag) && flag != 0u;
    }
    if (__ballot(!ready) != 0ull) {
      __builtin_amdgcn_s_sleep(1);
      continue;
    }
    const uint64_t prefixes = __ballot(flag == kFlagPrefix);
    if (prefixes != 0ull) {
      const uint32_t first = static_cast<uint32_t>(__builtin_ctzll(prefixes));
      excl += wave_sum64(lane <= first ? value : 0u);
      break;
    }
    excl += wave_sum64(value);
    k -= 64;
  }
  return excl > 0xffffffffull ? 0xffffffffu : static_cast<uint32_t>(excl);
}

// Publishes this block's aggregate, looks back, publishes its inclusive prefix; returns the
// exclusive prefix in every lane of wave 0 (callers use lane 0's).
__device__ uint32_t scan_across_blocks(uint64_t* status, uint32_t block, uint32_t tag, uint32_t aggregate) {
  uint32_t prefix = 0;
  if (block == 0) {
    if (lane_id() == 0) status_store(status, status_word(tag, kFlagPrefix, aggregate));
  } else {
    if (lane_id() == 0) status_store(status + block, status_word(tag, kFlagAggregate, aggregate));
    prefix = wave_lookback(status, block, tag);
    if (lane_id() == 0) status_store(status + block, status_word(tag, kFlagPrefix, sat_add(prefix, aggregate)));
  }
  return prefix;
}

__device__ __forceinline__ void view_point(const Camera& cam, const LodCluster& c, float v[3]) {
  const float* m = cam.view;
#pragma unroll
  for (int r = 0; r < 3; ++r)
    v[r] = ((m[r] * c.center[0] + m[4 + r] * c.center[1]) + m[8 + r] * c.center[2]) + m[12 + r] * 1.0f;
}

// lodOutside (SplatLOD.metal:26-32).
__device__ __forceinline__ bool lod_outside(const Camera& cam, const LodCluster& c) {
  float v[3];
  view_point(cam, c, v);
  const float z = -v[2], r = c.radius;
  const float tx = cam.tan_half_fov[0], ty = cam.tan_half_fov[1];
  const float guard = (2.0f * mx(z, 0.0f)) / mx(mn(cam.focal[0], cam.focal[1]), 1.0f);
  return z + r <= 0.0f || fabsf(v[0]) - z * tx > r * sqrtf(1.0f + tx * tx) + guard ||
         fabsf(v[1]) - z * ty > r * sqrtf(1.0f + ty * ty) + guard;
}

// lodErrorPixels (SplatLOD.metal:33-43).
__device__ __forceinline__ float lod_error_pixels(const Camera& cam, const LodCluster& c, float color_weight) {
  float v[3];
  view_point(cam, c, v);
  const float depth = mx(-v[2] - c.radius, 1e-4f);
  const float perspective = sqrtf(1.0f + (v[0] * v[0] + v[1] * v[1]) / (depth * depth));
  const float appearance = mn(sqrtf(c.color_variance), 1.0f);
  const float error = c.error + c.radius * appearance;
  return ((((error * mx(cam.focal[0], cam.focal[1])) / depth) * perspective) * mx(c.opacity, 0.1f)) *
         (1.0f + color_weight * sqrtf(c.color_variance));
}

__global__ void k_lod_prologue(LodArgs a) {
  const uint32_t tid = threadIdx.x;
  if (tid == 0) {
    a.ctrl->epoch = a.ctrl->epoch + 1u;
    a.ctrl->limited = 0;
    LodState s0{};
    s0.active = a.cluster_count > 0 ? 1u : 0u;
    a.state[0] = s0;
    a.frontier[0][0] = 0;
  }
  for (uint32_t k = tid; k < 2u * kLodMaxRounds; k += blockDim.x) a.ctrl->ticket[k] = 0;
  for (uint32_t r = 1 + tid; r <= a.rounds; r += blockDim.x) a.state[r] = LodState{};
}

// One frontier cluster's decision (SplatLOD.metal:45-80): drop, split (with the extra cost of its
// refinement) or keep.
struct LodDecision {
  uint32_t flags, extra, dropped;
};
__device__ __forceinline__ LodDecision lod_decide(const LodArgs& a, uint32_t round, uint32_t t) {
  LodDecision d{0u, 0u, 0u};
  const LodCluster node = a.clusters[a.frontier[round & 1u][t]];
  const Camera& cam = *a.cam;
  if (a.cull && lod_outside(cam, node)) {
    d.flags = kDrop;
    d.dropped = 1;
  } else if (node.child_count + node.leaf_count > 1 &&
             (a.pixel_limit == 0.0f || lod_error_pixels(cam, node, a.color_weight) > a.pixel_limit)) {
    d.flags = kSplit;
    d.extra = node.child_count + node.leaf_count - 1;
  } else if (node.child_count == 1 && node.leaf_count == 0) {
    d.flags = kSplit;
  }
  return d;
}

// One frontier cluster's compaction inputs (SplatLOD.metal:133-170): its flags after the budget's
// denial, and the children, leaves and packet it contributes.
struct LodEmit {
  uint32_t flags, kids, splats, packets, denied;
  LodCluster node;
};
__device__ __forceinline__ LodEmit lod_emit_counts(const LodArgs& a, uint32_t round, uint32_t t, uint32_t accepted) {
  LodEmit e{};
  e.flags = a.cost[2 * t] & ~kCost;
  const uint32_t extra = a.cost[2 * t] & kCost;
  if ((e.flags & kSplit) && extra > 0 && a.cost[2 * t + 1] > accepted) {
    e.flags = 0;
    e.denied = 1;
  }
  if (!(e.flags & kDrop)) {
    e.node = a.clusters[a.frontier[round & 1u][t]];
    e.kids = (e.flags & kSplit) ? e.node.child_count : 0u;
    e.splats = (e.flags & kSplit) ? e.node.leaf_count : 1u;
    e.packets = e.splats > 32u ? 1u : 0u;
  }
  return e;
}

// Writes one kept or split cluster's outputs (SplatLOD.metal:171-208): its children into the next
// frontier from kid_base, its index or leaves from dest, or a packet for more than 32 leaves.
__device__ __forceinline__ void lod_write(const LodArgs& a, uint32_t round, const LodEmit& e, uint32_t kid_base,
                                          uint32_t dest, uint32_t packet) {
  const bool split = (e.flags & kSplit) != 0;
  uint32_t* next = a.frontier[(round + 1) & 1u];
  for (uint32_t k = 0; k < e.kids; ++k) next[kid_base + k] = e.node.child_start + k;
  if (e.splats > 32u) {
    a.packets[packet] = make_uint4(e.node.leaf_start, e.splats, dest, 1u);
  } else if (!split) {
    a.indices[dest] = e.node.node;
  } else {
    for (uint32_t k = 0; k < e.splats; ++k) a.indices[dest + k] = a.leaves[e.node.leaf_start + k];
  }
}

__global__ __launch_bounds__(256) void k_lod_eval(LodArgs a, uint32_t round) {
  __shared__ uint32_t s_ticket, s_prefix[2], s_scan[4];
  const uint32_t tid = threadIdx.x;
  const uint32_t active = a.state[round].active;
  // At least one block runs, even for an empty frontier: it writes the round's totals. Blocks past
  // the frontier exit before they take a ticket.
  const uint32_t blocks = max((active + 255u) / 256u, 1u);
  if (blockIdx.x >= blocks) return;
  if (tid == 0) s_ticket = atomicAdd(&a.ctrl->ticket[2 * round], 1u);
  __syncthreads();
  const uint32_t b = s_ticket;
  const uint32_t t = b * 256u + tid;
  uint32_t extra = 0, flags = 0, dropped = 0;
  if (t < active) {
    const LodDecision d = lod_decide(a, round, t);
    flags = d.flags;
    extra = d.extra;
    dropped = d.dropped;
  }
  uint32_t agg_extra, agg_drop;
  const uint32_t ex = block_exclusive_scan(extra, s_scan, &agg_extra);
  block_exclusive_scan(dropped, s_scan, &agg_drop);
  if ((tid >> 6) == 0) {
    const uint32_t tag = make_tag(a.ctrl->epoch, 2 * round);
    const uint32_t p0 = scan_across_blocks(a.status[0], b, tag, agg_extra);
    const uint32_t p1 = scan_across_blocks(a.status[1], b, tag, agg_drop);
    if (lane_id() == 0) {
      s_prefix[0] = p0;
      s_prefix[1] = p1;
    }
  }
  __syncthreads();
  if (t < active) {
    a.cost[2 * t] = flags | extra;
    a.cost[2 * t + 1] = sat_add(s_prefix[0], ex + extra);  // inclusive prefix end
  }
  if (b == blocks - 1u && tid == 0) {
    a.state[round].total_extra = sat_add(s_prefix[0], agg_extra);
    a.state[round].drops = s_prefix[1] + agg_drop;
  }
}

// budgetSplatLOD (SplatLOD.metal:116-132), run by wave 0 of every compaction block (each gets the same
// answer; a launch of its own would add one dependent kernel to each of the cut's rounds): the extra
// cost the round accepts is all of it when it fits, else the inclusive prefix end of the longest run
// of refinements, in frontier order, that fits. The prefix ends are non-decreasing, so the count of
// those that fit comes from a 64-way search (one ballot of 64 probes per step).
__device__ __forceinline__ uint32_t lod_budget(const LodArgs& a, const LodState& s) {
  const uint32_t available = a.capacity - s.count - s.active + s.drops;
  if (s.total_extra <= available) return s.total_extra;
  const uint32_t lane = lane_id();
  uint32_t lo = 0, hi = s.active;  // prefix ends below lo fit, from hi on they do not
  while (hi > lo) {
    const uint32_t step = (hi - lo + 63u) / 64u;
    const uint32_t probe = lo + lane * step;
    const uint64_t fits = __ballot(probe < hi && a.cost[2 * probe + 1] <= available);
    const uint32_t m = static_cast<uint32_t>(__popcll(fits));  // a prefix of the lanes
    if (m == 0u) break;
    const uint32_t last = lo + (m - 1u) * step;
    const uint32_t next = lo + m * step;
    lo = last + 1u;
    hi = min(hi, next);
  }
  return lo == 0u ? 0u : a.cost[2 * (lo - 1u) + 1];
}

// compactSplatLOD + commitSplatLOD + scatterSplatLOD (SplatLOD.metal:133-208).
__global__ __launch_bounds__(256) void k_lod_compact(LodArgs a, uint32_t round) {
  __shared__ uint32_t s_ticket, s_accepted, s_prefix[3], s_scan[4];
  const uint32_t tid = threadIdx.x;
  const LodState s = a.state[round];
  // At least one block runs, even for an empty frontier: it carries the state to the next round.
  const uint32_t blocks = max((s.active + 255u) / 256u, 1u);
  if (blockIdx.x >= blocks) return;
  if (tid == 0) s_ticket = atomicAdd(&a.ctrl->ticket[2 * round + 1], 1u);
  if (tid < 64u) {
    const uint32_t accepted = lod_budget(a, s);
    if (tid == 0) s_accepted = accepted;
  }
  __syncthreads();
  const uint32_t b = s_ticket;
  const uint32_t accepted = s_accepted;
  const uint32_t t = b * 256u + tid;
  LodEmit e{};
  e.flags = kDrop;
  if (t < s.active) e = lod_emit_counts(a, round, t, accepted);
  const uint32_t kids = e.kids, splats = e.splats, packets = e.packets, denied = e.denied;
  uint32_t agg_kids, agg_splats, agg_packets, agg_denied;
  const uint32_t ek = block_exclusive_scan(kids, s_scan, &agg_kids);
  const uint32_t es = block_exclusive_scan(splats, s_scan, &agg_splats);
  const uint32_t ep = block_exclusive_scan(packets, s_scan, &agg_packets);
  block_exclusive_scan(denied, s_scan, &agg_denied);
  if ((tid >> 6) == 0) {
    const uint32_t tag = make_tag(a.ctrl->epoch, 2 * round + 1);
    const uint32_t p0 = scan_across_blocks(a.status[0], b, tag, agg_kids);
    const uint32_t p1 = scan_across_blocks(a.status[1], b, tag, agg_splats);
    const uint32_t p2 = scan_across_blocks(a.status[2], b, tag, agg_packets);
    if (lane_id() == 0) {
      s_prefix[0] = p0;
      s_prefix[1] = p1;
      s_prefix[2] = p2;
      if (agg_denied) atomicAdd(&a.ctrl->limited, agg_denied);
    }
  }
  __syncthreads();
  if (t < s.active && !(e.flags & kDrop))
    lod_write(a, round, e, s_prefix[0] + ek, s.count + s_prefix[1] + es, s.packets + s_prefix[2] + ep);
  if (b == blocks - 1u && tid == 0) {
    LodState n{};
    n.count = s.count + s_prefix[1] + agg_splats;
    n.active = s_prefix[0] + agg_kids;
    n.packets = s.packets + s_prefix[2] + agg_packets;
    n.accepted = accepted;
    n.evaluated = s.evaluated + s.active;
    a.state[round + 1] = n;
  }
}

// The cut's leading rounds in one block (with the prologue's work): while the frontier can only be
// small (round r evaluates clusters at depth r of the tree, and the host passes the number of leading
// depths with at most kLodHeadFrontier clusters), a round is a few 1024-element chunks, and two
// grid-wide launches per round (each a launch's latency plus look-back scans) would cost more than the
// work. The block runs each round exactly as k_lod_eval and k_lod_compact do, chunk by chunk in
// frontier order with carried prefixes in place of the look-back (the same saturating sums), and
// writes the same state records, costs, frontier, indices and packets.
__global__ __launch_bounds__(kLodHeadThreads) void k_lod_head(LodArgs a, uint32_t head_rounds) {
  constexpr uint32_t kWaves = kLodHeadThreads / 64u;
  __shared__ uint32_t s_scan[kWaves], s_accepted;
  const uint32_t tid = threadIdx.x;
  if (tid == 0) {
    a.ctrl->epoch = a.ctrl->epoch + 1u;
    a.ctrl->limited = 0;
    a.frontier[0][0] = 0;
  }
  for (uint32_t k = tid; k < 2u * kLodMaxRounds; k += kLodHeadThreads) a.ctrl->ticket[k] = 0;
  for (uint32_t r = 1 + tid; r <= a.rounds; r += kLodHeadThreads) a.state[r] = LodState{};
  LodState s{};
  s.active = a.cluster_count > 0 ? 1u : 0u;
  __syncthreads();
  for (uint32_t round = 0; round < head_rounds; ++round) {
    // k_lod_eval over the frontier.
    uint32_t carry_extra = 0, drops = 0;
    for (uint32_t c0 = 0; c0 < s.active; c0 += kLodHeadThreads) {
      const uint32_t t = c0 + tid;
      LodDecision d{0u, 0u, 0u};
      if (t < s.active) d = lod_decide(a, round, t);
      uint32_t agg_extra, agg_drop;
      const uint32_t ex = block_exclusive_scan_waves<kWaves>(d.extra, s_scan, &agg_extra);
      block_exclusive_scan_waves<kWaves>(d.dropped, s_scan, &agg_drop);
      if (t < s.active) {
        a.cost[2 * t] = d.flags | d.extra;
        a.cost[2 * t + 1] = sat_add(carry_extra, ex + d.extra);  // inclusive prefix end
      }
      carry_extra = sat_add(carry_extra, agg_extra);
      drops += agg_drop;
    }
    s.total_extra = carry_extra;
    s.drops = drops;
    if (tid == 0) a.state[round] = s;
    __syncthreads();  // the costs, written by every wave, are read by any
    if (tid < 64u) {
      const uint32_t accepted = lod_budget(a, s);
      if (tid == 0) s_accepted = accepted;
    }
    __syncthreads();
    const uint32_t accepted = s_accepted;
    // k_lod_compact over the frontier.
    uint32_t carry_kids = 0, carry_splats = 0, carry_packets = 0, denied = 0;
    for (uint32_t c0 = 0; c0 < s.active; c0 += kLodHeadThreads) {
      const uint32_t t = c0 + tid;
      LodEmit e{};
      e.flags = kDrop;
      if (t < s.active) e = lod_emit_counts(a, round, t, accepted);
      uint32_t agg_kids, agg_splats, agg_packets, agg_denied;
      const uint32_t ek = block_exclusive_scan_waves<kWaves>(e.kids, s_scan, &agg_kids);
      const uint32_t es = block_exclusive_scan_waves<kWaves>(e.splats, s_scan, &agg_splats);
      const uint32_t ep = block_exclusive_scan_waves<kWaves>(e.packets, s_scan, &agg_packets);
      block_exclusive_scan_waves<kWaves>(e.denied, s_scan, &agg_denied);
      if (t < s.active && !(e.flags & kDrop))
        lod_write(a, round, e, carry_kids + ek, s.count + carry_splats + es, s.packets + carry_packets + ep);
      carry_kids += agg_kids;
      carry_splats += agg_splats;
      carry_packets += agg_packets;
      denied += agg_denied;
    }
    LodState n{};
    n.count = s.count + carry_splats;
    n.active = carry_kids;
    n.packets = s.packets + carry_packets;
    n.accepted = accepted;
    n.evaluated = s.evaluated + s.active;
    if (tid == 0) {
      if (denied) atomicAdd(&a.ctrl->limited, denied);
      a.state[round + 1] = n;
    }
    s = n;
    __syncthreads();  // the next frontier, written by every wave, is read by any
  }
}

// emitSplatLOD (SplatLOD.metal:210-220): one packet per block iteration.
// One wave per packet (a packet holds 33 to a cluster's worth of leaves, 256 at most in the trees the
// builder makes), four lanes' worth of leaf references loaded before any index is stored, and the
// wave's next packet descriptor loaded while it copies the current one, so each wave waits for about
// one round trip per packet and thousands of packets are in flight at once.
__global__ __launch_bounds__(256) void k_lod_emit(LodArgs a) {
  const LodState s = a.state[a.rounds];
  const uint32_t waves = gridDim.x * (blockDim.x >> 6), lane = lane_id();
  uint32_t p = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  uint4 next = p < s.packets ? a.packets[p] : make_uint4(0u, 0u, 0u, 0u);
  for (; p < s.packets; p += waves) {
    const uint4 packet = next;
    if (p + waves < s.packets) next = a.packets[p + waves];
    for (uint32_t k0 = 0; k0 < packet.y; k0 += 256u) {
      uint32_t v[4];
#pragma unroll
      for (uint32_t j = 0; j < 4; ++j) {
        const uint32_t k = k0 + j * 64u + lane;
        v[j] = k < packet.y ? (packet.w ? a.leaves[packet.x + k] : packet.x) : 0u;
      }
#pragma unroll
      for (uint32_t j = 0; j < 4; ++j) {
        const uint32_t k = k0 + j * 64u + lane;
        if (k < packet.y) a.indices[packet.z + k] = v[j];
      }
    }
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) a.ctrl->final_limited = a.ctrl->limited;
}

}  // namespace

hipError_t launch_lod(const LodArgs& a, hipStream_t s) {
  const uint32_t head = min(a.head_rounds, a.rounds);
  if (head > 0)
    hipLaunchKernelGGL(k_lod_head, dim3(1), dim3(kLodHeadThreads), 0, s, a, head);
  else
    hipLaunchKernelGGL(k_lod_prologue, dim3(1), dim3(256), 0, s, a);
  const uint32_t blocks = (a.frontier_capacity + 255u) / 256u;
  for (uint32_t r = head; r < a.rounds; ++r) {
    hipLaunchKernelGGL(k_lod_eval, dim3(blocks ? blocks : 1u), dim3(256), 0, s, a, r);
    hipLaunchKernelGGL(k_lod_compact, dim3(blocks ? blocks : 1u), dim3(256), 0, s, a, r);
  }
  hipLaunchKernelGGL(k_lod_emit, dim3(1024), dim3(256), 0, s, a);
  return hipGetLastError();
}

}  // namespace skm
