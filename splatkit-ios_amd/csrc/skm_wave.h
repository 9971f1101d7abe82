// Wave64 and 256-thread block primitives shared by the kernels of the render path (gfx950).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace skm {

__device__ __forceinline__ uint32_t lane_id() { return threadIdx.x & 63u; }

// Number of set bits of `mask` below this lane.
__device__ __forceinline__ uint32_t rank_below(uint64_t mask) {
  return __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(mask >> 32),
                                   __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(mask), 0u));
}

// Inclusive sum over the wave in lane order, by data-parallel moves between lanes (no LDS round
// trips): shifts by 1, 2, 4 and 8 lanes within each row of 16, then row 0's last lane added to row 1
// and row 2's to row 3 (row_bcast:15), then lane 31 to rows 2 and 3 (row_bcast:31).
__device__ __forceinline__ uint32_t wave_inclusive_scan(uint32_t v) {
  const uint32_t lane = lane_id(), row_lane = lane & 15u;
  uint32_t t = __builtin_amdgcn_mov_dpp(v, 0x111, 0xf, 0xf, false);  // row_shr:1
  if (row_lane >= 1u) v += t;
  t = __builtin_amdgcn_mov_dpp(v, 0x112, 0xf, 0xf, false);  // row_shr:2
  if (row_lane >= 2u) v += t;
  t = __builtin_amdgcn_mov_dpp(v, 0x114, 0xf, 0xf, false);  // row_shr:4
  if (row_lane >= 4u) v += t;
  t = __builtin_amdgcn_mov_dpp(v, 0x118, 0xf, 0xf, false);  // row_shr:8
  if (row_lane >= 8u) v += t;
  t = __builtin_amdgcn_mov_dpp(v, 0x142, 0xf, 0xf, false);  // row_bcast:15
  if ((lane & 31u) >= 16u) v += t;
  t = __builtin_amdgcn_mov_dpp(v, 0x143, 0xf, 0xf, false);  // row_bcast:31
  if (lane >= 32u) v += t;
  return v;
}

// Inclusive maximum over the wave in lane order, by the same lane moves as wave_inclusive_scan.
__device__ __forceinline__ uint32_t wave_inclusive_max(uint32_t v) {
  const uint32_t lane = lane_id(), row_lane = lane & 15u;
  uint32_t t = __builtin_amdgcn_mov_dpp(v, 0x111, 0xf, 0xf, false);  // row_shr:1
  if (row_lane >= 1u) v = max(v, t);
  t = __builtin_amdgcn_mov_dpp(v, 0x112, 0xf, 0xf, false);  // row_shr:2
  if (row_lane >= 2u) v = max(v, t);
  t = __builtin_amdgcn_mov_dpp(v, 0x114, 0xf, 0xf, false);  // row_shr:4
  if (row_lane >= 4u) v = max(v, t);
  t = __builtin_amdgcn_mov_dpp(v, 0x118, 0xf, 0xf, false);  // row_shr:8
  if (row_lane >= 8u) v = max(v, t);
  t = __builtin_amdgcn_mov_dpp(v, 0x142, 0xf, 0xf, false);  // row_bcast:15
  if ((lane & 31u) >= 16u) v = max(v, t);
  t = __builtin_amdgcn_mov_dpp(v, 0x143, 0xf, 0xf, false);  // row_bcast:31
  if (lane >= 32u) v = max(v, t);
  return v;
}

// Exclusive scan of one value per thread over a block of WAVES waves; `scratch` holds WAVES words.
template <uint32_t WAVES>
__device__ __forceinline__ uint32_t block_exclusive_scan_waves(uint32_t v, uint32_t* scratch, uint32_t* total) {
  const uint32_t incl = wave_inclusive_scan(v);
  const uint32_t wave = threadIdx.x >> 6;
  if (lane_id() == 63) scratch[wave] = incl;
  __syncthreads();
  uint32_t before = 0, all = 0;
#pragma unroll
  for (uint32_t w = 0; w < WAVES; ++w) {
    const uint32_t c = scratch[w];
    before += (w < wave) ? c : 0u;
    all += c;
  }
  __syncthreads();
  *total = all;
  return before + incl - v;
}

// Exclusive scan of one value per thread over a 256-thread block; `scratch` holds 4 words.
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* scratch, uint32_t* total) {
  return block_exclusive_scan_waves<4>(v, scratch, total);
}

}  // namespace skm
