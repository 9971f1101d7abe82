// HIP kernels of the per-frame splat render path, written for gfx950 (wave64, 160 KB LDS per CU).
//
//   k_pack      SplatCloud floats -> resident SoA records (GpuLayout.cpp:92-137)
//   k_prologue  per-frame control block: epoch, counters, histograms, tile ranges
//   k_cull      visibility, projection, SH colour, depth key (SplatVisibility.metal:39-106,
//               SplatProjection.metalh:78-141), each cull tile's survivors in lane order from
//               its own base (a gapped layout the depth sort's first pass compacts)
//   k_hist      onesweep global digit histograms of a key array, all passes in one read
//   k_onesweep  one LSD radix pass, stable (MetalRadixSort.h:13-19): wave64 ballot ranking,
//               LDS-staged digit histograms and scatter, per-digit decoupled look-back
//   k_bin       tile rectangles of the depth-ordered survivors, one scan + emission pass
//   k_ranges    [start, end) per tile from the tile-sorted pairs
//   k_raster    16x16-tile front-to-back composite, LDS-staged batches, early termination
#include <algorithm>

#include "skm_internal.h"
#include "skm_lod_internal.h"
#include "skm_math.h"
#include "skm_wave.h"

namespace skm {
namespace {

// ---------------------------------------------------------------------------------------------
// Look-back status words: tag (30 bits) | flag (2 bits) | value (32 bits). The tag is derived from
// the frame epoch and the pass, so words left by an earlier frame or pass never read as ready.
constexpr uint32_t kFlagAggregate = 1;
constexpr uint32_t kFlagPrefix = 2;

__device__ __forceinline__ uint32_t make_tag(uint32_t epoch, uint32_t pass) {
  return (epoch * 16u + pass) & 0x3fffffffu;
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


// Workgroups reach the 8 XCDs of an MI355X round-robin in launch order (block i on XCD i % 8), and
// each XCD has its own L2. Block i of g takes logical block xcd_tile(i, g): XCD x works through one
// contiguous range of logical blocks in order. A radix pass's neighbouring tiles write neighbouring
// memory (the digit runs of consecutive tiles abut), so their partial lines meet in one L2 instead
// of eight; a batch's cull runs the views of one lane tile as consecutive logical blocks, so the
// tile's splat records come from HBM once and from that XCD's L2 for the other views.
constexpr uint32_t kXcds = 8;
__device__ __forceinline__ uint32_t xcd_tile(uint32_t i, uint32_t g) {
  const uint32_t x = i % kXcds, q = g / kXcds, r = g % kXcds;
  return x * q + min(x, r) + i / kXcds;
}
// Lanes of `live` whose BITS-bit digit equals this lane's d. Bit slice b: the lanes whose bit b is
// set (one ballot); a lane differs from lane L in bit b exactly where (ballot XOR its own bit
// spread over 32 bits) is set, so the lanes that differ in any bit are the OR of those words and the
// peers are the rest. Each slice is a sign-extending bit extract, the ballot's compare and one
// three-input logic op per half-word.
template <int BITS>
__device__ __forceinline__ uint64_t digit_peers(uint32_t d, uint64_t live) {
  uint32_t diff_lo = 0u, diff_hi = 0u;
#pragma unroll
  for (int bit = 0; bit < BITS; ++bit) {
    uint32_t spread = static_cast<uint32_t>(static_cast<int32_t>(d << (31 - bit)) >> 31);
    asm volatile("" : "+v"(spread));  // one extract serves the ballot and the XOR
    const uint64_t vote = __ballot(spread != 0u);
    diff_lo |= static_cast<uint32_t>(vote) ^ spread;
    diff_hi |= static_cast<uint32_t>(vote >> 32) ^ spread;
  }
  return live & ~((static_cast<uint64_t>(diff_hi) << 32) | diff_lo);
}

__device__ __forceinline__ void mat_vec(const float* m, float x, float y, float z, float w, float out[4]) {
#pragma unroll
  for (int r = 0; r < 4; ++r) out[r] = ((m[r] * x + m[4 + r] * y) + m[8 + r] * z) + m[12 + r] * w;
}

// ---------------------------------------------------------------------------------------------
// Projection of one splat (SplatProjection.metalh:78-141), the same specification as the CPU
// oracle, written for registers: SH words arrive preloaded.
template <int DEG>
__device__ __forceinline__ float sh_channel(const uint32_t* words, int ch, float x, float y, float z) {
  auto coef = [&](uint32_t k) {
    const uint32_t h = k * 3u + static_cast<uint32_t>(ch);
    const uint32_t w = words[h / 2u];
    return (h & 1u) == 0u ? lo_half(w) : hi_half(w);
  };
  const float C1 = 0.4886025119f;
  float c = ((-C1 * y) * coef(0) + (C1 * z) * coef(1)) - (C1 * x) * coef(2);
  if constexpr (DEG >= 2) {
    const float C2_0 = 1.0925484306f, C2_1 = -1.0925484306f, C2_2 = 0.3153915653f,
                C2_3 = -1.0925484306f, C2_4 = 0.5462742153f;
    const float xx = x * x, yy = y * y, zz = z * z, xy = x * y, yz = y * z, xz = x * z;
    const float t0 = (C2_0 * xy) * coef(3);
    const float t1 = (C2_1 * yz) * coef(4);
    const float t2 = (C2_2 * ((2.0f * zz - xx) - yy)) * coef(5);
    const float t3 = (C2_3 * xz) * coef(6);
    const float t4 = (C2_4 * (xx - yy)) * coef(7);
    c = c + ((((t0 + t1) + t2) + t3) + t4);
    if constexpr (DEG >= 3) {
      const float C3_0 = -0.5900435899f, C3_1 = 2.8906114426f, C3_2 = -0.4570457995f,
                  C3_3 = 0.3731763326f, C3_4 = -0.4570457995f, C3_5 = 1.4453057213f,
                  C3_6 = -0.5900435899f;
      const float u0 = ((C3_0 * y) * (3.0f * xx - yy)) * coef(8);
      const float u1 = ((C3_1 * xy) * z) * coef(9);
      const float u2 = ((C3_2 * y) * ((4.0f * zz - xx) - yy)) * coef(10);
      const float u3 = ((C3_3 * z) * ((2.0f * zz - 3.0f * xx) - 3.0f * yy)) * coef(11);
      const float u4 = ((C3_4 * x) * ((4.0f * zz - xx) - yy)) * coef(12);
      const float u5 = ((C3_5 * z) * (xx - yy)) * coef(13);
      const float u6 = ((C3_6 * x) * (xx - 3.0f * yy)) * coef(14);
      c = c + ((((((u0 + u1) + u2) + u3) + u4) + u5) + u6);
    }
  }
  return c;
}

// revealGlow (SplatProjection.metalh:177-189).
__device__ __forceinline__ void reveal_glow(const Camera& cam, float height, uint32_t rec[8]) {
  if (!(cam.reveal[1] > 0.0f)) return;
  float glow = mn(mx(1.0f - (cam.reveal[0] - height) / cam.reveal[1], 0.0f), 1.0f);
  glow = glow * glow;
  if (!(glow > 0.0f)) return;
  float rgb[3] = {lo_half(rec[4]), hi_half(rec[4]), lo_half(rec[5])};
  const float a = hi_half(rec[5]);
  const float t = cam.reveal[2] * glow;
  const float boost = 1.0f + cam.reveal[3] * glow;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    float v = rgb[c];
    if (cam.output_linear == 1u) v = pow_f(v, 1.0f / kDisplayGamma);
    v = (v + (cam.reveal_tint[c] - v) * t) * boost;
    if (cam.output_linear == 1u) v = pow_f(v, kDisplayGamma);
    rgb[c] = v;
  }
  rec[4] = pack_rne(rgb[0], rgb[1]);
  rec[5] = pack_rne(rgb[2], a);
}

template <int DEG, bool KEY16>
__device__ __forceinline__ bool cull_lane(const Camera& cam, const World& w, uint32_t index,
                                          uint32_t& key, uint32_t rec[8], const float* byte255) {
  const float px = w.px[index], py = w.py[index], pz = w.pz[index];
  const uint32_t rgba = w.rgba8[index];
  const uint32_t lodw = w.lod_alpha[index];
  float v[4];
  mat_vec(cam.view, px, py, pz, 1.0f, v);
  const float depth = -v[2];
  if constexpr (KEY16) {
    // quantizedDepthKey (SplatVisibility.metal:21-28).
    const float zNear = cam.proj[14] / cam.proj[10];
    const uint32_t steps = bits(mx(depth, zNear)) - bits(zNear);
    key = min(steps >> 11, 65535u);
  } else {
    key = bits(depth);
  }
  if (!(cam.reveal[0] - py >= 0.0f)) return false;
  const bool lod = lodw != 0u;
  const float ca = byte255[rgba >> 24];
  const float alpha = lod ? f32(lodw) : ca;
  if (alpha < kQuadMinAlpha) return false;
  if (v[2] >= 0.0f) return false;
  float clip[4];
  mat_vec(cam.proj, v[0], v[1], v[2], v[3], clip);
  if (clip[2] < 0.0f || clip[2] > clip[3]) return false;
  if (clip[3] <= cam.proj[14] / cam.proj[10]) return false;

  const uint32_t c0 = w.cov0[index], c1 = w.cov1[index], c2 = w.cov2[index];
  const float xx = lo_half(c0), xy = hi_half(c0), xz = lo_half(c1);
  const float yy = hi_half(c1), yz = lo_half(c2), zz = hi_half(c2);
  const float invZ = 1.0f / v[2];
  const float invZ2 = invZ * invZ;
  const float limx = kProjectionFovClamp * cam.tan_half_fov[0];
  const float limy = kProjectionFovClamp * cam.tan_half_fov[1];
  const float tx = mn(mx(v[0] * invZ, -limx), limx) * v[2];
  const float ty = mn(mx(v[1] * invZ, -limy), limy) * v[2];
  const float j00 = cam.focal[0] * invZ;
  const float j11 = cam.focal[1] * invZ;
  const float j02 = ((-cam.focal[0]) * tx) * invZ2;
  const float j12 = ((-cam.focal[1]) * ty) * invZ2;
  const float* V = cam.view;
  float T0[3], T1[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    T0[c] = j00 * V[c * 4 + 0] + j02 * V[c * 4 + 2];
    T1[c] = j11 * V[c * 4 + 1] + j12 * V[c * 4 + 2];
  }
  const float S[3][3] = {{xx, xy, xz}, {xy, yy, yz}, {xz, yz, zz}};
  float M0[3], M1[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    M0[c] = (T0[0] * S[0][c] + T0[1] * S[1][c]) + T0[2] * S[2][c];
    M1[c] = (T1[0] * S[0][c] + T1[1] * S[1][c]) + T1[2] * S[2][c];
  }
  const float c00 = (M0[0] * T0[0] + M0[1] * T0[1]) + M0[2] * T0[2];
  const float c10 = (M1[0] * T0[0] + M1[1] * T0[1]) + M1[2] * T0[2];
  const float c11 = (M1[0] * T1[0] + M1[1] * T1[1]) + M1[2] * T1[2];
  const float a = c00 + kLowPassVariance;
  const float b = c10;
  const float d = c11 + kLowPassVariance;

  const float det = a * d - b * b;
  const float mean = 0.5f * (a + d);
  const float dist = mx(0.1f, sqrtf(mx(mean * mean - det, 0.0f)));
  const float l1 = mean + dist;
  const float l2 = mean - dist;
  float e1x, e1y;
  if (b == 0.0f) {
    e1x = a > d ? 1.0f : 0.0f;
    e1y = a > d ? 0.0f : 1.0f;
  } else {
    const float nx = b, ny = d - l2;
    const float inv = 1.0f / sqrtf(nx * nx + ny * ny);
    e1x = nx * inv;
    e1y = ny * inv;
  }
  const float e2x = e1y, e2y = -e1x;
  const float s1 = sqrtf(l1);
  const float s2 = sqrtf(mx(l2, 0.0f));
  const float a1x = e1x * s1, a1y = e1y * s1;
  const float a2x = e2x * s2, a2y = e2y * s2;

  const float ratio = mx(alpha / kQuadMinAlpha, 1.0f);
  const float lg = log_f(ratio);
  float radius = mn(lod ? kLodBoundsRadius : kBoundsRadius, sqrtf(2.0f * lg));
  const float sigma = sqrtf(mx(a1x * a1x + a1y * a1y, a2x * a2x + a2y * a2y));
  radius = mn(radius, kMaxPixelRadius / mx(sigma, 1e-6f));

  const float extx = radius * (fabsf(round_half(a1x)) + fabsf(round_half(a2x)));
  const float exty = radius * (fabsf(round_half(a1y)) + fabsf(round_half(a2y)));
  const float marg_x = (((extx + 0.5f) * 2.0f) * clip[3]) / cam.screen_size[0];
  const float marg_y = (((exty + 0.5f) * 2.0f) * clip[3]) / cam.screen_size[1];
  if (fabsf(clip[0]) > clip[3] + marg_x || fabsf(clip[1]) > clip[3] + marg_y) return false;

  const float hd = 0.5f * (a - d);
  const float sv = (0.5f * (a + d) + sqrtf(mx(hd * hd + b * b, 0.0f))) - kLowPassVariance;
  if (!(sqrtf(mx(sv, 0.0f)) * radius >= cam.min_pixel_radius)) return false;

  float rgb[3] = {byte255[rgba & 0xffu], byte255[(rgba >> 8) & 0xffu], byte255[(rgba >> 16) & 0xffu]};
  if constexpr (DEG >= 1) {
    constexpr uint32_t kStride = ((DEG + 1) * (DEG + 1) - 1) * 3 / 2 + (((DEG + 1) * (DEG + 1) - 1) * 3) % 2;
    uint32_t words[kStride];
    // Non-temporal: a frame reads each survivor's SH words once, and the world is larger than the
    // last-level cache, so they should not displace the frame's sort and raster buffers from it.
#pragma unroll
    for (uint32_t k = 0; k < kStride; ++k)
      words[k] = __builtin_nontemporal_load(w.sh + static_cast<size_t>(k) * w.capacity + index);
    const float dx = px - cam.camera_position[0];
    const float dy = py - cam.camera_position[1];
    const float dz = pz - cam.camera_position[2];
    const float inv = 1.0f / sqrtf((dx * dx + dy * dy) + dz * dz);
    const float nx = dx * inv, ny = dy * inv, nz = dz * inv;
#pragma unroll
    for (int c = 0; c < 3; ++c) rgb[c] = mx(rgb[c] + sh_channel<DEG>(words, c, nx, ny, nz), 0.0f);
  }
  if (cam.output_linear == 1u) {
#pragma unroll
    for (int c = 0; c < 3; ++c) rgb[c] = pow_f(rgb[c], kDisplayGamma);
  }
  rec[0] = bits(clip[0] / clip[3]);
  rec[1] = bits(clip[1] / clip[3]);
  rec[2] = pack_rne(a1x, a1y);
  rec[3] = pack_rne(a2x, a2y);
  rec[4] = pack_rne(rgb[0], rgb[1]);
  rec[5] = pack_rne(rgb[2], alpha);
  rec[6] = bits(radius);
  rec[7] = index;
  reveal_glow(cam, py, rec);
  return isfinite(depth);
}

// findRange (SplatVisibility.metal:30-37).
__device__ __forceinline__ uint32_t find_range(const uint32_t* starts, uint32_t count, uint32_t t) {
  uint32_t lo = 0, hi = count;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) / 2;
    if (starts[mid] <= t) lo = mid; else hi = mid;
  }
  return lo;
}

// ---------------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_pack(WorldOut w, uint32_t offset, uint32_t count,
                                              const float* __restrict__ pos,
                                              const float* __restrict__ cov,
                                              const float* __restrict__ col,
                                              const float* __restrict__ alpha,
                                              const float* __restrict__ sh, int src_degree) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const uint32_t o = offset + i;
  w.px[o] = pos[3 * static_cast<size_t>(i)];
  w.py[o] = pos[3 * static_cast<size_t>(i) + 1];
  w.pz[o] = pos[3 * static_cast<size_t>(i) + 2];
  // packRgba8 (GpuLayout.cpp:22-27): lround(clamp(v, 0, 1) * 255) per channel.
  auto q = [](float v) {
    const float c = v < 0.0f ? 0.0f : (1.0f < v ? 1.0f : v);
    return static_cast<uint32_t>(roundf(c * 255.0f));
  };
  const float a = alpha[i];
  const float* c = col + 3 * static_cast<size_t>(i);
  w.rgba8[o] = q(c[0]) | (q(c[1]) << 8) | (q(c[2]) << 16) | (q(a) << 24);
  const float* v = cov + 6 * static_cast<size_t>(i);
  w.cov0[o] = half_sat(v[0]) | (half_sat(v[1]) << 16);
  w.cov1[o] = half_sat(v[2]) | (half_sat(v[3]) << 16);
  w.cov2[o] = half_sat(v[4]) | (half_sat(v[5]) << 16);
  w.lod_alpha[o] = a > 1.0f ? bits(a) : 0u;
  if (w.sh_degree > 0) {
    const uint32_t halves = static_cast<uint32_t>((w.sh_degree + 1) * (w.sh_degree + 1) - 1) * 3u;
    const uint32_t stride = (halves + 1u) / 2u;
    const size_t src = static_cast<size_t>((src_degree + 1) * (src_degree + 1) - 1) * 3u;
    const float* s = sh + src * i;
    for (uint32_t k = 0; k < stride; ++k) {
      const uint32_t h = 2u * k;
      const uint32_t lo = (h < halves && src_degree > 0) ? half_sat(s[h]) : 0u;
      const uint32_t hi = (h + 1u < halves && src_degree > 0) ? half_sat(s[h + 1]) : 0u;
      w.sh[static_cast<size_t>(k) * w.capacity + o] = lo | (hi << 16);
    }
  }
}

__global__ __launch_bounds__(256) void k_prologue(Ctrl* ctrl, uint32_t* hist, uint32_t hist_words,
                                                  uint32_t* ranges, uint32_t range_words,
                                                  uint32_t* extra, uint32_t extra_words,
                                                  const LodState* lod_final, const uint32_t* lod_limited) {
  const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
  if (gid == 0) {
    ctrl->epoch = ctrl->epoch + 1u;
    ctrl->visible = 0;
    ctrl->pairs = 0;
    ctrl->overflow = 0;
    ctrl->sort_pairs = 0;
    ctrl->cull_lanes = 0;
    for (int k = 0; k < 8; ++k) ctrl->raster_tests[k] = 0;
    for (int k = 0; k < 16; ++k) ctrl->ticket[k] = 0;
    // The frame's LOD cut, kept with the frame's counters: the selection buffers are free for the
    // next frame's cut once this frame's cull has read the indices.
    static_assert(sizeof(LodState) == 8 * sizeof(uint32_t), "LodState fits ctrl->lod[0..7]");
    const uint32_t* st = reinterpret_cast<const uint32_t*>(lod_final);
    for (int k = 0; k < 8; ++k) ctrl->lod[k] = st != nullptr ? st[k] : 0u;
    ctrl->lod[8] = lod_limited != nullptr ? *lod_limited : 0u;
    for (int k = 9; k < 16; ++k) ctrl->lod[k] = 0;
  }
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t i = gid; i < hist_words; i += stride) hist[i] = 0;
  for (uint32_t i = gid; i < range_words; i += stride) ranges[i] = 0;
  for (uint32_t i = gid; i < extra_words; i += stride) extra[i] = 0;
}

// log2 of a record's half peak opacity, clamped like splatDepth (SplatProjection.metalh:169-171).
__device__ __forceinline__ float log2_peak(uint32_t w5) {
  const float a0 = hi_half(w5);
  const float l2 = log_f(mx(a0, kHalfMin)) * kInvLn2F;
  return mn(mx(l2, -kMaxLog2Alpha), kMaxLog2Alpha);
}

// Footprint half-extents in pixels (the oracle's footprint, same specification): the smaller of the
// quad's box r(|a1| + |a2|) and the cutoff disc's box sqrt(L / K^2) sqrt(a1^2 + a2^2) per axis, the
// latter with slack. False when no pixel can pass the cutoff.
__device__ __forceinline__ bool footprint(const uint32_t* p, float log2a0, float& ex, float& ey) {
  const float cut = log2a0 - kLog2MinAlpha;
  if (!(cut >= 0.0f)) return false;
  const float rho = sqrtf(cut * kInvGaussianScale2) * kFootprintSlack;
  const float a1x = lo_half(p[2]), a1y = hi_half(p[2]);
  const float a2x = lo_half(p[3]), a2y = hi_half(p[3]);
  const float r = f32(p[6]);
  const float qx = r * (fabsf(a1x) + fabsf(a2x));
  const float qy = r * (fabsf(a1y) + fabsf(a2y));
  const float sx = rho * sqrtf(a1x * a1x + a2x * a2x);
  const float sy = rho * sqrtf(a1y * a1y + a2y * a2y);
  ex = mn(qx, sx);
  ey = mn(qy, sy);
  return true;
}

// Tile rectangle of a projected record (the CPU oracle's orc_tile_rect, same specification).
struct Rect {
  int32_t tx0, ty0, tx1, ty1;
};
__device__ __forceinline__ uint32_t tile_rect(const uint32_t* p, uint32_t width, uint32_t height,
                                              uint32_t tile, Rect& r) {
  const float cx = (f32(p[0]) * 0.5f + 0.5f) * static_cast<float>(width);
  const float cy = (0.5f - f32(p[1]) * 0.5f) * static_cast<float>(height);
  r.tx0 = r.ty0 = 0;
  r.tx1 = r.ty1 = -1;
  float ex = 0.0f, ey = 0.0f;
  if (!footprint(p, log2_peak(p[5]), ex, ey)) return 0;
  float x0 = ((cx - ex) - 0.5f) - kBinPad, x1 = ((cx + ex) - 0.5f) + kBinPad;
  float y0 = ((cy - ey) - 0.5f) - kBinPad, y1 = ((cy + ey) - 0.5f) + kBinPad;
  if (!(x0 <= x1) || !(y0 <= y1)) return 0;
  x0 = mx(x0, -1.0f);
  x1 = mn(x1, static_cast<float>(width));
  y0 = mx(y0, -1.0f);
  y1 = mn(y1, static_cast<float>(height));
  int32_t px0 = static_cast<int32_t>(ceilf(x0)), px1 = static_cast<int32_t>(floorf(x1));
  int32_t py0 = static_cast<int32_t>(ceilf(y0)), py1 = static_cast<int32_t>(floorf(y1));
  px0 = max(px0, 0);
  py0 = max(py0, 0);
  px1 = min(px1, static_cast<int32_t>(width) - 1);
  py1 = min(py1, static_cast<int32_t>(height) - 1);
  if (px0 > px1 || py0 > py1) return 0;
  // 0 <= px0 <= px1 and 0 <= py0 <= py1 here and the tile is a power of two (8, 16 or 32), so the
  // quotient by the tile is a shift by its (wave-uniform) logarithm.
  const uint32_t shift = static_cast<uint32_t>(__builtin_ctz(tile));
  r.tx0 = px0 >> shift;
  r.ty0 = py0 >> shift;
  r.tx1 = px1 >> shift;
  r.ty1 = py1 >> shift;
  return static_cast<uint32_t>((r.tx1 - r.tx0 + 1) * (r.ty1 - r.ty0 + 1));
}

// A survivor's tile rectangle packed for the binning: tx0 | ty0 << 12 | tx1 << 24 | ty1 << 36 |
// view << 48 (12 bits per tile coordinate: up to 4096 tiles per axis), all ones when it covers none.
constexpr uint64_t kNoRect = ~0ull;
__device__ __forceinline__ uint64_t pack_rect(const Rect& r, uint32_t count, uint32_t view) {
  if (count == 0u) return kNoRect;
  return static_cast<uint64_t>(r.tx0) | (static_cast<uint64_t>(r.ty0) << 12) |
         (static_cast<uint64_t>(r.tx1) << 24) | (static_cast<uint64_t>(r.ty1) << 36) |
         (static_cast<uint64_t>(view) << 48);
}

// 256 threads, kCullItems lanes each: lane t = tile * kCullTile + k * 256 + thread. Logical block b
// (xcd_tile of the launch index) culls lane tile b / views for view b % views; its
// survivors are ranked in lane order (k, wave, lane) and written from position b * kCullTile, with
// their count in cull_counts[b]: no block waits for another, and ordering by position is lane order
// (compaction order). Block 0 publishes the layout's extent, every block adds its count to
// ctrl->visible.
template <int DEG, bool INDEXED, bool KEY16>
__global__ __launch_bounds__(256) void k_cull(FrameArgs a) {
  constexpr uint32_t kItems = kCullItems;
  constexpr uint32_t kTile = 256u * kItems;
  __shared__ uint32_t s_seg[kItems * 4];
  __shared__ float s_byte255[256];  // b / 255 for every byte: the divisions of packRgba8's inverse
  const uint32_t tid = threadIdx.x, wave = tid >> 6;
  const uint32_t total = INDEXED ? min(*a.active_count, a.capacity) : a.total;
  const uint32_t view_tiles = (total + kTile - 1u) / kTile;
  const uint32_t blocks = view_tiles * a.view_count;
  if (blockIdx.x >= blocks) return;
  s_byte255[tid] = static_cast<float>(tid) / 255.0f;
  __syncthreads();
  const uint32_t b = xcd_tile(blockIdx.x, blocks);
  const uint32_t view = a.view_count > 1u ? b % a.view_count : 0u;
  const uint32_t lt = a.view_count > 1u ? b / a.view_count : b;  // the lane tile
  const Camera& cam = a.cam[view];
  uint32_t key[kItems], rec[kItems][8];
  uint64_t ballot[kItems];
#pragma unroll
  for (uint32_t k = 0; k < kItems; ++k) {
    const uint32_t t = lt * kTile + k * 256u + tid;
    bool visible = false;
    key[k] = 0;
#pragma unroll
    for (int w = 0; w < 8; ++w) rec[k][w] = 0;
    if (t < total) {
      uint32_t index;
      if constexpr (INDEXED) {
        index = a.active[t];
      } else {
        const uint32_t r = find_range(a.range_starts, a.range_count, t);
        index = a.ranges[2 * r] + (t - a.range_starts[r]);
      }
      visible = cull_lane<DEG, KEY16>(cam, a.world, index, key[k], rec[k], s_byte255);
    }
    ballot[k] = __ballot(visible);
    if (lane_id() == 0) s_seg[k * 4 + wave] = static_cast<uint32_t>(__popcll(ballot[k]));
  }
  __syncthreads();
  uint32_t aggregate = 0;
#pragma unroll
  for (uint32_t g = 0; g < kItems * 4; ++g) aggregate += s_seg[g];
  if (tid == 0) {
    a.cull_counts[b] = aggregate;
    if (b == 0) a.ctrl->cull_lanes = blocks * kTile;
    if (aggregate != 0u) {
      atomicAdd(&a.ctrl->visible, aggregate);
      if (a.view_visible != nullptr) atomicAdd(&a.view_visible[view], aggregate);
    }
  }
  uint32_t before = b * kTile;
#pragma unroll
  for (uint32_t k = 0; k < kItems; ++k) {
    uint32_t seg = before;
#pragma unroll
    for (uint32_t w = 0; w < 4; ++w) seg += (w < wave) ? s_seg[k * 4 + w] : 0u;
    if ((ballot[k] >> lane_id()) & 1ull) {
      const uint32_t off = seg + rank_below(ballot[k]);
      a.cull_keys[off] = key[k];
      uint4* dst = reinterpret_cast<uint4*>(a.projected + static_cast<size_t>(off) * 8u);
      dst[0] = make_uint4(rec[k][0], rec[k][1], rec[k][2], rec[k][3]);
      dst[1] = make_uint4(rec[k][4], rec[k][5], rec[k][6], rec[k][7]);
      // The binning's tile rectangle, from the record just written (read back in depth order, a
      // gather of 8 bytes instead of the 32-byte record).
      Rect r;
      const uint32_t tiles = tile_rect(rec[k], static_cast<uint32_t>(cam.screen_size[0]),
                                       static_cast<uint32_t>(cam.screen_size[1]), a.tile, r);
      a.rects[off] = pack_rect(r, tiles, view);
    }
#pragma unroll
    for (uint32_t w = 0; w < 4; ++w) before += s_seg[k * 4 + w];
  }
}

// Global digit histograms of every pass in one read of the keys (onesweep's upfront histogram).
// Sort inputs in the cull's gapped layout (seg non-null: survivors per cull tile): position pos holds
// a key when pos < n and pos % kCullTile is below its cull tile's count. A wave's 64 consecutive
// positions from `row` lie in one cull tile (kCullTile is a multiple of 64), so the count is one
// scalar load per row; it is ~0 for a packed input and for rows at or past n (which hold no key).
static_assert(kCullTile % 64u == 0u, "a wave row lies in one cull tile");
__device__ __forceinline__ uint32_t row_limit(const uint32_t* seg, uint32_t row, uint32_t n) {
  const uint32_t r = __builtin_amdgcn_readfirstlane(row);
  if (seg == nullptr || r >= n) return ~0u;
  return seg[r / kCullTile];
}
__device__ __forceinline__ bool sort_valid(uint32_t pos, uint32_t n, uint32_t limit) {
  return pos < n && pos % kCullTile < limit;
}

__global__ __launch_bounds__(256) void k_hist(const uint32_t* __restrict__ keys, const uint32_t* n_ptr,
                                              const uint32_t* __restrict__ seg, uint32_t* hist, DigitPlan plan) {
  __shared__ uint32_t s_h[4][4][256];
  const uint32_t tid = threadIdx.x, wave = tid >> 6;
  const int passes = plan.passes;
  for (uint32_t i = tid; i < 4 * 4 * 256; i += 256) (&s_h[0][0][0])[i] = 0;
  __syncthreads();
  const uint32_t n = *n_ptr;
  for (uint32_t i = blockIdx.x * 256u + tid; i < n; i += gridDim.x * 256u) {
    const uint32_t limit = row_limit(seg, i - lane_id(), n);
    const uint32_t k = keys[i];
    if (!sort_valid(i, n, limit)) continue;
    for (int p = 0; p < passes; ++p)
      atomicAdd(&s_h[wave][p][(k >> plan.shift[p]) & ((1u << plan.bits[p]) - 1u)], 1u);
  }
  __syncthreads();
  for (int p = 0; p < passes; ++p) {
    const uint32_t v = s_h[0][p][tid] + s_h[1][p][tid] + s_h[2][p][tid] + s_h[3][p][tid];
    if (v != 0u) atomicAdd(&hist[p * 256 + tid], v);
  }
}

constexpr uint32_t kLookback = 16;


// One stable LSD pass over kSortTile-key tiles. Keys of tile b sit at b*kSortTile + wave*kSortItems*64 +
// j*64 + lane (positions sort_valid rejects hold none: the first depth pass reads the cull's gapped
// layout, with a key's position as its value, and writes the survivors packed):
// each wave ranks its kSortItems rows in order with 8 ballots per row (lanes with equal digits form a
// peer mask), so the in-tile order is position order and the pass is stable. The tile then learns
// where each of its digits starts in the output:
//   LOOKBACK (onesweep): tiles take tickets in order and look back across predecessors' per-digit
//     status words (decoupled look-back), one kernel per pass;
//   !LOOKBACK (reduce then scan): k_tile_counts and k_digit_scan have already written the offset of
//     digit d of tile t within digit d at offsets[d * stride + t] and digit d's total at hist[d].
#ifndef SKM_SCATTER_WAVES
#define SKM_SCATTER_WAVES 4  // waves per SIMD the scatter's register budget is held to
#endif
template <bool LOOKBACK, int BITS>
__global__ __launch_bounds__(256, SKM_SCATTER_WAVES) void k_onesweep(const uint32_t* __restrict__ keys_in,
                                                  const uint32_t* __restrict__ vals_in,
                                                  uint32_t* __restrict__ keys_out,
                                                  uint32_t* __restrict__ vals_out, const uint32_t* n_ptr,
                                                  const uint32_t* __restrict__ hist, uint32_t shift,
                                                  uint32_t mask, uint64_t* status, Ctrl* ctrl, uint32_t ticket,
                                                  uint32_t tag_pass, const uint32_t* __restrict__ offsets,
                                                  uint32_t stride, const uint32_t* __restrict__ seg,
                                                  uint32_t view_keys) {
  __shared__ uint32_t s_ticket;
  __shared__ uint32_t s_wave_count[4][kRadix];
  __shared__ uint32_t s_base[kRadix];
  __shared__ uint32_t s_scan[4];
  __shared__ uint32_t s_keys[kSortTile];
  __shared__ uint32_t s_vals[kSortTile];
  const uint32_t n = *n_ptr;
  const uint32_t blocks = (n + kSortTile - 1u) / kSortTile;
  // Tiles i = blockIdx.x, blockIdx.x + gridDim.x, ... < blocks: a grid sized for the capacity takes one
  // tile per block; a smaller grid (a multiple of the XCD count) walks the tiles. The look-back
  // variant is launched with a block per tile of the capacity (its tickets order the tiles).
  for (uint32_t i = blockIdx.x; i < blocks; i += gridDim.x) {
    // The thread's index passes an empty asm each tile, so lane-dependent offsets are computed per
    // tile rather than hoisted out of the loop, where they would hold registers across the tile.
    uint32_t tid = threadIdx.x;
    asm volatile("" : "+v"(tid));
    const uint32_t wave = tid >> 6, lane = tid & 63u;
    if (LOOKBACK && tid == 0) s_ticket = atomicAdd(&ctrl->ticket[ticket], 1u);
#pragma unroll
    for (uint32_t w = 0; w < 4; ++w) s_wave_count[w][tid] = 0;
    __syncthreads();
    const uint32_t b = LOOKBACK ? s_ticket : xcd_tile(i, blocks);
    const uint32_t base = b * kSortTile;

    uint32_t key[kSortItems], val[kSortItems], rank[kSortItems];
    // Every load first (positions clamped to n - 1), so a thread waits for one round trip. The inputs
    // are read once per pass, so they are loaded non-temporally: they then displace fewer of the
    // scatter's partially written lines from the XCD's L2 before those lines are merged.
#pragma unroll
    for (uint32_t j = 0; j < kSortItems; ++j) {
      const uint32_t pos = min(base + wave * (kSortItems * 64u) + j * 64u + lane, n - 1u);
      key[j] = __builtin_nontemporal_load(keys_in + pos);
      val[j] = vals_in != nullptr ? __builtin_nontemporal_load(vals_in + pos) : pos;
    }
    uint32_t valid_rows = 0;  // bit j: row j's position holds a key
#pragma unroll
    for (uint32_t j = 0; j < kSortItems; ++j) {
      const uint32_t row = base + wave * (kSortItems * 64u) + j * 64u;
      const bool valid = sort_valid(row + lane, n, row_limit(seg, row, n));
      valid_rows |= valid ? 1u << j : 0u;
      key[j] = valid ? key[j] : 0u;
      val[j] = valid ? val[j] : 0u;
    }
#pragma unroll
    for (uint32_t j = 0; j < kSortItems; ++j) {
      const bool valid = (valid_rows >> j) & 1u;
      const uint32_t d = (key[j] >> shift) & mask;  // mask < 2^BITS
      const uint64_t peers = digit_peers<BITS>(d, __ballot(valid));
      const uint32_t below = rank_below(peers);
      const uint32_t before = s_wave_count[wave][d];
      rank[j] = before + below;
      // The lowest lane of each digit (no peer below it) stores the digit's new count.
      if (valid && below == 0u) s_wave_count[wave][d] = before + static_cast<uint32_t>(__popcll(peers));
    }
    __syncthreads();
    // Per digit (thread = digit): offsets of each wave, the tile's total, and its start in the tile.
    const uint32_t c0 = s_wave_count[0][tid], c1 = s_wave_count[1][tid];
    const uint32_t c2 = s_wave_count[2][tid], c3 = s_wave_count[3][tid];
    const uint32_t count = c0 + c1 + c2 + c3;
    uint32_t tile_total;
    const uint32_t start = block_exclusive_scan(count, s_scan, &tile_total);
    // Where wave w's keys of digit `tid` start in the staged tile.
    s_wave_count[0][tid] = start;
    s_wave_count[1][tid] = start + c0;
    s_wave_count[2][tid] = start + c0 + c1;
    s_wave_count[3][tid] = start + c0 + c1 + c2;
    if constexpr (LOOKBACK) {
      // Global start of each digit: exclusive scan of the pass histogram.
      uint32_t all;
      const uint32_t gstart = block_exclusive_scan(hist[tid], s_scan, &all);
      const uint32_t tag = make_tag(ctrl->epoch, tag_pass);
      // Per-digit look-back across tiles.
      uint64_t* st = status + static_cast<size_t>(b) * kRadix;
      uint32_t excl = 0;
      if (b == 0) {
        status_store(st + tid, status_word(tag, kFlagPrefix, count));
      } else {
        status_store(st + tid, status_word(tag, kFlagAggregate, count));
        // kLookback predecessors per round trip (loads in flight together), nearest first, until one
        // holds an inclusive prefix.
        int32_t k = static_cast<int32_t>(b) - 1;
        while (k >= 0) {
          uint64_t sw[kLookback];
#pragma unroll
          for (int32_t j = 0; j < static_cast<int32_t>(kLookback); ++j)
            sw[j] = (k - j >= 0) ? status_load(status + static_cast<size_t>(k - j) * kRadix + tid) : 0ull;
          int32_t j = 0;
          bool done = false;
#pragma unroll
          for (int32_t i = 0; i < static_cast<int32_t>(kLookback); ++i) {
            if (done || j != i || k - i < 0) continue;
            const uint32_t flag = static_cast<uint32_t>(sw[i] >> 32) & 3u;
            if (static_cast<uint32_t>(sw[i] >> 34) != tag || flag == 0u) continue;  // not published yet
            excl += static_cast<uint32_t>(sw[i]);
            j = i + 1;
            if (flag == kFlagPrefix) done = true;
          }
          if (done) break;
          k -= j;
        }
        status_store(st + tid, status_word(tag, kFlagPrefix, excl + count));
      }
      s_base[tid] = gstart + excl - start;
    } else {
      // Global start of each digit from the digit totals (zero for digits above the mask).
      uint32_t all;
      const uint32_t gstart = block_exclusive_scan(tid <= mask ? hist[tid] : 0u, s_scan, &all);
      s_base[tid] = gstart + offsets[static_cast<size_t>(tid) * stride + b] - start;
    }
    __syncthreads();
    // Stage the tile in digit order in LDS.
#pragma unroll
    for (uint32_t j = 0; j < kSortItems; ++j) {
      if ((valid_rows >> j) & 1u) {
        const uint32_t d = (key[j] >> shift) & mask;
        const uint32_t local = s_wave_count[wave][d] + rank[j];
        s_keys[local] = key[j];
        s_vals[local] = val[j];
      }
    }
    __syncthreads();
    // Write the staged tile: every LDS read of the thread's keys first, then their bases, then the
    // stores (one round trip of each instead of two per key).
#pragma unroll
    for (uint32_t j = 0; j < kSortItems; ++j) {
      key[j] = s_keys[j * 256u + tid];
      val[j] = s_vals[j * 256u + tid];
    }
#pragma unroll
    for (uint32_t j = 0; j < kSortItems; ++j) rank[j] = s_base[(key[j] >> shift) & mask] + j * 256u + tid;
    // view_keys (a batch's view count): the keys written are the values' views instead, the key of the
    // sort on the view that follows (a survivor's position lies in a cull tile of its view).
    if (view_keys != 0u) {
#pragma unroll
      for (uint32_t j = 0; j < kSortItems; ++j) key[j] = (val[j] / kCullTile) % view_keys;
    }
    if (tile_total == kSortTile) {  // a full tile (every tile of a packed input but the last): no masks
      if (keys_out != nullptr) {
#pragma unroll
        for (uint32_t j = 0; j < kSortItems; ++j) keys_out[rank[j]] = key[j];
      }
#pragma unroll
      for (uint32_t j = 0; j < kSortItems; ++j) vals_out[rank[j]] = val[j];
    } else {
      if (keys_out != nullptr) {
#pragma unroll
        for (uint32_t j = 0; j < kSortItems; ++j)
          if (j * 256u + tid < tile_total) keys_out[rank[j]] = key[j];
      }
#pragma unroll
      for (uint32_t j = 0; j < kSortItems; ++j)
        if (j * 256u + tid < tile_total) vals_out[rank[j]] = val[j];
    }
  }
}

// Reduce then scan, step 1: the digit counts of each kSortTile-key tile, digit-major
// (counts[d * stride + tile]), from per-wave LDS histograms. Low digits are close to random in a
// pass's input order (the pairs of one survivor have distinct tiles, and every later pass reads keys
// in the previous digit's order), so each key adds itself. A key's top digit takes few values (the
// exponent of a depth, the row band of a tile), and many lanes of a row adding to one LDS word are
// serialised, so with RUNS a row adds once per run of equal digits instead: a lane starts a run when
// its digit differs from the previous lane's, and the run ends at the next start or at the row's
// last key (a row's keys are a prefix of its lanes, in both the packed and the gapped layout).
template <bool RUNS>
__global__ __launch_bounds__(256) void k_tile_counts(const uint32_t* __restrict__ keys, const uint32_t* n_ptr,
                                                     uint32_t shift, uint32_t mask, uint32_t* __restrict__ counts,
                                                     uint32_t stride, const uint32_t* __restrict__ seg) {
  __shared__ uint32_t s_h[4][kRadix];
  const uint32_t n = *n_ptr;
  const uint32_t blocks = (n + kSortTile - 1u) / kSortTile;
  // Tiles i = blockIdx.x, blockIdx.x + gridDim.x, ... < blocks (a grid smaller than the capacity's
  // tiles is a multiple of the XCD count).
  for (uint32_t i = blockIdx.x; i < blocks; i += gridDim.x) {
    uint32_t tid = threadIdx.x;
    asm volatile("" : "+v"(tid));  // lane-dependent offsets per tile, not held across the loop
    const uint32_t wave = tid >> 6, lane = tid & 63u;
#pragma unroll
    for (uint32_t w = 0; w < 4; ++w) s_h[w][tid] = 0;  // column tid: read by this thread only below
    __syncthreads();
    const uint32_t b = xcd_tile(i, blocks);
    const uint32_t base = b * kSortTile;
    uint32_t key[kSortItems];
#pragma unroll
    for (uint32_t j = 0; j < kSortItems; ++j) key[j] = keys[min(base + j * 256u + tid, n - 1u)];
    // Every row's validity first: the gapped layout's per-row scalar loads then share one wait, and the
    // LDS adds below are not held behind it row by row (a scalar load and an LDS op share a counter).
    uint32_t valid_rows = 0;
#pragma unroll
    for (uint32_t j = 0; j < kSortItems; ++j) {
      const uint32_t row = base + j * 256u + wave * 64u;
      valid_rows |= sort_valid(row + lane, n, row_limit(seg, row, n)) ? 1u << j : 0u;
    }
    // The previous lane's digit of every row, all the lane moves in flight together.
    uint32_t prev[kSortItems];
    if constexpr (RUNS) {
#pragma unroll
      for (uint32_t j = 0; j < kSortItems; ++j) prev[j] = __shfl_up((key[j] >> shift) & mask, 1u, 64);
    }
#pragma unroll
    for (uint32_t j = 0; j < kSortItems; ++j) {
      const bool valid = (valid_rows >> j) & 1u;
      const uint32_t d = (key[j] >> shift) & mask;
      if constexpr (RUNS) {
        const bool start = valid && (lane == 0u || prev[j] != d);
        const uint64_t starts = __ballot(start);
        const uint32_t keys_in_row = static_cast<uint32_t>(__popcll(__ballot(valid)));
        const uint64_t later = starts & ~((2ull << lane) - 1ull);  // starts after this lane
        const uint32_t end = later != 0ull ? static_cast<uint32_t>(__builtin_ctzll(later)) : keys_in_row;
        if (start) atomicAdd(&s_h[wave][d], end - lane);
      } else {
        if (valid) atomicAdd(&s_h[wave][d], 1u);
      }
    }
    __syncthreads();
    counts[static_cast<size_t>(tid) * stride + b] = s_h[0][tid] + s_h[1][tid] + s_h[2][tid] + s_h[3][tid];
  }
}

// Reduce then scan, step 2: block d scans digit d's row of tile counts in place (exclusive, from 0)
// and stores the row's total in totals[d]. Per round, wave w takes kPer rows of 64 consecutive tiles
// (lanes read consecutive words, every load in flight together) and scans them in order, one wave scan
// (data-parallel lane moves, no LDS) per row; the waves' runs are then offset by the runs before them.
__global__ __launch_bounds__(256) void k_digit_scan(uint32_t* __restrict__ counts, const uint32_t* n_ptr,
                                                    uint32_t* __restrict__ totals, uint32_t stride) {
  constexpr uint32_t kPer = 16;
  constexpr uint32_t kRound = 256u * kPer;
  __shared__ uint32_t s_run[4];
  const uint32_t tid = threadIdx.x, wave = tid >> 6, lane = lane_id(), d = blockIdx.x;
  const uint32_t n = *n_ptr;
  const uint32_t tiles = (n + kSortTile - 1u) / kSortTile;
  uint32_t carry = 0;
  uint32_t* row = counts + static_cast<size_t>(d) * stride;
  for (uint32_t t0 = 0; t0 < tiles; t0 += kRound) {
    const uint32_t w0 = t0 + wave * (kPer * 64u);
    uint32_t v[kPer];
#pragma unroll
    for (uint32_t i = 0; i < kPer; ++i) {
      const uint32_t t = w0 + i * 64u + lane;
      v[i] = t < tiles ? row[t] : 0u;
    }
    uint32_t run = 0;  // this wave's sum over its rows so far
#pragma unroll
    for (uint32_t i = 0; i < kPer; ++i) {
      const uint32_t incl = wave_inclusive_scan(v[i]);
      const uint32_t excl = run + incl - v[i];
      run += static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(incl), 63));
      v[i] = excl;
    }
    if (lane == 0u) s_run[wave] = run;
    __syncthreads();
    uint32_t before = carry, all = 0;
#pragma unroll
    for (uint32_t w = 0; w < 4; ++w) {
      before += w < wave ? s_run[w] : 0u;
      all += s_run[w];
    }
    __syncthreads();
#pragma unroll
    for (uint32_t i = 0; i < kPer; ++i) {
      const uint32_t t = w0 + i * 64u + lane;
      if (t < tiles) row[t] = before + v[i];
    }
    carry += all;
  }
  if (tid == 0) totals[d] = carry;
}

// Tile pairs of a packed rectangle (k_cull's pack_rect): none for kNoRect.
__device__ __forceinline__ uint32_t rect_pairs(uint64_t rc) {
  if (rc == kNoRect) return 0u;
  const uint32_t w = static_cast<uint32_t>((rc >> 24) & 4095u) - static_cast<uint32_t>(rc & 4095u) + 1u;
  const uint32_t h = static_cast<uint32_t>((rc >> 36) & 4095u) - static_cast<uint32_t>((rc >> 12) & 4095u) + 1u;
  return w * h;
}

__device__ __forceinline__ uint64_t wave_inclusive_scan64(uint64_t v) {
  const uint32_t lane = lane_id();
#pragma unroll
  for (uint32_t d = 1; d < 64; d <<= 1) {
    const uint32_t lo = __shfl_up(static_cast<uint32_t>(v), d, 64);
    const uint32_t hi = __shfl_up(static_cast<uint32_t>(v >> 32), d, 64);
    if (lane >= d) v += (static_cast<uint64_t>(hi) << 32) | lo;
  }
  return v;
}

// Binning, first pass (reduce then scan) over the survivors in depth order, kBinItems * 256 per
// survivor block (blocks b = blockIdx.x, blockIdx.x + gridDim.x, ...): a block gathers its survivors'
// tile rectangles into depth order (low words in keys[0], high words in keys[1]: the depth sort is
// done with both) and stores its pair total in bin_starts[b].
__global__ __launch_bounds__(256) void k_bin_count(FrameArgs a, const uint32_t* __restrict__ order) {
  constexpr uint32_t kTile = 256u * kBinItems;
  __shared__ uint32_t s_scan[4];
  const uint32_t tid = threadIdx.x;
  const uint32_t n = a.ctrl->visible;
  const uint32_t blocks = (n + kTile - 1u) / kTile;
  for (uint32_t b = blockIdx.x; b < blocks; b += gridDim.x) {
    // All loads first (indices clamped to the last survivor), so a thread waits for two round trips
    // rather than two per item.
    uint32_t value[kBinItems];
    uint64_t rc[kBinItems];
#pragma unroll
    for (uint32_t k = 0; k < kBinItems; ++k) value[k] = order[min(b * kTile + k * 256u + tid, n - 1u)];
#pragma unroll
    for (uint32_t k = 0; k < kBinItems; ++k) rc[k] = a.rects[value[k]];
    uint32_t sum = 0;
#pragma unroll
    for (uint32_t k = 0; k < kBinItems; ++k) {
      const uint32_t i = b * kTile + k * 256u + tid;
      if (i < n) {
        a.keys[0][i] = static_cast<uint32_t>(rc[k]);
        a.keys[1][i] = static_cast<uint32_t>(rc[k] >> 32);
        sum += rect_pairs(rc[k]);
      }
    }
    uint32_t total;
    (void)block_exclusive_scan(sum, s_scan, &total);
    if (tid == 0) a.bin_starts[b] = total;
  }
}

// Binning, the scan: one block turns the block totals into block starts in place (exclusive,
// saturating at 2^32 - 1) and publishes the frame's pair count.
__global__ __launch_bounds__(256) void k_bin_scan(FrameArgs a) {
  constexpr uint32_t kTile = 256u * kBinItems, kPer = 8;
  __shared__ uint64_t s_wave[4];
  const uint32_t tid = threadIdx.x, wave = tid >> 6;
  const uint32_t n = a.ctrl->visible;
  const uint32_t blocks = (n + kTile - 1u) / kTile;
  uint64_t carry = 0;
  for (uint32_t t0 = 0; t0 < blocks; t0 += 256u * kPer) {
    uint32_t v[kPer];
    uint64_t sum = 0;
#pragma unroll
    for (uint32_t i = 0; i < kPer; ++i) {
      const uint32_t t = t0 + tid * kPer + i;
      v[i] = t < blocks ? a.bin_starts[t] : 0u;
      sum += v[i];
    }
    const uint64_t incl = wave_inclusive_scan64(sum);
    if (lane_id() == 63) s_wave[wave] = incl;
    __syncthreads();
    uint64_t before = 0, all = 0;
#pragma unroll
    for (uint32_t w = 0; w < 4; ++w) {
      before += w < wave ? s_wave[w] : 0ull;
      all += s_wave[w];
    }
    __syncthreads();
    uint64_t run = carry + before + incl - sum;
#pragma unroll
    for (uint32_t i = 0; i < kPer; ++i) {
      const uint32_t t = t0 + tid * kPer + i;
      if (t < blocks) a.bin_starts[t] = run > 0xffffffffull ? 0xffffffffu : static_cast<uint32_t>(run);
      run += v[i];
    }
    carry += all;
  }
  if (tid == 0) {
    const uint32_t pairs = carry > 0xffffffffull ? 0xffffffffu : static_cast<uint32_t>(carry);
    const bool over = carry > a.pair_capacity;
    a.ctrl->pairs = pairs;
    a.ctrl->overflow = over ? 1u : 0u;
    // The tile sort's count: zero after an overflow (the frame is dropped).
    a.ctrl->sort_pairs = over ? 0u : pairs;
  }
}


// Binning, second pass, balanced by output: emission block k writes pairs [k C, (k + 1) C) of the
// frame (C = kEmitChunk), whatever survivors they come from (the nearest survivors come first in
// depth order and have the largest footprints). It finds the survivor block holding its first pair
// (a 64-way search over the block starts), then, for each survivor block its range meets, reads
// that block's rectangles (in depth order) and values, scans their pair counts into slot offsets in
// LDS and writes its share. Pair e of a survivor block belongs to the last slot whose offset is <= e,
// and its tile is that slot's rectangle cell k = e - offset, row-major, so pairs come out in depth
// order, survivor by survivor. Each wave takes a contiguous run of 64-pair chunks: the slots of a
// chunk are a window of 64 slots from the previous chunk's last owner, whose starts are marked in
// LDS and spread by a wave max-scan (a binary search covers the first chunk and any pair the window
// misses).
//
// An emission block's pairs are exactly one tile of the tile sort's first radix pass, so with the
// reduce-then-scan passes (`digit_mask` != 0) it also counts that pass's digits (the tile id's low
// bits) in LDS and writes the tile's counts, which the pass would otherwise read every key again for.
constexpr uint32_t kEmitChunk = kSortTile;
__global__ __launch_bounds__(256) void k_bin(FrameArgs a, const uint32_t* __restrict__ order, uint32_t digit_mask,
                                             uint32_t stride) {
  constexpr uint32_t kItems = kBinItems;
  constexpr uint32_t kTile = 256u * kItems;
  __shared__ uint32_t s_scan[4];
  __shared__ uint32_t s_excl[kTile + 1u], s_val[kTile], s_origin[kTile], s_width[kTile], s_base[kTile];
  __shared__ uint32_t s_mark[4][64];
  __shared__ uint32_t s_h[4][kRadix];
  const uint32_t tid = threadIdx.x, wave = tid >> 6, lane = tid & 63u;
  const uint32_t total = a.ctrl->sort_pairs;
  const uint32_t chunks = (total + kEmitChunk - 1u) / kEmitChunk;
  const uint32_t n = a.ctrl->visible;
  const uint32_t blocks = (n + kTile - 1u) / kTile;
  const uint32_t* starts = a.bin_starts;  // exclusive pair start of each survivor block
  // Chunks blockIdx.x, blockIdx.x + gridDim.x, ... (a grid smaller than the capacity's chunks walks
  // them); the tile counts go to counts[d * stride + chunk], as k_tile_counts lays them out (stride:
  // the pair capacity's tiles).
  for (uint32_t chunk = blockIdx.x; chunk < chunks; chunk += gridDim.x) {
    const uint32_t e0 = chunk * kEmitChunk;
    const uint32_t e1 = min(e0 + kEmitChunk, total);
#pragma unroll
    for (uint32_t w = 0; w < 4; ++w) s_h[w][tid] = 0u;  // published by the first __syncthreads below
    // The survivor block of pair e0: the last block whose start is <= e0.
    uint32_t lo = 0, hi = blocks;
    while (hi - lo > 1u) {
      const uint32_t step = (hi - lo + 63u) / 64u;
      const uint32_t probe = lo + lane * step;
      const uint64_t ok = __ballot(probe < hi && starts[probe] <= e0);
      lo += (63u - static_cast<uint32_t>(__builtin_clzll(ok))) * step;
      hi = min(hi, lo + step);
    }
    s_mark[wave][lane] = 0u;
    for (uint32_t b = lo; b < blocks; ++b) {
      const uint32_t block_start = starts[b];
      if (block_start >= e1) break;
      uint32_t values[kItems], lo_words[kItems], hi_words[kItems];
#pragma unroll
      for (uint32_t k = 0; k < kItems; ++k) {
        const uint32_t i = min(b * kTile + k * 256u + tid, n - 1u);
        values[k] = order[i];
        lo_words[k] = a.keys[0][i];
        hi_words[k] = a.keys[1][i];
      }
      __syncthreads();  // the previous block's table is no longer read
      uint32_t aggregate = 0;
#pragma unroll
      for (uint32_t k = 0; k < kItems; ++k) {
        const uint32_t slot = k * 256u + tid;
        const uint32_t i = b * kTile + slot;
        uint32_t count = 0, value = 0, base = 0, tiles_x = a.tiles_x;
        Rect r{0, 0, -1, -1};
        if (i < n) {
          value = values[k];
          const uint64_t rc = static_cast<uint64_t>(lo_words[k]) | (static_cast<uint64_t>(hi_words[k]) << 32);
          if (rc != kNoRect) {
            r.tx0 = static_cast<int32_t>(rc & 4095u);
            r.ty0 = static_cast<int32_t>((rc >> 12) & 4095u);
            r.tx1 = static_cast<int32_t>((rc >> 24) & 4095u);
            r.ty1 = static_cast<int32_t>((rc >> 36) & 4095u);
            count = static_cast<uint32_t>((r.tx1 - r.tx0 + 1) * (r.ty1 - r.ty0 + 1));
            if (a.views != nullptr) {
              const uint32_t view = static_cast<uint32_t>(rc >> 48);
              const ViewInfo& vi = a.views[view];
              base = a.local_tile_bits != 0u ? view << a.local_tile_bits : vi.tile_base;
              tiles_x = vi.tiles_x;
            }
          }
        }
        uint32_t row_total;
        const uint32_t excl = block_exclusive_scan(count, s_scan, &row_total);
        s_excl[slot] = aggregate + excl;
        s_val[slot] = value;
        s_origin[slot] = static_cast<uint32_t>(r.tx0) | (static_cast<uint32_t>(r.ty0) << 16);
        s_width[slot] = static_cast<uint32_t>(r.tx1 - r.tx0 + 1) | (tiles_x << 16);
        s_base[slot] = base;
        aggregate += row_total;
      }
      if (tid == 0) s_excl[kTile] = aggregate;
      __syncthreads();
      auto find = [&](uint32_t e) {  // last slot with s_excl[slot] <= e
        uint32_t l = 0, h = kTile - 1u;
        while (l < h) {
          const uint32_t mid = (l + h + 1u) / 2u;
          if (s_excl[mid] <= e) l = mid; else h = mid - 1u;
        }
        return l;
      };
      // This block's pairs in [e0, e1), as block-local indices [first_pair, end_pair).
      const uint32_t first_pair = e0 > block_start ? e0 - block_start : 0u;
      const uint32_t end_pair = min(e1 - block_start, aggregate);
      const uint32_t c_first = first_pair / 64u, c_end = (end_pair + 63u) / 64u;
      const uint32_t chunks = c_end > c_first ? c_end - c_first : 0u;
      const uint32_t c0 = c_first + chunks * wave / 4u, c1 = c_first + chunks * (wave + 1u) / 4u;
      uint32_t s0 = c0 < c1 ? find(c0 * 64u) : 0u;
      for (uint32_t c = c0; c < c1; ++c) {
        const uint32_t first = c * 64u;
        const uint32_t s = s0 + lane;
        if (s < kTile) {
          const uint32_t ex = s_excl[s], cnt = s_excl[s + 1u] - ex;
          if (cnt != 0u && ex < first + 64u && ex + cnt > first)
            s_mark[wave][ex > first ? ex - first : 0u] = s + 1u;
        }
        const uint32_t m = wave_inclusive_max(s_mark[wave][lane]);
        s_mark[wave][lane] = 0u;
        const uint32_t e = first + lane;
        uint32_t owner = m != 0u ? m - 1u : 0u;
        if (e >= first_pair && e < end_pair) {
          if (m == 0u || e < s_excl[owner] || e >= s_excl[owner + 1u]) owner = find(e);
          const uint32_t k = e - s_excl[owner];
          const uint32_t wdt = s_width[owner] & 0xffffu, stride = s_width[owner] >> 16;
          // k / wdt from a reciprocal estimate, corrected to the exact quotient.
          uint32_t q = static_cast<uint32_t>(static_cast<float>(k) * __builtin_amdgcn_rcpf(static_cast<float>(wdt)));
          int32_t rem = static_cast<int32_t>(k - q * wdt);
          while (rem < 0) {
            --q;
            rem += static_cast<int32_t>(wdt);
          }
          while (rem >= static_cast<int32_t>(wdt)) {
            ++q;
            rem -= static_cast<int32_t>(wdt);
          }
          const uint32_t tx = (s_origin[owner] & 0xffffu) + static_cast<uint32_t>(rem);
          const uint32_t ty = (s_origin[owner] >> 16) + q;
          const uint32_t idx = block_start + e;  // < total <= pair capacity
          const uint32_t tile = s_base[owner] + ty * stride + tx;
          a.pair_keys[0][idx] = tile;
          a.pair_vals[0][idx] = s_val[owner];
          if (digit_mask != 0u) atomicAdd(&s_h[wave][tile & digit_mask], 1u);
        }
        s0 = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(owner), 63));
      }
    }
    if (digit_mask != 0u) {
      __syncthreads();
      if (tid <= digit_mask)
        a.sort_counts[static_cast<size_t>(tid) * stride + chunk] =
            s_h[0][tid] + s_h[1][tid] + s_h[2][tid] + s_h[3][tid];
    }
  }
}

// [start, end) of each tile's pairs by boundary detection: thread t takes keys [4t, 4t + 4) as one
// 16-byte load (the key arrays are 256-byte aligned) and its two neighbours from the cache, and the
// next group of 4 a grid's threads further while any are left (the grid comes from earlier frames). With
// view-local keys (local_bits != 0) a key is view << local_bits | the tile's index in its view, and
// its range goes to the view's tile_base + that index; equal keys are adjacent (the pairs are sorted
// on the index, and each index's pairs come in view order).
__global__ __launch_bounds__(256) void k_ranges(const uint32_t* __restrict__ keys, const uint32_t* n_ptr,
                                                uint32_t* ranges, const ViewInfo* __restrict__ views,
                                                uint32_t local_bits) {
  const uint32_t n = *n_ptr;
  for (uint32_t i0 = (blockIdx.x * 256u + threadIdx.x) * 4u; i0 < n; i0 += gridDim.x * 1024u) {
    uint32_t k[6];
    if (i0 + 4u <= n) {
      const uint4 v = *reinterpret_cast<const uint4*>(keys + i0);
      k[1] = v.x;
      k[2] = v.y;
      k[3] = v.z;
      k[4] = v.w;
    } else {
#pragma unroll
      for (uint32_t j = 0; j < 4; ++j) k[1 + j] = i0 + j < n ? keys[i0 + j] : ~0u;
    }
    k[0] = i0 > 0u ? keys[i0 - 1u] : ~0u;
    k[5] = i0 + 4u < n ? keys[i0 + 4u] : ~0u;
#pragma unroll
    for (uint32_t j = 0; j < 4; ++j) {
      const uint32_t i = i0 + j;
      if (i >= n) break;
      const uint32_t t = k[1 + j];
      const bool first = i == 0u || k[j] != t, last = i + 1u == n || k[2 + j] != t;
      if (!first && !last) continue;
      const uint32_t g = local_bits != 0u ? views[t >> local_bits].tile_base + (t & ((1u << local_bits) - 1u)) : t;
      if (first) ranges[2 * g] = i;
      if (last) ranges[2 * g + 1] = i + 1u;
    }
  }
}

// The raster's tile order, heaviest first: a tile's bucket is the count of leading zero bits of its
// pairs + 1, so bucket 0 holds the tiles with the most pairs and bucket 31 the empty ones. Counts per
// bucket go to counters[0, 32), then each tile takes a slot in its bucket's range through the
// cursors at counters[32, 64) (the order within a bucket does not matter: a tile's pixels do not
// depend on when it runs).
__device__ __forceinline__ uint32_t order_bucket(const uint32_t* ranges, uint32_t t) {
  return static_cast<uint32_t>(__clz(static_cast<int>(ranges[2 * t + 1] - ranges[2 * t] + 1u)));
}
__global__ __launch_bounds__(256) void k_order_count(const uint32_t* __restrict__ ranges, uint32_t tiles,
                                                     uint32_t* __restrict__ counters) {
  __shared__ uint32_t s_count[32];
  const uint32_t tid = threadIdx.x, t = blockIdx.x * 256u + tid;
  if (tid < 32u) s_count[tid] = 0u;
  __syncthreads();
  if (t < tiles) atomicAdd(&s_count[order_bucket(ranges, t)], 1u);
  __syncthreads();
  if (tid < 32u && s_count[tid] != 0u) atomicAdd(&counters[tid], s_count[tid]);
}
__global__ __launch_bounds__(256) void k_order_scatter(const uint32_t* __restrict__ ranges, uint32_t tiles,
                                                       uint32_t* __restrict__ counters, uint32_t* __restrict__ order,
                                                       uint32_t split_cap) {
  __shared__ uint32_t s_start[32], s_count[32], s_base[32];
  const uint32_t tid = threadIdx.x, t = blockIdx.x * 256u + tid;
  if (tid < 64u) {
    const uint32_t v = tid < 32u ? counters[tid] : 0u;
    const uint32_t incl = wave_inclusive_scan(v);
    if (tid < 32u) {
      s_start[tid] = incl - v;
      s_count[tid] = 0u;
    }
    // The raster's split tiles: the heaviest whole buckets holding at most split_cap tiles (the scan
    // is non-decreasing, so its largest value within the cap).
    const uint32_t within = __builtin_amdgcn_readfirstlane(
        static_cast<uint32_t>(__popcll(__ballot(tid < 32u && incl <= split_cap))));
    if (blockIdx.x == 0 && tid == 0)
      counters[64] = within == 0u ? 0u : __builtin_amdgcn_readlane(incl, within - 1u);
  }
  __syncthreads();
  const uint32_t bucket = t < tiles ? order_bucket(ranges, t) : 0u;
  const uint32_t local = t < tiles ? atomicAdd(&s_count[bucket], 1u) : 0u;
  __syncthreads();
  if (tid < 32u && s_count[tid] != 0u) s_base[tid] = s_start[tid] + atomicAdd(&counters[32u + tid], s_count[tid]);
  __syncthreads();
  if (t < tiles) order[s_base[bucket] + local] = t;
}

// One coverage-and-blend step of a pixel against a prepared record (the oracle's decide and fp32
// composite): the scaled quad coordinates are affine in the pixel centre, gu = g00 fx + h01 fy + u0
// (two fused multiply-adds per coordinate), log2 alpha = log2a0 - gu^2 - gv^2 (two more), the quad
// and cutoff tests select alpha = 2^min(log2 alpha, 0) or 0, then w = T alpha, C += w c, T -= w. A
// pixel whose transmittance is already below kTransmittanceStop takes no more splats, so a pixel's
// value does not depend on which other pixels share its wave. No branch: a wave runs the same
// instructions whatever its lanes decide. A = (g00, h01, u0, g10), B = (h11, v0, K r, log2a0).
// shape_test is the quad and cutoff part, which does not depend on the transmittance.
__device__ __forceinline__ bool shape_test(const float4 A, const float4 B, float fx, float fy, float& e) {
  const float gu = fmaf(A.x, fx, fmaf(A.y, fy, A.z));
  const float gv = fmaf(A.w, fx, fmaf(B.x, fy, B.y));
  e = fmaf(-gu, gu, fmaf(-gv, gv, B.w));
  // One compare for the quad: a NaN or infinite coordinate fails the cutoff test through e.
  return static_cast<int>(fmaxf(fabsf(gu), fabsf(gv)) <= B.z) & static_cast<int>(e >= kLog2MinAlpha);
}
// 2^x clamped to [0, 1] by the instruction's output clamp. 2^x is monotonic with 2^0 = 1, so this is
// 2^min(x, 0) for every x, and 0 for x = -inf.
__device__ __forceinline__ float exp2_clamped(float x) {
  float r;
  asm("v_exp_f32_e64 %0, %1 clamp" : "=v"(r) : "v"(x));
  return r;
}
__device__ __forceinline__ void blend(bool take, float e, const float4 Cc, float& T, float& C0, float& C1,
                                      float& C2) {
  const float alpha = exp2_clamped(take ? e : -__builtin_inff());
  const float w = T * alpha;
  C0 = fmaf(w, Cc.x, C0);
  C1 = fmaf(w, Cc.y, C1);
  C2 = fmaf(w, Cc.z, C2);
  T = T - w;
}

// Records the raster tests per LDS round trip (a build may set it; it divides every batch size).
// Four keeps k_raster<16, 1> at 73 VGPRs (six waves per SIMD); eight tests more records per round
// trip at 113 VGPRs (four waves), which is faster for one frame alone and slower once frames
// overlap or views are batched.
#ifndef SKM_RASTER_GROUP
#define SKM_RASTER_GROUP 4
#endif
constexpr uint32_t kGroup = SKM_RASTER_GROUP;
static_assert(64u % kGroup == 0u, "group size");

// Raster block shapes: a TILE x TILE tile per block, cut into square regions of 8 * PX pixels, one
// wave per region, PX x PX pixels per lane (lane (x, y) of the 8 x 8 lane grid takes the pixels
// (x + 8 i, y + 8 j)).
template <int TILE, int PX>
struct RasterShape {
  static constexpr uint32_t kRegion = 8u * PX;
  static constexpr uint32_t kSide = TILE / kRegion;     // regions per tile side
  static constexpr uint32_t kWaves = kSide * kSide;
  static constexpr uint32_t kThreads = 64u * kWaves;   // also the records prepared per batch
  static constexpr uint32_t kPixels = PX * PX;
  // Waves per SIMD the register budget is held to: six (80 VGPRs) for one pixel per lane, five for
  // four.
  static constexpr uint32_t kMinWaves = PX == 1 ? 6u : 5u;
  static_assert(TILE % kRegion == 0 && kWaves <= 4, "shape");
};

// TILE x TILE tile per block. Each batch of up to kThreads pairs is prepared once into LDS (centre in
// pixels, the inverse axis matrix scaled by kGaussianScale, the scaled radius, log2 peak, clamped
// colour), and every region gets the list, in batch order, of the records whose footprint box
// reaches one of its pixel centres. A wave walks its region's list front to back; it leaves the list
// once all of its pixels are below kTransmittanceStop, and the tile stops when every pixel is.
//
// Split tiles (16 x 16 tiles, one pixel per lane): the heaviest whole buckets of the tile order, at
// most `split_cap` tiles (a choice that does not depend on the order within a bucket), are each
// drawn by four blocks, one per 8 x 8 quadrant, whose waves take 4 x 4 pixels with four lanes per
// pixel: lane 4 p + r tests record r of each group of four against pixel p, then the group's blends
// run in order, each with record r's result broadcast across the pixel's four lanes. A wave thus
// walks a 4 x 4 region's list, four records per test, instead of an 8 x 8 region's list one record
// per test: the heaviest tiles, whose walks run long after the others have finished, finish in about
// a third of the time. Every blend is the same operation on the same operands, so the pixels are those
// of an unsplit tile bit for bit.
template <int TILE, int PX>
__device__ __forceinline__ uint32_t region_mask(uint32_t size, float bx, float by, float cx, float cy, float ex,
                                                float ey, float g00, float g01, float g10, float g11, float rk) {
  const float x0 = (cx - ex) - kBinPad, x1 = (cx + ex) + kBinPad;
  const float y0 = (cy - ey) - kBinPad, y1 = (cy + ey) + kBinPad;
  // The quad's two slabs |gu|, |gv| <= K r against the region: gu varies by at most
  // (|g00| + |g01|) h around the region's centre, h its half span plus the binning pad. The
  // pad's share, (|g00| + |g01|) / 16, and the slack cover the fp32 rounding of the per-pixel
  // rule, at most about 3 W 2^-24 (|g00| + |g01|) at image width W.
  const float span = static_cast<float>(size - 1u);
  const float half = 0.5f * span + kBinPad;
  const float hu = (fabsf(g00) + fabsf(g01)) * half, hv = (fabsf(g10) + fabsf(g11)) * half;
  const float lim_u = (rk + hu) * kFootprintSlack, lim_v = (rk + hv) * kFootprintSlack;
  uint32_t regions = 0;
#pragma unroll
  for (uint32_t q = 0; q < RasterShape<TILE, PX>::kWaves; ++q) {
    const float qx = bx + static_cast<float>((q % RasterShape<TILE, PX>::kSide) * size) + 0.5f;
    const float qy = by + static_cast<float>((q / RasterShape<TILE, PX>::kSide) * size) + 0.5f;
    const float dxc = (qx + 0.5f * span) - cx, dyc = cy - (qy + 0.5f * span);
    const float uc = fmaf(g00, dxc, g01 * dyc), vc = fmaf(g10, dxc, g11 * dyc);
    if (x1 >= qx && x0 <= qx + span && y1 >= qy && y0 <= qy + span && fabsf(uc) <= lim_u && fabsf(vc) <= lim_v)
      regions |= 1u << q;
  }
  return regions;
}

// Step U of a split wave's group: record U's result, held by lane 4 p + U, broadcast to pixel p's four
// lanes (a quad permutation), then blended as in the unsplit walk.
template <int U>
__device__ __forceinline__ void quad_blend(float x, const float4& Cc, float& T, float& C0, float& C1, float& C2) {
  const float xb = __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), U * 0x55, 0xf, 0xf, false));
  const float alpha = exp2_clamped(T >= kTransmittanceStop ? xb : -__builtin_inff());
  const float w = T * alpha;
  C0 = fmaf(w, Cc.x, C0);
  C1 = fmaf(w, Cc.y, C1);
  C2 = fmaf(w, Cc.z, C2);
  T = T - w;
}

// The LDS of a raster block. s_rec[0..2][k]: record k's A, B (coverage) and colour. The region lists
// hold the byte offset of s_rec[0][k] (16 k), which the LDS reads take as their address with no
// arithmetic. Slot kB of each record array is a null record that no pixel takes (radius -1): region
// lists are padded with it to whole groups.
template <int TILE, int PX>
struct RasterLds {
  static constexpr uint32_t kB = RasterShape<TILE, PX>::kThreads, kW = RasterShape<TILE, PX>::kWaves;
  static constexpr uint32_t kS = kB + 1u;
  float4 rec[3][kS];
  __align__(16) uint32_t list[kW][kB + kGroup];
  uint32_t count[kW][kW];  // [region][wave]
  uint32_t walked[kW];
};

// One block's work: a whole tile, or (QUAD) one quadrant of a split tile.
template <int TILE, int PX, bool QUAD>
__device__ __forceinline__ void raster_block(const FrameArgs& a, const uint32_t* __restrict__ pair_vals,
                                             const uint32_t* __restrict__ ranges, uint32_t pos,
                                             RasterLds<TILE, PX>& L) {
  using S = RasterShape<TILE, PX>;
  constexpr uint32_t kB = S::kThreads, kW = S::kWaves, kP = S::kPixels;
  constexpr uint32_t kS = kB + 1u;
  constexpr bool quad = QUAD;
  auto& s_rec = L.rec;
  auto& s_list = L.list;
  auto& s_count = L.count;
  const uint32_t tid = threadIdx.x, wave = tid >> 6, lane = tid & 63u;
  const uint32_t tile_id = a.tile_order != nullptr ? a.tile_order[pos] : pos;  // global tile id
  uint32_t width = a.width, height = a.height, tiles_x = a.tiles_x, local = tile_id;
  uint64_t pixel_base = 0;
  if (a.tile_view != nullptr) {
    const ViewInfo& vi = a.views[a.tile_view[tile_id]];
    width = vi.width;
    height = vi.height;
    tiles_x = vi.tiles_x;
    local = tile_id - vi.tile_base;
    pixel_base = vi.pixel_base;
  }
  const uint32_t ox = (local % tiles_x) * TILE, oy = (local / tiles_x) * TILE;
  // The block's regions: the tile's 8 x 8 quadrants, or a split block's quadrant's 4 x 4 quarters.
  const uint32_t rsize = quad ? S::kRegion / 2u : S::kRegion;
  const uint32_t bx = quad ? ox + (blockIdx.x & 1u) * S::kRegion : ox;
  const uint32_t by = quad ? oy + ((blockIdx.x >> 1) & 1u) * S::kRegion : oy;
  const uint32_t rx = bx + (wave % S::kSide) * rsize, ry = by + (wave / S::kSide) * rsize;
  // The lane's pixel: (lane % 8, lane / 8) of an 8 x 8 region, or pixel lane / 4 of a 4 x 4 one.
  const uint32_t lx = quad ? (lane >> 2) & 3u : lane & 7u, ly = quad ? lane >> 4 : lane >> 3;
  const uint32_t start = ranges[2 * tile_id], end = ranges[2 * tile_id + 1];
  const float fw = static_cast<float>(width), fh = static_cast<float>(height);
  float fx[kP], fy[kP], T[kP], C0[kP], C1[kP], C2[kP];
  bool in_image[kP];
#pragma unroll
  for (uint32_t q = 0; q < kP; ++q) {
    const uint32_t px = rx + lx + 8u * (q % PX), py = ry + ly + 8u * (q / PX);
    in_image[q] = px < width && py < height;
    fx[q] = static_cast<float>(px) + 0.5f;
    fy[q] = static_cast<float>(py) + 0.5f;
    T[q] = in_image[q] ? 1.0f : 0.0f;
    C0[q] = C1[q] = C2[q] = 0.0f;
  }
  auto live = [&]() {
    bool any = false;
#pragma unroll
    for (uint32_t q = 0; q < kP; ++q) any |= T[q] >= kTransmittanceStop;
    return any;
  };
  if (tid == 0) {
    s_rec[0][kB] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    s_rec[1][kB] = make_float4(0.0f, 0.0f, -1.0f, -__builtin_inff());
    s_rec[2][kB] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  }
  // Two-stage prefetch of the pair records: while batch b is prepared and walked, the record of batch
  // b + 1 and the pair value of batch b + 2 are in flight.
  auto record = [&](uint32_t idx, uint4& w0, uint4& w1) {
    const uint4* src = reinterpret_cast<const uint4*>(a.projected + static_cast<size_t>(idx) * 8u);
    w0 = src[0];
    w1 = src[1];
  };
  uint32_t walked = 0;  // records this wave tested (wave-uniform)
  uint4 next0 = make_uint4(0u, 0u, 0u, 0u), next1 = next0;
  if (start + tid < end) record(pair_vals[start + tid], next0, next1);
  uint32_t next_idx = start + kB + tid < end ? pair_vals[start + kB + tid] : 0u;
  for (uint32_t s0 = start; s0 < end; s0 += kB) {
    const uint32_t j = s0 + tid;
    uint32_t regions = 0;
    const uint4 w0 = next0, w1 = next1;
    if (j + kB < end) record(next_idx, next0, next1);
    next_idx = j + 2u * kB < end ? pair_vals[j + 2u * kB] : 0u;
    if (j < end) {
      const uint32_t p[8] = {w0.x, w0.y, w0.z, w0.w, w1.x, w1.y, w1.z, w1.w};
      const float cx = (f32(w0.x) * 0.5f + 0.5f) * fw;
      const float cy = (0.5f - f32(w0.y) * 0.5f) * fh;
      const float a1x = lo_half(w0.z), a1y = hi_half(w0.z);
      const float a2x = lo_half(w0.w), a2y = hi_half(w0.w);
      const float det = a1x * a2y - a2x * a1y;
      const float sc = kGaussianScale / det;
      const float log2a0 = log2_peak(w1.y);
      const float g00 = a2y * sc, g01 = -(a2x * sc), g10 = -(a1y * sc), g11 = a1x * sc;
      const float rk = kGaussianScale * f32(w1.z);
      float ex = 0.0f, ey = 0.0f;
      if (footprint(p, log2a0, ex, ey))
        regions = region_mask<TILE, PX>(rsize, static_cast<float>(bx), static_cast<float>(by), cx, cy, ex, ey, g00,
                                        g01, g10, g11, rk);
      // gu = g00 (fx - cx) + g01 (cy - fy) = g00 fx + h01 fy + u0 with h01 = -g01 and
      // u0 = g01 cy - g00 cx; likewise gv.
      const float u0 = fmaf(g01, cy, -(g00 * cx)), v0 = fmaf(g11, cy, -(g10 * cx));
      s_rec[0][tid] = make_float4(g00, -g01, u0, g10);
      s_rec[1][tid] = make_float4(-g11, v0, rk, log2a0);
      s_rec[2][tid] = make_float4(mn(mx(lo_half(w1.x), 0.0f), kMaxColorScale),
                             mn(mx(hi_half(w1.x), 0.0f), kMaxColorScale),
                             mn(mx(lo_half(w1.y), 0.0f), kMaxColorScale), 0.0f);
    }
    uint64_t ball[kW];
#pragma unroll
    for (uint32_t q = 0; q < kW; ++q) {
      ball[q] = __ballot((regions >> q) & 1u);
      if (lane == 0) s_count[q][wave] = static_cast<uint32_t>(__popcll(ball[q]));
    }
    __syncthreads();
#pragma unroll
    for (uint32_t q = 0; q < kW; ++q) {
      if ((regions >> q) & 1u) {
        uint32_t off = rank_below(ball[q]);
#pragma unroll
        for (uint32_t w = 0; w + 1 < kW; ++w) off += (w < wave) ? s_count[q][w] : 0u;
        s_list[q][off] = tid * static_cast<uint32_t>(sizeof(float4));
      }
    }
    uint32_t n = 0;
#pragma unroll
    for (uint32_t w = 0; w < kW; ++w) n += s_count[wave][w];
    n = __builtin_amdgcn_readfirstlane(n);
    __syncthreads();
    if (lane < kGroup) s_list[wave][n + lane] = kB * static_cast<uint32_t>(sizeof(float4));  // the null record
    if (quad && __ballot(live()) != 0) {
      if constexpr (QUAD) {
        static_assert(kGroup == 4 && kP == 1, "a split wave tests one group of four records per lane quad");
        const char* base = reinterpret_cast<const char*>(&s_rec[0][0]);
        const uint32_t r = lane & 3u;
        for (uint32_t i = 0; i < n; i += kGroup) {
          walked += kGroup;
          // Lane 4 p + r tests record i + r (past n: the null record padding) against pixel p.
          const uint32_t mine = s_list[wave][i + r];
          const float4* rec = reinterpret_cast<const float4*>(base + mine);
          float e;
          const bool hit = shape_test(rec[0], rec[kS], fx[0], fy[0], e);
          const float x = hit ? e : -__builtin_inff();
          // Records that no live pixel of the wave hits leave every pixel as it is: skipped.
          const uint64_t takes = __ballot(hit && T[0] >= kTransmittanceStop);
          auto colour = [&](uint32_t u) {
            const uint32_t at = __builtin_amdgcn_readlane(mine, u);
            return reinterpret_cast<const float4*>(base + at)[2 * kS];
          };
          if (takes & 0x1111111111111111ull) quad_blend<0>(x, colour(0), T[0], C0[0], C1[0], C2[0]);
          if (takes & 0x2222222222222222ull) quad_blend<1>(x, colour(1), T[0], C0[0], C1[0], C2[0]);
          if (takes & 0x4444444444444444ull) quad_blend<2>(x, colour(2), T[0], C0[0], C1[0], C2[0]);
          if (takes & 0x8888888888888888ull) quad_blend<3>(x, colour(3), T[0], C0[0], C1[0], C2[0]);
          if (__ballot(live()) == 0) break;
        }
      }
    } else if (__ballot(live()) != 0) {
      for (uint32_t i = 0; i < n; i += kGroup) {
        walked += kGroup;
        // A group of kGroup records: their geometry does not depend on the transmittance, so all of
        // them are read and tested together (one LDS latency per group); the blends then run in order.
        uint32_t addr[kGroup];
#pragma unroll
        for (uint32_t u = 0; u < kGroup; ++u) addr[u] = s_list[wave][i + u];
        float e[kGroup][kP];
        bool hit[kGroup][kP];
#pragma unroll
        for (uint32_t u = 0; u < kGroup; ++u) {
          const float4* rec =
              reinterpret_cast<const float4*>(reinterpret_cast<const char*>(&s_rec[0][0]) + addr[u]);
          const float4 A = rec[0], B = rec[kS];
#pragma unroll
          for (uint32_t q = 0; q < kP; ++q) hit[u][q] = shape_test(A, B, fx[q], fy[q], e[u][q]);
        }
#pragma unroll
        for (uint32_t u = 0; u < kGroup; ++u) {
          bool take[kP], any = false;
#pragma unroll
          for (uint32_t q = 0; q < kP; ++q) {
            take[q] = hit[u][q] && T[q] >= kTransmittanceStop;
            any |= take[q];
          }
          // A record no pixel of the wave takes leaves the wave's pixels exactly as they are: skip
          // its blend (a wave-uniform branch).
          if (__ballot(any) == 0) continue;
          // The colour is read only for a record the wave blends.
          const float4 Cc =
              reinterpret_cast<const float4*>(reinterpret_cast<const char*>(&s_rec[0][0]) + addr[u])[2 * kS];
#pragma unroll
          for (uint32_t q = 0; q < kP; ++q) blend(take[q], e[u][q], Cc, T[q], C0[q], C1[q], C2[q]);
        }
        if (__ballot(live()) == 0) break;
      }
    }
    if (__syncthreads_count(live() ? 1 : 0) == 0) break;
  }
  // The block's coverage tests, when the renderer counts them: one add per block, to one of 8 slots.
  if (a.count_tests != 0u) {
    auto& s_walked = L.walked;
    if (lane == 0) s_walked[wave] = walked;
    __syncthreads();
    if (tid == 0) {
      uint32_t sum = 0;
#pragma unroll
      for (uint32_t w = 0; w < kW; ++w) sum += s_walked[w];
      // A split wave tests 16 pixels per record, an unsplit one 64 kP.
      if (sum != 0u)
        atomicAdd(&a.ctrl->raster_tests[blockIdx.x & 7u], static_cast<unsigned long long>(sum) * (quad ? 16u : 64u * kP));
    }
  }
#pragma unroll
  for (uint32_t q = 0; q < kP; ++q) {
    if (!in_image[q] || (quad && (lane & 3u) != 0u)) continue;  // a split pixel's first lane writes it
    const uint32_t px = rx + lx + 8u * (q % PX), py = ry + ly + 8u * (q / PX);
    float4* out = reinterpret_cast<float4*>(a.image) + pixel_base + static_cast<size_t>(py) * width + px;
    *out = make_float4(C0[q] + T[q] * a.background[0], C1[q] + T[q] * a.background[1],
                       C2[q] + T[q] * a.background[2], 1.0f);
  }
}

template <int TILE, int PX>
__global__ __launch_bounds__((RasterShape<TILE, PX>::kThreads), (RasterShape<TILE, PX>::kMinWaves)) void k_raster(
    FrameArgs a, const uint32_t* __restrict__ pair_vals, const uint32_t* __restrict__ ranges, uint32_t tiles,
    uint32_t split_cap) {
  constexpr bool kSplit = TILE == 16 && PX == 1;
  __shared__ RasterLds<TILE, PX> L;
  // Blocks [0, 4 split) draw the quadrants of the first `split` tiles of the order, the next ones a
  // tile each; the grid has 3 split_cap spare blocks, which exit.
  const uint32_t split = kSplit && split_cap != 0u ? a.hist[kOrderCounters + 64u] : 0u;
  if (blockIdx.x >= tiles + 3u * split) return;
  if constexpr (kSplit) {
    if (blockIdx.x < 4u * split) {
      raster_block<TILE, PX, true>(a, pair_vals, ranges, blockIdx.x >> 2, L);
      return;
    }
  }
  raster_block<TILE, PX, false>(a, pair_vals, ranges, blockIdx.x - 3u * split, L);
}

// The raster block shapes on offer: (tile, pixels per lane side).
hipError_t launch_raster(const FrameArgs& a, const uint32_t* pair_vals, uint32_t tiles, hipStream_t s) {
  const int shape = a.tile * 4 + a.raster_px;
  switch (shape) {
    case 8 * 4 + 1:
      hipLaunchKernelGGL((k_raster<8, 1>), dim3(tiles), dim3(64), 0, s, a, pair_vals, a.tile_ranges, tiles, 0u);
      break;
    case 16 * 4 + 2:
      hipLaunchKernelGGL((k_raster<16, 2>), dim3(tiles), dim3(64), 0, s, a, pair_vals, a.tile_ranges, tiles, 0u);
      break;
    case 32 * 4 + 2:
      hipLaunchKernelGGL((k_raster<32, 2>), dim3(tiles), dim3(256), 0, s, a, pair_vals, a.tile_ranges, tiles, 0u);
      break;
    case 16 * 4 + 1: {
      // The heaviest tiles of the order are split (the order is heaviest first).
      const uint32_t cap = a.tile_order != nullptr ? std::min(tiles, a.split_cap) : 0u;
      hipLaunchKernelGGL((k_raster<16, 1>), dim3(tiles + 3u * cap), dim3(256), 0, s, a, pair_vals, a.tile_ranges,
                         tiles, cap);
      break;
    }
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

// MetalFrameCapture's BGRA8 (MetalFrameCapture.h:18) of the RGBA float composite: each channel
// clamped to [0, 1], times 255, rounded half away from zero (the host conversion's lround).
__global__ __launch_bounds__(256) void k_to_bgra8(const float4* __restrict__ rgba, uint32_t* __restrict__ out,
                                                  uint64_t pixels) {
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < pixels; i += stride) {
    const float4 c = rgba[i];
    auto q = [](float v) {
      const float x = v < 0.0f ? 0.0f : (v > 1.0f ? 1.0f : v);
      return static_cast<uint32_t>(roundf(x * 255.0f));
    };
    out[i] = q(c.z) | (q(c.y) << 8) | (q(c.x) << 16) | (q(c.w) << 24);
  }
}

template <int DEG, bool INDEXED>
void launch_cull_key(const FrameArgs& a, uint32_t blocks, hipStream_t s) {
  if (a.key16) hipLaunchKernelGGL((k_cull<DEG, INDEXED, true>), dim3(blocks), dim3(256), 0, s, a);
  else hipLaunchKernelGGL((k_cull<DEG, INDEXED, false>), dim3(blocks), dim3(256), 0, s, a);
}
template <int DEG>
void launch_cull_deg(const FrameArgs& a, uint32_t blocks, hipStream_t s) {
  if (a.active != nullptr) launch_cull_key<DEG, true>(a, blocks, s);
  else launch_cull_key<DEG, false>(a, blocks, s);
}

uint32_t grid_for(uint32_t items, uint32_t per_block) {
  const uint32_t b = (items + per_block - 1u) / per_block;
  return b == 0u ? 1u : b;
}

// The grid of a kernel whose blocks walk their work in strides of the grid: `estimate` items (0: the
// capacity) in blocks of `per`, rounded up to a multiple of the XCD count, at most the capacity's
// blocks. Scratch is sized for the largest frame (the whole world, or a tree cut's capacity), and a
// block per unit of it spends microseconds dispatching blocks that find no work; a grid from the
// counts of earlier frames keeps the hardware's in-order dispatch of the blocks that do work, and a
// frame with more than the estimate loops.
uint32_t work_grid(uint32_t capacity, uint32_t estimate, uint32_t per) {
  const uint32_t full = grid_for(capacity, per);
  if (estimate == 0u || estimate >= capacity) return full;
  const uint32_t b = (grid_for(estimate, per) + kXcds - 1u) / kXcds * kXcds;
  return std::min(b, full);
}

// Blocks of the histogram kernel: enough to stream the keys, few enough that the per-block global
// adds into each of the histogram's words stay short (each word takes one add per block).
uint32_t hist_grid(uint32_t capacity) { return std::min(std::max(grid_for(capacity, 16384u), 16u), 2048u); }

// One stable LSD pass over the digit of `bits` bits at `shift`, in the chosen variant. The kernel
// ranks with BITS >= bits ballots (the digit's missing high bits are zero in every lane); `counts`
// holds 256 words per kSortTile-key tile of `capacity` (reduce then scan only).
template <bool LOOKBACK, int BITS>
void sort_pass_bits(const uint32_t* kin, const uint32_t* vin, uint32_t* kout, uint32_t* vout,
                    const uint32_t* n_ptr, uint32_t tiles, uint32_t grid, uint32_t* hist, uint32_t shift,
                    uint32_t bits, uint64_t* status, Ctrl* ctrl, uint32_t ticket, uint32_t tag_pass,
                    uint32_t* counts, const uint32_t* seg, bool runs, bool counted, uint32_t view_keys,
                    hipStream_t s) {
  const uint32_t mask = (1u << bits) - 1u;
  if constexpr (LOOKBACK) {
    hipLaunchKernelGGL((k_onesweep<true, BITS>), dim3(tiles), dim3(256), 0, s, kin, vin, kout, vout, n_ptr, hist,
                       shift, mask, status, ctrl, ticket, tag_pass, nullptr, 0u, seg, view_keys);
  } else {
    if (counted) {
      // The producer of the keys has written this pass's tile counts.
    } else if (runs)
      hipLaunchKernelGGL(k_tile_counts<true>, dim3(grid), dim3(256), 0, s, kin, n_ptr, shift, mask, counts, tiles,
                         seg);
    else
      hipLaunchKernelGGL(k_tile_counts<false>, dim3(grid), dim3(256), 0, s, kin, n_ptr, shift, mask, counts, tiles,
                         seg);
    hipLaunchKernelGGL(k_digit_scan, dim3(1u << bits), dim3(256), 0, s, counts, n_ptr, hist, tiles);
    hipLaunchKernelGGL((k_onesweep<false, BITS>), dim3(grid), dim3(256), 0, s, kin, vin, kout, vout, n_ptr, hist,
                       shift, mask, status, ctrl, ticket, tag_pass, counts, tiles, seg, view_keys);
  }
}

template <bool LOOKBACK>
void sort_pass_variant(const uint32_t* kin, const uint32_t* vin, uint32_t* kout, uint32_t* vout,
                       const uint32_t* n_ptr, uint32_t tiles, uint32_t grid, uint32_t* hist, uint32_t shift,
                       uint32_t bits, uint64_t* status, Ctrl* ctrl, uint32_t ticket, uint32_t tag_pass,
                       uint32_t* counts, const uint32_t* seg, bool runs, bool counted, uint32_t view_keys,
                       hipStream_t s) {
  if (bits <= 6)
    sort_pass_bits<LOOKBACK, 6>(kin, vin, kout, vout, n_ptr, tiles, grid, hist, shift, bits, status, ctrl, ticket,
                                tag_pass, counts, seg, runs, counted, view_keys, s);
  else if (bits == 7)
    sort_pass_bits<LOOKBACK, 7>(kin, vin, kout, vout, n_ptr, tiles, grid, hist, shift, bits, status, ctrl, ticket,
                                tag_pass, counts, seg, runs, counted, view_keys, s);
  else
    sort_pass_bits<LOOKBACK, 8>(kin, vin, kout, vout, n_ptr, tiles, grid, hist, shift, bits, status, ctrl, ticket,
                                tag_pass, counts, seg, runs, counted, view_keys, s);
}

// Pass p of a sort under its digit plan; `seg` is the cull's gapped layout (first depth pass) or
// null. The last pass, on the key's top digit, counts runs of equal digits unless `top_runs` is
// false; `counted`: the keys' producer has written the pass's tile counts; `view_keys`: the pass
// writes the values' views as keys (k_onesweep); `estimate`: the keys expected (0: the capacity),
// which sizes the reduce-then-scan kernels' grids (the look-back variant takes a block per tile).
void sort_pass(const uint32_t* kin, const uint32_t* vin, uint32_t* kout, uint32_t* vout, const uint32_t* n_ptr,
               uint32_t capacity, uint32_t* hist, const DigitPlan& plan, int p, uint64_t* status, Ctrl* ctrl,
               uint32_t ticket, uint32_t tag_pass, uint32_t* counts, int variant, const uint32_t* seg,
               hipStream_t s, bool counted = false, bool top_runs = true, uint32_t view_keys = 0u,
               uint32_t estimate = 0u) {
  const uint32_t tiles = grid_for(capacity, kSortTile);
  const bool runs = top_runs && p + 1 == plan.passes;
  if (variant == kSortOnesweep)
    sort_pass_variant<true>(kin, vin, kout, vout, n_ptr, tiles, tiles, hist, plan.shift[p], plan.bits[p], status,
                            ctrl, ticket, tag_pass, counts, seg, runs, counted, view_keys, s);
  else
    sort_pass_variant<false>(kin, vin, kout, vout, n_ptr, tiles, work_grid(capacity, estimate, kSortTile), hist,
                             plan.shift[p], plan.bits[p], status, ctrl, ticket, tag_pass, counts, seg, runs, counted,
                             view_keys, s);
}

}  // namespace

hipError_t launch_pack(const WorldOut& out, uint32_t offset, uint32_t count, const float* positions,
                       const float* covariances, const float* colors, const float* alphas,
                       const float* sh, int src_degree, hipStream_t stream) {
  if (count == 0) return hipSuccess;
  hipLaunchKernelGGL(k_pack, dim3(grid_for(count, 256)), dim3(256), 0, stream, out, offset, count,
                     positions, covariances, colors, alphas, sh, src_degree);
  return hipGetLastError();
}

namespace {
__global__ __launch_bounds__(256) void k_morton(const float* __restrict__ pos, uint32_t n, float lo0, float lo1,
                                                float lo2, float hi0, float hi1, float hi2,
                                                uint32_t* __restrict__ codes) {
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i >= n) return;
  const float p[3] = {pos[3 * static_cast<size_t>(i)], pos[3 * static_cast<size_t>(i) + 1],
                      pos[3 * static_cast<size_t>(i) + 2]};
  const float lo[3] = {lo0, lo1, lo2}, hi[3] = {hi0, hi1, hi2};
  codes[i] = morton_code(p, lo, hi);
}

__global__ __launch_bounds__(256) void k_gather_rows(const float* __restrict__ in, float* __restrict__ out,
                                                     const uint32_t* __restrict__ order, uint32_t n,
                                                     uint32_t stride) {
  const uint64_t total = static_cast<uint64_t>(n) * stride;
  for (uint64_t e = static_cast<uint64_t>(blockIdx.x) * 256u + threadIdx.x; e < total;
       e += static_cast<uint64_t>(gridDim.x) * 256u) {
    const uint64_t row = e / stride, col = e - row * stride;
    out[e] = in[static_cast<uint64_t>(order[row]) * stride + col];
  }
}
}  // namespace

hipError_t launch_morton(const float* positions, uint32_t n, const float lo[3], const float hi[3], uint32_t* codes,
                         hipStream_t stream) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_morton, dim3(grid_for(n, 256)), dim3(256), 0, stream, positions, n, lo[0], lo[1], lo[2],
                     hi[0], hi[1], hi[2], codes);
  return hipGetLastError();
}

hipError_t launch_gather_rows(const float* in, float* out, const uint32_t* order, uint32_t n, uint32_t stride,
                              hipStream_t stream) {
  if (n == 0 || stride == 0) return hipSuccess;
  const uint64_t total = static_cast<uint64_t>(n) * stride;
  const uint32_t blocks = static_cast<uint32_t>(std::min<uint64_t>((total + 255) / 256, 8192));
  hipLaunchKernelGGL(k_gather_rows, dim3(blocks), dim3(256), 0, stream, in, out, order, n, stride);
  return hipGetLastError();
}

hipError_t launch_to_bgra8(const float* rgba, uint32_t* out, uint64_t pixels, hipStream_t stream) {
  if (pixels == 0) return hipSuccess;
  const uint32_t blocks = static_cast<uint32_t>(std::min<uint64_t>((pixels + 255) / 256, 4096));
  hipLaunchKernelGGL(k_to_bgra8, dim3(blocks), dim3(256), 0, stream, reinterpret_cast<const float4*>(rgba), out,
                     pixels);
  return hipGetLastError();
}

hipError_t launch_argsort(Ctrl* ctrl, uint32_t* keys[2], uint32_t* vals[2], const uint32_t* n_dev,
                          uint32_t capacity, const DigitPlan& plan, uint32_t* hist, uint64_t* status,
                          uint32_t* counts, int variant, hipStream_t stream) {
  const int passes = plan.passes;
  // keys[0] holds the input; passes alternate keys[0] -> keys[1] -> keys[0] ...; the order ends in
  // vals[(passes - 1) & 1].
  // The reduce-then-scan passes take each digit's total from their own tile counts and use neither
  // the control block nor the histograms; the onesweep passes need a fresh epoch, tickets and every
  // pass's histogram first.
  if (variant == kSortOnesweep) {
    hipLaunchKernelGGL(k_prologue, dim3(1), dim3(256), 0, stream, ctrl, hist, 4u * 256u, nullptr, 0u, nullptr,
                       0u, nullptr, nullptr);
    hipLaunchKernelGGL(k_hist, dim3(hist_grid(capacity)), dim3(256), 0, stream, keys[0], n_dev, nullptr, hist,
                       plan);
  }
  for (int p = 0; p < passes; ++p) {
    const uint32_t* kin = keys[p & 1];
    uint32_t* kout = (p + 1 < passes) ? keys[(p + 1) & 1] : nullptr;
    const uint32_t* vin = p == 0 ? nullptr : vals[(p - 1) & 1];
    sort_pass(kin, vin, kout, vals[p & 1], n_dev, capacity, hist + p * 256, plan, p, status, ctrl,
              kTicketDepthSort + p, 2u + p, counts, variant, nullptr, stream);
  }
  return hipGetLastError();
}

hipError_t launch_frame(const FrameArgs& a, const LodArgs* lod, const DigitPlan& depth_plan,
                        const DigitPlan& tile_plan, hipEvent_t* marks, hipStream_t s) {
  const int depth_passes = depth_plan.passes, tile_passes = tile_plan.passes;
  const uint32_t tiles = a.total_tiles;
  if (marks) (void)hipEventRecord(marks[0], s);
  if (lod != nullptr) {
    const hipError_t e = launch_lod(*lod, s);
    if (e != hipSuccess) return e;
  }
  if (marks) (void)hipEventRecord(marks[1], s);
  hipLaunchKernelGGL(k_prologue, dim3(min(grid_for(tiles * 2u, 256), 256u)), dim3(256), 0, s, a.ctrl,
                     a.hist, kHistWords, a.tile_ranges, tiles * 2u, a.view_visible,
                     a.view_visible != nullptr ? a.view_count : 0u,
                     lod != nullptr ? lod->state + lod->rounds : nullptr,
                     lod != nullptr ? &lod->ctrl->final_limited : nullptr);
  const uint32_t view_lanes = grid_for(a.total, 256u * kCullItems) * 256u * kCullItems;
  const uint32_t lanes = a.active != nullptr ? a.capacity : view_lanes * a.view_count;
  // A tree cut's cull takes a block per cull tile of the cut's capacity (the blocks past the cut's
  // lanes exit at once): a loop over tiles past an estimated grid made the kernel spill scalar
  // registers and cost more than the empty blocks.
  const uint32_t cull_blocks = grid_for(lanes, 256u * kCullItems);
  switch (a.world.sh_degree) {
    case 0: launch_cull_deg<0>(a, cull_blocks, s); break;
    case 1: launch_cull_deg<1>(a, cull_blocks, s); break;
    case 2: launch_cull_deg<2>(a, cull_blocks, s); break;
    default: launch_cull_deg<3>(a, cull_blocks, s); break;
  }
  if (marks) (void)hipEventRecord(marks[2], s);
  // Depth sort of the survivors: cull keys -> keys[0] -> keys[1] -> keys[0] ... The first pass reads
  // the cull's gapped layout (its extent and per-tile counts) and writes the survivors packed, with
  // their positions as values.
  const uint32_t* visible = &a.ctrl->visible;
  const uint32_t* cull_lanes = &a.ctrl->cull_lanes;
  if (a.sort_variant == kSortOnesweep)
    hipLaunchKernelGGL(k_hist, dim3(hist_grid(a.capacity)), dim3(256), 0, s, a.cull_keys, cull_lanes, a.cull_counts,
                       a.hist, depth_plan);
  // A batch with view-local tile keys sorts the survivors by view after depth (stable: depth order
  // within each view); the last depth pass writes the views as the keys of those passes.
  const bool view_sort = a.local_tile_bits != 0u;
  for (int p = 0; p < depth_passes; ++p) {
    const bool last = p + 1 == depth_passes;
    const uint32_t* kin = p == 0 ? a.cull_keys : a.keys[(p - 1) & 1];
    uint32_t* kout = (!last || view_sort) ? a.keys[p & 1] : nullptr;
    const uint32_t* vin = p == 0 ? nullptr : a.order[(p - 1) & 1];
    sort_pass(kin, vin, kout, a.order[p & 1], p == 0 ? cull_lanes : visible, a.capacity, a.hist + p * 256, depth_plan, p,
              a.sort_status, a.ctrl, kTicketDepthSort + p, 2u + p, a.sort_counts, a.sort_variant,
              p == 0 ? a.cull_counts : nullptr, s, false, true, last && view_sort ? a.view_count : 0u,
              p == 0 ? a.lane_estimate : a.survivor_estimate);
  }
  int survivor_passes = depth_passes;
  if (view_sort) {
    // Reduce then scan whatever the variant (the depth sort's histogram slots are free again). The
    // views of neighbouring survivors differ, so the counts add key by key.
    const DigitPlan view_plan = make_digit_plan(key_bits_for(a.view_count - 1u), 8u);
    for (int v = 0; v < view_plan.passes; ++v, ++survivor_passes) {
      const int q = survivor_passes;
      sort_pass(a.keys[(q - 1) & 1], a.order[(q - 1) & 1], v + 1 < view_plan.passes ? a.keys[q & 1] : nullptr,
                a.order[q & 1], visible, a.capacity, a.hist + v * 256, view_plan, v, a.sort_status, a.ctrl,
                kTicketDepthSort, 2u, a.sort_counts, kSortReduceScan, nullptr, s, false, false, 0u,
                a.survivor_estimate);
    }
  }
  const uint32_t* order = a.order[(survivor_passes - 1) & 1];
  if (marks) (void)hipEventRecord(marks[3], s);
  hipLaunchKernelGGL(k_bin_count, dim3(work_grid(a.capacity, a.survivor_estimate, 256u * kBinItems)), dim3(256), 0, s,
                     a, order);
  hipLaunchKernelGGL(k_bin_scan, dim3(1), dim3(256), 0, s, a);
  // With reduce-then-scan passes the emission counts the tile sort's first digits itself.
  const bool bin_counts = a.sort_variant != kSortOnesweep && tile_passes > 0;
  hipLaunchKernelGGL(k_bin, dim3(work_grid(a.pair_capacity, a.pair_estimate, kEmitChunk)), dim3(256), 0, s, a, order,
                     bin_counts ? (1u << tile_plan.bits[0]) - 1u : 0u, grid_for(a.pair_capacity, kSortTile));
  const uint32_t* sort_n = &a.ctrl->sort_pairs;
  uint32_t* tile_hist = a.hist + 4 * 256;
  if (a.sort_variant == kSortOnesweep)
    hipLaunchKernelGGL(k_hist, dim3(hist_grid(a.pair_capacity)), dim3(256), 0, s, a.pair_keys[0], sort_n, nullptr,
                       tile_hist, tile_plan);
  for (int p = 0; p < tile_passes; ++p) {
    sort_pass(a.pair_keys[p & 1], a.pair_vals[p & 1], a.pair_keys[(p + 1) & 1], a.pair_vals[(p + 1) & 1],
              sort_n, a.pair_capacity, tile_hist + p * 256, tile_plan, p, a.sort_status, a.ctrl, kTicketTileSort + p,
              6u + p, a.sort_counts, a.sort_variant, nullptr, s, p == 0 && bin_counts, true, 0u, a.pair_estimate);
  }
  const int fin = tile_passes & 1;
  hipLaunchKernelGGL(k_ranges, dim3(work_grid(a.pair_capacity, a.pair_estimate, 1024u)), dim3(256), 0, s,
                     a.pair_keys[fin], sort_n, a.tile_ranges, a.views, a.local_tile_bits);
  if (a.tile_order != nullptr) {
    uint32_t* counters = a.hist + kOrderCounters;
    hipLaunchKernelGGL(k_order_count, dim3(grid_for(tiles, 256)), dim3(256), 0, s, a.tile_ranges, tiles, counters);
    hipLaunchKernelGGL(k_order_scatter, dim3(grid_for(tiles, 256)), dim3(256), 0, s, a.tile_ranges, tiles, counters,
                       a.tile_order, a.split_cap);
  }
  if (marks) (void)hipEventRecord(marks[4], s);
  if (const hipError_t e = launch_raster(a, a.pair_vals[fin], tiles, s); e != hipSuccess) return e;
  if (marks) (void)hipEventRecord(marks[5], s);
  return hipGetLastError();
}

}  // namespace skm
