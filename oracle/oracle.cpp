// CPU oracle for the per-frame splat render path. Test infrastructure only: tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg load it as the checker; the
// product never does.
//
// Floating point: every fp32 expression below is written with an explicit evaluation order and
// this file is compiled with -ffp-contract=off and without fast-math, so each operation is one
// IEEE-754 rounding. The reference compiles its MSL with fast-math (MetalShaderLibrary.h:15-20)
// and fixes no order; the order written here is the build's specification of the reference's
// formulas, and the HIP kernels implement the same specification.
#include "oracle.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <numeric>
#include <thread>
#include <vector>

namespace {

inline float f32(uint32_t b) {
  float f;
  std::memcpy(&f, &b, 4);
  return f;
}
inline uint32_t u32(float f) {
  uint32_t b;
  std::memcpy(&b, &f, 4);
  return b;
}
// max/min with a fixed answer for NaN (the other operand) and for equal operands (the first), so
// signed zeros come out the same on every machine; the kernels use the same rule.
inline float mx(float a, float b) { return (a >= b || b != b) ? a : b; }
inline float mn(float a, float b) { return (a <= b || b != b) ? a : b; }

// Constants of the path (SplatTypes.metalh:83-110, MetalShaderTypes.h:35).
constexpr float kQuadMinAlpha = 2.0f / 255.0f;
constexpr float kBoundsRadius = 2.5f;
constexpr float kLodBoundsRadius = 5.0f;
constexpr float kMaxPixelRadius = 512.0f;
constexpr float kLowPassVariance = 0.3f;
constexpr float kProjectionFovClamp = 1.3f;
constexpr float kDisplayGamma = 2.2f;
constexpr float kGaussianScale = 0.8493218f;
constexpr float kLog2MinAlpha = -7.994353f;
constexpr float kMaxLog2Alpha = 16.0f;
constexpr float kMaxColorScale = 4.0f;
constexpr float kHalfMin = 6.103515625e-05f;
constexpr uint32_t kDepthKeyShift = 11;
constexpr uint32_t kDepthKeyMax = 65535;
// Tile binning pads the quad's bounding box by this many pixels so the fp32 coverage rule can
// never accept a pixel outside the listed tiles.
constexpr float kBinPad = 1.0f / 16.0f;
// 1 / kGaussianScale^2 = 2 ln 2: the footprint disc's radius^2 in axis units per unit of log2 alpha
// above the cutoff, and the slack its bound carries over the fp32 rounding of the coverage rule.
constexpr float kInvGaussianScale2 = 1.38629436f;
constexpr float kFootprintSlack = 1.0009765625f;
constexpr double kLn2 = 0.69314718055994530942;
constexpr double kLn2Hi = 6.93147180369123816490e-01;
constexpr double kLn2Lo = 1.90821492927058770002e-10;

// ---------------------------------------------------------------------------------------------
// Transcendentals from basic IEEE double operations (frexp/ldexp are exact).
// log: x = m 2^e with m in [sqrt(1/2), sqrt(2)), log m = 2 atanh(s), s = (m-1)/(m+1), summed as
// the odd series of atanh (Abramowitz & Stegun 4.5.64); |s| <= 0.1716, 12 terms reach 1e-18.
// exp: x = k ln2 + r, |r| <= ln2/2, e^r by its Taylor series to r^20/20!, then scaled by 2^k.
double log_d(double x) {
  int e = 0;
  double m = std::frexp(x, &e);
  if (m < 0.70710678118654752440) {
    m = m * 2.0;
    e = e - 1;
  }
  const double s = (m - 1.0) / (m + 1.0);
  const double z = s * s;
  double p = 1.0 / 23.0;
  p = 1.0 / 21.0 + z * p;
  p = 1.0 / 19.0 + z * p;
  p = 1.0 / 17.0 + z * p;
  p = 1.0 / 15.0 + z * p;
  p = 1.0 / 13.0 + z * p;
  p = 1.0 / 11.0 + z * p;
  p = 1.0 / 9.0 + z * p;
  p = 1.0 / 7.0 + z * p;
  p = 1.0 / 5.0 + z * p;
  p = 1.0 / 3.0 + z * p;
  p = 1.0 + z * p;
  return static_cast<double>(e) * kLn2 + 2.0 * s * p;
}

double exp_d(double x) {
  if (x < -1000.0) return 0.0;
  if (x > 1000.0) return HUGE_VAL;
  const double k = std::nearbyint(x / kLn2);
  // Cody-Waite: ln2 split so that k * kLn2Hi is exact for |k| < 2^20.
  const double r = (x - k * kLn2Hi) - k * kLn2Lo;
  double p = 1.0;
  for (int n = 20; n >= 1; --n) p = 1.0 + (r / static_cast<double>(n)) * p;
  return std::ldexp(p, static_cast<int>(k));
}

// Natural logarithm in fp32 with a fixed operation order (the kernels' log_f): x = m 2^e,
// m in [sqrt(1/2), sqrt(2)), log m = 2 atanh(s), s = (m-1)/(m+1), by the odd series of atanh to
// s^11 (Abramowitz & Stegun 4.5.64), e ln 2 with a split constant. For x normal and positive.
constexpr float kLn2HiF = 0.693145751953125f;
constexpr float kLn2LoF = 1.428606765330187e-06f;
constexpr float kInvLn2F = 1.44269504088896341f;
float log_f(float x) {
  int e = 0;
  float m = std::frexp(x, &e);
  if (m < 0.707106781f) {
    m = m * 2.0f;
    e = e - 1;
  }
  const float s = (m - 1.0f) / (m + 1.0f);
  const float z = s * s;
  float p = 0.0909090936f;
  p = 0.111111112f + z * p;
  p = 0.142857149f + z * p;
  p = 0.200000003f + z * p;
  p = 0.333333343f + z * p;
  p = 1.0f + z * p;
  const float fe = static_cast<float>(e);
  return fe * kLn2HiF + (fe * kLn2LoF + (2.0f * s) * p);
}

float pow_f(float x, float y) {
  if (!(x > 0.0f)) return 0.0f;
  return static_cast<float>(exp_d(static_cast<double>(y) * log_d(static_cast<double>(x))));
}

// ---------------------------------------------------------------------------------------------
// Halves.

// Saturating binary16 with round to nearest even: the host packer's rule (Half.h:12-38).
uint16_t to_half_sat(float value) {
  const uint32_t bits = u32(value);
  const uint32_t sign = (bits >> 16) & 0x8000u;
  const uint32_t biased = (bits >> 23) & 0xffu;
  uint32_t mant = bits & 0x7fffffu;
  if (biased == 0xffu) return static_cast<uint16_t>(sign | (mant ? 0x7e00u : 0x7bffu));
  const int32_t exp = static_cast<int32_t>(biased) - 112;  // rebias 127 -> 15
  if (exp >= 31) return static_cast<uint16_t>(sign | 0x7bffu);
  if (exp <= 0) {
    if (exp < -10) return static_cast<uint16_t>(sign);
    mant |= 0x800000u;
    const uint32_t shift = static_cast<uint32_t>(14 - exp);
    const uint32_t q = mant >> shift;
    const uint32_t rem = mant & ((1u << shift) - 1u);
    const uint32_t mid = 1u << (shift - 1);
    const uint32_t up = (rem > mid || (rem == mid && (q & 1u))) ? 1u : 0u;
    return static_cast<uint16_t>(sign | (q + up));
  }
  uint32_t h = (static_cast<uint32_t>(exp) << 10) | (mant >> 13);
  const uint32_t rem = mant & 0x1fffu;
  if (rem > 0x1000u || (rem == 0x1000u && (h & 1u))) ++h;
  if (h >= 0x7c00u) h = 0x7bffu;
  return static_cast<uint16_t>(sign | h);
}

// IEEE binary16 conversion with round to nearest even and overflow to infinity: the shader's
// half(float) and half2(float2) (SplatProjection.metalh:252-255).
uint16_t half_rne(float value) {
  const uint32_t bits = u32(value);
  const uint32_t sign = (bits >> 16) & 0x8000u;
  const uint32_t biased = (bits >> 23) & 0xffu;
  uint32_t mant = bits & 0x7fffffu;
  if (biased == 0xffu) return static_cast<uint16_t>(sign | (mant ? 0x7e00u : 0x7c00u));
  const int32_t exp = static_cast<int32_t>(biased) - 112;
  if (exp >= 31) return static_cast<uint16_t>(sign | 0x7c00u);
  if (exp <= 0) {
    if (exp < -10) return static_cast<uint16_t>(sign);
    mant |= 0x800000u;
    const uint32_t shift = static_cast<uint32_t>(14 - exp);
    const uint32_t q = mant >> shift;
    const uint32_t rem = mant & ((1u << shift) - 1u);
    const uint32_t mid = 1u << (shift - 1);
    const uint32_t up = (rem > mid || (rem == mid && (q & 1u))) ? 1u : 0u;
    return static_cast<uint16_t>(sign | (q + up));
  }
  uint32_t h = (static_cast<uint32_t>(exp) << 10) | (mant >> 13);
  const uint32_t rem = mant & 0x1fffu;
  if (rem > 0x1000u || (rem == 0x1000u && (h & 1u))) ++h;  // may carry into infinity: IEEE
  return static_cast<uint16_t>(sign | h);
}

// Half.h:40-65.
float from_half(uint16_t half) {
  const uint32_t sign = (static_cast<uint32_t>(half) & 0x8000u) << 16;
  uint32_t exp = (half >> 10) & 0x1fu;
  uint32_t mant = half & 0x3ffu;
  uint32_t bits;
  if (exp == 0) {
    if (mant == 0) {
      bits = sign;
    } else {
      exp = 1;
      while ((mant & 0x400u) == 0) {
        mant <<= 1;
        --exp;
      }
      mant &= 0x3ffu;
      bits = sign | ((exp + 112) << 23) | (mant << 13);
    }
  } else if (exp == 31) {
    bits = sign | 0x7f800000u | (mant << 13);
  } else {
    bits = sign | ((exp + 112) << 23) | (mant << 13);
  }
  return f32(bits);
}

inline float lo_half(uint32_t w) { return from_half(static_cast<uint16_t>(w & 0xffffu)); }
inline float hi_half(uint32_t w) { return from_half(static_cast<uint16_t>(w >> 16)); }
inline uint32_t pack_rne(float a, float b) {
  return static_cast<uint32_t>(half_rne(a)) | (static_cast<uint32_t>(half_rne(b)) << 16);
}
inline float round_half(float v) { return from_half(half_rne(v)); }

uint32_t sh_coefficients(int degree) { return static_cast<uint32_t>((degree + 1) * (degree + 1) - 1); }

// ---------------------------------------------------------------------------------------------
// Projection (SplatProjection.metalh:78-141 with :6-24, :26-37, :50-76).

struct Proj {
  uint32_t w[8];
};

inline void mat_vec(const float* m, float x, float y, float z, float w, float out[4]) {
  for (int r = 0; r < 4; ++r) {
    out[r] = ((m[r] * x + m[4 + r] * y) + m[8 + r] * z) + m[12 + r] * w;
  }
}

// shColor (SplatProjection.metalh:50-76), one channel.
float sh_channel(const uint32_t* words, int degree, int ch, float x, float y, float z) {
  auto coef = [&](uint32_t k) {
    const uint32_t h = k * 3u + static_cast<uint32_t>(ch);
    const uint32_t word = words[h / 2u];
    return (h & 1u) == 0u ? lo_half(word) : hi_half(word);
  };
  const float C1 = 0.4886025119f;
  const float C2_0 = 1.0925484306f, C2_1 = -1.0925484306f, C2_2 = 0.3153915653f,
              C2_3 = -1.0925484306f, C2_4 = 0.5462742153f;
  const float C3_0 = -0.5900435899f, C3_1 = 2.8906114426f, C3_2 = -0.4570457995f,
              C3_3 = 0.3731763326f, C3_4 = -0.4570457995f, C3_5 = 1.4453057213f,
              C3_6 = -0.5900435899f;
  float c = ((-C1 * y) * coef(0) + (C1 * z) * coef(1)) - (C1 * x) * coef(2);
  if (degree >= 2) {
    const float xx = x * x, yy = y * y, zz = z * z, xy = x * y, yz = y * z, xz = x * z;
    const float t0 = (C2_0 * xy) * coef(3);
    const float t1 = (C2_1 * yz) * coef(4);
    const float t2 = (C2_2 * ((2.0f * zz - xx) - yy)) * coef(5);
    const float t3 = (C2_3 * xz) * coef(6);
    const float t4 = (C2_4 * (xx - yy)) * coef(7);
    c = c + ((((t0 + t1) + t2) + t3) + t4);
    if (degree >= 3) {
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
void reveal_glow(const orc_camera& cam, float height, Proj& p) {
  if (!(cam.reveal[1] > 0.0f)) return;
  float glow = mn(mx(1.0f - (cam.reveal[0] - height) / cam.reveal[1], 0.0f), 1.0f);
  glow = glow * glow;
  if (!(glow > 0.0f)) return;
  float rgb[3] = {lo_half(p.w[4]), hi_half(p.w[4]), lo_half(p.w[5])};
  const float a = hi_half(p.w[5]);
  const float t = cam.reveal[2] * glow;
  const float boost = 1.0f + cam.reveal[3] * glow;
  for (int c = 0; c < 3; ++c) {
    float v = rgb[c];
    if (cam.output_linear == 1u) v = pow_f(v, 1.0f / kDisplayGamma);
    v = (v + (cam.reveal_tint[c] - v) * t) * boost;
    if (cam.output_linear == 1u) v = pow_f(v, kDisplayGamma);
    rgb[c] = v;
  }
  p.w[4] = pack_rne(rgb[0], rgb[1]);
  p.w[5] = pack_rne(rgb[2], a);
}

// projectSplat. Returns true when visible; view-space z is reported for the depth key.
bool project(const orc_camera& cam, const uint32_t* s, const uint32_t* shw, int degree,
             uint32_t index, Proj& out, float* view_z) {
  const float px = f32(s[0]), py = f32(s[1]), pz = f32(s[2]);
  float v[4];
  mat_vec(cam.view, px, py, pz, 1.0f, v);
  *view_z = v[2];
  if (!(cam.reveal[0] - py >= 0.0f)) return false;
  const uint32_t rgba = s[3];
  const float cr = static_cast<float>(rgba & 0xffu) / 255.0f;
  const float cg = static_cast<float>((rgba >> 8) & 0xffu) / 255.0f;
  const float cb = static_cast<float>((rgba >> 16) & 0xffu) / 255.0f;
  const float ca = static_cast<float>(rgba >> 24) / 255.0f;
  const bool lod = s[7] != 0u;
  const float alpha = lod ? f32(s[7]) : ca;
  if (alpha < kQuadMinAlpha) return false;
  if (v[2] >= 0.0f) return false;
  float clip[4];
  mat_vec(cam.proj, v[0], v[1], v[2], v[3], clip);
  if (clip[2] < 0.0f || clip[2] > clip[3]) return false;
  if (clip[3] <= cam.proj[14] / cam.proj[10]) return false;

  // projectCovariance (:6-24).
  const float xx = lo_half(s[4]), xy = hi_half(s[4]), xz = lo_half(s[5]);
  const float yy = hi_half(s[5]), yz = lo_half(s[6]), zz = hi_half(s[6]);
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
  const float* V = cam.view;  // W(r,c) = V[c*4 + r]
  float T[2][3];
  for (int c = 0; c < 3; ++c) {
    T[0][c] = j00 * V[c * 4 + 0] + j02 * V[c * 4 + 2];
    T[1][c] = j11 * V[c * 4 + 1] + j12 * V[c * 4 + 2];
  }
  const float S[3][3] = {{xx, xy, xz}, {xy, yy, yz}, {xz, yz, zz}};
  float M[2][3];
  for (int r = 0; r < 2; ++r)
    for (int c = 0; c < 3; ++c) M[r][c] = (T[r][0] * S[0][c] + T[r][1] * S[1][c]) + T[r][2] * S[2][c];
  const float c00 = (M[0][0] * T[0][0] + M[0][1] * T[0][1]) + M[0][2] * T[0][2];
  const float c10 = (M[1][0] * T[0][0] + M[1][1] * T[0][1]) + M[1][2] * T[0][2];
  const float c11 = (M[1][0] * T[1][0] + M[1][1] * T[1][1]) + M[1][2] * T[1][2];
  const float a = c00 + kLowPassVariance;
  const float b = c10;
  const float d = c11 + kLowPassVariance;

  // ellipseAxes (:26-37).
  const float det = a * d - b * b;
  const float mean = 0.5f * (a + d);
  const float dist = mx(0.1f, std::sqrt(mx(mean * mean - det, 0.0f)));
  const float l1 = mean + dist;
  const float l2 = mean - dist;
  float e1x, e1y;
  if (b == 0.0f) {
    e1x = a > d ? 1.0f : 0.0f;
    e1y = a > d ? 0.0f : 1.0f;
  } else {
    const float nx = b, ny = d - l2;
    const float inv = 1.0f / std::sqrt(nx * nx + ny * ny);
    e1x = nx * inv;
    e1y = ny * inv;
  }
  const float e2x = e1y, e2y = -e1x;
  const float s1 = std::sqrt(l1);
  const float s2 = std::sqrt(mx(l2, 0.0f));
  const float a1x = e1x * s1, a1y = e1y * s1;
  const float a2x = e2x * s2, a2y = e2y * s2;

  // Quad radius in sigma (:104-108).
  const float ratio = mx(alpha / kQuadMinAlpha, 1.0f);
  const float lg = log_f(ratio);
  float radius = mn(lod ? kLodBoundsRadius : kBoundsRadius, std::sqrt(2.0f * lg));
  const float sigma = std::sqrt(mx(a1x * a1x + a1y * a1y, a2x * a2x + a2y * a2y));
  radius = mn(radius, kMaxPixelRadius / mx(sigma, 1e-6f));

  // Footprint on the half-packed axes (:112-116).
  const float extx = radius * (std::fabs(round_half(a1x)) + std::fabs(round_half(a2x)));
  const float exty = radius * (std::fabs(round_half(a1y)) + std::fabs(round_half(a2y)));
  const float marg_x = (((extx + 0.5f) * 2.0f) * clip[3]) / cam.screen_size[0];
  const float marg_y = (((exty + 0.5f) * 2.0f) * clip[3]) / cam.screen_size[1];
  if (std::fabs(clip[0]) > clip[3] + marg_x || std::fabs(clip[1]) > clip[3] + marg_y) return false;

  // Sub-pixel cutoff on the source covariance (:119-124); NaN fails.
  const float hd = 0.5f * (a - d);
  const float sv = (0.5f * (a + d) + std::sqrt(mx(hd * hd + b * b, 0.0f))) - kLowPassVariance;
  if (!(std::sqrt(mx(sv, 0.0f)) * radius >= cam.min_pixel_radius)) return false;

  // Colour (:126-131).
  float rgb[3] = {cr, cg, cb};
  if (degree >= 1) {
    const float dx = px - cam.camera_position[0];
    const float dy = py - cam.camera_position[1];
    const float dz = pz - cam.camera_position[2];
    const float inv = 1.0f / std::sqrt((dx * dx + dy * dy) + dz * dz);
    const float nx = dx * inv, ny = dy * inv, nz = dz * inv;
    for (int c = 0; c < 3; ++c) rgb[c] = mx(rgb[c] + sh_channel(shw, degree, c, nx, ny, nz), 0.0f);
  }
  if (cam.output_linear == 1u)
    for (int c = 0; c < 3; ++c) rgb[c] = pow_f(rgb[c], kDisplayGamma);

  out.w[0] = u32(clip[0] / clip[3]);
  out.w[1] = u32(clip[1] / clip[3]);
  out.w[2] = pack_rne(a1x, a1y);
  out.w[3] = pack_rne(a2x, a2y);
  out.w[4] = pack_rne(rgb[0], rgb[1]);
  out.w[5] = pack_rne(rgb[2], alpha);
  out.w[6] = u32(radius);
  out.w[7] = index;
  return true;
}

// quantizedDepthKey (SplatVisibility.metal:21-28).
uint32_t depth_key16(const orc_camera& cam, float depth) {
  const float zNear = cam.proj[14] / cam.proj[10];
  const uint32_t steps = u32(mx(depth, zNear)) - u32(zNear);
  return std::min(steps >> kDepthKeyShift, kDepthKeyMax);
}

// One visibility lane (SplatVisibility.metal:61-93): projection, reveal glow, key.
bool visibility_lane(const orc_camera& cam, const uint32_t* splats, const uint32_t* sh,
                     int degree, uint32_t index, int key16, uint32_t& key, Proj& p) {
  const uint32_t stride = orc_sh_stride(degree);
  const uint32_t* s = splats + size_t{index} * 8u;
  float vz = 0.0f;
  bool visible = project(cam, s, degree > 0 ? sh + size_t{index} * stride : nullptr, degree, index, p, &vz);
  if (visible) reveal_glow(cam, f32(s[1]), p);
  const float depth = -vz;
  key = key16 ? depth_key16(cam, depth) : u32(depth);
  return visible && std::isfinite(depth);
}

// ---------------------------------------------------------------------------------------------
// Raster rule. Per record: centre in pixels, the inverse of the half-packed axis matrix scaled by
// kGaussianScale (so the quad coordinates come out in units where the Gaussian's log2 falloff is
// their squared length), the radius in the same units, log2 of the half peak opacity (clamped like
// splatDepth, SplatProjection.metalh:169-171) and the colour clamped to packColor's range
// [0, kMaxColorScale] (:146-149).
struct Prep {
  float cx, cy, g00, g01, g10, g11, rk, log2a0;
  // The quad coordinates are affine in the pixel centre: gu = g00 fx + h01 fy + u0 (likewise gv).
  float h01, h11, u0, v0;
  float c[3];
  // FP64 copies for the oracle's values.
  double cx64, cy64, a1x, a1y, a2x, a2y, a0;
};

float log2_peak(const uint32_t* p) {
  const float a0 = hi_half(p[5]);
  const float l2 = log_f(mx(a0, kHalfMin)) * kInvLn2F;
  return mn(mx(l2, -kMaxLog2Alpha), kMaxLog2Alpha);
}

Prep prep(const uint32_t* p, uint32_t width, uint32_t height) {
  Prep q;
  const float ncx = f32(p[0]), ncy = f32(p[1]);
  q.cx = (ncx * 0.5f + 0.5f) * static_cast<float>(width);
  q.cy = (0.5f - ncy * 0.5f) * static_cast<float>(height);
  const float a1x = lo_half(p[2]), a1y = hi_half(p[2]);
  const float a2x = lo_half(p[3]), a2y = hi_half(p[3]);
  const float det = a1x * a2y - a2x * a1y;
  const float s = kGaussianScale / det;
  q.g00 = a2y * s;
  q.g01 = -(a2x * s);
  q.g10 = -(a1y * s);
  q.g11 = a1x * s;
  q.h01 = -q.g01;
  q.h11 = -q.g11;
  q.u0 = std::fmaf(q.g01, q.cy, -(q.g00 * q.cx));
  q.v0 = std::fmaf(q.g11, q.cy, -(q.g10 * q.cx));
  q.rk = kGaussianScale * f32(p[6]);
  q.log2a0 = log2_peak(p);
  const float a0 = hi_half(p[5]);
  const float col[3] = {lo_half(p[4]), hi_half(p[4]), lo_half(p[5])};
  for (int c = 0; c < 3; ++c) q.c[c] = mn(mx(col[c], 0.0f), kMaxColorScale);
  q.cx64 = (static_cast<double>(ncx) * 0.5 + 0.5) * width;
  q.cy64 = (0.5 - static_cast<double>(ncy) * 0.5) * height;
  q.a1x = a1x;
  q.a1y = a1y;
  q.a2x = a2x;
  q.a2y = a2y;
  q.a0 = a0;
  return q;
}

// The fp32 coverage rule: the pixel centre lies in the quad c + (u a1 + v a2), |u|,|v| <= r
// (expandQuad, SplatProjection.metalh:155-165) and the Gaussian's log2 alpha is at or above
// kLog2MinAlpha (splatFragmentUnder, SplatRaster.metal:40-44). In scaled units gu = K u,
// gv = K v: |gu|,|gv| <= K r and log2 alpha = log2a0 - gu^2 - gv^2. With dx = fx - cx and
// dy = cy - fy, gu = g00 dx + g01 dy = g00 fx + h01 fy + u0: two fused multiply-adds per
// coordinate from the pixel centre (fx, fy), and two for log2 alpha.
inline bool decide(const Prep& q, uint32_t px, uint32_t py, float* power) {
  const float fx = static_cast<float>(px) + 0.5f;
  const float fy = static_cast<float>(py) + 0.5f;
  const float gu = std::fmaf(q.g00, fx, std::fmaf(q.h01, fy, q.u0));
  const float gv = std::fmaf(q.g10, fx, std::fmaf(q.h11, fy, q.v0));
  *power = std::fmaf(-gu, gu, std::fmaf(-gv, gv, q.log2a0));
  if (!(std::fabs(gu) <= q.rk && std::fabs(gv) <= q.rk)) return false;
  return *power >= kLog2MinAlpha;
}

// Ideal alpha at the pixel centre in FP64 over the same packed record.
inline double alpha64(const Prep& q, uint32_t px, uint32_t py) {
  const double dx = (px + 0.5) - q.cx64;
  const double dy = q.cy64 - (py + 0.5);
  const double det = q.a1x * q.a2y - q.a2x * q.a1y;
  const double u = (dx * q.a2y - q.a2x * dy) / det;
  const double v = (q.a1x * dy - dx * q.a1y) / det;
  const double a = q.a0 * std::exp(-0.5 * (u * u + v * v));
  return a < 1.0 ? a : 1.0;
}

template <typename F>
void parallel_for(uint32_t n, uint32_t threads, F&& body) {
  if (threads <= 1 || n < 2) {
    body(0u, n);
    return;
  }
  threads = std::min(threads, n);
  std::vector<std::thread> pool;
  for (uint32_t t = 0; t < threads; ++t) {
    const uint32_t lo = static_cast<uint32_t>((uint64_t{n} * t) / threads);
    const uint32_t hi = static_cast<uint32_t>((uint64_t{n} * (t + 1)) / threads);
    pool.emplace_back([&body, lo, hi] { body(lo, hi); });
  }
  for (auto& th : pool) th.join();
}

template <bool kIdeal>
void raster_tiles(const uint32_t* projected, const uint32_t* pair_values, const uint32_t* ranges,
                  uint32_t width, uint32_t height, uint32_t tile, const float* bg, float t_stop,
                  uint32_t threads, float* out) {
  const uint32_t tiles_x = (width + tile - 1) / tile;
  const uint32_t tiles_y = (height + tile - 1) / tile;
  parallel_for(tiles_x * tiles_y, threads, [&](uint32_t lo, uint32_t hi) {
    std::vector<Prep> preps;
    for (uint32_t t = lo; t < hi; ++t) {
      const uint32_t start = ranges[2 * t], end = ranges[2 * t + 1];
      preps.clear();
      for (uint32_t j = start; j < end; ++j)
        preps.push_back(prep(projected + size_t{pair_values[j]} * 8u, width, height));
      const uint32_t x0 = (t % tiles_x) * tile, y0 = (t / tiles_x) * tile;
      for (uint32_t py = y0; py < std::min(y0 + tile, height); ++py) {
        for (uint32_t px = x0; px < std::min(x0 + tile, width); ++px) {
          float* o = out + (size_t{py} * width + px) * 4u;
          if (kIdeal) {
            double T = 1.0, C[3] = {0, 0, 0};
            for (const Prep& q : preps) {
              float power;
              if (!decide(q, px, py, &power)) continue;
              const double a = alpha64(q, px, py);
              for (int c = 0; c < 3; ++c) C[c] += T * a * q.c[c];
              T *= (1.0 - a);
            }
            for (int c = 0; c < 3; ++c) o[c] = static_cast<float>(C[c] + T * bg[c]);
          } else {
            float T = 1.0f, C[3] = {0, 0, 0};
            for (const Prep& q : preps) {
              float power;
              if (!decide(q, px, py, &power)) continue;
              const float a = std::exp2(mn(power, 0.0f));
              const float w = T * a;
              for (int c = 0; c < 3; ++c) C[c] = std::fmaf(w, q.c[c], C[c]);
              T = T - w;
              if (T < t_stop) break;
            }
            for (int c = 0; c < 3; ++c) o[c] = C[c] + T * bg[c];
          }
          o[3] = 1.0f;
        }
      }
    }
  });
}

// A LodCluster (LodTree.h:28-35) as 16 words.
struct Cluster {
  float center[3], radius, extent[3], error, color_variance, opacity;
  uint32_t node, child_start, child_count, leaf_start, leaf_count, subtree_leaves;
};
static_assert(sizeof(Cluster) == 64, "cluster layout");

// lodOutside (SplatLOD.metal:26-32): sphere against the frustum's side planes, with a guard band
// of two pixels at the cluster's depth.
bool lod_outside(const orc_camera& cam, const Cluster& c) {
  float v[4];
  mat_vec(cam.view, c.center[0], c.center[1], c.center[2], 1.0f, v);
  const float z = -v[2], r = c.radius;
  const float tx = cam.tan_half_fov[0], ty = cam.tan_half_fov[1];
  const float guard = (2.0f * mx(z, 0.0f)) / mx(mn(cam.focal[0], cam.focal[1]), 1.0f);
  return z + r <= 0.0f || std::fabs(v[0]) - z * tx > r * std::sqrt(1.0f + tx * tx) + guard ||
         std::fabs(v[1]) - z * ty > r * std::sqrt(1.0f + ty * ty) + guard;
}

// lodErrorPixels (SplatLOD.metal:33-43).
float lod_error_pixels(const orc_camera& cam, const Cluster& c, float color_weight) {
  float v[4];
  mat_vec(cam.view, c.center[0], c.center[1], c.center[2], 1.0f, v);
  const float depth = mx(-v[2] - c.radius, 1e-4f);
  const float perspective = std::sqrt(1.0f + (v[0] * v[0] + v[1] * v[1]) / (depth * depth));
  const float appearance = mn(std::sqrt(c.color_variance), 1.0f);
  const float error = c.error + c.radius * appearance;
  return ((((error * mx(cam.focal[0], cam.focal[1])) / depth) * perspective) * mx(c.opacity, 0.1f)) *
         (1.0f + color_weight * std::sqrt(c.color_variance));
}

}  // namespace

extern "C" {

int orc_lod_cluster(const orc_camera* cam, const uint32_t* cluster16, float color_weight, float* error_out) {
  Cluster c;
  std::memcpy(&c, cluster16, 64);
  if (error_out != nullptr) *error_out = lod_error_pixels(*cam, c, color_weight);
  return lod_outside(*cam, c) ? 1 : 0;
}

// Rounds of evaluate, budget, compact, commit and scatter (SplatLOD.metal:54-208) over the frontier
// in order, then the packet emission (:210-220). The budget keeps every refinement of the longest
// frontier prefix whose extra cost fits (:116-132) and denies the rest (:148-151).
uint32_t orc_lod_select(const orc_camera* cam, const uint32_t* clusters16, uint32_t cluster_count,
                        const uint32_t* leaves, uint32_t rounds, uint32_t capacity, float pixel_limit,
                        float color_weight, int cull, uint32_t* indices, uint32_t* stats) {
  constexpr uint32_t kSplit = 0x80000000u, kDrop = 0x40000000u;
  const Cluster* nodes = reinterpret_cast<const Cluster*>(clusters16);
  std::vector<uint32_t> frontier;
  if (cluster_count > 0) frontier.push_back(0);
  struct Packet {
    uint32_t leaf_start, count, dest;
  };
  std::vector<Packet> packets;
  uint32_t count = 0, limited = 0, evaluated = 0, accepted = 0;
  for (uint32_t level = 0; level < rounds; ++level) {
    const uint32_t active = static_cast<uint32_t>(frontier.size());
    std::vector<uint32_t> flags(active, 0), extra(active, 0);
    uint32_t drops = 0;
    uint64_t total = 0;
    for (uint32_t t = 0; t < active; ++t) {
      const Cluster& node = nodes[frontier[t]];
      if (cull && lod_outside(*cam, node)) {
        flags[t] = kDrop;
        ++drops;
      } else if (node.child_count + node.leaf_count > 1 &&
                 (pixel_limit == 0.0f || lod_error_pixels(*cam, node, color_weight) > pixel_limit)) {
        flags[t] = kSplit;
        extra[t] = node.child_count + node.leaf_count - 1;
      } else if (node.child_count == 1 && node.leaf_count == 0) {
        flags[t] = kSplit;
      }
      total += extra[t];
    }
    const uint32_t available = capacity - count - active + drops;
    std::vector<uint64_t> end(active);
    uint64_t run = 0;
    for (uint32_t t = 0; t < active; ++t) end[t] = (run += extra[t]);
    if (total <= available) {
      accepted = static_cast<uint32_t>(total);
    } else {
      uint32_t low = 0;
      while (low < active && end[low] <= available) ++low;
      accepted = low == 0 ? 0u : static_cast<uint32_t>(end[low - 1]);
    }
    std::vector<uint32_t> next;
    uint32_t emitted = 0;
    for (uint32_t t = 0; t < active; ++t) {
      uint32_t f = flags[t];
      if ((f & kSplit) && extra[t] > 0 && end[t] > accepted) {
        f = 0;
        ++limited;
      }
      if (f & kDrop) continue;
      const Cluster& node = nodes[frontier[t]];
      const bool split = (f & kSplit) != 0;
      if (split)
        for (uint32_t k = 0; k < node.child_count; ++k) next.push_back(node.child_start + k);
      const uint32_t splats = split ? node.leaf_count : 1u;
      const uint32_t dest = count + emitted;
      if (splats > 32) {
        packets.push_back({node.leaf_start, splats, dest});
      } else if (!split) {
        indices[dest] = node.node;
      } else {
        for (uint32_t k = 0; k < splats; ++k) indices[dest + k] = leaves[node.leaf_start + k];
      }
      emitted += splats;
    }
    evaluated += active;
    count += emitted;
    frontier.swap(next);
  }
  for (const Packet& p : packets)
    for (uint32_t k = 0; k < p.count; ++k) indices[p.dest + k] = leaves[p.leaf_start + k];
  if (stats != nullptr) {
    stats[0] = count;
    stats[1] = static_cast<uint32_t>(frontier.size());
    stats[2] = static_cast<uint32_t>(packets.size());
    stats[3] = accepted;
    stats[4] = limited;
    stats[5] = evaluated;
  }
  return count;
}

double orc_log_d(double x) { return log_d(x); }
float orc_log_f(float x) { return log_f(x); }
double orc_exp_d(double x) { return exp_d(x); }
uint16_t orc_to_half(float v) { return to_half_sat(v); }
uint16_t orc_half_rne(float v) { return half_rne(v); }
float orc_from_half(uint16_t h) { return from_half(h); }

// packSplatRange + packRgba8 (GpuLayout.cpp:22-27,122-137).
void orc_pack_splats(const float* positions, const float* covariances, const float* colors,
                     const float* alphas, uint32_t n, uint32_t* out) {
  auto q = [](float v) {
    return static_cast<uint32_t>(std::lround(std::clamp(v, 0.0f, 1.0f) * 255.0f));
  };
  for (uint32_t i = 0; i < n; ++i) {
    uint32_t* g = out + size_t{i} * 8u;
    std::memcpy(g, positions + size_t{i} * 3u, 12);
    const float a = alphas[i];
    const float* c = colors + size_t{i} * 3u;
    g[3] = q(c[0]) | (q(c[1]) << 8) | (q(c[2]) << 16) | (q(a) << 24);
    const float* v = covariances + size_t{i} * 6u;
    g[4] = static_cast<uint32_t>(to_half_sat(v[0])) | (static_cast<uint32_t>(to_half_sat(v[1])) << 16);
    g[5] = static_cast<uint32_t>(to_half_sat(v[2])) | (static_cast<uint32_t>(to_half_sat(v[3])) << 16);
    g[6] = static_cast<uint32_t>(to_half_sat(v[4])) | (static_cast<uint32_t>(to_half_sat(v[5])) << 16);
    g[7] = a > 1.0f ? u32(a) : 0u;
  }
}

// shStride (GpuLayout.cpp:66-70).
uint32_t orc_sh_stride(int degree) {
  if (degree <= 0) return 0;
  return (sh_coefficients(degree) * 3u + 1u) / 2u;
}

// packShRange, float16 storage (GpuLayout.cpp:92-113).
void orc_pack_sh(const float* sh, uint32_t n, int src_degree, int degree, uint32_t* out) {
  if (degree <= 0) return;
  const size_t src = size_t{sh_coefficients(src_degree)} * 3u;
  const uint32_t halves = sh_coefficients(degree) * 3u;
  const uint32_t stride = orc_sh_stride(degree);
  for (uint32_t i = 0; i < n; ++i) {
    const float* s = sh + size_t{i} * src;
    for (uint32_t w = 0; w < stride; ++w) {
      const uint32_t h = w * 2u;
      const uint32_t lo = h < halves ? to_half_sat(s[h]) : 0u;
      const uint32_t hi = h + 1 < halves ? to_half_sat(s[h + 1]) : 0u;
      out[size_t{i} * stride + w] = lo | (hi << 16);
    }
  }
}

// MetalFrameCompute::updateCamera (MetalFrameCompute.mm:141-164).
void orc_make_camera(const float* view16, const float* proj16, uint32_t width, uint32_t height,
                     int output_linear, float min_pixel_radius, const float* camera_position3,
                     const float* reveal8, orc_camera* out) {
  orc_camera u;
  std::memset(&u, 0, sizeof(u));
  std::memcpy(u.view, view16, 64);
  std::memcpy(u.proj, proj16, 64);
  u.screen_size[0] = static_cast<float>(width);
  u.screen_size[1] = static_cast<float>(height);
  u.focal[0] = u.screen_size[0] * proj16[0] / 2;
  u.focal[1] = u.screen_size[1] * proj16[5] / 2;
  u.tan_half_fov[0] = 1 / proj16[0];
  u.tan_half_fov[1] = 1 / proj16[5];
  u.output_linear = output_linear ? 1u : 0u;
  u.min_pixel_radius = min_pixel_radius;
  u.camera_position[0] = camera_position3[0];
  u.camera_position[1] = camera_position3[1];
  u.camera_position[2] = camera_position3[2];
  u.camera_position[3] = 1;
  if (reveal8 != nullptr) {
    std::memcpy(u.reveal, reveal8, 16);
    std::memcpy(u.reveal_tint, reveal8 + 4, 16);
  } else {
    u.reveal[0] = HUGE_VALF;  // RevealUniform idle (Reveal.h:25-30)
    u.reveal_tint[0] = u.reveal_tint[1] = u.reveal_tint[2] = 1.0f;
  }
  *out = u;
}

int orc_project(const orc_camera* cam, const uint32_t* splat8, const uint32_t* sh_words,
                int sh_degree, uint32_t index, uint32_t* projected8) {
  Proj p{};
  float vz;
  const bool vis = project(*cam, splat8, sh_words, sh_degree, index, p, &vz);
  if (vis) reveal_glow(*cam, f32(splat8[1]), p);
  std::memcpy(projected8, p.w, 32);
  return vis ? 1 : 0;
}

uint32_t orc_depth_key16(const orc_camera* cam, float depth) { return depth_key16(*cam, depth); }

// The visibility kernel (SplatVisibility.metal:39-106) with its range mapping (:30-37,62-67),
// compacted in lane order t, which is the deterministic order this build defines in place of
// the reference's per-SIMD-group atomic (:95-105).
uint32_t orc_cull(const orc_camera* cam, const uint32_t* splats, const uint32_t* sh, int sh_degree,
                  const uint32_t* ranges, uint32_t range_count, const uint32_t* active,
                  uint32_t active_count, int key16, uint32_t* keys, uint32_t* projected) {
  uint32_t count = 0;
  auto lane = [&](uint32_t index) {
    uint32_t key;
    Proj p{};
    if (visibility_lane(*cam, splats, sh, sh_degree, index, key16, key, p)) {
      keys[count] = key;
      std::memcpy(projected + size_t{count} * 8u, p.w, 32);
      ++count;
    }
  };
  if (active != nullptr) {
    for (uint32_t t = 0; t < active_count; ++t) lane(active[t]);
  } else {
    for (uint32_t r = 0; r < range_count; ++r)
      for (uint32_t k = 0; k < ranges[2 * r + 1]; ++k) lane(ranges[2 * r] + k);
  }
  return count;
}

// The sort contract (MetalRadixSort.h:13-19): ascending keys, equal keys by index; the oracle is
// the reference test's own std::stable_sort (MetalRadixSortTest.mm:58-62).
void orc_argsort(const uint32_t* keys, uint32_t n, uint32_t* order) {
  std::iota(order, order + n, 0u);
  std::stable_sort(order, order + n, [keys](uint32_t a, uint32_t b) { return keys[a] < keys[b]; });
}

// spreadBits (SpatialOrder.cpp:12-18): the low 10 bits of v, one every third bit.
static uint32_t spread_bits(uint32_t v) {
  v = (v | (v << 16)) & 0x030000ffu;
  v = (v | (v << 8)) & 0x0300f00fu;
  v = (v | (v << 4)) & 0x030c30c3u;
  v = (v | (v << 2)) & 0x09249249u;
  return v;
}

// quantise (SpatialOrder.cpp:20-26): 1024 cells over [lo, hi], 0 for a flat axis.
static uint32_t quantise(float value, float lo, float hi) {
  const float extent = hi - lo;
  if (!(extent > 0.0f)) return 0u;
  const float t = (value - lo) / extent;
  const float scaled = std::floor(t * 1024.0f);
  return static_cast<uint32_t>(std::clamp(scaled, 0.0f, 1023.0f));
}

uint32_t orc_morton_code(const float* p, const float* lo, const float* hi) {
  return spread_bits(quantise(p[0], lo[0], hi[0])) | (spread_bits(quantise(p[1], lo[1], hi[1])) << 1) |
         (spread_bits(quantise(p[2], lo[2], hi[2])) << 2);
}

// reorderSpatially's permutation (SpatialOrder.cpp:48-84): LSD radix sort of the 30-bit codes in
// three stable counting passes of 10 bits.
void orc_spatial_order(const float* positions, uint32_t n, const float* lo, const float* hi, uint32_t* order) {
  std::vector<uint32_t> keys(n), keys2(n), ord2(n);
  for (uint32_t i = 0; i < n; ++i) {
    keys[i] = orc_morton_code(positions + size_t{i} * 3, lo, hi);
    order[i] = i;
  }
  uint32_t *kin = keys.data(), *kout = keys2.data(), *oin = order, *oout = ord2.data();
  for (int shift = 0; shift < 30; shift += 10) {
    std::vector<uint32_t> start(1025, 0u);
    for (uint32_t i = 0; i < n; ++i) ++start[((kin[i] >> shift) & 0x3ffu) + 1];
    for (int d = 0; d < 1024; ++d) start[d + 1] += start[d];
    for (uint32_t i = 0; i < n; ++i) {
      const uint32_t dst = start[(kin[i] >> shift) & 0x3ffu]++;
      kout[dst] = kin[i];
      oout[dst] = oin[i];
    }
    std::swap(kin, kout);
    std::swap(oin, oout);
  }
  if (oin != order) std::copy(oin, oin + n, order);
}

// Footprint half-extents in pixels: every pixel centre the coverage rule accepts lies within
// (ex, ey) of the centre. The quad bounds the offset by r(|a1| + |a2|) per axis (expandQuad,
// SplatProjection.metalh:155-165); the cutoff bounds gu^2 + gv^2 by L = log2a0 - kLog2MinAlpha,
// so u^2 + v^2 <= L / K^2 and |u a1x + v a2x| <= sqrt(L / K^2) sqrt(a1x^2 + a2x^2) (Cauchy-Schwarz).
// Each axis takes the smaller bound, the disc's with slack over the rule's fp32 rounding. False when
// L < 0: no pixel passes the cutoff.
static bool footprint(const uint32_t* p, float* ex, float* ey) {
  const float cut = log2_peak(p) - kLog2MinAlpha;
  if (!(cut >= 0.0f)) return false;
  const float rho = std::sqrt(cut * kInvGaussianScale2) * kFootprintSlack;
  const float a1x = lo_half(p[2]), a1y = hi_half(p[2]);
  const float a2x = lo_half(p[3]), a2y = hi_half(p[3]);
  const float r = f32(p[6]);
  const float qx = r * (std::fabs(a1x) + std::fabs(a2x));
  const float qy = r * (std::fabs(a1y) + std::fabs(a2y));
  const float sx = rho * std::sqrt(a1x * a1x + a2x * a2x);
  const float sy = rho * std::sqrt(a1y * a1y + a2y * a2y);
  *ex = mn(qx, sx);
  *ey = mn(qy, sy);
  return true;
}

// Tile rectangle of a record: the footprint box around the centre, in pixel indices whose centres
// it may contain, padded by kBinPad, in 16x16 (or `tile`) tiles.
uint32_t orc_tile_rect(const uint32_t* p, uint32_t width, uint32_t height, uint32_t tile,
                       int32_t* rect) {
  const float cx = (f32(p[0]) * 0.5f + 0.5f) * static_cast<float>(width);
  const float cy = (0.5f - f32(p[1]) * 0.5f) * static_cast<float>(height);
  rect[0] = rect[1] = 0;
  rect[2] = rect[3] = -1;
  float ex = 0.0f, ey = 0.0f;
  if (!footprint(p, &ex, &ey)) return 0;
  float x0 = ((cx - ex) - 0.5f) - kBinPad, x1 = ((cx + ex) - 0.5f) + kBinPad;
  float y0 = ((cy - ey) - 0.5f) - kBinPad, y1 = ((cy + ey) - 0.5f) + kBinPad;
  if (!(x0 <= x1) || !(y0 <= y1)) return 0;
  x0 = mx(x0, -1.0f);
  x1 = mn(x1, static_cast<float>(width));
  y0 = mx(y0, -1.0f);
  y1 = mn(y1, static_cast<float>(height));
  int32_t px0 = static_cast<int32_t>(std::ceil(x0)), px1 = static_cast<int32_t>(std::floor(x1));
  int32_t py0 = static_cast<int32_t>(std::ceil(y0)), py1 = static_cast<int32_t>(std::floor(y1));
  px0 = std::max(px0, 0);
  py0 = std::max(py0, 0);
  px1 = std::min(px1, static_cast<int32_t>(width) - 1);
  py1 = std::min(py1, static_cast<int32_t>(height) - 1);
  if (px0 > px1 || py0 > py1) return 0;
  rect[0] = px0 / static_cast<int32_t>(tile);
  rect[1] = py0 / static_cast<int32_t>(tile);
  rect[2] = px1 / static_cast<int32_t>(tile);
  rect[3] = py1 / static_cast<int32_t>(tile);
  return static_cast<uint32_t>((rect[2] - rect[0] + 1) * (rect[3] - rect[1] + 1));
}

// (tile | depth) order: pairs emitted in depth order (row-major over each rectangle), then a
// stable sort by tile; equal tiles keep depth order, equal depths keep compaction order. Tile
// ranges are [start, end) per tile, zero for an empty tile. The sort is a counting sort: per-tile
// counts, their exclusive scan, then each pair at its tile's cursor.
static void bin_pairs(const uint32_t* projected, const uint32_t* order, uint32_t n, uint32_t width,
                      uint32_t height, uint32_t tile, std::vector<uint32_t>* pair_tiles,
                      std::vector<uint32_t>* pair_values, uint32_t* ranges) {
  const uint32_t tiles_x = (width + tile - 1) / tile;
  const uint32_t tiles_y = (height + tile - 1) / tile;
  const uint32_t T = tiles_x * tiles_y;
  std::vector<int32_t> rects(size_t{n} * 4);
  std::vector<uint64_t> cursor(size_t{T} + 1, 0);
  for (uint32_t i = 0; i < n; ++i) {
    int32_t* rect = &rects[size_t{i} * 4];
    if (!orc_tile_rect(projected + size_t{order[i]} * 8u, width, height, tile, rect)) {
      rect[0] = 0, rect[1] = 0, rect[2] = -1, rect[3] = -1;
      continue;
    }
    for (int32_t ty = rect[1]; ty <= rect[3]; ++ty)
      for (int32_t tx = rect[0]; tx <= rect[2]; ++tx)
        ++cursor[size_t{static_cast<uint32_t>(ty) * tiles_x + static_cast<uint32_t>(tx)} + 1];
  }
  for (uint32_t t = 0; t < T; ++t) cursor[t + 1] += cursor[t];
  if (ranges != nullptr) {
    std::memset(ranges, 0, size_t{T} * 8u);
    for (uint32_t t = 0; t < T; ++t) {
      if (cursor[t + 1] == cursor[t]) continue;
      ranges[2 * t] = static_cast<uint32_t>(cursor[t]);
      ranges[2 * t + 1] = static_cast<uint32_t>(cursor[t + 1]);
    }
  }
  pair_tiles->resize(cursor[T]);
  pair_values->resize(cursor[T]);
  for (uint32_t i = 0; i < n; ++i) {
    const int32_t* rect = &rects[size_t{i} * 4];
    for (int32_t ty = rect[1]; ty <= rect[3]; ++ty)
      for (int32_t tx = rect[0]; tx <= rect[2]; ++tx) {
        const uint32_t t = static_cast<uint32_t>(ty) * tiles_x + static_cast<uint32_t>(tx);
        const uint64_t j = cursor[t]++;
        (*pair_tiles)[j] = t;
        (*pair_values)[j] = order[i];
      }
  }
}

uint64_t orc_bin(const uint32_t* projected, const uint32_t* order, uint32_t n, uint32_t width,
                 uint32_t height, uint32_t tile, uint32_t* pair_tiles, uint32_t* pair_values,
                 uint64_t capacity, uint32_t* ranges) {
  if (pair_tiles == nullptr) {
    uint64_t total = 0;
    for (uint32_t i = 0; i < n; ++i) {
      int32_t rect[4];
      total += orc_tile_rect(projected + size_t{order[i]} * 8u, width, height, tile, rect);
    }
    if (ranges != nullptr) {
      const uint32_t T = ((width + tile - 1) / tile) * ((height + tile - 1) / tile);
      std::memset(ranges, 0, size_t{T} * 8u);
    }
    return total;
  }
  std::vector<uint32_t> pt, pv;
  bin_pairs(projected, order, n, width, height, tile, &pt, &pv, ranges);
  const uint64_t total = pt.size();
  if (total > capacity) {
    if (ranges != nullptr) {
      const uint32_t T = ((width + tile - 1) / tile) * ((height + tile - 1) / tile);
      std::memset(ranges, 0, size_t{T} * 8u);
    }
    return total;
  }
  std::copy(pt.begin(), pt.end(), pair_tiles);
  std::copy(pv.begin(), pv.end(), pair_values);
  return total;
}

// Front-to-back "under" compositing onto zero, then the background under the rest
// (MetalSplatPipeline.mm:46-63,153-161; SplatRaster.metal:40-44,67-69), in FP64, with no early
// termination: the ideal image for the fp32 coverage rule.
void orc_raster64(const uint32_t* projected, const uint32_t* pair_values, const uint32_t* ranges,
                  uint32_t width, uint32_t height, uint32_t tile, const float* background3,
                  float* out) {
  raster_tiles<true>(projected, pair_values, ranges, width, height, tile, background3, 0.0f,
                     std::max(1u, std::thread::hardware_concurrency()), out);
}

void orc_raster32(const uint32_t* projected, const uint32_t* pair_values, const uint32_t* ranges,
                  uint32_t width, uint32_t height, uint32_t tile, const float* background3,
                  float t_stop, uint32_t threads, float* out) {
  raster_tiles<false>(projected, pair_values, ranges, width, height, tile, background3, t_stop,
                      std::max(1u, threads), out);
}

uint64_t orc_missed_coverage(const uint32_t* projected, uint32_t count, uint32_t width,
                             uint32_t height, uint32_t tile, const uint32_t* pair_values,
                             const uint32_t* ranges, uint64_t* accepted) {
  const uint32_t tiles_x = (width + tile - 1) / tile;
  uint64_t missed = 0, taken = 0;
  for (uint32_t i = 0; i < count; ++i) {
    const uint32_t* p = projected + size_t{i} * 8u;
    const Prep q = prep(p, width, height);
    const double r = f32(p[6]);
    const double ex = r * (std::fabs(q.a1x) + std::fabs(q.a2x)) + 2.0;
    const double ey = r * (std::fabs(q.a1y) + std::fabs(q.a2y)) + 2.0;
    if (!(ex == ex) || !(ey == ey)) continue;
    const int64_t x0 = std::max<int64_t>(0, static_cast<int64_t>(std::floor(q.cx - ex)));
    const int64_t x1 = std::min<int64_t>(width - 1, static_cast<int64_t>(std::ceil(q.cx + ex)));
    const int64_t y0 = std::max<int64_t>(0, static_cast<int64_t>(std::floor(q.cy - ey)));
    const int64_t y1 = std::min<int64_t>(height - 1, static_cast<int64_t>(std::ceil(q.cy + ey)));
    for (int64_t py = y0; py <= y1; ++py)
      for (int64_t px = x0; px <= x1; ++px) {
        float power;
        if (!decide(q, static_cast<uint32_t>(px), static_cast<uint32_t>(py), &power)) continue;
        ++taken;
        const uint32_t t = static_cast<uint32_t>(py / tile) * tiles_x + static_cast<uint32_t>(px / tile);
        const uint32_t* b = pair_values + ranges[2 * t];
        const uint32_t* e = pair_values + ranges[2 * t + 1];
        if (std::find(b, e, i) == e) ++missed;
      }
  }
  if (accepted != nullptr) *accepted = taken;
  return missed;
}

// The raster's region test (the kernel's per-region record lists), restated to check that it never
// drops a record the coverage rule would take in the region: the footprint box against the region's
// pixel centres [x0, x0 + span] x [y0, y0 + span], then the quad's slabs |gu|, |gv| <= K r widened by
// gu's variation over the region.
int orc_region_keep(const uint32_t* p, uint32_t width, uint32_t height, uint32_t x0, uint32_t y0,
                    uint32_t span) {
  const Prep q = prep(p, width, height);
  float ex = 0.0f, ey = 0.0f;
  if (!footprint(p, &ex, &ey)) return 0;
  const float bx0 = (q.cx - ex) - kBinPad, bx1 = (q.cx + ex) + kBinPad;
  const float by0 = (q.cy - ey) - kBinPad, by1 = (q.cy + ey) + kBinPad;
  const float qx = static_cast<float>(x0) + 0.5f, qy = static_cast<float>(y0) + 0.5f;
  const float fs = static_cast<float>(span);
  const float half = 0.5f * fs + kBinPad;
  const float hu = (std::fabs(q.g00) + std::fabs(q.g01)) * half, hv = (std::fabs(q.g10) + std::fabs(q.g11)) * half;
  const float lim_u = (q.rk + hu) * kFootprintSlack, lim_v = (q.rk + hv) * kFootprintSlack;
  const float dxc = (qx + 0.5f * fs) - q.cx, dyc = q.cy - (qy + 0.5f * fs);
  const float uc = std::fmaf(q.g00, dxc, q.g01 * dyc), vc = std::fmaf(q.g10, dxc, q.g11 * dyc);
  return bx1 >= qx && bx0 <= qx + fs && by1 >= qy && by0 <= qy + fs && std::fabs(uc) <= lim_u &&
         std::fabs(vc) <= lim_v;
}

// Brute force over every pixel of every record's quad box: counts accepted (pixel, record) pairs
// whose region (the aligned `region` x `region` square holding the pixel) the region test drops.
uint64_t orc_region_misses(const uint32_t* projected, uint32_t count, uint32_t width, uint32_t height,
                           uint32_t region, uint64_t* accepted) {
  uint64_t missed = 0, taken = 0;
  for (uint32_t i = 0; i < count; ++i) {
    const uint32_t* p = projected + size_t{i} * 8u;
    const Prep q = prep(p, width, height);
    const double r = f32(p[6]);
    const double ex = r * (std::fabs(q.a1x) + std::fabs(q.a2x)) + 2.0;
    const double ey = r * (std::fabs(q.a1y) + std::fabs(q.a2y)) + 2.0;
    if (!(ex == ex) || !(ey == ey)) continue;
    const int64_t x0 = std::max<int64_t>(0, static_cast<int64_t>(std::floor(q.cx - ex)));
    const int64_t x1 = std::min<int64_t>(width - 1, static_cast<int64_t>(std::ceil(q.cx + ex)));
    const int64_t y0 = std::max<int64_t>(0, static_cast<int64_t>(std::floor(q.cy - ey)));
    const int64_t y1 = std::min<int64_t>(height - 1, static_cast<int64_t>(std::ceil(q.cy + ey)));
    for (int64_t py = y0; py <= y1; ++py)
      for (int64_t px = x0; px <= x1; ++px) {
        float power;
        if (!decide(q, static_cast<uint32_t>(px), static_cast<uint32_t>(py), &power)) continue;
        ++taken;
        const uint32_t rx = static_cast<uint32_t>(px) / region * region, ry = static_cast<uint32_t>(py) / region * region;
        if (!orc_region_keep(p, width, height, rx, ry, region - 1)) ++missed;
      }
  }
  if (accepted != nullptr) *accepted = taken;
  return missed;
}

int orc_decide(const uint32_t* projected8, uint32_t width, uint32_t height, uint32_t px,
               uint32_t py, float* power_out) {
  const Prep q = prep(projected8, width, height);
  float power = 0.0f;
  const bool take = decide(q, px, py, &power);
  if (power_out != nullptr) *power_out = power;
  return take ? 1 : 0;
}

int orc_render_frame(const orc_camera* cam, const uint32_t* splats, const uint32_t* sh,
                     int sh_degree, uint32_t n, int key16, uint32_t tile, const float* background3,
                     uint32_t threads, float* out_rgba, double* stage_seconds,
                     uint32_t* visible_out, uint64_t* pairs_out) {
  using clock = std::chrono::steady_clock;
  threads = std::max(1u, threads);
  const uint32_t width = static_cast<uint32_t>(cam->screen_size[0]);
  const uint32_t height = static_cast<uint32_t>(cam->screen_size[1]);
  auto t0 = clock::now();
  // Cull: chunks compact locally and concatenate in order.
  const uint32_t chunks = threads;
  std::vector<std::vector<uint32_t>> ck(chunks), cp(chunks);
  parallel_for(chunks, threads, [&](uint32_t lo, uint32_t hi) {
    for (uint32_t c = lo; c < hi; ++c) {
      const uint32_t b = static_cast<uint32_t>((uint64_t{n} * c) / chunks);
      const uint32_t e = static_cast<uint32_t>((uint64_t{n} * (c + 1)) / chunks);
      for (uint32_t i = b; i < e; ++i) {
        uint32_t key;
        Proj p{};
        if (visibility_lane(*cam, splats, sh, sh_degree, i, key16, key, p)) {
          ck[c].push_back(key);
          cp[c].insert(cp[c].end(), p.w, p.w + 8);
        }
      }
    }
  });
  std::vector<uint32_t> keys, projected;
  for (uint32_t c = 0; c < chunks; ++c) {
    keys.insert(keys.end(), ck[c].begin(), ck[c].end());
    projected.insert(projected.end(), cp[c].begin(), cp[c].end());
  }
  const uint32_t visible = static_cast<uint32_t>(keys.size());
  auto t1 = clock::now();
  std::vector<uint32_t> order(visible);
  orc_argsort(keys.data(), visible, order.data());
  auto t2 = clock::now();
  const uint32_t tiles_x = (width + tile - 1) / tile, tiles_y = (height + tile - 1) / tile;
  std::vector<uint32_t> ranges(size_t{tiles_x} * tiles_y * 2u);
  std::vector<uint32_t> pt, pv;
  bin_pairs(projected.data(), order.data(), visible, width, height, tile, &pt, &pv, ranges.data());
  const uint64_t pairs = pt.size();
  auto t3 = clock::now();
  orc_raster32(projected.data(), pv.data(), ranges.data(), width, height, tile, background3,
               1.0f / 4096.0f, threads, out_rgba);
  auto t4 = clock::now();
  if (stage_seconds != nullptr) {
    stage_seconds[0] = std::chrono::duration<double>(t1 - t0).count();
    stage_seconds[1] = std::chrono::duration<double>(t2 - t1).count();
    stage_seconds[2] = std::chrono::duration<double>(t3 - t2).count();
    stage_seconds[3] = std::chrono::duration<double>(t4 - t3).count();
  }
  if (visible_out != nullptr) *visible_out = visible;
  if (pairs_out != nullptr) *pairs_out = pairs;
  return 0;
}

// LodThreshold (LodThreshold.h:20-36) over a sequence of completed selections: starts at the policy's
// threshold; each step multiplies it by 1.1 when refinements were denied (limited > 0), capped at
// max(64, policy), else divides it by 1.03 when selected < limit / 10 * 9 (integer division), floored
// at the policy; out[i] is the threshold after step i.
void orc_lod_threshold(float policy, uint32_t steps, const uint32_t* selected, const uint32_t* limited,
                       const uint32_t* limit, float* out) {
  float pixels = policy;
  for (uint32_t i = 0; i < steps; ++i) {
    if (limited[i] > 0) {
      const float raised = pixels * 1.1f;
      const float ceiling = policy > 64.0f ? policy : 64.0f;
      pixels = raised < ceiling ? raised : ceiling;
    } else if (selected[i] < limit[i] / 10u * 9u) {
      const float relaxed = pixels / 1.03f;
      pixels = relaxed > policy ? relaxed : policy;
    }
    out[i] = pixels;
  }
}

}  // extern "C"
