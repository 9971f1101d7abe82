// Device arithmetic of the render path. Every fp32 expression has an explicit evaluation order and
// the library is compiled with -ffp-contract=off and without fast-math, so each operation is one
// IEEE-754 rounding and the results are reproducible bit for bit on any IEEE machine that
// evaluates the same expressions. The reference (MSL, compiled with fast-math,
// MetalShaderLibrary.h:15-20) fixes no order; the order here is this library's specification of
// the reference's formulas.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace skm {

__device__ __forceinline__ float f32(uint32_t b) { return __uint_as_float(b); }
__device__ __forceinline__ uint32_t bits(float f) { return __float_as_uint(f); }

// max/min with a fixed answer for NaN (the other operand) and for equal operands (the first),
// so signed zeros come out the same on every machine.
__device__ __forceinline__ float mx(float a, float b) { return (a >= b || b != b) ? a : b; }
__device__ __forceinline__ float mn(float a, float b) { return (a <= b || b != b) ? a : b; }

__device__ __forceinline__ float half_to_float(uint32_t h16) {
  _Float16 h;
  const uint16_t v = static_cast<uint16_t>(h16);
  __builtin_memcpy(&h, &v, 2);
  return static_cast<float>(h);
}
__device__ __forceinline__ float lo_half(uint32_t w) { return half_to_float(w & 0xffffu); }
__device__ __forceinline__ float hi_half(uint32_t w) { return half_to_float(w >> 16); }
// IEEE binary16, round to nearest even, overflow to infinity: the shader's half(float), by the
// conversion instruction. The empty asm makes the fp32 value a register first: without it the
// compiler may fuse a preceding multiply into v_fma_mix (one rounding from the exact product
// instead of two), which changes the result at fp32 values that lie exactly between two halves.
// NaN becomes the quiet NaN 0x7e00 with its sign (the oracle's half_rne).
__device__ __forceinline__ uint32_t half_rne(float value) {
  if (value != value) return ((bits(value) >> 16) & 0x8000u) | 0x7e00u;
  asm volatile("" : "+v"(value));
  const _Float16 h = static_cast<_Float16>(value);
  uint16_t out;
  __builtin_memcpy(&out, &h, 2);
  return out;
}
__device__ __forceinline__ uint32_t pack_rne(float a, float b) { return half_rne(a) | (half_rne(b) << 16); }
__device__ __forceinline__ float round_half(float f) { return half_to_float(half_rne(f)); }

// Saturating binary16 with round to nearest even: the host packer's conversion (Half.h:12-38),
// restated with integer operations.
__device__ __forceinline__ uint32_t half_sat(float value) {
  const uint32_t b = bits(value);
  const uint32_t sign = (b >> 16) & 0x8000u;
  const uint32_t biased = (b >> 23) & 0xffu;
  uint32_t mant = b & 0x7fffffu;
  if (biased == 0xffu) return sign | (mant ? 0x7e00u : 0x7bffu);
  const int32_t e = static_cast<int32_t>(biased) - 112;
  if (e >= 31) return sign | 0x7bffu;
  if (e <= 0) {
    if (e < -10) return sign;
    mant |= 0x800000u;
    const uint32_t shift = static_cast<uint32_t>(14 - e);
    const uint32_t q = mant >> shift;
    const uint32_t rem = mant & ((1u << shift) - 1u);
    const uint32_t mid = 1u << (shift - 1u);
    return sign | (q + ((rem > mid || (rem == mid && (q & 1u))) ? 1u : 0u));
  }
  uint32_t h = (static_cast<uint32_t>(e) << 10) | (mant >> 13);
  const uint32_t rem = mant & 0x1fffu;
  if (rem > 0x1000u || (rem == 0x1000u && (h & 1u))) ++h;
  if (h >= 0x7c00u) h = 0x7bffu;
  return sign | h;
}

constexpr double kLn2 = 0.69314718055994530942;
constexpr double kLn2Hi = 6.93147180369123816490e-01;
constexpr double kLn2Lo = 1.90821492927058770002e-10;

// Natural logarithm from IEEE double basic operations: x = m 2^e, m in [sqrt(1/2), sqrt(2)),
// log m = 2 atanh((m-1)/(m+1)) by its odd series (Abramowitz & Stegun 4.5.64), 12 terms.
__device__ __forceinline__ double log_d(double x) {
  int e = 0;
  double m = __builtin_frexp(x, &e);
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

// e^x: x = k ln2 + r, e^r by its Taylor series to r^20/20!, scaled by 2^k.
__device__ __forceinline__ double exp_d(double x) {
  if (x < -1000.0) return 0.0;
  if (x > 1000.0) return __builtin_huge_val();
  const double k = __builtin_nearbyint(x / kLn2);
  // Cody-Waite: ln2 split so that k * kLn2Hi is exact for |k| < 2^20.
  const double r = (x - k * kLn2Hi) - k * kLn2Lo;
  double p = 1.0;
  for (int n = 20; n >= 1; --n) p = 1.0 + (r / static_cast<double>(n)) * p;
  return __builtin_ldexp(p, static_cast<int>(k));
}

// Natural logarithm in fp32 with a fixed operation order (the oracle's log_f, same
// specification): x = m 2^e, m in [sqrt(1/2), sqrt(2)), log m = 2 atanh(s), s = (m-1)/(m+1),
// |s| <= 0.1716, by the odd series of atanh to s^11 (Abramowitz & Stegun 4.5.64; the next term is
// below 2^-33 relative), e ln 2 with a split constant. For x normal and positive.
constexpr float kLn2HiF = 0.693145751953125f;   // 15 significant bits: e * kLn2HiF is exact
constexpr float kLn2LoF = 1.428606765330187e-06f;
constexpr float kInvLn2F = 1.44269504088896341f;
__device__ __forceinline__ float log_f(float x) {
  int e = 0;
  float m = frexpf(x, &e);
  if (m < 0.707106781f) {
    m = m * 2.0f;
    e = e - 1;
  }
  const float s = (m - 1.0f) / (m + 1.0f);
  const float z = s * s;
  float p = 0.0909090936f;  // 1/11
  p = 0.111111112f + z * p;   // 1/9
  p = 0.142857149f + z * p;   // 1/7
  p = 0.200000003f + z * p;   // 1/5
  p = 0.333333343f + z * p;   // 1/3
  p = 1.0f + z * p;
  const float fe = static_cast<float>(e);
  return fe * kLn2HiF + (fe * kLn2LoF + (2.0f * s) * p);
}

__device__ __forceinline__ float pow_f(float x, float y) {
  if (!(x > 0.0f)) return 0.0f;
  return static_cast<float>(exp_d(static_cast<double>(y) * log_d(static_cast<double>(x))));
}

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
constexpr float kBinPad = 1.0f / 16.0f;
// 1 / kGaussianScale^2 = 2 ln 2, and the slack of the footprint disc over the coverage rule's fp32
// rounding.
constexpr float kInvGaussianScale2 = 1.38629436f;
constexpr float kFootprintSlack = 1.0009765625f;

}  // namespace skm
