// Synthetic scenes and camera matrices (declared in include/skm_scene.h).
#include "skm_scene.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <thread>
#include <vector>

namespace {

uint64_t splitmix64(uint64_t& x) {
  uint64_t z = (x += 0x9e3779b97f4a7c15ull);
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

// xoshiro256** (Blackman and Vigna, 2018) with Box-Muller normals.
struct Rng {
  uint64_t s[4];
  bool has_spare = false;
  double spare = 0.0;

  Rng(uint64_t seed, uint64_t stream) {
    uint64_t x = seed ^ (stream * 0xd1b54a32d192ed03ull);
    for (auto& w : s) w = splitmix64(x);
  }
  uint64_t next() {
    const uint64_t result = rotl(s[1] * 5, 7) * 9;
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return result;
  }
  // Uniform in (0, 1): 53 random bits, offset by half an ulp so zero never occurs.
  double uniform() { return (static_cast<double>(next() >> 11) + 0.5) * (1.0 / 9007199254740992.0); }
  double normal() {
    if (has_spare) {
      has_spare = false;
      return spare;
    }
    const double u1 = uniform(), u2 = uniform();
    const double r = std::sqrt(-2.0 * std::log(u1));
    const double a = 6.283185307179586476925 * u2;
    spare = r * std::sin(a);
    has_spare = true;
    return r * std::cos(a);
  }
};

struct Box {
  float lo[3], hi[3];
  float face_cdf[6];
};

std::vector<Box> make_boxes(const skm_scene_spec& spec) {
  Rng rng(spec.seed, 0xb0b0b0b0ull);
  std::vector<Box> boxes(6);
  const float a = spec.area;
  for (Box& b : boxes) {
    const float cx = static_cast<float>((rng.uniform() - 0.5) * 0.8 * a);
    const float cz = static_cast<float>((rng.uniform() - 0.5) * 0.8 * a);
    float size[3];
    for (float& s : size) s = static_cast<float>((0.05 + 0.15 * rng.uniform()) * a);
    b.lo[0] = cx - size[0] / 2;
    b.hi[0] = cx + size[0] / 2;
    b.lo[1] = 0.0f;
    b.hi[1] = size[1];
    b.lo[2] = cz - size[2] / 2;
    b.hi[2] = cz + size[2] / 2;
    const float yz = size[1] * size[2], xz = size[0] * size[2], xy = size[0] * size[1];
    const float areas[6] = {yz, yz, xz, xz, xy, xy};
    float run = 0.0f, total = 2.0f * (yz + xz + xy);
    for (int f = 0; f < 6; ++f) {
      run += areas[f];
      b.face_cdf[f] = run / total;
    }
  }
  return boxes;
}

void fill_range(const skm_scene_spec& spec, const std::vector<Box>& boxes, uint32_t lo,
                uint32_t hi, float* positions, float* covariances, float* colors, float* alphas,
                float* sh) {
  const int deg = std::clamp(spec.sh_degree, 0, 3);
  const uint32_t sh_floats = static_cast<uint32_t>((deg + 1) * (deg + 1) - 1) * 3u;
  for (uint32_t i = lo; i < hi; ++i) {
    Rng rng(spec.seed, uint64_t{i} + 1);
    float p[3];
    if (spec.layout == SKM_SCENE_SURFACES && rng.uniform() < spec.surface_share) {
      const float j0 = static_cast<float>(rng.normal() * spec.surface_sigma);
      const float j1 = static_cast<float>(rng.normal() * spec.surface_sigma);
      const float j2 = static_cast<float>(rng.normal() * spec.surface_sigma);
      if (rng.uniform() < 0.5) {  // ground square y = 0
        p[0] = static_cast<float>((rng.uniform() - 0.5) * spec.area);
        p[1] = 0.0f;
        p[2] = static_cast<float>((rng.uniform() - 0.5) * spec.area);
      } else {  // a box face, chosen by area
        const Box& b = boxes[static_cast<size_t>(rng.uniform() * boxes.size()) % boxes.size()];
        const float f = static_cast<float>(rng.uniform());
        int face = 0;
        while (face < 5 && f > b.face_cdf[face]) ++face;
        const int axis = face / 2;
        for (int k = 0; k < 3; ++k)
          p[k] = b.lo[k] + static_cast<float>(rng.uniform()) * (b.hi[k] - b.lo[k]);
        p[axis] = (face % 2 == 0) ? b.lo[axis] : b.hi[axis];
      }
      p[0] += j0;
      p[1] += j1;
      p[2] += j2;
    } else {
      for (float& v : p) v = static_cast<float>(rng.normal() * spec.spread);
    }
    float scale[3];
    for (float& s : scale)
      s = std::exp(static_cast<float>(spec.log_scale_mean + rng.normal() * spec.log_scale_sigma));
    float q[4];
    double qn = 0.0;
    for (float& v : q) {
      v = static_cast<float>(rng.normal());
      qn += double{v} * v;
    }
    const float inv = static_cast<float>(1.0 / std::sqrt(qn));
    for (float& v : q) v *= inv;
    const float logit = static_cast<float>(rng.normal() * spec.opacity_sigma);
    float dc[3];
    for (float& v : dc) v = static_cast<float>(rng.normal() * spec.dc_sigma);

    // Covariance R diag(s)^2 R^T from the unit quaternion (x, y, z, w): Mat4::rotation(q)
    // (Mat4.h:55-71), then M = R diag(s), Sigma = M M^T (Covariance.cpp:7-17).
    const float x = q[0], y = q[1], z = q[2], w = q[3];
    const float R[3][3] = {{1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)},
                           {2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)},
                           {2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)}};
    float m[3][3];
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) m[r][c] = R[r][c] * scale[c];
    auto dotr = [&m](int a, int b) { return m[a][0] * m[b][0] + m[a][1] * m[b][1] + m[a][2] * m[b][2]; };
    float* cov = covariances + size_t{i} * 6u;
    cov[0] = dotr(0, 0);
    cov[1] = dotr(0, 1);
    cov[2] = dotr(0, 2);
    cov[3] = dotr(1, 1);
    cov[4] = dotr(1, 2);
    cov[5] = dotr(2, 2);

    for (int k = 0; k < 3; ++k) positions[size_t{i} * 3u + k] = p[k];
    for (int k = 0; k < 3; ++k) colors[size_t{i} * 3u + k] = std::clamp(0.5f + 0.282095f * dc[k], 0.0f, 1.0f);
    alphas[i] = 1.0f / (1.0f + std::exp(-logit));
    if (sh != nullptr)
      for (uint32_t k = 0; k < sh_floats; ++k)
        sh[size_t{i} * sh_floats + k] = static_cast<float>(rng.normal() * spec.sh_sigma);
  }
}

}  // namespace

extern "C" {

int skm_scene_fill(const skm_scene_spec* spec, float* positions, float* covariances, float* colors,
                   float* alphas, float* sh, uint32_t threads) {
  if (spec == nullptr || positions == nullptr || covariances == nullptr || colors == nullptr ||
      alphas == nullptr || spec->sh_degree < 0 || spec->sh_degree > 3 ||
      (spec->sh_degree > 0 && sh == nullptr))
    return -1;
  const auto boxes = make_boxes(*spec);
  const uint32_t n = spec->count;
  threads = std::max(1u, std::min(threads, 64u));
  std::vector<std::thread> pool;
  for (uint32_t t = 0; t < threads; ++t) {
    const uint32_t lo = static_cast<uint32_t>(uint64_t{n} * t / threads);
    const uint32_t hi = static_cast<uint32_t>(uint64_t{n} * (t + 1) / threads);
    pool.emplace_back([&, lo, hi] { fill_range(*spec, boxes, lo, hi, positions, covariances, colors, alphas, sh); });
  }
  for (auto& th : pool) th.join();
  return 0;
}

void skm_look_at(const float* eye, const float* target, const float* up, float* view16) {
  float f[3] = {target[0] - eye[0], target[1] - eye[1], target[2] - eye[2]};
  float fl = std::sqrt(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]);
  for (float& v : f) v /= fl;
  float s[3] = {f[1] * up[2] - f[2] * up[1], f[2] * up[0] - f[0] * up[2], f[0] * up[1] - f[1] * up[0]};
  float sl = std::sqrt(s[0] * s[0] + s[1] * s[1] + s[2] * s[2]);
  for (float& v : s) v /= sl;
  const float u[3] = {s[1] * f[2] - s[2] * f[1], s[2] * f[0] - s[0] * f[2], s[0] * f[1] - s[1] * f[0]};
  const float rows[3][3] = {{s[0], s[1], s[2]}, {u[0], u[1], u[2]}, {-f[0], -f[1], -f[2]}};
  for (int k = 0; k < 16; ++k) view16[k] = 0.0f;
  for (int r = 0; r < 3; ++r) {
    for (int c = 0; c < 3; ++c) view16[c * 4 + r] = rows[r][c];
    view16[12 + r] = -(rows[r][0] * eye[0] + rows[r][1] * eye[1] + rows[r][2] * eye[2]);
  }
  view16[15] = 1.0f;
}

void skm_perspective(float fovy, float aspect, float near_plane, float far_plane, float* proj16) {
  const float ys = 1 / std::tan(fovy * 0.5f);
  const float xs = ys / aspect;
  const float zs = far_plane / (near_plane - far_plane);
  for (int k = 0; k < 16; ++k) proj16[k] = 0.0f;
  proj16[0] = xs;
  proj16[5] = ys;
  proj16[10] = zs;
  proj16[14] = near_plane * zs;
  proj16[11] = -1;
}

}  // extern "C"
