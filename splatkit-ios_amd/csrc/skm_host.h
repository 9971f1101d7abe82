// Host-side helpers shared by the renderer, the view batches and the server: RAII device and
// pinned buffers, the camera uniform and the BGRA8 conversion.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstring>

#include "skm.h"
#include "skm_internal.h"

namespace skm {

inline size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

inline uint32_t sh_stride(int degree) {
  if (degree <= 0) return 0;
  return (static_cast<uint32_t>((degree + 1) * (degree + 1) - 1) * 3u + 1u) / 2u;
}

// One device allocation; freed on destruction. Moves only.
struct DeviceBuffer {
  void* ptr = nullptr;
  size_t bytes = 0;
  DeviceBuffer() = default;
  DeviceBuffer(const DeviceBuffer&) = delete;
  DeviceBuffer& operator=(const DeviceBuffer&) = delete;
  DeviceBuffer(DeviceBuffer&& o) noexcept : ptr(o.ptr), bytes(o.bytes) { o.ptr = nullptr, o.bytes = 0; }
  DeviceBuffer& operator=(DeviceBuffer&& o) noexcept {
    if (this != &o) {
      reset();
      ptr = o.ptr;
      bytes = o.bytes;
      o.ptr = nullptr;
      o.bytes = 0;
    }
    return *this;
  }
  ~DeviceBuffer() { reset(); }
  void reset() {
    if (ptr != nullptr) (void)hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
  }
  hipError_t allocate(size_t n) {
    reset();
    if (n == 0) n = 256;
    const hipError_t e = hipMalloc(&ptr, n);
    if (e != hipSuccess) {
      ptr = nullptr;
      return e;
    }
    bytes = n;
    return hipSuccess;
  }
  template <class T>
  T* as(size_t byte_offset = 0) const {
    return reinterpret_cast<T*>(static_cast<char*>(ptr) + byte_offset);
  }
};

struct PinnedBuffer {
  void* ptr = nullptr;
  PinnedBuffer() = default;
  PinnedBuffer(const PinnedBuffer&) = delete;
  PinnedBuffer& operator=(const PinnedBuffer&) = delete;
  ~PinnedBuffer() {
    if (ptr != nullptr) (void)hipHostFree(ptr);
  }
  hipError_t allocate(size_t n) {
    if (ptr != nullptr) (void)hipHostFree(ptr);
    ptr = nullptr;
    return hipHostMalloc(&ptr, n, hipHostMallocDefault);
  }
  template <class T>
  T* as() const {
    return static_cast<T*>(ptr);
  }
};

inline void fill_camera(Camera& cam, const float* view, const float* proj, uint32_t W, uint32_t H, bool linear,
                 float min_px, const float* eye, const skm_reveal* reveal) {
  // MetalFrameCompute::updateCamera (MetalFrameCompute.mm:141-164).
  std::memset(&cam, 0, sizeof(cam));
  std::memcpy(cam.view, view, 64);
  std::memcpy(cam.proj, proj, 64);
  cam.screen_size[0] = static_cast<float>(W);
  cam.screen_size[1] = static_cast<float>(H);
  cam.focal[0] = cam.screen_size[0] * proj[0] / 2;
  cam.focal[1] = cam.screen_size[1] * proj[5] / 2;
  cam.tan_half_fov[0] = 1 / proj[0];
  cam.tan_half_fov[1] = 1 / proj[5];
  cam.output_linear = linear ? 1u : 0u;
  cam.min_pixel_radius = min_px;
  if (eye != nullptr) {
    cam.camera_position[0] = eye[0];
    cam.camera_position[1] = eye[1];
    cam.camera_position[2] = eye[2];
  }
  cam.camera_position[3] = 1;
  if (reveal != nullptr) {
    std::memcpy(cam.reveal, reveal, 16);
    std::memcpy(cam.reveal_tint, reveal->tint, 16);
  } else {
    cam.reveal[0] = HUGE_VALF;
    cam.reveal_tint[0] = cam.reveal_tint[1] = cam.reveal_tint[2] = 1.0f;
  }
}

// MetalFrameCapture's BGRA8 (MetalFrameCapture.h:18) from the RGBA float composite.
inline void to_bgra8(const float* rgba, size_t pixels, uint8_t* out) {
  auto q = [](float v) {
    const float c = v < 0.0f ? 0.0f : (v > 1.0f ? 1.0f : v);
    return static_cast<uint8_t>(std::lround(c * 255.0f));
  };
  for (size_t i = 0; i < pixels; ++i) {
    out[4 * i + 0] = q(rgba[4 * i + 2]);
    out[4 * i + 1] = q(rgba[4 * i + 1]);
    out[4 * i + 2] = q(rgba[4 * i + 0]);
    out[4 * i + 3] = q(rgba[4 * i + 3]);
  }
}

}  // namespace skm
