// The radix sort of device keys without any allocation (skm_argsort_u32_device_static, include/skm.h):
// the same LSD passes as skm_argsort_u32_device, with their scratch (control block, digit totals,
// per-tile digit counts, the key and value ping-pong halves and the key count) in one file-scope
// device array instead of renderer-owned buffers. Each GPU holds its own copy of the array; its
// address on a device comes from hipGetSymbolAddress once, at renderer creation with that device
// current, and host code only ever passes that address on.
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <mutex>

#include "skm.h"
#include "skm_host.h"
#include "skm_internal.h"

namespace skm {

// renderer.cpp
hipStream_t renderer_stream(skm_renderer* r);
int renderer_device(const skm_renderer* r);
skm_status renderer_fail(skm_renderer* r, skm_status s, const char* message);

namespace {

constexpr uint32_t kStaticKeys = 1u << 20;  // most keys one call sorts
constexpr uint32_t kStaticTiles = (kStaticKeys + kSortTile - 1) / kSortTile;
constexpr size_t kCtrlAt = 0;
constexpr size_t kCountAt = 256;  // the key count the passes read
constexpr size_t kHistAt = 512;
constexpr size_t kTileCountsAt = kHistAt + ((size_t{kHistWords} * 4 + 255) / 256) * 256;
constexpr size_t kKeysAt = kTileCountsAt + ((size_t{kStaticTiles} * kRadix * 4 + 255) / 256) * 256;
constexpr size_t kValsAt = kKeysAt + size_t{kStaticKeys} * 4;
constexpr size_t kScratchBytes = kValsAt + size_t{kStaticKeys} * 4;
static_assert(sizeof(Ctrl) <= kCountAt - kCtrlAt, "control block fits its slot");

// 8.7 MB per GPU, held until the process ends.
__device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_scratch[kMaxDevices] = {};
hipError_t g_setup[kMaxDevices] = {};

__global__ void k_store_word(uint32_t* dst, uint32_t value) { *dst = value; }

}  // namespace

// Called by skm_renderer_create with `device` current: resolves this device's copy of the array.
void static_scratch_setup(int device) {
  if (device < 0 || device >= kMaxDevices) return;
  std::lock_guard<std::mutex> g(g_mu);
  if (g_scratch[device] != nullptr) return;
  void* ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  g_setup[device] = e;
  g_scratch[device] = e == hipSuccess ? ptr : nullptr;
}

}  // namespace skm

extern "C" skm_status skm_argsort_u32_device_static(skm_renderer* r, void* keys_dev, uint32_t n, int key_bits,
                                                    void* order_dev) {
  using namespace skm;
  if (r == nullptr || (n > 0 && (keys_dev == nullptr || order_dev == nullptr))) return SKM_ERR_INVALID;
  if (key_bits < 1 || key_bits > 32) return renderer_fail(r, SKM_ERR_INVALID, "key_bits must be 1 to 32");
  if (n > kStaticKeys) {
    char msg[160];
    std::snprintf(msg, sizeof msg, "%u keys exceed the static scratch's maximum of %u", n, kStaticKeys);
    return renderer_fail(r, SKM_ERR_INVALID, msg);
  }
  if (n == 0) return SKM_OK;
  const int device = renderer_device(r);
  char* base = nullptr;
  {
    std::lock_guard<std::mutex> g(g_mu);
    if (device >= 0 && device < kMaxDevices) base = static_cast<char*>(g_scratch[device]);
    if (base == nullptr) {
      char msg[200];
      std::snprintf(msg, sizeof msg, "static scratch unavailable on device %d: %s", device,
                    hipGetErrorString(device >= 0 && device < kMaxDevices ? g_setup[device] : hipErrorInvalidDevice));
      return renderer_fail(r, SKM_ERR_STATE, msg);
    }
  }
  hipError_t e = hipSetDevice(device);
  hipStream_t stream = renderer_stream(r);
  Ctrl* ctrl = reinterpret_cast<Ctrl*>(base + kCtrlAt);
  uint32_t* n_dev = reinterpret_cast<uint32_t*>(base + kCountAt);
  uint32_t* hist = reinterpret_cast<uint32_t*>(base + kHistAt);
  uint32_t* counts = reinterpret_cast<uint32_t*>(base + kTileCountsAt);
  // keys[0] holds the input and is overwritten as the ping-pong's first half, as in
  // skm_argsort_u32_device; the order ends in vals[(passes - 1) & 1], the caller's buffer.
  uint32_t* keys[2] = {static_cast<uint32_t*>(keys_dev), reinterpret_cast<uint32_t*>(base + kKeysAt)};
  const DigitPlan plan = make_digit_plan(static_cast<uint32_t>(key_bits), 8u);
  uint32_t* vals[2] = {reinterpret_cast<uint32_t*>(base + kValsAt), reinterpret_cast<uint32_t*>(base + kValsAt)};
  vals[(plan.passes - 1) & 1] = static_cast<uint32_t*>(order_dev);
  if (e == hipSuccess) {
    // The count goes in by a kernel: no host buffer has to outlive the call, and no copy node is made.
    hipLaunchKernelGGL(k_store_word, dim3(1), dim3(1), 0, stream, n_dev, n);
    e = hipGetLastError();
  }
  // The reduce-then-scan passes: every word they read is written earlier in the same call (each
  // pass writes all its tile counts, and its digit scan the totals of every digit the pass reads),
  // so nothing the previous call left matters.
  if (e == hipSuccess)
    e = launch_argsort(ctrl, keys, vals, n_dev, n, plan, hist, nullptr, counts, kSortReduceScan, stream);
  if (e == hipSuccess) e = hipStreamSynchronize(stream);
  if (e != hipSuccess) {
    char msg[200];
    std::snprintf(msg, sizeof msg, "static sort: %s", hipGetErrorString(e));
    return renderer_fail(r, e == hipErrorOutOfMemory ? SKM_ERR_MEMORY : SKM_ERR_DEVICE, msg);
  }
  return SKM_OK;
}
