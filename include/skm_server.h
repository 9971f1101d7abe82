/*
 * skm server: a request loop over several device-resident scenes on one MI355X.
 *
 * No counterpart in the reference, which draws one frame per engine tick
 * (packages/splatkit-engine/src/engine/SplatEngine.cpp:559-638). Its anchors: the per-splat device
 * footprint of a world (gpuWorldBytes, packages/splatkit-engine/src/rendering/GpuLayout.cpp:72-77)
 * priced against available memory like the engine's MemoryGuard
 * (packages/splat-core/include/splat/loading/SplatWorldLoader.h:64), and p50/p99 latency like
 * summarizeTimings (packages/splat-core/src/diagnostics/TimingSummary.cpp:7-22): p50 = sorted[n / 2],
 * p99 = sorted[n * 99 / 100].
 *
 * Scenes are uploaded once and stay resident. Requests (scene, camera, resolution) are queued as
 * they arrive; a scheduler thread admits them into launches as executors and the HBM budget free
 * up: continuous batching, each launch one graph replay of the view-batch pipeline over the
 * requests of one scene that are waiting, with launches of different scenes overlapping on the
 * device. Pixels reach host memory as BGRA8 (MetalFrameCapture.h:18). A request's pixels do not
 * depend on what it was batched with.
 *
 * Thread safety: submit, poll, image, release and stats may be called from any thread.
 */
#pragma once
#include <stddef.h>
#include <stdint.h>

#include "skm.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct skm_server skm_server;

typedef struct skm_server_config {
  int32_t device;
  uint32_t max_batch_views; /* most requests per launch; 1 = one request per launch */
  uint64_t hbm_budget_bytes; /* scenes + launch scratch + outputs; 0 = 90% of free device memory */
  uint32_t sort_depth;       /* depth key bits: 16 or 32 (0 = 16, the policy default) */
  uint32_t max_in_flight;    /* launches on the device at once (0 = one per scene) */
} skm_server_config;

typedef struct skm_request {
  uint32_t scene;
  uint32_t reserved;
  uint64_t user_data;
  skm_view view; /* camera and resolution */
} skm_request;

typedef struct skm_completion {
  uint64_t ticket;
  uint64_t user_data;
  int32_t status;       /* SKM_OK, or why the request was not rendered */
  uint32_t batch_views; /* requests in the launch that rendered it */
  double queue_ms;      /* submit -> launch */
  double latency_ms;    /* submit -> pixels in host memory */
} skm_completion;

typedef struct skm_server_stats {
  uint64_t submitted;
  uint64_t completed;
  uint64_t rejected;
  uint64_t launches;
  uint64_t resident_bytes; /* scenes */
  uint64_t scratch_bytes;  /* launch scratch and outputs currently allocated */
  uint64_t budget_bytes;
  double p50_ms;           /* latency percentiles over the completions since the last reset */
  double p99_ms;
  double mean_ms;
  double mean_batch_views;
  double frames_per_s;     /* completions per second of wall time since the last reset */
} skm_server_stats;

skm_status skm_server_create(const skm_server_config* config, skm_server** out);
/* Waits for the launches in flight, then frees everything. */
void skm_server_destroy(skm_server* s);
const char* skm_server_last_error(const skm_server* s);

/* Uploads a scene; it stays resident until the server is destroyed. Fails with SKM_ERR_MEMORY when
 * the scene does not fit the budget next to the scenes already resident. */
skm_status skm_server_add_scene(skm_server* s, const skm_cloud_view* cloud, int max_sh_degree, uint32_t* scene);

/* Queues a request; returns at once. A request that can never fit the budget completes with
 * SKM_ERR_MEMORY. */
skm_status skm_server_submit(skm_server* s, const skm_request* request, uint64_t* ticket);

/* Takes up to `capacity` completions, waiting up to timeout_ms for the first one. */
skm_status skm_server_poll(skm_server* s, skm_completion* out, uint32_t capacity, uint32_t* count,
                           double timeout_ms);

/* The BGRA8 pixels of a completed request (top-down rows, width * height * 4 bytes), valid until
 * skm_server_release. */
skm_status skm_server_image(skm_server* s, uint64_t ticket, const uint8_t** pixels, uint32_t* width,
                            uint32_t* height);
skm_status skm_server_release(skm_server* s, uint64_t ticket);

/* Counters and latency percentiles; reset restarts the window. */
skm_status skm_server_stats_read(skm_server* s, skm_server_stats* out, int reset);

#ifdef __cplusplus
}
#endif
