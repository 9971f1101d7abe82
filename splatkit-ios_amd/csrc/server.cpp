// The request loop of include/skm_server.h: resident scenes, each with its own renderer (world,
// stream, view-batch scratch); a scheduler thread that retires finished launches and admits waiting
// requests into new ones under the HBM budget; BGRA8 outputs converted on the device and copied
// into pinned host blocks.
#include "skm_server.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstdarg>
#include <cstdio>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "skm.h"
#include "skm_batch.h"
#include "skm_host.h"

static_assert(sizeof(skm_request) == 168 && sizeof(skm_completion) == 40 && sizeof(skm_server_config) == 24 &&
                  sizeof(skm_server_stats) == 96,
              "server ABI layouts");

namespace {

using Clock = std::chrono::steady_clock;

// What one launch may take past its first request: 2^26 cull lanes (views x the scene's lanes per
// view; about 4 GB of survivor scratch and as much again of tile pairs) and 2^25 pixels (four
// 3840x2160 views; 512 MB of image).
constexpr uint64_t kLaunchLanes = uint64_t{1} << 26;
constexpr uint64_t kLaunchPixels = uint64_t{1} << 25;
// Launches a request may take part in whose tile pairs overflowed (each grows the pair capacity to
// what the last one needed, so the next fits unless the scene's pairs keep growing); then it
// completes with SKM_ERR_CAPACITY.
constexpr uint32_t kMaxAttempts = 3;

double ms_between(Clock::time_point a, Clock::time_point b) {
  return std::chrono::duration<double, std::milli>(b - a).count();
}

// Pinned host blocks for outputs, reused by size.
class PinnedPool {
 public:
  ~PinnedPool() {
    for (auto& [bytes, list] : free_)
      for (void* p : list) (void)hipHostFree(p);
  }
  void* take(size_t bytes) {
    auto it = free_.find(bytes);
    if (it != free_.end() && !it->second.empty()) {
      void* p = it->second.back();
      it->second.pop_back();
      return p;
    }
    void* p = nullptr;
    if (hipHostMalloc(&p, bytes, hipHostMallocDefault) != hipSuccess) return nullptr;
    return p;
  }
  void give(void* p, size_t bytes) {
    if (p != nullptr) free_[bytes].push_back(p);
  }

 private:
  std::map<size_t, std::vector<void*>> free_;
};

struct Request {
  uint64_t ticket = 0;
  skm_request req{};
  Clock::time_point submitted, launched;
  uint8_t* pixels = nullptr;
  size_t bytes = 0;
  int32_t status = SKM_OK;
  uint32_t batch_views = 0;
  uint32_t attempts = 0;  // launches that dropped it (tile pairs overflowed)
  uint64_t retry_of = 0;  // the dropped launch it is to be launched again with (0: none)
  bool done = false;
};

struct Scene {
  skm_renderer* r = nullptr;
  uint32_t count = 0;
  uint64_t world_bytes = 0;
  skm::DeviceBuffer bgra;  // BGRA8 of the launch's views
  std::vector<Request*> running;
  hipEvent_t done = nullptr;
  bool busy = false;
};

}  // namespace

struct skm_server {
  skm_server_config cfg{};
  // Eager launches: a served batch's shape (request count, each request's resolution) changes from
  // launch to launch, so a captured graph would be captured again almost every time.
  skm_batch_tuning tuning{16, 8, 0, 0, 1, 0};
  int device = 0;
  uint64_t budget = 0;
  std::string error;

  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::vector<std::unique_ptr<Scene>> scenes;
  std::deque<Request*> pending;
  std::unordered_map<uint64_t, std::unique_ptr<Request>> requests;
  std::deque<skm_completion> completions;
  PinnedPool pool;
  uint64_t next_ticket = 1;
  bool stop = false;
  std::thread worker;

  // Stats window.
  uint64_t submitted = 0, completed = 0, rejected = 0, launches = 0, launched_views = 0;
  uint64_t retry_groups = 0;  // dropped launches so far (the marks of their requests)
  std::vector<double> latencies;
  Clock::time_point window_start = Clock::now();

  skm_status fail(skm_status s, const char* fmt, ...) {
    char buf[512];
    va_list args;
    va_start(args, fmt);
    std::vsnprintf(buf, sizeof(buf), fmt, args);
    va_end(args);
    error = buf;
    return s;
  }

  uint64_t resident() const {
    uint64_t b = 0;
    for (const auto& s : scenes) b += s->world_bytes;
    return b;
  }
  uint64_t scratch() const {
    uint64_t b = 0;
    for (const auto& s : scenes) b += skm::renderer_batch(s->r)->device_bytes() + s->bgra.bytes;
    return b;
  }
  uint32_t in_flight() const {
    uint32_t n = 0;
    for (const auto& s : scenes) n += s->busy ? 1u : 0u;
    return n;
  }

  void finish_request(Request* q, int32_t status, Clock::time_point now) {
    q->status = status;
    q->done = true;
    skm_completion c{};
    c.ticket = q->ticket;
    c.user_data = q->req.user_data;
    c.status = status;
    c.batch_views = q->batch_views;
    c.queue_ms = q->batch_views ? ms_between(q->submitted, q->launched) : 0.0;
    c.latency_ms = ms_between(q->submitted, now);
    completions.push_back(c);
    if (status == SKM_OK) {
      ++completed;
      latencies.push_back(c.latency_ms);
    } else {
      ++rejected;
    }
  }

  // Bytes a launch of these requests of scene `s` adds to what its executor already holds.
  uint64_t growth(Scene& sc, const std::vector<Request*>& batch) const {
    skm::Batch* b = skm::renderer_batch(sc.r);
    uint64_t tiles = 0, pixels = 0;
    for (const Request* q : batch) {
      const uint64_t w = q->req.view.width, h = q->req.view.height;
      tiles += ((w + 15) / 16) * ((h + 15) / 16);
      pixels += w * h;
    }
    const uint64_t lanes = (uint64_t{sc.count} + skm::kCullTile - 1) / skm::kCullTile * skm::kCullTile * batch.size();
    const uint64_t need = b->bytes_for(static_cast<uint32_t>(batch.size()), lanes, static_cast<uint32_t>(tiles),
                                       pixels) + std::max<uint64_t>(pixels * 4, sc.bgra.bytes);
    const uint64_t have = b->device_bytes() + sc.bgra.bytes;
    return need > have ? need - have : 0;
  }

  // Launches `batch` on the scene's stream: the view batch, the BGRA8 conversion, one copy per
  // request into its pinned block, and the completion event. Every host block and the output buffer
  // are taken before any work is queued; if queueing fails part-way, the stream is drained before
  // returning, so no copy is still writing a block its request gives back.
  skm_status launch(Scene& sc, std::vector<Request*>& batch) {
    std::vector<skm_view> views(batch.size());
    uint64_t pixels = 0;
    for (size_t k = 0; k < batch.size(); ++k) {
      Request* q = batch[k];
      views[k] = q->req.view;
      const size_t bytes = size_t{q->req.view.width} * q->req.view.height * 4;
      if (q->pixels != nullptr && q->bytes != bytes) {
        pool.give(q->pixels, q->bytes);
        q->pixels = nullptr;
      }
      q->bytes = bytes;
      if (q->pixels == nullptr) q->pixels = static_cast<uint8_t*>(pool.take(q->bytes));
      if (q->pixels == nullptr) return fail(SKM_ERR_MEMORY, "pinned output");
      pixels += uint64_t{q->req.view.width} * q->req.view.height;
    }
    if (sc.bgra.bytes < pixels * 4 && sc.bgra.allocate(pixels * 4) != hipSuccess)
      return fail(SKM_ERR_MEMORY, "output buffer");
    skm_batch_tuning t = tuning;
    const skm_range whole{0, sc.count};
    skm_status st = skm_draw_views(sc.r, views.data(), static_cast<uint32_t>(views.size()), &whole, 1, &t);
    if (st != SKM_OK) return fail(st, "%s", skm_last_error(sc.r));
    skm::Batch* b = skm::renderer_batch(sc.r);
    hipStream_t stream = skm::renderer_stream(sc.r);
    hipError_t e = skm::launch_to_bgra8(b->image_device(), sc.bgra.as<uint32_t>(), b->pixels(), stream);
    const auto now = Clock::now();
    for (size_t k = 0; k < batch.size() && e == hipSuccess; ++k) {
      Request* q = batch[k];
      const skm::ViewInfo& vi = b->info(static_cast<uint32_t>(k));
      e = hipMemcpyAsync(q->pixels, sc.bgra.as<uint32_t>() + vi.pixel_base, q->bytes, hipMemcpyDeviceToHost, stream);
      q->launched = now;
      q->batch_views = static_cast<uint32_t>(batch.size());
    }
    if (e == hipSuccess) e = hipEventRecord(sc.done, stream);
    if (e != hipSuccess) {
      (void)hipStreamSynchronize(stream);
      return fail(SKM_ERR_DEVICE, "launch: %s", hipGetErrorString(e));
    }
    sc.running = batch;
    sc.busy = true;
    ++launches;
    launched_views += batch.size();
    return SKM_OK;
  }

  // Retires finished launches. The requests of a launch whose tile pairs overflowed go back to the
  // head of the queue in arrival order, marked with that launch: admission launches them again as
  // the same group and nothing else (the scene's pair capacity has grown to fit exactly those views;
  // requests that arrived meanwhile would make the launch bigger than that) and prices the grown
  // scratch against the budget again. After kMaxAttempts drops, or when the capacity is at its 32-bit
  // limit, they complete with SKM_ERR_CAPACITY.
  void retire() {
    for (auto& up : scenes) {
      Scene& sc = *up;
      if (!sc.busy || hipEventQuery(sc.done) != hipSuccess) continue;
      sc.busy = false;
      const skm_status st = skm_finish(sc.r);
      const auto now = Clock::now();
      if (st == SKM_ERR_CAPACITY) {
        const bool limit = skm::renderer_batch(sc.r)->at_pair_limit();
        const uint64_t group = ++retry_groups;
        std::vector<Request*> again;
        for (Request* q : sc.running) {
          if (++q->attempts >= kMaxAttempts || limit) {
            finish_request(q, SKM_ERR_CAPACITY, now);
          } else {
            q->retry_of = group;
            again.push_back(q);
          }
        }
        pending.insert(pending.begin(), again.begin(), again.end());
      } else {
        for (Request* q : sc.running) finish_request(q, st, now);
      }
      sc.running.clear();
      cv_done.notify_all();
    }
  }

  // Forms launches from the waiting requests: the oldest request's scene first, then its other
  // waiting requests in arrival order, as many as the budget and max_batch_views allow.
  void admit() {
    const uint32_t max_flight = cfg.max_in_flight ? cfg.max_in_flight : static_cast<uint32_t>(scenes.size());
    std::vector<bool> tried(scenes.size(), false);
    while (!pending.empty() && in_flight() < max_flight) {
      Request* head = nullptr;
      for (Request* q : pending)
        if (!scenes[q->req.scene]->busy && !tried[q->req.scene]) {
          head = q;
          break;
        }
      if (head == nullptr) return;
      Scene& sc = *scenes[head->req.scene];
      tried[head->req.scene] = true;
      // At most kLaunchLanes cull lanes and kLaunchPixels pixels per launch (past the first request):
      // a launch of that size already fills the device for many milliseconds, and bigger ones would
      // only grow every executor's scratch (each reallocation waits for the whole device).
      const uint64_t lanes_per_view = (uint64_t{sc.count} + skm::kCullTile - 1) / skm::kCullTile * skm::kCullTile;
      std::vector<Request*> batch;
      uint64_t pixels = 0;
      for (Request* q : pending) {
        if (q->req.scene != head->req.scene || q->retry_of != head->retry_of ||
            batch.size() >= std::max(1u, cfg.max_batch_views))
          continue;
        const uint64_t px = uint64_t{q->req.view.width} * q->req.view.height;
        if (!batch.empty() &&
            ((batch.size() + 1) * lanes_per_view > kLaunchLanes || pixels + px > kLaunchPixels))
          break;
        batch.push_back(q);
        pixels += px;
      }
      const uint64_t base = resident() + scratch();
      while (!batch.empty() && base + growth(sc, batch) > budget) batch.pop_back();
      if (batch.empty()) {
        // Not even one fits next to the other executors' scratch: release idle executors and
        // retry once everything is idle; a request that cannot fit alone is rejected.
        if (in_flight() > 0) continue;
        for (auto& other : scenes) {
          skm::renderer_batch(other->r)->release();
          other->bgra.reset();
        }
        std::vector<Request*> one{head};
        if (resident() + growth(sc, one) > budget) {
          pending.erase(std::find(pending.begin(), pending.end(), head));
          finish_request(head, SKM_ERR_MEMORY, Clock::now());
          cv_done.notify_all();
        }
        continue;
      }
      for (Request* q : batch) pending.erase(std::find(pending.begin(), pending.end(), q));
      if (const skm_status st = launch(sc, batch); st != SKM_OK) {
        const auto now = Clock::now();
        for (Request* q : batch) finish_request(q, st, now);
        cv_done.notify_all();
      }
    }
  }

  void run() {
    (void)hipSetDevice(device);
    std::unique_lock<std::mutex> lock(mu);
    while (!stop) {
      retire();
      admit();
      if (in_flight() > 0) {
        // Launches are in flight: poll their events at a fine grain.
        lock.unlock();
        std::this_thread::sleep_for(std::chrono::microseconds(20));
        lock.lock();
      } else {
        cv_work.wait_for(lock, std::chrono::milliseconds(5), [this] { return stop || !pending.empty(); });
      }
    }
  }
};

extern "C" {

skm_status skm_server_create(const skm_server_config* config, skm_server** out) {
  if (out == nullptr) return SKM_ERR_INVALID;
  *out = nullptr;
  auto s = std::make_unique<skm_server>();
  if (config) s->cfg = *config;
  if (s->cfg.max_batch_views == 0) s->cfg.max_batch_views = 64;
  if (s->cfg.sort_depth != 0 && s->cfg.sort_depth != 16 && s->cfg.sort_depth != 32) return SKM_ERR_INVALID;
  s->device = s->cfg.device;
  if (hipSetDevice(s->device) != hipSuccess) return SKM_ERR_DEVICE;
  size_t free_b = 0, total_b = 0;
  if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) return SKM_ERR_DEVICE;
  s->budget = s->cfg.hbm_budget_bytes ? s->cfg.hbm_budget_bytes : free_b / 10 * 9;
  skm_server* raw = s.release();
  raw->worker = std::thread([raw] { raw->run(); });
  *out = raw;
  return SKM_OK;
}

void skm_server_destroy(skm_server* s) {
  if (s == nullptr) return;
  {
    std::lock_guard<std::mutex> g(s->mu);
    s->stop = true;
  }
  s->cv_work.notify_all();
  if (s->worker.joinable()) s->worker.join();
  (void)hipSetDevice(s->device);
  for (auto& sc : s->scenes) {
    (void)skm_finish(sc->r);
    if (sc->done) (void)hipEventDestroy(sc->done);
    sc->bgra.reset();
    skm_renderer_destroy(sc->r);
  }
  for (auto& [t, q] : s->requests) s->pool.give(q->pixels, q->bytes);
  s->requests.clear();
  delete s;
}

const char* skm_server_last_error(const skm_server* s) { return s ? s->error.c_str() : "null server"; }

skm_status skm_server_add_scene(skm_server* s, const skm_cloud_view* cloud, int max_sh_degree, uint32_t* scene) {
  if (s == nullptr || cloud == nullptr || scene == nullptr) return SKM_ERR_INVALID;
  std::lock_guard<std::mutex> g(s->mu);
  const uint64_t bytes = uint64_t{cloud->count} * (32 + 4 * skm::sh_stride(std::min(cloud->sh_degree, max_sh_degree)));
  if (s->resident() + s->scratch() + bytes > s->budget)
    return s->fail(SKM_ERR_MEMORY, "scene of %llu bytes exceeds the budget", static_cast<unsigned long long>(bytes));
  (void)hipSetDevice(s->device);
  skm_config c{s->device, 16, 8, 0};
  auto sc = std::make_unique<Scene>();
  if (skm_renderer_create(&c, &sc->r) != SKM_OK) return s->fail(SKM_ERR_DEVICE, "renderer");
  skm_render_policy p{};
  skm_device_capabilities caps{};
  (void)skm_get_device_capabilities(sc->r, &caps);
  p = caps.policy.fallback;
  p.sort_depth = s->cfg.sort_depth ? s->cfg.sort_depth : 16;
  char why[256] = {};
  skm_status st = skm_apply_render_policy(sc->r, &p, why, sizeof why);
  if (st == SKM_OK) st = skm_upload_world(sc->r, cloud, max_sh_degree);
  if (st == SKM_OK && hipEventCreateWithFlags(&sc->done, hipEventDisableTiming) != hipSuccess) st = SKM_ERR_DEVICE;
  if (st != SKM_OK) {
    s->fail(st, "scene: %s", skm_last_error(sc->r));
    skm_renderer_destroy(sc->r);
    return st;
  }
  sc->count = cloud->count;
  sc->world_bytes = skm::renderer_world_bytes(sc->r);
  *scene = static_cast<uint32_t>(s->scenes.size());
  s->scenes.push_back(std::move(sc));
  return SKM_OK;
}

skm_status skm_server_submit(skm_server* s, const skm_request* request, uint64_t* ticket) {
  if (s == nullptr || request == nullptr || ticket == nullptr) return SKM_ERR_INVALID;
  const auto now = Clock::now();
  std::lock_guard<std::mutex> g(s->mu);
  if (request->scene >= s->scenes.size()) return s->fail(SKM_ERR_INVALID, "unknown scene %u", request->scene);
  const skm_view& v = request->view;
  if (v.width == 0 || v.height == 0 || v.width > 16384 || v.height > 16384)
    return s->fail(SKM_ERR_INVALID, "view extent out of range");
  auto q = std::make_unique<Request>();
  q->ticket = s->next_ticket++;
  q->req = *request;
  q->submitted = now;
  *ticket = q->ticket;
  s->pending.push_back(q.get());
  s->requests.emplace(q->ticket, std::move(q));
  ++s->submitted;
  s->cv_work.notify_one();
  return SKM_OK;
}

skm_status skm_server_poll(skm_server* s, skm_completion* out, uint32_t capacity, uint32_t* count,
                           double timeout_ms) {
  if (s == nullptr || count == nullptr || (capacity > 0 && out == nullptr)) return SKM_ERR_INVALID;
  std::unique_lock<std::mutex> lock(s->mu);
  if (s->completions.empty() && timeout_ms > 0)
    s->cv_done.wait_for(lock, std::chrono::duration<double, std::milli>(timeout_ms),
                        [s] { return !s->completions.empty(); });
  uint32_t n = 0;
  while (n < capacity && !s->completions.empty()) {
    out[n++] = s->completions.front();
    s->completions.pop_front();
  }
  *count = n;
  return SKM_OK;
}

skm_status skm_server_image(skm_server* s, uint64_t ticket, const uint8_t** pixels, uint32_t* width,
                            uint32_t* height) {
  if (s == nullptr || pixels == nullptr) return SKM_ERR_INVALID;
  std::lock_guard<std::mutex> g(s->mu);
  auto it = s->requests.find(ticket);
  if (it == s->requests.end()) return s->fail(SKM_ERR_INVALID, "unknown ticket");
  const Request& q = *it->second;
  if (!q.done) return s->fail(SKM_ERR_STATE, "request %llu not complete", static_cast<unsigned long long>(ticket));
  if (q.status != SKM_OK) return s->fail(static_cast<skm_status>(q.status), "request failed");
  *pixels = q.pixels;
  if (width) *width = q.req.view.width;
  if (height) *height = q.req.view.height;
  return SKM_OK;
}

skm_status skm_server_release(skm_server* s, uint64_t ticket) {
  if (s == nullptr) return SKM_ERR_INVALID;
  std::lock_guard<std::mutex> g(s->mu);
  auto it = s->requests.find(ticket);
  if (it == s->requests.end()) return s->fail(SKM_ERR_INVALID, "unknown ticket");
  if (!it->second->done) return s->fail(SKM_ERR_STATE, "request in flight");
  s->pool.give(it->second->pixels, it->second->bytes);
  s->requests.erase(it);
  return SKM_OK;
}

skm_status skm_server_stats_read(skm_server* s, skm_server_stats* out, int reset) {
  if (s == nullptr || out == nullptr) return SKM_ERR_INVALID;
  std::lock_guard<std::mutex> g(s->mu);
  *out = skm_server_stats{};
  out->submitted = s->submitted;
  out->completed = s->completed;
  out->rejected = s->rejected;
  out->launches = s->launches;
  out->resident_bytes = s->resident();
  out->scratch_bytes = s->scratch();
  out->budget_bytes = s->budget;
  std::vector<double> l = s->latencies;
  if (!l.empty()) {
    // summarizeTimings (TimingSummary.cpp:7-22): sorted, p50 = [n/2], p99 = [n*99/100].
    std::sort(l.begin(), l.end());
    double sum = 0;
    for (double x : l) sum += x;
    out->p50_ms = l[l.size() / 2];
    out->p99_ms = l[(l.size() * 99) / 100];
    out->mean_ms = sum / static_cast<double>(l.size());
  }
  out->mean_batch_views = s->launches ? static_cast<double>(s->launched_views) / s->launches : 0.0;
  const double secs = std::chrono::duration<double>(Clock::now() - s->window_start).count();
  out->frames_per_s = secs > 0 ? static_cast<double>(l.size()) / secs : 0.0;
  if (reset) {
    s->latencies.clear();
    s->submitted = s->completed = s->rejected = s->launches = s->launched_views = 0;
    s->window_start = Clock::now();
  }
  return SKM_OK;
}

}  // extern "C"
