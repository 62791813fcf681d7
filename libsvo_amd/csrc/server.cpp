// Render server: several resident scenes on one device, requests admitted
// by HBM budget and batched into multi-view launches as they arrive, with
// up to max_inflight launches overlapping on their own streams.
//
// The reference renders one camera per frame from its GLUT loop
// (projects/raycast_serialized_opencl/src/raycast_serialized_opencl.cpp:
// 684-749); this layer has no counterpart there.  Every launch it makes is
// the multi-view launch of render.cpp, so a request's pixels do not depend
// on which other requests share its launch.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <deque>
#include <map>
#include <memory>
#include <vector>

#include "svo_internal.h"
#include "trace.h"

using namespace svo_impl;

#define HIP_TRY(expr)                                                   \
  do {                                                                  \
    hipError_t e_ = (expr);                                             \
    if (e_ != hipSuccess) {                                             \
      set_error("%s failed: %s", #expr, hipGetErrorString(e_));         \
      return SVO_ERR_HIP;                                               \
    }                                                                   \
  } while (0)

namespace {

constexpr uint32_t kMaxInflight = 16;
constexpr uint32_t kMaxBatchViews = 1u << 16;
constexpr uint64_t kAlign = 256;

using Clock = std::chrono::steady_clock;

// First-fit arena over one device allocation, coalescing on free.
class Arena {
 public:
  void init(uint64_t bytes) {
    free_.clear();
    if (bytes) free_[0] = bytes;
    size_ = bytes;
    used_ = peak_ = 0;
  }
  bool alloc(uint64_t bytes, uint64_t* off) {
    bytes = (bytes + kAlign - 1) / kAlign * kAlign;
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second < bytes) continue;
      *off = it->first;
      const uint64_t rest = it->second - bytes, at = it->first + bytes;
      free_.erase(it);
      if (rest) free_[at] = rest;
      used_ += bytes;
      peak_ = std::max(peak_, used_);
      return true;
    }
    return false;
  }
  void release(uint64_t off, uint64_t bytes) {
    bytes = (bytes + kAlign - 1) / kAlign * kAlign;
    used_ -= bytes;
    auto it = free_.emplace(off, bytes).first;
    auto nx = std::next(it);
    if (nx != free_.end() && it->first + it->second == nx->first) {
      it->second += nx->second;
      free_.erase(nx);
    }
    if (it != free_.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        free_.erase(it);
      }
    }
  }
  uint64_t size() const { return size_; }
  uint64_t peak() const { return peak_; }
  void reset_peak() { peak_ = used_; }

 private:
  std::map<uint64_t, uint64_t> free_;
  uint64_t size_ = 0, used_ = 0, peak_ = 0;
};

struct Request {
  uint64_t id;
  svo_request req;
  double arrival_s;
  bool pooled = false;
  uint64_t pool_off = 0, pool_bytes = 0;
  double launch_s = 0;
};

struct Flight {
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;
  TraceFrame* d = nullptr;
  TraceFrame* h = nullptr;  // pinned
  uint32_t cap = 0;
  bool busy = false;
  uint32_t batch = 0;
  std::vector<Request> reqs;
};

uint64_t frame_bytes(const svo_camera& c) { return (uint64_t)c.width * c.height * 16; }

}  // namespace

struct svo_server {
  svo_renderer* r = nullptr;  // launches take its current launch shape
  int device = 0;             // the renderer's, kept so that teardown never reads the renderer
  svo_server_config cfg{};
  std::vector<svo_mm*> scenes;
  uint64_t scene_bytes = 0;
  char* pool = nullptr;
  Arena arena;
  Flight flights[kMaxInflight];
  std::deque<Request> waiting;
  uint64_t next_id = 0;
  uint32_t batches = 0;
  uint32_t max_views_seen = 0, max_inflight_seen = 0;
  uint64_t views_launched = 0;
  uint64_t expired = 0;
  Clock::time_point t0 = Clock::now();
  std::vector<svo_view> views;   // scratch
  StagedLaunch staged;           // scratch
  std::vector<svo_completion> held;  // completions not yet returned by poll

  ~svo_server() {
    hipSetDevice(device);
    for (Flight& f : flights) {
      if (f.stream) hipStreamSynchronize(f.stream);
      if (f.done) hipEventDestroy(f.done);
      if (f.stream) hipStreamDestroy(f.stream);
      hipFree(f.d);
      if (f.h) hipHostFree(f.h);
    }
    hipFree(pool);
  }

  double now() const { return std::chrono::duration<double>(Clock::now() - t0).count(); }

  uint32_t inflight() const {
    uint32_t n = 0;
    for (uint32_t i = 0; i < cfg.max_inflight; ++i) n += flights[i].busy;
    return n;
  }

  svo_status enqueue(const svo_request& q, double arrival, uint64_t* id) {
    if (q.scene >= scenes.size()) { set_error("request for unknown scene %u", q.scene); return SVO_ERR_ARG; }
    const svo_camera& c = q.camera;
    if (c.width == 0 || c.height == 0 || c.width > 16384 || c.height > 16384) {
      set_error("request frame %ux%u out of range", c.width, c.height);
      return SVO_ERR_ARG;
    }
    if (cfg.max_waiting && waiting.size() >= cfg.max_waiting) {
      set_error("server queue full (%u requests waiting)", cfg.max_waiting);
      return SVO_ERR_CAPACITY;
    }
    Request r;
    r.id = next_id++;
    r.req = q;
    r.arrival_s = arrival;
    r.pooled = !(q.out.t || q.out.node || q.out.treelet || q.out.rgba || q.out.pushes);
    if (r.pooled) {
      r.pool_bytes = frame_bytes(c);
      if (r.pool_bytes > arena.size()) {
        set_error("request frame of %llu bytes exceeds the frame pool (%llu bytes)",
                  (unsigned long long)r.pool_bytes, (unsigned long long)arena.size());
        return SVO_ERR_CAPACITY;
      }
    }
    waiting.push_back(r);
    if (id) *id = r.id;
    return SVO_OK;
  }

  // Finished launches -> completions; frames go back to the pool.
  svo_status retire(std::vector<svo_completion>* out) {
    for (uint32_t i = 0; i < cfg.max_inflight; ++i) {
      Flight& f = flights[i];
      if (!f.busy) continue;
      hipError_t e = hipEventQuery(f.done);
      if (e == hipErrorNotReady) continue;
      if (e != hipSuccess) { set_error("render launch failed: %s", hipGetErrorString(e)); return SVO_ERR_HIP; }
      const double t = now();
      for (const Request& q : f.reqs) {
        if (q.pooled) arena.release(q.pool_off, q.pool_bytes);
        out->push_back(svo_completion{q.id, q.arrival_s, q.launch_s, t, f.batch, (uint32_t)f.reqs.size(),
                                      SVO_REQ_DONE, 0});
      }
      f.reqs.clear();
      f.busy = false;
    }
    return SVO_OK;
  }

  // Starts launches on free streams while requests wait and frames fit.
  svo_status launch_ready() {
    const LaunchShape shape = renderer_shape(r);
    for (uint32_t i = 0; i < cfg.max_inflight && !waiting.empty(); ++i) {
      Flight& f = flights[i];
      if (f.busy) continue;
      uint64_t rays = 0;
      while (!waiting.empty() && f.reqs.size() < cfg.max_batch_views) {
        Request& q = waiting.front();
        const uint64_t qr = (uint64_t)q.req.camera.width * q.req.camera.height;
        if (!f.reqs.empty() && cfg.max_batch_rays && rays + qr > cfg.max_batch_rays) break;
        if (q.pooled && !arena.alloc(q.pool_bytes, &q.pool_off)) break;  // waits for memory
        rays += qr;
        f.reqs.push_back(q);
        waiting.pop_front();
      }
      if (f.reqs.empty()) break;  // head request waits for memory
      views.resize(f.reqs.size());
      for (size_t k = 0; k < f.reqs.size(); ++k) {
        const Request& q = f.reqs[k];
        svo_frame_out o = q.req.out;
        if (q.pooled) {
          const uint64_t px = (uint64_t)q.req.camera.width * q.req.camera.height;
          char* base = pool + q.pool_off;
          o.t = (float*)base;
          o.node = (uint32_t*)(base + 4 * px);
          o.treelet = (uint32_t*)(base + 8 * px);
          o.rgba = (uint32_t*)(base + 12 * px);
          o.pushes = nullptr;
        }
        views[k] = svo_view{scenes[q.req.scene], q.req.camera, o};
      }
      svo_status st = stage_views(views.data(), (uint32_t)views.size(), cfg.policy, shape, device, &staged);
      if (st != SVO_OK) {  // e.g. a scene with slots not uploaded: the requests wait
        for (auto it = f.reqs.rbegin(); it != f.reqs.rend(); ++it) {
          if (it->pooled) arena.release(it->pool_off, it->pool_bytes);
          waiting.push_front(*it);
        }
        f.reqs.clear();
        return st;
      }
      const uint32_t n = (uint32_t)staged.frames.size();
      if (n > f.cap) {
        HIP_TRY(hipFree(f.d));
        if (f.h) HIP_TRY(hipHostFree(f.h));
        f.d = nullptr;
        f.h = nullptr;
        f.cap = 0;
        HIP_TRY(hipMalloc(&f.d, n * sizeof(TraceFrame)));
        HIP_TRY(hipHostMalloc(&f.h, n * sizeof(TraceFrame), hipHostMallocDefault));
        f.cap = n;
      }
      std::copy(staged.frames.begin(), staged.frames.end(), f.h);
      HIP_TRY(hipMemcpyAsync(f.d, f.h, n * sizeof(TraceFrame), hipMemcpyHostToDevice, f.stream));
      if ((st = sync_staged_normals(staged, f.stream)) != SVO_OK) return st;
      int e = launch_trace(shape, f.d, n, staged.num_tiles, staged.stack_levels, cfg.policy, staged.shade, f.stream);
      if (e != 0) { set_error("render launch failed: %s", hipGetErrorString((hipError_t)e)); return SVO_ERR_HIP; }
      HIP_TRY(hipEventRecord(f.done, f.stream));
      const double t = now();
      for (Request& q : f.reqs) q.launch_s = t;
      f.batch = batches++;
      f.busy = true;
      views_launched += n;
      max_views_seen = std::max(max_views_seen, n);
      max_inflight_seen = std::max(max_inflight_seen, inflight());
    }
    return SVO_OK;
  }

  // Waiting requests past their timeout complete unrendered.
  void expire(std::vector<svo_completion>* out) {
    const double t = now();
    for (auto it = waiting.begin(); it != waiting.end();) {
      if (it->req.timeout_us && (t - it->arrival_s) * 1e6 > (double)it->req.timeout_us) {
        out->push_back(svo_completion{it->id, it->arrival_s, t, t, 0, 0, SVO_REQ_EXPIRED, 0});
        ++expired;
        it = waiting.erase(it);
      } else {
        ++it;
      }
    }
  }

  svo_status progress(std::vector<svo_completion>* out) {
    svo_status st = retire(out);
    if (st != SVO_OK) return st;
    expire(out);
    return launch_ready();
  }
};

extern "C" svo_status svo_server_create(svo_renderer* r, const svo_server_config* cfg, svo_server** out) {
  if (!r || !cfg || !out) { set_error("svo_server_create: null argument"); return SVO_ERR_ARG; }
  *out = nullptr;
  if (cfg->max_inflight < 1 || cfg->max_inflight > kMaxInflight) {
    set_error("max_inflight must be 1..%u", kMaxInflight);
    return SVO_ERR_ARG;
  }
  if (cfg->max_batch_views < 1 || cfg->max_batch_views > kMaxBatchViews) {
    set_error("max_batch_views must be 1..%u", kMaxBatchViews);
    return SVO_ERR_ARG;
  }
  if (cfg->policy != SVO_POLICY_TREELET && cfg->policy != SVO_POLICY_CPU3 && cfg->policy != SVO_POLICY_SINGLE) {
    set_error("unknown traversal policy %d", cfg->policy);
    return SVO_ERR_ARG;
  }
  if (cfg->hbm_budget_bytes && cfg->frame_pool_bytes > cfg->hbm_budget_bytes) {
    set_error("frame pool of %llu bytes exceeds the HBM budget of %llu bytes",
              (unsigned long long)cfg->frame_pool_bytes, (unsigned long long)cfg->hbm_budget_bytes);
    return SVO_ERR_CAPACITY;
  }
  HIP_TRY(hipSetDevice(renderer_device(r)));
  std::unique_ptr<svo_server> s(new svo_server());
  s->r = r;
  s->device = renderer_device(r);
  s->cfg = *cfg;
  if (cfg->frame_pool_bytes) HIP_TRY(hipMalloc(&s->pool, cfg->frame_pool_bytes));
  s->arena.init(cfg->frame_pool_bytes);
  for (uint32_t i = 0; i < cfg->max_inflight; ++i) {
    Flight& f = s->flights[i];
    HIP_TRY(hipStreamCreateWithFlags(&f.stream, hipStreamNonBlocking));
    HIP_TRY(hipEventCreateWithFlags(&f.done, hipEventDisableTiming));
    const uint32_t cap = std::min(cfg->max_batch_views, 256u);
    HIP_TRY(hipMalloc(&f.d, cap * sizeof(TraceFrame)));
    HIP_TRY(hipHostMalloc(&f.h, cap * sizeof(TraceFrame), hipHostMallocDefault));
    f.cap = cap;
    f.reqs.reserve(cap);
  }
  *out = s.release();
  return SVO_OK;
}

extern "C" void svo_server_destroy(svo_server* s) { delete s; }

extern "C" svo_status svo_server_add_scene(svo_server* s, svo_mm* mm, uint32_t* scene_id) {
  if (!s || !mm || !scene_id) { set_error("svo_server_add_scene: null argument"); return SVO_ERR_ARG; }
  svo_mm_info info;
  svo_status st = svo_mm_get_info(mm, &info);
  if (st != SVO_OK) return st;
  if ((int)info.device != s->device || info.device_bytes == 0) {
    set_error("scene is not resident on the server's device %d", s->device);
    return SVO_ERR_ARG;
  }
  const uint64_t need = s->scene_bytes + info.device_bytes + s->cfg.frame_pool_bytes;
  if (s->cfg.hbm_budget_bytes && need > s->cfg.hbm_budget_bytes) {
    set_error("scene of %llu device bytes does not fit the HBM budget (%llu of %llu bytes in use)",
              (unsigned long long)info.device_bytes, (unsigned long long)(s->scene_bytes + s->cfg.frame_pool_bytes),
              (unsigned long long)s->cfg.hbm_budget_bytes);
    return SVO_ERR_CAPACITY;
  }
  s->scene_bytes += info.device_bytes;
  s->scenes.push_back(mm);
  *scene_id = (uint32_t)s->scenes.size() - 1;
  return SVO_OK;
}

extern "C" svo_status svo_server_submit(svo_server* s, const svo_request* req, uint64_t* id) {
  if (!s || !req) { set_error("svo_server_submit: null argument"); return SVO_ERR_ARG; }
  return s->enqueue(*req, s->now(), id);
}

extern "C" svo_status svo_server_poll(svo_server* s, svo_completion* done, uint32_t cap, uint32_t* n_done) {
  if (!s || !n_done || (cap && !done)) { set_error("svo_server_poll: bad arguments"); return SVO_ERR_ARG; }
  HIP_TRY(hipSetDevice(s->device));
  *n_done = 0;
  // completions beyond `cap` are held for the next poll
  std::vector<svo_completion>& held = s->held;
  svo_status st = s->progress(&held);
  const uint32_t k = (uint32_t)std::min<size_t>(cap, held.size());
  std::copy(held.begin(), held.begin() + k, done);
  held.erase(held.begin(), held.begin() + k);
  *n_done = k;
  return st;
}

extern "C" svo_status svo_server_pending(const svo_server* s, uint32_t* n) {
  if (!s || !n) { set_error("svo_server_pending: null argument"); return SVO_ERR_ARG; }
  uint32_t k = (uint32_t)s->waiting.size();
  for (uint32_t i = 0; i < s->cfg.max_inflight; ++i) k += (uint32_t)s->flights[i].reqs.size();
  *n = k;
  return SVO_OK;
}

extern "C" svo_status svo_server_replay(svo_server* s, const svo_request* reqs, uint32_t n, svo_completion* done,
                                        svo_server_stats* stats) {
  if (!s || (n && !reqs)) { set_error("svo_server_replay: null argument"); return SVO_ERR_ARG; }
  uint32_t pend = 0;
  svo_server_pending(s, &pend);
  pend += (uint32_t)s->held.size();
  if (pend) { set_error("svo_server_replay needs an idle server (%u requests pending)", pend); return SVO_ERR_STATE; }
  for (uint32_t i = 1; i < n; ++i)
    if (!(reqs[i].arrival_s >= reqs[i - 1].arrival_s)) { set_error("arrival_s must be ascending"); return SVO_ERR_ARG; }
  HIP_TRY(hipSetDevice(s->device));
  for (Flight& f : s->flights)
    if (f.stream) HIP_TRY(hipStreamSynchronize(f.stream));  // outstanding caller work is not timed

  const uint64_t id0 = s->next_id;
  const uint64_t e0 = s->expired;
  uint64_t rejected = 0;
  const uint32_t b0 = s->batches;
  const uint64_t v0 = s->views_launched;
  s->max_views_seen = s->max_inflight_seen = 0;
  s->arena.reset_peak();
  std::vector<svo_completion> comp;
  comp.reserve(n);
  std::vector<svo_completion> fresh;
  s->t0 = Clock::now();
  uint32_t next = 0;
  while (comp.size() < n) {
    const double t = s->now();
    while (next < n && reqs[next].arrival_s <= t) {
      svo_status st = s->enqueue(reqs[next], reqs[next].arrival_s, nullptr);
      if (st == SVO_ERR_CAPACITY && s->cfg.max_waiting && s->waiting.size() >= s->cfg.max_waiting) {
        // queue full: the request completes as rejected (its id is the one it would have had)
        clear_error();
        comp.push_back(svo_completion{s->next_id++, reqs[next].arrival_s, t, t, 0, 0, SVO_REQ_REJECTED, 0});
        ++rejected;
      } else if (st != SVO_OK) {
        return st;
      }
      ++next;
    }
    fresh.clear();
    svo_status st = s->progress(&fresh);
    if (st != SVO_OK) return st;
    comp.insert(comp.end(), fresh.begin(), fresh.end());
  }
  std::vector<double> lat;
  lat.reserve(n);
  double first = n ? reqs[0].arrival_s : 0, last = 0, sum = 0;
  uint64_t rays = 0;
  for (const svo_completion& c : comp) {
    const uint64_t i = c.id - id0;
    if (c.status == SVO_REQ_DONE) {
      lat.push_back((c.done_s - c.arrival_s) * 1e3);
      sum += lat.back();
      rays += (uint64_t)reqs[i].camera.width * reqs[i].camera.height;
    }
    last = std::max(last, c.done_s);
    if (done) {
      done[i] = c;
      done[i].id = i;
    }
  }
  const uint64_t rendered = lat.size();
  if (stats) {
    svo_server_stats& S = *stats;
    S = svo_server_stats{};
    S.requests = n;
    S.rays = rays;
    S.batches = s->batches - b0;
    S.peak_pool_bytes = s->arena.peak();
    S.max_batch_views_seen = s->max_views_seen;
    S.max_inflight_seen = s->max_inflight_seen;
    S.wall_s = last - first;
    if (rendered && S.wall_s > 0) {
      S.frames_per_s = rendered / S.wall_s;
      S.mrays_per_s = rays / S.wall_s / 1e6;
    }
    S.expired = s->expired - e0;
    S.rejected = rejected;
    std::vector<double> sorted = lat;
    std::sort(sorted.begin(), sorted.end());
    auto pct = [&](double p) {
      if (sorted.empty()) return 0.0;
      // nearest-rank percentile
      size_t k = (size_t)std::ceil(p / 100.0 * sorted.size());
      return sorted[std::min(sorted.size(), std::max<size_t>(k, 1)) - 1];
    };
    S.p50_ms = pct(50);
    S.p99_ms = pct(99);
    S.mean_ms = rendered ? sum / rendered : 0;
    S.max_ms = sorted.empty() ? 0 : sorted.back();
    S.mean_batch_views = S.batches ? (double)(s->views_launched - v0) / S.batches : 0;
  }
  return SVO_OK;
}
