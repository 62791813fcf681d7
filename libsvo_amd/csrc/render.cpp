// Launch side of the renderer: view descriptors staged through pinned memory
// into a small device ring, multi-view launches, event timing, the launch
// shape autotuner and hipGraph-captured batches.
//
// Replaces the per-frame renderToBuffer enqueue of the OpenCL app
// (projects/raycast_serialized_opencl/src/raycast_serialized_opencl.cpp:
// 729-749), generalised to many cameras and scenes per launch.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
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
constexpr uint32_t kRing = 8;         // launches in flight per renderer
constexpr uint32_t kMaxViews = 1u << 16;
}  // namespace

namespace svo_impl {

svo_status check_view_mode(uint32_t mode, int policy, uint32_t i) {
  if (mode > SVO_VIEW_FILL_SHADOW) { set_error("view %u: unknown view mode %u", i, mode); return SVO_ERR_ARG; }
  if ((mode == SVO_VIEW_FILL_SHADOW || mode == SVO_VIEW_FILL_REFLECTION) && policy != SVO_POLICY_SINGLE) {
    set_error("view %u: view mode %u traces single-SVO secondary rays and needs SVO_POLICY_SINGLE", i, mode);
    return SVO_ERR_ARG;
  }
  return SVO_OK;
}

bool mode_needs_normals(uint32_t mode) {
  return mode == SVO_VIEW_NORMAL || mode == SVO_VIEW_PHONG || mode == SVO_VIEW_FILL_SHADOW ||
         mode == SVO_VIEW_FILL_REFLECTION;
}

svo_status stage_views_into(const svo_view* views, uint32_t n, int policy, const LaunchShape& shape, int device,
                            TraceFrame* frames, StageTotals* totals) {
  if (!views || n == 0 || n > kMaxViews) { set_error("need 1..%u views", kMaxViews); return SVO_ERR_ARG; }
  const uint32_t tw = shape.tile_w, th = 64u / shape.tile_w;
  uint64_t tiles = 0;
  uint32_t levels = 1;
  bool shade = false;
  for (uint32_t i = 0; i < n; ++i) {
    const svo_view& v = views[i];
    if (!v.mm) { set_error("view %u has no residency handle", i); return SVO_ERR_ARG; }
    if (mm_device(v.mm) != device) { set_error("view %u lives on device %d, renderer on %d", i, mm_device(v.mm), device); return SVO_ERR_ARG; }
    const svo_camera& c = v.camera;
    if (c.width == 0 || c.height == 0 || c.width > 16384 || c.height > 16384) {
      set_error("view %u: frame %ux%u out of range", i, c.width, c.height);
      return SVO_ERR_ARG;
    }
    TraceFrame& F = frames[i];
    svo_status r = mm_trace_pools(v.mm, policy, &F.pools);
    if (r != SVO_OK) return r;
    F.cam = c;
    F.out = v.out;
    F.tile_begin = (uint32_t)tiles;
    F.tiles_x = (c.width + tw - 1) / tw;
    r = check_view_mode(c.view_mode, policy, i);
    if (r != SVO_OK) return r;
    if (v.feedback_stride > 65535u) { set_error("view %u: feedback stride %u too large", i, v.feedback_stride); return SVO_ERR_ARG; }
    F.fb_stride = v.feedback_stride;
    F.pad = 0;
    tiles += (uint64_t)F.tiles_x * ((c.height + th - 1) / th);
    levels = std::max(levels, F.pools.stack_levels);
    shade = shade || v.out.shade != nullptr;
  }
  if (tiles > 0x7FFFFFFFull) { set_error("launch of %llu tiles too large", (unsigned long long)tiles); return SVO_ERR_ARG; }
  *totals = StageTotals{(uint32_t)tiles, levels, shade};
  return SVO_OK;
}

bool view_needs_normals(const svo_view& v) {
  return v.out.shade && mode_needs_normals(v.camera.view_mode);
}

svo_status stage_views(const svo_view* views, uint32_t n, int policy, const LaunchShape& shape, int device,
                       StagedLaunch* out) {
  if (!views || n == 0 || n > kMaxViews) { set_error("need 1..%u views", kMaxViews); return SVO_ERR_ARG; }
  out->frames.resize(n);
  out->normal_handles.clear();
  StageTotals t{};
  const svo_status r = stage_views_into(views, n, policy, shape, device, out->frames.data(), &t);
  if (r != SVO_OK) return r;
  for (uint32_t i = 0; i < n; ++i)
    if (view_needs_normals(views[i]) &&
        std::find(out->normal_handles.begin(), out->normal_handles.end(), views[i].mm) == out->normal_handles.end())
      out->normal_handles.push_back(views[i].mm);
  out->num_tiles = t.num_tiles;
  out->stack_levels = t.stack_levels;
  out->shade = t.shade;
  return SVO_OK;
}

svo_status sync_staged_normals(const StagedLaunch& sg, void* stream) {
  for (svo_mm* mm : sg.normal_handles) {
    const svo_status r = mm_sync_normals(mm, stream);
    if (r != SVO_OK) return r;
  }
  return SVO_OK;
}

}  // namespace svo_impl

namespace {

struct Slot {
  TraceFrame* d = nullptr;
  TraceFrame* h = nullptr;  // pinned
  uint32_t cap = 0;
  hipEvent_t done = nullptr;  // recorded after the launch that read d
  bool pending = false;
};

svo_status shape_from_config(const svo_launch_config* c, LaunchShape* s) {
  LaunchShape t{c->tile_width, c->waves_per_group, c->stack_layout};
  if (!shape_supported(t)) {
    set_error("launch config (tile %u, waves %u, stack %u) not supported", c->tile_width, c->waves_per_group,
              c->stack_layout);
    return SVO_ERR_ARG;
  }
  *s = t;
  return SVO_OK;
}

}  // namespace

struct svo_renderer {
  int device = 0;
  LaunchShape shape{8, 1, SVO_STACK_LDS};
  Slot ring[kRing];
  uint32_t next = 0;

  ~svo_renderer() {
    int cur = 0;
    hipGetDevice(&cur);
    hipSetDevice(device);
    for (Slot& s : ring) {
      if (s.pending) hipEventSynchronize(s.done);
      if (s.done) hipEventDestroy(s.done);
      hipFree(s.d);
      if (s.h) hipHostFree(s.h);
    }
    hipSetDevice(cur);
  }

  // Copies the staged descriptors into the next ring slot on `st` and
  // returns its device pointer; the slot is reused only after the launch that
  // read it has finished.
  svo_status upload(const StagedLaunch& sg, hipStream_t st, Slot** used) {
    Slot& s = ring[next];
    next = (next + 1) % kRing;
    if (s.pending) {
      HIP_TRY(hipEventSynchronize(s.done));
      s.pending = false;
    }
    const uint32_t n = (uint32_t)sg.frames.size();
    if (n > s.cap) {
      if (s.d) HIP_TRY(hipFree(s.d));
      if (s.h) HIP_TRY(hipHostFree(s.h));
      s.d = nullptr;
      s.h = nullptr;
      s.cap = 0;
      const uint32_t cap = std::max(n, 64u);
      HIP_TRY(hipMalloc(&s.d, cap * sizeof(TraceFrame)));
      HIP_TRY(hipHostMalloc(&s.h, cap * sizeof(TraceFrame), hipHostMallocDefault));
      s.cap = cap;
    }
    if (!s.done) HIP_TRY(hipEventCreateWithFlags(&s.done, hipEventDisableTiming));
    std::copy(sg.frames.begin(), sg.frames.end(), s.h);
    HIP_TRY(hipMemcpyAsync(s.d, s.h, n * sizeof(TraceFrame), hipMemcpyHostToDevice, st));
    *used = &s;
    return SVO_OK;
  }

  svo_status launch(const LaunchShape& sh, const StagedLaunch& sg, Slot* s, int policy, hipStream_t st) {
    const svo_status ns = sync_staged_normals(sg, st);
    if (ns != SVO_OK) return ns;
    int e = launch_trace(sh, s->d, (uint32_t)sg.frames.size(), sg.num_tiles, sg.stack_levels, policy, sg.shade, st);
    if (e != 0) { set_error("trace launch failed: %s", hipGetErrorString((hipError_t)e)); return SVO_ERR_HIP; }
    HIP_TRY(hipEventRecord(s->done, st));
    s->pending = true;
    return SVO_OK;
  }

  svo_status timed(const LaunchShape& sh, const svo_view* views, uint32_t n, int policy, hipStream_t st,
                   uint32_t warmup, uint32_t iters, float* ms) {
    StagedLaunch sg;
    svo_status r = stage_views(views, n, policy, sh, device, &sg);
    if (r != SVO_OK) return r;
    Slot* s = nullptr;
    r = upload(sg, st, &s);
    if (r != SVO_OK) return r;
    for (uint32_t i = 0; i < warmup; ++i)
      if ((r = launch(sh, sg, s, policy, st)) != SVO_OK) return r;
    hipEvent_t a, b;
    HIP_TRY(hipEventCreate(&a));
    HIP_TRY(hipEventCreate(&b));
    HIP_TRY(hipEventRecord(a, st));
    for (uint32_t i = 0; i < iters; ++i)
      if ((r = launch(sh, sg, s, policy, st)) != SVO_OK) return r;
    HIP_TRY(hipEventRecord(b, st));
    HIP_TRY(hipEventSynchronize(b));
    float total = 0;
    HIP_TRY(hipEventElapsedTime(&total, a, b));
    hipEventDestroy(a);
    hipEventDestroy(b);
    *ms = total / (float)iters;
    return SVO_OK;
  }
};

namespace svo_impl {
LaunchShape renderer_shape(const svo_renderer* r) { return r->shape; }
int renderer_device(const svo_renderer* r) { return r->device; }
}  // namespace svo_impl

extern "C" svo_status svo_renderer_create(int device, svo_renderer** out) {
  if (!out) { set_error("svo_renderer_create: out is null"); return SVO_ERR_ARG; }
  *out = nullptr;
  int ndev = 0;
  HIP_TRY(hipGetDeviceCount(&ndev));
  if (device < 0 || device >= ndev) { set_error("svo_renderer_create: device %d not present (%d devices)", device, ndev); return SVO_ERR_ARG; }
  svo_renderer* r = new svo_renderer();
  r->device = device;
  *out = r;
  return SVO_OK;
}

extern "C" void svo_renderer_destroy(svo_renderer* r) { delete r; }

extern "C" svo_status svo_renderer_set_config(svo_renderer* r, const svo_launch_config* c) {
  if (!r) { set_error("svo_renderer_set_config: null renderer"); return SVO_ERR_ARG; }
  if (!c) { r->shape = LaunchShape{8, 1, SVO_STACK_LDS}; return SVO_OK; }
  return shape_from_config(c, &r->shape);
}

extern "C" svo_status svo_renderer_get_config(const svo_renderer* r, svo_launch_config* c) {
  if (!r || !c) { set_error("svo_renderer_get_config: null argument"); return SVO_ERR_ARG; }
  *c = svo_launch_config{r->shape.tile_w, r->shape.waves, r->shape.stack, 0};
  return SVO_OK;
}

extern "C" svo_status svo_render_views(svo_renderer* r, const svo_view* views, uint32_t n, int policy, void* stream) {
  if (!r) { set_error("svo_render_views: null renderer"); return SVO_ERR_ARG; }
  HIP_TRY(hipSetDevice(r->device));
  StagedLaunch sg;
  svo_status st = stage_views(views, n, policy, r->shape, r->device, &sg);
  if (st != SVO_OK) return st;
  Slot* s = nullptr;
  if ((st = r->upload(sg, (hipStream_t)stream, &s)) != SVO_OK) return st;
  return r->launch(r->shape, sg, s, policy, (hipStream_t)stream);
}

extern "C" svo_status svo_render_views_timed(svo_renderer* r, const svo_view* views, uint32_t n, int policy,
                                             void* stream, uint32_t iters, float* mean_ms) {
  if (!r || !mean_ms || iters == 0) { set_error("svo_render_views_timed: bad arguments"); return SVO_ERR_ARG; }
  HIP_TRY(hipSetDevice(r->device));
  return r->timed(r->shape, views, n, policy, (hipStream_t)stream, 0, iters, mean_ms);
}

extern "C" svo_status svo_autotune(svo_renderer* r, const svo_view* views, uint32_t n, int policy, void* stream,
                                   uint32_t iters, svo_launch_config* best, float* best_ms) {
  if (!r || iters == 0) { set_error("svo_autotune: bad arguments"); return SVO_ERR_ARG; }
  HIP_TRY(hipSetDevice(r->device));
  const uint32_t tiles[] = {8, 16, 32}, waves[] = {1, 2, 4},
                 stacks[] = {SVO_STACK_LDS, SVO_STACK_PRIVATE, SVO_STACK_LDS_NODE};
  float bm = 0;
  LaunchShape bs{};
  bool found = false;
  for (uint32_t t : tiles)
    for (uint32_t w : waves)
      for (uint32_t k : stacks) {
        const LaunchShape sh{t, w, k};
        float ms = 0;
        svo_status st = r->timed(sh, views, n, policy, (hipStream_t)stream, 1, iters, &ms);
        if (st != SVO_OK) return st;
        if (!found || ms < bm) { bm = ms; bs = sh; found = true; }
      }
  r->shape = bs;
  if (best) *best = svo_launch_config{bs.tile_w, bs.waves, bs.stack, 0};
  if (best_ms) *best_ms = bm;
  return SVO_OK;
}

// ------------------------------------------------------------ autotune cache

namespace {
const char* const kStackNames[] = {"lds", "private", "lds_node"};
const char kCacheHeader[] = "# libsvo_amd autotune cache v1: key<TAB>tile_width<TAB>waves<TAB>stack<TAB>ms";

uint64_t pow2_at_least(uint64_t v) {
  uint64_t p = 1;
  while (p < v) p <<= 1;
  return p;
}
}  // namespace

struct svo_tune_cache {
  struct Entry {
    LaunchShape shape;
    float ms;
  };
  std::mutex lock;
  std::map<std::string, Entry> entries;
};

extern "C" svo_status svo_tune_cache_create(svo_tune_cache** out) {
  if (!out) { set_error("svo_tune_cache_create: null out"); return SVO_ERR_ARG; }
  *out = new svo_tune_cache();
  return SVO_OK;
}

extern "C" void svo_tune_cache_destroy(svo_tune_cache* c) { delete c; }

extern "C" uint32_t svo_tune_cache_size(const svo_tune_cache* c) {
  if (!c) return 0;
  std::lock_guard<std::mutex> g(const_cast<svo_tune_cache*>(c)->lock);
  return (uint32_t)c->entries.size();
}

extern "C" svo_status svo_tune_cache_load(svo_tune_cache* c, const char* path) {
  if (!c || !path) { set_error("svo_tune_cache_load: bad arguments"); return SVO_ERR_ARG; }
  FILE* f = fopen(path, "r");
  if (!f) return SVO_OK;  // no file yet: nothing cached
  std::map<std::string, svo_tune_cache::Entry> read;
  char line[1024];
  unsigned lineno = 0;
  svo_status st = SVO_OK;
  while (fgets(line, sizeof line, f)) {
    ++lineno;
    size_t len = strlen(line);
    if (len == sizeof line - 1 && line[len - 1] != '\n') {
      set_error("%s:%u: line too long", path, lineno);
      st = SVO_ERR_FORMAT;
      break;
    }
    while (len && (line[len - 1] == '\n' || line[len - 1] == '\r')) line[--len] = 0;
    if (len == 0 || line[0] == '#') continue;
    char* tab = strchr(line, '\t');
    unsigned tw = 0, w = 0;
    char stack[16] = {0};
    float ms = 0;
    if (!tab || sscanf(tab + 1, "%u\t%u\t%15s\t%f", &tw, &w, stack, &ms) != 4) {
      set_error("%s:%u: expected key, tile width, waves, stack and ms separated by tabs", path, lineno);
      st = SVO_ERR_FORMAT;
      break;
    }
    uint32_t k = 3;
    for (uint32_t i = 0; i < 3; ++i)
      if (strcmp(stack, kStackNames[i]) == 0) k = i;
    const svo_launch_config cfg{tw, w, k, 0};
    LaunchShape sh{};
    if (k == 3 || shape_from_config(&cfg, &sh) != SVO_OK) {
      set_error("%s:%u: no launch shape (%u, %u, %s)", path, lineno, tw, w, stack);
      st = SVO_ERR_FORMAT;
      break;
    }
    read[std::string(line, tab)] = svo_tune_cache::Entry{sh, ms};
  }
  fclose(f);
  if (st != SVO_OK) return st;
  std::lock_guard<std::mutex> g(c->lock);
  for (auto& kv : read) c->entries[kv.first] = kv.second;
  return SVO_OK;
}

extern "C" svo_status svo_tune_cache_save(svo_tune_cache* c, const char* path) {
  if (!c || !path) { set_error("svo_tune_cache_save: bad arguments"); return SVO_ERR_ARG; }
  const std::string tmp = std::string(path) + ".tmp";
  FILE* f = fopen(tmp.c_str(), "w");
  if (!f) { set_error("svo_tune_cache_save: cannot write %s", tmp.c_str()); return SVO_ERR_IO; }
  bool ok = fprintf(f, "%s\n", kCacheHeader) > 0;
  {
    std::lock_guard<std::mutex> g(c->lock);
    for (const auto& kv : c->entries)
      ok = ok && fprintf(f, "%s\t%u\t%u\t%s\t%.5f\n", kv.first.c_str(), kv.second.shape.tile_w, kv.second.shape.waves,
                         kStackNames[kv.second.shape.stack], kv.second.ms) > 0;
  }
  ok = (fclose(f) == 0) && ok;
  if (!ok || rename(tmp.c_str(), path) != 0) {
    remove(tmp.c_str());
    set_error("svo_tune_cache_save: cannot write %s", path);
    return SVO_ERR_IO;
  }
  return SVO_OK;
}

extern "C" svo_status svo_tune_cache_key(svo_renderer* r, const svo_view* views, uint32_t n, int policy, char* key,
                                         size_t cap) {
  if (!r || !views || n == 0 || !key || cap == 0) { set_error("svo_tune_cache_key: bad arguments"); return SVO_ERR_ARG; }
  hipDeviceProp_t prop;
  HIP_TRY(hipGetDeviceProperties(&prop, r->device));
  std::string arch(prop.gcnArchName);
  arch = arch.substr(0, arch.find(':'));
  // one class per distinct scene shape: scaleMax, stack levels, nodes per
  // treelet and whether the LOD stop is on
  std::vector<std::string> scenes;
  uint64_t rays = 0;
  for (uint32_t i = 0; i < n; ++i) {
    if (!views[i].mm) { set_error("view %u has no residency handle", i); return SVO_ERR_ARG; }
    TracePools P;
    const svo_status st = mm_trace_pools(views[i].mm, policy, &P);
    if (st != SVO_OK) return st;
    char s[96];
    snprintf(s, sizeof s, "%u/%u/%u/%s", P.scale_max, P.stack_levels, P.nodes_per_treelet,
             views[i].camera.t_scale_ratio != 0.0f ? "lod" : "full");
    scenes.emplace_back(s);
    rays += (uint64_t)views[i].camera.width * views[i].camera.height;
  }
  std::sort(scenes.begin(), scenes.end());
  scenes.erase(std::unique(scenes.begin(), scenes.end()), scenes.end());
  std::string k = "arch=" + arch + ";cus=" + std::to_string(prop.multiProcessorCount) +
                  ";policy=" + std::to_string(policy) + ";views=" + std::to_string(pow2_at_least(n)) +
                  ";rays=" + std::to_string(pow2_at_least(rays)) + ";scenes=";
  for (size_t i = 0; i < scenes.size(); ++i) k += (i ? "," : "") + scenes[i];
  if (k.size() + 1 > cap) { set_error("svo_tune_cache_key: key needs %zu bytes", k.size() + 1); return SVO_ERR_ARG; }
  memcpy(key, k.c_str(), k.size() + 1);
  return SVO_OK;
}

extern "C" svo_status svo_autotune_cached(svo_renderer* r, svo_tune_cache* c, const svo_view* views, uint32_t n,
                                          int policy, void* stream, uint32_t iters, svo_launch_config* best,
                                          float* best_ms, int* cache_hit) {
  if (!r || !c) { set_error("svo_autotune_cached: bad arguments"); return SVO_ERR_ARG; }
  char key[4096];
  svo_status st = svo_tune_cache_key(r, views, n, policy, key, sizeof key);
  if (st != SVO_OK) return st;
  {
    std::lock_guard<std::mutex> g(c->lock);
    auto it = c->entries.find(key);
    if (it != c->entries.end()) {
      r->shape = it->second.shape;
      if (best) *best = svo_launch_config{r->shape.tile_w, r->shape.waves, r->shape.stack, 0};
      if (best_ms) *best_ms = it->second.ms;
      if (cache_hit) *cache_hit = 1;
      return SVO_OK;
    }
  }
  float ms = 0;
  st = svo_autotune(r, views, n, policy, stream, iters, best, &ms);
  if (st != SVO_OK) return st;
  {
    std::lock_guard<std::mutex> g(c->lock);
    c->entries[key] = svo_tune_cache::Entry{r->shape, ms};
  }
  if (best_ms) *best_ms = ms;
  if (cache_hit) *cache_hit = 0;
  return SVO_OK;
}

// ------------------------------------------------------------ allocation-free launches

namespace {
// Host image of the descriptors of one allocation-free launch (static
// storage: staging allocates nothing), serialised by a lock.
std::mutex g_noalloc_lock;
TraceFrame g_noalloc_stage[kNoAllocMaxViews];
}  // namespace

extern "C" svo_status svo_render_views_noalloc(const svo_view* views, uint32_t n, int policy,
                                               const svo_launch_config* config, void* stream) {
  if (!views || n == 0) { set_error("svo_render_views_noalloc: no views"); return SVO_ERR_ARG; }
  if (n > kNoAllocMaxViews) {
    set_error("svo_render_views_noalloc: %u views, at most %u (its descriptor scratch holds %u)", n, kNoAllocMaxViews,
              kNoAllocMaxViews);
    return SVO_ERR_ARG;
  }
  if (!views[0].mm) { set_error("svo_render_views_noalloc: view 0 has no residency handle"); return SVO_ERR_ARG; }
  LaunchShape sh{8, 1, SVO_STACK_LDS};
  if (config) {
    const svo_status r = shape_from_config(config, &sh);
    if (r != SVO_OK) return r;
  }
  const int device = mm_device(views[0].mm);
  if (device < 0) { set_error("svo_render_views_noalloc: handle has no device pools"); return SVO_ERR_STATE; }
  HIP_TRY(hipSetDevice(device));
  TraceFrame* d = nullptr;
  int e = noalloc_scratch(device, &d);
  if (e != 0) { set_error("svo_render_views_noalloc: scratch address: %s", hipGetErrorString((hipError_t)e)); return SVO_ERR_HIP; }
  std::lock_guard<std::mutex> g(g_noalloc_lock);
  StageTotals t{};
  svo_status st = stage_views_into(views, n, policy, sh, device, g_noalloc_stage, &t);
  if (st != SVO_OK) return st;
  for (uint32_t i = 0; i < n; ++i)
    if (view_needs_normals(views[i]) && (st = mm_sync_normals(views[i].mm, stream)) != SVO_OK) return st;
  // pageable source: the copy has consumed the host image when it returns
  HIP_TRY(hipMemcpyAsync(d, g_noalloc_stage, (size_t)n * sizeof(TraceFrame), hipMemcpyHostToDevice, (hipStream_t)stream));
  e = launch_trace(sh, d, n, t.num_tiles, t.stack_levels, policy, t.shade, stream);
  if (e != 0) { set_error("trace launch failed: %s", hipGetErrorString((hipError_t)e)); return SVO_ERR_HIP; }
  return SVO_OK;
}

// ------------------------------------------------------------ batches

struct svo_batch {
  int device = 0;  // the renderer's at creation (the batch does not keep the renderer)
  int policy = 0;
  LaunchShape shape{};
  StagedLaunch sg;
  std::vector<svo_mm*> mms;  // residency handle of each view
  TraceFrame* d = nullptr;
  TraceFrame* h = nullptr;  // pinned
  hipEvent_t copied = nullptr;
  bool pending = false;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;

  ~svo_batch() {
    int cur = 0;
    hipGetDevice(&cur);
    hipSetDevice(device);
    if (pending) hipEventSynchronize(copied);
    hipDeviceSynchronize();
    if (exec) hipGraphExecDestroy(exec);
    if (graph) hipGraphDestroy(graph);
    if (copied) hipEventDestroy(copied);
    hipFree(d);
    if (h) hipHostFree(h);
    hipSetDevice(cur);
  }
};

extern "C" svo_status svo_batch_create(svo_renderer* r, const svo_view* views, uint32_t n, int policy, void* stream,
                                       svo_batch** out) {
  if (!r || !out) { set_error("svo_batch_create: null argument"); return SVO_ERR_ARG; }
  *out = nullptr;
  HIP_TRY(hipSetDevice(r->device));
  std::unique_ptr<svo_batch> b(new svo_batch());
  b->device = r->device;
  b->policy = policy;
  b->shape = r->shape;
  svo_status st = stage_views(views, n, policy, b->shape, r->device, &b->sg);
  if (st != SVO_OK) return st;
  b->mms.resize(n);
  for (uint32_t i = 0; i < n; ++i) b->mms[i] = views[i].mm;
  const size_t bytes = n * sizeof(TraceFrame);
  HIP_TRY(hipMalloc(&b->d, bytes));
  HIP_TRY(hipHostMalloc(&b->h, bytes, hipHostMallocDefault));
  HIP_TRY(hipEventCreateWithFlags(&b->copied, hipEventDisableTiming));
  std::copy(b->sg.frames.begin(), b->sg.frames.end(), b->h);
  HIP_TRY(hipMemcpyAsync(b->d, b->h, bytes, hipMemcpyHostToDevice, (hipStream_t)stream));
  HIP_TRY(hipStreamSynchronize((hipStream_t)stream));
  // capture the launch on a private stream (the default stream cannot capture)
  hipStream_t cap;
  HIP_TRY(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking));
  hipError_t e = hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal);
  if (e == hipSuccess) {
    int le = launch_trace(b->shape, b->d, n, b->sg.num_tiles, b->sg.stack_levels, policy, b->sg.shade, cap);
    hipError_t e2 = hipStreamEndCapture(cap, &b->graph);
    e = le != 0 ? (hipError_t)le : e2;
  }
  hipStreamDestroy(cap);
  if (e != hipSuccess) { set_error("graph capture of the batch launch failed: %s", hipGetErrorString(e)); return SVO_ERR_HIP; }
  HIP_TRY(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
  *out = b.release();
  return SVO_OK;
}

extern "C" svo_status svo_batch_set_cameras(svo_batch* b, const svo_camera* cams, uint32_t n, void* stream) {
  if (!b || !cams) { set_error("svo_batch_set_cameras: null argument"); return SVO_ERR_ARG; }
  if (n != b->sg.frames.size()) { set_error("batch has %zu views, got %u cameras", b->sg.frames.size(), n); return SVO_ERR_ARG; }
  for (uint32_t i = 0; i < n; ++i) {
    if (cams[i].width != b->sg.frames[i].cam.width || cams[i].height != b->sg.frames[i].cam.height) {
      set_error("camera %u changes the frame size of the batch", i);
      return SVO_ERR_ARG;
    }
    const svo_status r = check_view_mode(cams[i].view_mode, b->policy, i);
    if (r != SVO_OK) return r;
  }
  HIP_TRY(hipSetDevice(b->device));
  if (b->pending) {
    HIP_TRY(hipEventSynchronize(b->copied));
    b->pending = false;
  }
  // the handles whose normals the next replays shade with follow the new
  // view modes
  b->sg.normal_handles.clear();
  for (uint32_t i = 0; i < n; ++i) {
    b->h[i].cam = cams[i];
    b->sg.frames[i].cam = cams[i];
    if (b->sg.frames[i].out.shade && mode_needs_normals(cams[i].view_mode) &&
        std::find(b->sg.normal_handles.begin(), b->sg.normal_handles.end(), b->mms[i]) == b->sg.normal_handles.end())
      b->sg.normal_handles.push_back(b->mms[i]);
  }
  HIP_TRY(hipMemcpyAsync(b->d, b->h, n * sizeof(TraceFrame), hipMemcpyHostToDevice, (hipStream_t)stream));
  HIP_TRY(hipEventRecord(b->copied, (hipStream_t)stream));
  b->pending = true;
  return SVO_OK;
}

extern "C" svo_status svo_batch_launch(svo_batch* b, void* stream) {
  if (!b) { set_error("svo_batch_launch: null batch"); return SVO_ERR_ARG; }
  HIP_TRY(hipSetDevice(b->device));
  for (svo_mm* mm : b->mms)
    if (mm_dirty_slots(mm) != 0) {
      set_error("svo_batch_launch: %u incore slots not uploaded; call svo_mm_update_device first", mm_dirty_slots(mm));
      return SVO_ERR_STATE;
    }
  const svo_status ns = sync_staged_normals(b->sg, stream);
  if (ns != SVO_OK) return ns;
  HIP_TRY(hipGraphLaunch(b->exec, (hipStream_t)stream));
  return SVO_OK;
}

extern "C" void svo_batch_destroy(svo_batch* b) { delete b; }
