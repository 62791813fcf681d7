// Treelet residency: the host incore buffer with the reference's slot, stitch
// and evict semantics (src/TreeletMemoryManager.cpp), backed by HBM pools
// that replace the OpenCL buffers of TreeletMemoryManagerCl
// (src/TreeletMemoryManagerCl.cpp).
//
// HBM layout of one handle (all pools indexed by incore node index
// n = slot * nodes_per_treelet + local):
//   nodes    : 8 B per node {first_child, masks}     (device copy of
//              getIncoreBuffer())
//   colors   : 4 B per node packed_color             (the half of the
//   normals  : 4 B per node packed_normal             reference's 8 B Attribs
//              that the colour epilogue reads is a separate, dense array)
//   slot_gid : 4 B per slot, gid held by the slot or -1
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cstdlib>
#include <type_traits>
#include <cstring>
#include <map>
#include <memory>
#include <queue>
#include <set>
#include <unordered_map>
#include <vector>

#include "svo_internal.h"
#include "trace.h"

using namespace svo_impl;

namespace {

constexpr int kMaxVisibility = 8;  // MAX_VISIBILITY, src/TreeletMemoryManager.cpp:47

struct SlotInfo {  // TreeletMemoryManager::SlotInfo (src/TreeletMemoryManager.h:72-89)
  int32_t gid = -1;
  int visibility = 0;
  float error = 100000.0f;
};

#define HIP_TRY(expr)                                                                   \
  do {                                                                                  \
    hipError_t e_ = (expr);                                                             \
    if (e_ != hipSuccess) {                                                             \
      set_error("%s failed: %s", #expr, hipGetErrorString(e_));                         \
      return SVO_ERR_HIP;                                                               \
    }                                                                                   \
  } while (0)

// Host incore arrays are page-locked (hipHostMalloc) when the handle has a
// device, so that uploads are DMA transfers that need no staging copy;
// host-only handles use ordinary memory.
template <typename T>
struct HostAlloc {
  using value_type = T;
  using propagate_on_container_move_assignment = std::true_type;
  using propagate_on_container_copy_assignment = std::true_type;
  using propagate_on_container_swap = std::true_type;
  using is_always_equal = std::false_type;
  bool pinned = false;
  HostAlloc() = default;
  explicit HostAlloc(bool p) : pinned(p) {}
  template <typename U>
  HostAlloc(const HostAlloc<U>& o) : pinned(o.pinned) {}
  T* allocate(size_t n) {
    void* p = nullptr;
    if (pinned) {
      if (hipHostMalloc(&p, n * sizeof(T), hipHostMallocDefault) != hipSuccess) throw std::bad_alloc();
    } else {
      p = std::malloc(n * sizeof(T));
      if (!p) throw std::bad_alloc();
    }
    return static_cast<T*>(p);
  }
  void deallocate(T* p, size_t) {
    if (pinned) hipHostFree(p);
    else std::free(p);
  }
  template <typename U>
  bool operator==(const HostAlloc<U>& o) const { return pinned == o.pinned; }
  template <typename U>
  bool operator!=(const HostAlloc<U>& o) const { return pinned != o.pinned; }
};
template <typename T>
using HostVec = std::vector<T, HostAlloc<T>>;

}  // namespace

struct svo_mm {
  const svo_scene* scene = nullptr;
  uint32_t N = 0;           // nodes per treelet
  uint32_t num_slots = 0;
  int device = 0;
  // host incore buffer (SoA attributes)
  HostVec<svo_node> incore;
  HostVec<uint32_t> normals, colors;
  std::vector<SlotInfo> slots;
  std::vector<int32_t> treelet_slot;                 // Treelet::getSlotGid, -1 = not resident
  std::vector<std::set<uint32_t>> children_resident;  // _childTreeletsInIncoreBuffer
  std::vector<uint32_t> free_slots;                  // stack, top = back()
  std::set<uint32_t> dirty;                          // _incoreSlotsToUpload
  uint32_t search_pos = 1;                           // freeIncorePosition's probe cursor
  uint32_t resident = 0;
  // Per slot, the number of leading nodes that may be non-zero (a high-water
  // mark; everything after it is zero) in the host buffers and on the
  // device: fills and uploads skip the zero padding of partial treelets.
  std::vector<uint32_t> extent, dev_extent;
  // Normals go up separately: only the view modes that shade with them
  // read them on the device.  A refilled slot's normals are stale there
  // (nrm_stale, listed in nrm_slots) until a launch with such a view, or a
  // device readback of attributes, sends them (sync_normals);
  // dev_nrm_extent is their device high-water mark.
  std::vector<uint32_t> dev_nrm_extent;
  std::vector<uint8_t> nrm_stale;
  std::vector<uint32_t> nrm_slots;
  // Set once a caller has taken the device pools (svo_mm_get_device_pools):
  // from then on every update sends the normals too, since kernels outside
  // this library may read them.
  bool eager_normals = false;
  // What the next upload must send: slots whose content was replaced (a
  // fill or a reset) go up within their extents; otherwise only the nodes
  // a stitch or unstitch changed (global node indices).
  std::vector<uint8_t> refilled;
  std::vector<uint64_t> patched;
  // HBM
  uint64_t* d_nodes = nullptr;
  uint32_t* d_colors = nullptr;
  uint32_t* d_normals = nullptr;
  int32_t* d_slot_gid = nullptr;
  svo_renderer* renderer = nullptr;  // default renderer of svo_render
  // analysis feedback tables (see TracePools) and their pinned readback
  unsigned long long* d_fb_slot = nullptr;
  unsigned long long* d_fb_gid = nullptr;
  uint32_t* d_fb_queue = nullptr;
  uint32_t* d_fb_state = nullptr;            // [0] queue count, [1] the cycle's stamp
  unsigned long long* h_fb_slot = nullptr;
  unsigned long long* h_fb_gid = nullptr;  // key of each asked gid, by queue position
  uint32_t* h_zero = nullptr;               // a pinned 0 that resets the queue count
  uint32_t* h_fb_queue = nullptr;
  uint32_t* h_fb_count = nullptr;
  uint32_t fb_stamp = 1;  // host copy of the device stamp of the current cycle
  uint64_t evictions = 0;  // treelets freed by freeIncorePosition (incl. descendants)
  // uploads: span descriptors (pinned, and their device copy read by the
  // upload kernel) and the pinned slot table
  UploadSpan* h_spans = nullptr;
  UploadSpan* d_spans = nullptr;
  uint32_t span_cap = 0;
  int32_t* h_slot_tab = nullptr;
  // the records read back by the last table read, and whether the sets made
  // from them (fb_sets) wait for the next apply
  bool fb_cached = false;
  // Readback states of a cycle: an export enqueued by svo_mm_fetch_feedback
  // (fb_pending, finished by fb_event), its tables parsed into fb_seen /
  // fb_asked (fb_parsed), the visible sets built from them (fb_cached).
  bool fb_pending = false;
  bool fb_parsed = false;
  uint32_t fb_export_stamp = 0;  // stamp of the cycle the export reads
  hipEvent_t fb_event = nullptr;
  std::vector<std::pair<uint32_t, float>> fb_seen, fb_asked;

  bool host_only() const { return device < 0; }

  ~svo_mm() {
    if (host_only()) return;
    int cur = 0;
    hipGetDevice(&cur);
    hipSetDevice(device);
    if (renderer) svo_renderer_destroy(renderer);
    hipDeviceSynchronize();
    if (fb_event) hipEventDestroy(fb_event);
    hipFree(d_nodes);
    hipFree(d_colors);
    hipFree(d_normals);
    hipFree(d_slot_gid);
    hipFree(d_fb_slot);
    hipFree(d_fb_gid);
    hipFree(d_fb_queue);
    hipFree(d_fb_state);
    if (h_fb_slot) hipHostFree(h_fb_slot);
    if (h_fb_gid) hipHostFree(h_fb_gid);
    if (h_zero) hipHostFree(h_zero);
    if (h_fb_queue) hipHostFree(h_fb_queue);
    if (h_fb_count) hipHostFree(h_fb_count);
    hipFree(d_spans);
    if (h_spans) hipHostFree(h_spans);
    if (h_slot_tab) hipHostFree(h_slot_tab);
    hipSetDevice(cur);
  }

  void mark(uint32_t slot) { dirty.insert(slot); }

  // resetIncoreBuffer (src/TreeletMemoryManager.cpp:303-380), host part.
  void reset_host() {
    // zero what the slots' extents say may be non-zero (fresh buffers are
    // zero already)
    for (uint32_t sl = 0; sl < (uint32_t)extent.size(); ++sl) {
      const uint64_t pos = (uint64_t)sl * N;
      std::fill(incore.begin() + pos, incore.begin() + pos + extent[sl], svo_node{0, 0});
      std::fill(normals.begin() + pos, normals.begin() + pos + extent[sl], 0u);
      std::fill(colors.begin() + pos, colors.begin() + pos + extent[sl], 0u);
    }
    slots.assign(num_slots, SlotInfo());
    treelet_slot.assign(scene->treelets.size(), -1);
    children_resident.assign(scene->treelets.size(), std::set<uint32_t>());
    free_slots.clear();
    for (uint32_t i = 1; i != num_slots + 1; ++i) free_slots.push_back(num_slots - i);  // top = slot 0
    dirty.clear();
    extent.assign(num_slots, 0u);
    refilled.assign(num_slots, 1u);  // everything goes up after a reset
    patched.clear();
    resident = 0;  // search_pos survives a reset, like the reference's static cursor
  }

  // removeTreeletFromIncoreBuffer (src/TreeletMemoryManager.cpp:665-700).
  void remove_one(uint32_t gid) {
    const svo_treelet_header& h = scene->treelets[gid].h;
    const int32_t slot = treelet_slot[gid];
    const int32_t pslot = treelet_slot[h.parent_gid];
    data_op(DataOp{DataOp::kUnstitch, (uint32_t)pslot, gid, 0});
    mark((uint32_t)pslot);
    slots[slot] = SlotInfo();
    free_slots.push_back((uint32_t)slot);
    treelet_slot[gid] = -1;
    children_resident[h.parent_gid].erase(gid);
    --resident;
  }

  // Frees the resident sub-treelets below `gid` (the descendant sweep of
  // freeIncorePosition, src/TreeletMemoryManager.cpp:603-648).
  void free_descendants(uint32_t gid) {
    std::queue<uint32_t> q;
    q.push(gid);
    std::vector<uint32_t> clear;
    while (!q.empty()) {
      const uint32_t p = q.front();
      q.pop();
      for (uint32_t c : children_resident[p]) { q.push(c); clear.push_back(c); }
      children_resident[p].clear();
    }
    for (uint32_t c : clear) {
      const int32_t s = treelet_slot[c];
      if (s <= 0) continue;  // slot 0 (the root) is never freed
      free_slots.push_back((uint32_t)s);
      treelet_slot[c] = -1;
      slots[s] = SlotInfo();
      --resident;
      ++evictions;
    }
  }

  // freeIncorePosition (src/TreeletMemoryManager.cpp:560-653): probes up to
  // 100 slots from the cursor for an invisible treelet whose parent is
  // invisible too, then frees it and everything resident below it.
  // Returns the treelet freed (its descendants went with it), 0 if none.
  uint32_t free_position() {
    uint32_t freeable = 0;
    uint32_t tests = 0;
    const uint32_t probes = 100;
    for (uint32_t i = 0; i != num_slots; ++i) {
      const SlotInfo& s = slots[search_pos];
      if (s.visibility < 0 && s.gid != -1) {
        const uint32_t parent = scene->treelets[s.gid].h.parent_gid;
        const int32_t ps = treelet_slot[parent];
        if (ps >= 0 && slots[ps].visibility < 0) { freeable = (uint32_t)s.gid; break; }
      }
      ++search_pos;
      ++tests;
      if (search_pos == num_slots) search_pos = 1;
      if (tests == probes) break;
    }
    if (freeable == 0) return 0;
    free_descendants(freeable);
    remove_one(freeable);
    ++evictions;
    return freeable;
  }

  // The host half of the analyse pass and the client-side update, from
  // aggregated records: `seen` holds (slot, largest error) of the slots that
  // hold a record's node, `asked` (gid, largest error) of the sub-treelets
  // that records ask for, each gid once.
  //  * RenderPassAnalyse::performAnalysePass (src/RenderPassAnalyse.cpp:
  //    248-356): every seen treelet and its ancestors are visible ("old"),
  //    the root with error 99; the asked treelets ("new") are ordered by
  //    descending error (ties in ascending gid, the order in which the
  //    reference's map feeds its multiset) and capped at 12 MiB of treelets;
  //  * updateClientSideIncoreBuffer (src/TreeletMemoryManager.cpp:392-443):
  //    every slot but the root's loses one visibility, visible treelets get
  //    MAX_VISIBILITY, then the new treelets are inserted in order
  //    (insertTreeletIntoIncoreBuffer, evicting through freeIncorePosition).
  // The visible sets (performAnalysePass's host half): `old` gets every
  // seen treelet and its ancestors with the largest error, `asked` is sorted
  // and capped.  Returns the number of seen slots that hold a treelet.
  // Scratch of visible_sets: the largest error per gid, valid where
  // vis_mark[gid] == vis_epoch.
  mutable std::vector<float> vis_err;
  mutable std::vector<uint32_t> vis_mark;
  mutable uint32_t vis_epoch = 0;

  uint32_t visible_sets(const std::vector<std::pair<uint32_t, float>>& seen,
                        std::vector<std::pair<uint32_t, float>>* asked,
                        std::vector<std::pair<uint32_t, float>>* old) const {
    const size_t ng = scene->treelets.size();
    if (vis_mark.size() != ng) {
      vis_err.assign(ng, 0.0f);
      vis_mark.assign(ng, 0u);
      vis_epoch = 0;
    }
    if (++vis_epoch == 0) {  // wrapped: start the marks afresh
      std::fill(vis_mark.begin(), vis_mark.end(), 0u);
      vis_epoch = 1;
    }
    std::vector<uint32_t> touched;
    bool root = false;
    uint32_t visible_slots = 0;
    for (const auto& e : seen) {
      if (e.first >= num_slots || slots[e.first].gid < 0) continue;
      ++visible_slots;
      root = true;  // the root with error 99 (RenderPassAnalyse.cpp)
      // the treelet and its ancestors keep the largest error seen below
      // them; an ancestor that already holds at least this error passes it
      // on to its own ancestors already, so the walk stops there
      for (uint32_t g = (uint32_t)slots[e.first].gid; g != 0; g = scene->treelets[g].h.parent_gid) {
        if (vis_mark[g] != vis_epoch) {
          vis_mark[g] = vis_epoch;
          vis_err[g] = e.second;
          touched.push_back(g);
        } else if (e.second > vis_err[g]) {
          vis_err[g] = e.second;
        } else if (vis_err[g] >= e.second) {
          break;
        }
      }
    }
    std::sort(touched.begin(), touched.end());
    old->clear();
    old->reserve(touched.size() + 1);
    if (root) old->push_back({0u, 99.0f});
    for (uint32_t g : touched) old->push_back({g, vis_err[g]});
    std::stable_sort(asked->begin(), asked->end(), [](const std::pair<uint32_t, float>& a, const std::pair<uint32_t, float>& b) {
      return a.first < b.first;
    });
    std::stable_sort(asked->begin(), asked->end(), [](const std::pair<uint32_t, float>& a, const std::pair<uint32_t, float>& b) {
      return a.second > b.second;
    });
    const uint32_t cap = (uint32_t)((12.0 * 1048576.0f) / (double)scene->treelet_bytes);
    if (asked->size() > cap) asked->resize(cap);
    return visible_slots;
  }

  // The sets one analysis cycle yields, computed against the slot table of
  // the moment (as performAnalysePass does), applied later.
  struct FeedbackSets {
    std::vector<std::pair<uint32_t, float>> old;  // ascending gid
    std::vector<std::pair<uint32_t, float>> asked;  // sorted, capped
    uint32_t visible_slots = 0;
    uint32_t requests = 0;                          // before the cap
  };
  FeedbackSets fb_sets;  // computed by svo_mm_read_feedback (valid while fb_cached)

  FeedbackSets make_sets(const std::vector<std::pair<uint32_t, float>>& seen,
                         std::vector<std::pair<uint32_t, float>> asked) const {
    FeedbackSets F;
    F.requests = (uint32_t)asked.size();
    F.visible_slots = visible_sets(seen, &asked, &F.old);
    F.asked = std::move(asked);
    return F;
  }

  void apply_sets(const FeedbackSets& F, svo_feedback_stats* st) {
    svo_feedback_stats S{};
    S.requests = F.requests;
    S.visible_slots = F.visible_slots;
    S.visible_treelets = (uint32_t)F.old.size();
    S.requests_kept = (uint32_t)F.asked.size();
    for (uint32_t i = 1; i < num_slots; ++i) --slots[i].visibility;
    for (const auto& e : F.old) {
      const int32_t sl = treelet_slot[e.first];
      if (sl >= 0) slots[sl].visibility = kMaxVisibility;
    }
    const uint64_t ev0 = evictions;
    deferring = true;
    for (const auto& e : F.asked) {
      if (insert(e.first, e.second) == SVO_OK) ++S.inserted;
      else ++S.insert_failed;
    }
    replay_oplog();
    clear_error();
    S.evicted = (uint32_t)(evictions - ev0);
    S.resident_treelets = resident;
    if (st) *st = S;
  }

  void apply_records(const std::vector<std::pair<uint32_t, float>>& seen,
                     std::vector<std::pair<uint32_t, float>> asked, svo_feedback_stats* st) {
    apply_sets(make_sets(seen, std::move(asked)), st);
  }

  // insertTreeletIntoIncoreBuffer (src/TreeletMemoryManager.cpp:455-548).
  svo_status insert(uint32_t gid, float error) {
    if (gid >= scene->treelets.size()) { set_error("insert: treelet gid %u out of range", gid); return SVO_ERR_ARG; }
    if (treelet_slot[gid] >= 0) { set_error("insert: treelet %u is already resident", gid); return SVO_ERR_STATE; }
    const svo_treelet_header& h = scene->treelets[gid].h;
    if (gid != 0 && treelet_slot[h.parent_gid] < 0) {
      set_error("insert: parent treelet %u of %u is not resident", h.parent_gid, gid);
      return SVO_ERR_STATE;
    }
    if (free_slots.empty() && !free_position()) { set_error("insert: no free or freeable incore slot for treelet %u", gid); return SVO_ERR_CAPACITY; }
    // eviction may have removed the parent
    if (gid != 0 && treelet_slot[h.parent_gid] < 0) {
      set_error("insert: parent treelet %u of %u was evicted", h.parent_gid, gid);
      return SVO_ERR_CAPACITY;
    }
    const uint32_t slot = free_slots.back();
    free_slots.pop_back();
    slots[slot] = SlotInfo{(int32_t)gid, kMaxVisibility, error};
    treelet_slot[gid] = (int32_t)slot;
    ++resident;
    data_op(DataOp{DataOp::kFill, slot, gid, 0});
    mark(slot);
    if (gid == 0) return SVO_OK;
    const int32_t pslot = treelet_slot[h.parent_gid];
    data_op(DataOp{DataOp::kStitch, (uint32_t)pslot, gid, slot});
    mark((uint32_t)pslot);
    children_resident[h.parent_gid].insert(gid);
    return SVO_OK;
  }

  // ---- data movement of inserts and removes.  Outside apply_records each
  // operation runs at once; inside, they are logged and replayed at its end:
  // the last fill of every slot in parallel (a fill rewrites the whole slot),
  // then, in log order, every stitch/unstitch that no later fill of its slot
  // overwrites.  Each slot ends as the operations in log order leave it.
  struct DataOp {
    enum : uint32_t { kFill, kStitch, kUnstitch };
    uint32_t kind;
    uint32_t slot;       // the slot written (for a stitch/unstitch: the parent's)
    uint32_t gid;        // the treelet filled, stitched or unstitched
    uint32_t src_slot;   // stitch: the slot holding the child
  };
  bool deferring = false;
  std::vector<DataOp> oplog;

  void data_op(const DataOp& op) {
    if (deferring) oplog.push_back(op);
    else run_op(op);
  }

  void run_op(const DataOp& op) {
    const svo_treelet_header& h = scene->treelets[op.gid].h;
    if (op.kind == DataOp::kFill) {
      const uint64_t pos = (uint64_t)op.slot * N;
      const TreeletData& T = scene->treelets[op.gid];
      const uint32_t n = T.h.num_nodes, old = extent[op.slot];
      std::copy(T.nodes.begin(), T.nodes.begin() + n, incore.begin() + pos);
      for (uint32_t i = 0; i < n; ++i) {
        normals[pos + i] = T.attribs[2 * (size_t)i];
        colors[pos + i] = T.attribs[2 * (size_t)i + 1];
      }
      if (old > n) {  // the previous occupant's nodes beyond this one's
        std::fill(incore.begin() + pos + n, incore.begin() + pos + old, svo_node{0, 0});
        std::fill(normals.begin() + pos + n, normals.begin() + pos + old, 0u);
        std::fill(colors.begin() + pos + n, colors.begin() + pos + old, 0u);
      }
      extent[op.slot] = n;
      refilled[op.slot] = 1u;
      return;
    }
    const uint64_t leaf = (uint64_t)op.slot * N + h.parent_leaf_pos;
    const uint64_t leaf_parent = (uint64_t)op.slot * N + h.parent_leaf_parent_pos;
    if (op.kind == DataOp::kStitch) {
      // the child root replaces the parent's leaf, which stops being a leaf
      // and points (relatively, modulo 2^32) at the child's level 1; a child
      // is stitched right after its own fill, so its root is the pristine one
      const uint64_t pos = (uint64_t)op.src_slot * N;
      incore[leaf] = scene->treelets[op.gid].nodes[0];
      incore[leaf_parent].masks &= ~(1u << (kLeafShift + h.parent_leaf_idx));
      incore[leaf].first_child = (uint32_t)(pos + 1) - (uint32_t)leaf;
    } else {
      incore[leaf] = scene->treelets[h.parent_gid].nodes[h.parent_leaf_pos];  // pristine leaf
      incore[leaf_parent].masks |= 1u << (kLeafShift + h.parent_leaf_idx);
    }
    if (!host_only()) {  // a host-only handle uploads nothing
      patched.push_back(leaf);
      patched.push_back(leaf_parent);
    }
  }

  void replay_oplog() {
    deferring = false;
    std::unordered_map<uint32_t, size_t> last_fill;
    for (size_t i = 0; i < oplog.size(); ++i)
      if (oplog[i].kind == DataOp::kFill) last_fill[oplog[i].slot] = i;
    std::vector<size_t> fills;
    fills.reserve(last_fill.size());
    for (const auto& e : last_fill) fills.push_back(e.second);
#pragma omp parallel for schedule(dynamic, 1) if (fills.size() > 1)
    for (size_t k = 0; k < fills.size(); ++k) run_op(oplog[fills[k]]);
    for (size_t i = 0; i < oplog.size(); ++i) {
      if (oplog[i].kind == DataOp::kFill) continue;
      const auto it = last_fill.find(oplog[i].slot);
      if (it != last_fill.end() && it->second > i) continue;  // overwritten by a later fill
      run_op(oplog[i]);
    }
    oplog.clear();
  }

  // Uploads runs of slots {first, count} (nodes, colours, normals) and the
  // slot table in one upload-kernel launch reading the pinned host buffers,
  // then waits for it: the host buffers may change again once this returns.
  svo_status upload_runs(const std::vector<std::pair<uint32_t, uint32_t>>& runs, hipStream_t st, uint64_t* bytes,
                         uint64_t* bytes_moved) {
    for (const auto& r : runs)
      if (bytes) *bytes += (uint64_t)r.second * N * 16;
    if (host_only()) return SVO_OK;
    uint32_t n = 0;
    uint64_t chunks = 0, moved = 0;
    auto add = [&](const void* src, void* dst, uint64_t b) {
      h_spans[n++] = UploadSpan{(uint64_t)(uintptr_t)src, (uint64_t)(uintptr_t)dst, b, chunks};
      chunks += (b + 15) / 16;
    };
    uint64_t run_slots = 0;
    for (const auto& r : runs) run_slots += r.second;
    if (run_slots * 3 + patched.size() + 1 > span_cap) {
      // grow the span tables (pinned and device) for this upload
      const uint64_t cap = run_slots * 3 + patched.size() + 1;
      if (cap > 0xFFFFFFFFull) { set_error("upload: %llu spans", (unsigned long long)cap); return SVO_ERR_STATE; }
      HIP_TRY(hipStreamSynchronize(st));
      HIP_TRY(hipFree(d_spans));
      HIP_TRY(hipHostFree(h_spans));
      d_spans = nullptr;
      h_spans = nullptr;
      span_cap = 0;
      if (hipHostMalloc(&h_spans, cap * sizeof(UploadSpan), 0) != hipSuccess ||
          hipMalloc(&d_spans, cap * sizeof(UploadSpan)) != hipSuccess) {
        set_error("upload: span tables of %llu entries could not be allocated", (unsigned long long)cap);
        return SVO_ERR_NOMEM;
      }
      span_cap = (uint32_t)cap;
    }
    // The merged ranges name what the reference would copy; what moves is
    // the refilled slots within their extents (the host's non-zero part and
    // whatever the device still holds beyond it) and the stitched nodes of
    // slots that were not refilled.  The device ends equal to the host.
    for (const auto& r : runs) {
      for (uint32_t sl = r.first; sl < r.first + r.second; ++sl) {
        if (!refilled[sl]) continue;
        const size_t n0 = (size_t)sl * N, nn = std::max(extent[sl], dev_extent[sl]);
        if (nn == 0) continue;
        add(incore.data() + n0, d_nodes + n0, nn * 8);
        add(colors.data() + n0, d_colors + n0, nn * 4);
        moved += nn * 12;
      }
    }
    std::sort(patched.begin(), patched.end());
    patched.erase(std::unique(patched.begin(), patched.end()), patched.end());
    for (const uint64_t node : patched) {
      if (refilled[node / N]) continue;
      add(incore.data() + node, d_nodes + node, 8);
      moved += 8;
    }
    for (uint32_t s = 0; s < num_slots; ++s) h_slot_tab[s] = slots[s].gid;
    add(h_slot_tab, d_slot_gid, (uint64_t)num_slots * 4);
    if (bytes_moved) *bytes_moved += moved + (uint64_t)num_slots * 4;
    HIP_TRY(hipMemcpyAsync(d_spans, h_spans, (size_t)n * sizeof(UploadSpan), hipMemcpyHostToDevice, st));
    const int e = launch_upload(d_spans, n, chunks, st);
    if (e != 0) { set_error("upload launch failed: %s", hipGetErrorString((hipError_t)e)); return SVO_ERR_HIP; }
    HIP_TRY(hipStreamSynchronize(st));
    // the device now holds what was sent: record it
    for (const auto& r : runs)
      for (uint32_t sl = r.first; sl < r.first + r.second; ++sl) {
        if (!refilled[sl]) continue;
        refilled[sl] = 0;
        dev_extent[sl] = extent[sl];
        if (!nrm_stale[sl]) {
          nrm_stale[sl] = 1;
          nrm_slots.push_back(sl);
        }
      }
    patched.clear();
    return SVO_OK;
  }

  // Sends the stale normals (see nrm_stale) in one upload-kernel launch on
  // `st` and waits for it; the slots count as current once it has ended.
  svo_status sync_normals(hipStream_t st) {
    if (nrm_slots.empty() || host_only()) return SVO_OK;
    if (nrm_slots.size() + 1 > span_cap) { set_error("normals upload: %zu spans", nrm_slots.size()); return SVO_ERR_STATE; }
    uint32_t n = 0;
    uint64_t chunks = 0;
    for (const uint32_t sl : nrm_slots) {
      const size_t n0 = (size_t)sl * N, nn = std::max(extent[sl], dev_nrm_extent[sl]);
      if (nn == 0) continue;
      h_spans[n++] = UploadSpan{(uint64_t)(uintptr_t)(normals.data() + n0), (uint64_t)(uintptr_t)(d_normals + n0), nn * 4,
                                chunks};
      chunks += (nn * 4 + 15) / 16;
    }
    if (n != 0) {
      HIP_TRY(hipSetDevice(device));
      HIP_TRY(hipMemcpyAsync(d_spans, h_spans, (size_t)n * sizeof(UploadSpan), hipMemcpyHostToDevice, st));
      const int e = launch_upload(d_spans, n, chunks, st);
      if (e != 0) { set_error("normals upload launch failed: %s", hipGetErrorString((hipError_t)e)); return SVO_ERR_HIP; }
      HIP_TRY(hipStreamSynchronize(st));
    }
    for (const uint32_t sl : nrm_slots) {
      nrm_stale[sl] = 0;
      dev_nrm_extent[sl] = extent[sl];
    }
    nrm_slots.clear();
    return SVO_OK;
  }

  // updateDeviceMemory (src/TreeletMemoryManagerCl.cpp:155-251): sorted
  // dirty slots are merged while (treelets+1)/(distance+1) >= 0.6*2.
  svo_status update_device(hipStream_t st, svo_upload_stats* stats) {
    svo_upload_stats s{0, 0, 0, 0};
    s.slots_uploaded = (uint32_t)dirty.size();
    if (!dirty.empty()) {
      const float limit = 0.6f * 2.0f;
      std::vector<std::pair<uint32_t, uint32_t>> runs;
      auto it = dirty.begin();
      while (it != dirty.end()) {
        const uint32_t start = *it;
        uint32_t num_slots_r = 1, num_treelets = 1;
        while (it != dirty.end()) {
          const int dist = (int)*it - (int)start;
          const float ratio = (float)(num_treelets + 1) / (float)(dist + 1);
          if (ratio >= limit) {
            num_slots_r = (uint32_t)dist + 1;
            ++num_treelets;
            ++it;
          } else {
            break;
          }
        }
        ++s.copy_calls;
        runs.push_back({start, num_slots_r});
      }
      svo_status r = upload_runs(runs, st, &s.bytes, &s.bytes_moved);
      if (r != SVO_OK) return r;
      dirty.clear();
      if (eager_normals && (r = sync_normals(st)) != SVO_OK) return r;
    }
    if (stats) *stats = s;
    return SVO_OK;
  }
};

extern "C" svo_status svo_mm_create(const svo_scene* scene, uint64_t incore_bytes, int device, svo_mm** out) {
  if (!scene || !out) { set_error("svo_mm_create: null argument"); return SVO_ERR_ARG; }
  *out = nullptr;
  const uint64_t tb = scene->treelet_bytes;
  const uint64_t slots = incore_bytes / tb;
  if (slots == 0) { set_error("svo_mm_create: incore buffer of %llu B holds no %llu B treelet", (unsigned long long)incore_bytes, (unsigned long long)tb); return SVO_ERR_ARG; }
  const uint64_t pool_nodes = slots * (tb / 8);
  if (pool_nodes >= 0xFFFFFFFFull) { set_error("svo_mm_create: pool of %llu nodes exceeds 32-bit node indices", (unsigned long long)pool_nodes); return SVO_ERR_ARG; }
  if (device < -1) { set_error("svo_mm_create: device %d invalid", device); return SVO_ERR_ARG; }
  if (device >= 0) {
    int ndev = 0;
    HIP_TRY(hipGetDeviceCount(&ndev));
    if (device >= ndev) { set_error("svo_mm_create: device %d not present (%d devices)", device, ndev); return SVO_ERR_ARG; }
  }
  std::unique_ptr<svo_mm> mm(new svo_mm());
  mm->scene = scene;
  mm->N = (uint32_t)(tb / 8);
  mm->num_slots = (uint32_t)slots;
  mm->device = device;
  try {
    if (device >= 0) HIP_TRY(hipSetDevice(device));
    const bool pinned = device >= 0;
    mm->incore = HostVec<svo_node>(HostAlloc<svo_node>(pinned));
    mm->normals = HostVec<uint32_t>(HostAlloc<uint32_t>(pinned));
    mm->colors = HostVec<uint32_t>(HostAlloc<uint32_t>(pinned));
    mm->incore.resize(pool_nodes);
    mm->normals.resize(pool_nodes);
    mm->colors.resize(pool_nodes);
  } catch (const std::bad_alloc&) {
    set_error("svo_mm_create: host incore buffer of %llu nodes", (unsigned long long)pool_nodes);
    return SVO_ERR_NOMEM;
  }
  mm->reset_host();
  svo_status r = mm->insert(0, 999.0f);  // root, SlotInfo(0, MAX_VISIBILITY, 999)
  if (r != SVO_OK) return r;
  if (mm->host_only()) {
    mm->dirty.clear();
    *out = mm.release();
    return SVO_OK;
  }
  HIP_TRY(hipSetDevice(device));
  if (hipMalloc(&mm->d_nodes, pool_nodes * 8) != hipSuccess || hipMalloc(&mm->d_colors, pool_nodes * 4) != hipSuccess ||
      hipMalloc(&mm->d_normals, pool_nodes * 4) != hipSuccess || hipMalloc(&mm->d_slot_gid, slots * 4) != hipSuccess) {
    set_error("svo_mm_create: hipMalloc of %llu B of pools failed", (unsigned long long)(pool_nodes * 16 + slots * 4));
    return SVO_ERR_NOMEM;
  }
  const size_t ng = scene->treelets.size();
  if (hipMalloc(&mm->d_fb_slot, slots * 8) != hipSuccess || hipMalloc(&mm->d_fb_gid, ng * 8) != hipSuccess ||
      hipMalloc(&mm->d_fb_queue, ng * 4) != hipSuccess ||
      hipMalloc(&mm->d_fb_state, 8) != hipSuccess || hipHostMalloc(&mm->h_fb_slot, slots * 8, 0) != hipSuccess ||
      hipHostMalloc(&mm->h_fb_gid, ng * 8, 0) != hipSuccess || hipHostMalloc(&mm->h_fb_queue, ng * 4, 0) != hipSuccess ||
      hipHostMalloc(&mm->h_zero, 4, 0) != hipSuccess ||
      hipHostMalloc(&mm->h_fb_count, 4, 0) != hipSuccess) {
    set_error("svo_mm_create: feedback tables for %zu treelets could not be allocated", ng);
    return SVO_ERR_NOMEM;
  }
  HIP_TRY(hipMemset(mm->d_fb_slot, 0, slots * 8));
  HIP_TRY(hipMemset(mm->d_fb_gid, 0, ng * 8));
  {
    const uint32_t state[2] = {0u, mm->fb_stamp};
    HIP_TRY(hipMemcpy(mm->d_fb_state, state, sizeof state, hipMemcpyHostToDevice));
  }
  HIP_TRY(hipEventCreateWithFlags(&mm->fb_event, hipEventDisableTiming));
  *mm->h_zero = 0u;
  mm->span_cap = (uint32_t)(3 * slots + 1);  // three spans per slot, plus the slot table
  if (hipHostMalloc(&mm->h_spans, (size_t)mm->span_cap * sizeof(UploadSpan), 0) != hipSuccess ||
      hipMalloc(&mm->d_spans, (size_t)mm->span_cap * sizeof(UploadSpan)) != hipSuccess ||
      hipHostMalloc(&mm->h_slot_tab, slots * 4, 0) != hipSuccess) {
    set_error("svo_mm_create: upload tables for %llu slots could not be allocated", (unsigned long long)slots);
    return SVO_ERR_NOMEM;
  }
  // initClBuffer: the whole incore buffer goes up once; beyond the slots'
  // extents it is zero, which the device gets from a memset
  HIP_TRY(hipMemset(mm->d_nodes, 0, pool_nodes * 8));
  HIP_TRY(hipMemset(mm->d_colors, 0, pool_nodes * 4));
  HIP_TRY(hipMemset(mm->d_normals, 0, pool_nodes * 4));
  mm->dev_extent.assign(mm->num_slots, 0u);
  mm->dev_nrm_extent.assign(mm->num_slots, 0u);
  mm->nrm_stale.assign(mm->num_slots, 0u);
  r = mm->upload_runs({{0u, mm->num_slots}}, nullptr, nullptr, nullptr);
  if (r != SVO_OK) return r;
  mm->dirty.clear();
  *out = mm.release();
  return SVO_OK;
}

extern "C" void svo_mm_destroy(svo_mm* mm) { delete mm; }

extern "C" svo_status svo_mm_reset(svo_mm* mm) {
  if (!mm) { set_error("svo_mm_reset: null handle"); return SVO_ERR_ARG; }
  mm->reset_host();
  if (mm->fb_pending) hipEventSynchronize(mm->fb_event);  // its pinned buffers stay in use until then
  mm->fb_cached = false;
  mm->fb_pending = false;
  mm->fb_parsed = false;
  for (uint32_t s = 0; s < mm->num_slots; ++s) mm->mark(s);
  return mm->insert(0, 999.0f);
}

extern "C" svo_status svo_mm_insert(svo_mm* mm, uint32_t gid, float error) {
  if (!mm) { set_error("svo_mm_insert: null handle"); return SVO_ERR_ARG; }
  return mm->insert(gid, error);
}

extern "C" svo_status svo_mm_remove(svo_mm* mm, uint32_t gid) {
  if (!mm) { set_error("svo_mm_remove: null handle"); return SVO_ERR_ARG; }
  if (gid == 0 || gid >= mm->scene->treelets.size()) { set_error("svo_mm_remove: gid %u cannot be removed", gid); return SVO_ERR_ARG; }
  if (mm->treelet_slot[gid] < 0) { set_error("svo_mm_remove: treelet %u is not resident", gid); return SVO_ERR_STATE; }
  mm->free_descendants(gid);
  mm->remove_one(gid);
  return SVO_OK;
}

extern "C" svo_status svo_mm_make_resident_all(svo_mm* mm) {
  if (!mm) { set_error("svo_mm_make_resident_all: null handle"); return SVO_ERR_ARG; }
  for (uint32_t g = 1; g < mm->scene->treelets.size(); ++g) {
    if (mm->treelet_slot[g] >= 0) continue;
    svo_status r = mm->insert(g, 0.0f);
    if (r != SVO_OK) return r;
  }
  return SVO_OK;
}

extern "C" svo_status svo_mm_update_device(svo_mm* mm, void* stream, svo_upload_stats* stats) {
  if (!mm) { set_error("svo_mm_update_device: null handle"); return SVO_ERR_ARG; }
  if (!mm->host_only()) HIP_TRY(hipSetDevice(mm->device));
  return mm->update_device((hipStream_t)stream, stats);
}

extern "C" svo_status svo_mm_get_info(const svo_mm* mm, svo_mm_info* out) {
  if (!mm || !out) { set_error("svo_mm_get_info: null argument"); return SVO_ERR_ARG; }
  out->num_slots = mm->num_slots;
  out->nodes_per_treelet = mm->N;
  out->free_slots = (uint32_t)mm->free_slots.size();
  out->dirty_slots = (uint32_t)mm->dirty.size();
  out->resident_treelets = mm->resident;
  out->device = (uint32_t)mm->device;
  out->pool_nodes = (uint64_t)mm->num_slots * mm->N;
  out->device_bytes = mm->host_only() ? 0
                                      : out->pool_nodes * 16 + (uint64_t)mm->num_slots * 12 +
                                            (uint64_t)mm->scene->treelets.size() * 20 + 8;
  return SVO_OK;
}

extern "C" svo_status svo_mm_slot_treelet(const svo_mm* mm, uint32_t slot, int32_t* gid) {
  if (!mm || !gid) { set_error("svo_mm_slot_treelet: null argument"); return SVO_ERR_ARG; }
  if (slot >= mm->num_slots) { set_error("slot %u out of range (%u slots)", slot, mm->num_slots); return SVO_ERR_ARG; }
  *gid = mm->slots[slot].gid;
  return SVO_OK;
}

extern "C" svo_status svo_mm_treelet_slot(const svo_mm* mm, uint32_t gid, int32_t* slot) {
  if (!mm || !slot) { set_error("svo_mm_treelet_slot: null argument"); return SVO_ERR_ARG; }
  if (gid >= mm->treelet_slot.size()) { set_error("treelet gid %u out of range", gid); return SVO_ERR_ARG; }
  *slot = mm->treelet_slot[gid];
  return SVO_OK;
}

extern "C" svo_status svo_mm_read_host_incore(const svo_mm* mm, uint64_t first, uint64_t count, svo_node* nodes,
                                              uint32_t* attribs) {
  if (!mm) { set_error("svo_mm_read_host_incore: null handle"); return SVO_ERR_ARG; }
  if (first > mm->incore.size() || count > mm->incore.size() - first) { set_error("incore range out of bounds"); return SVO_ERR_ARG; }
  if (nodes) std::copy(mm->incore.begin() + first, mm->incore.begin() + first + count, nodes);
  if (attribs)
    for (uint64_t i = 0; i < count; ++i) {
      attribs[2 * i] = mm->normals[first + i];
      attribs[2 * i + 1] = mm->colors[first + i];
    }
  return SVO_OK;
}

extern "C" svo_status svo_mm_read_device_incore(const svo_mm* mm, uint64_t first, uint64_t count, svo_node* nodes,
                                                uint32_t* attribs) {
  if (!mm) { set_error("svo_mm_read_device_incore: null handle"); return SVO_ERR_ARG; }
  if (first > mm->incore.size() || count > mm->incore.size() - first) { set_error("incore range out of bounds"); return SVO_ERR_ARG; }
  if (mm->host_only()) { set_error("svo_mm_read_device_incore: handle has no device pools"); return SVO_ERR_STATE; }
  HIP_TRY(hipSetDevice(mm->device));
  if (nodes) HIP_TRY(hipMemcpy(nodes, mm->d_nodes + first, count * 8, hipMemcpyDeviceToHost));
  if (attribs) {
    // bring the device normals up to date first (the handle's observable
    // state does not change)
    svo_status r = const_cast<svo_mm*>(mm)->sync_normals(nullptr);
    if (r != SVO_OK) return r;
    std::vector<uint32_t> n(count), c(count);
    HIP_TRY(hipMemcpy(n.data(), mm->d_normals + first, count * 4, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(c.data(), mm->d_colors + first, count * 4, hipMemcpyDeviceToHost));
    for (uint64_t i = 0; i < count; ++i) { attribs[2 * i] = n[i]; attribs[2 * i + 1] = c[i]; }
  }
  return SVO_OK;
}

extern "C" svo_status svo_mm_get_treelets(const svo_mm* mm, uint32_t first, uint32_t count, svo_treelet_header* headers,
                                          int32_t* slots) {
  if (!mm) { set_error("svo_mm_get_treelets: null handle"); return SVO_ERR_ARG; }
  const size_t nt = mm->scene->treelets.size();
  if (first > nt || count > nt - first) { set_error("svo_mm_get_treelets: treelets [%u, %u+%u) out of range (%zu)", first, first, count, nt); return SVO_ERR_ARG; }
  for (uint32_t i = 0; i < count; ++i) {
    if (headers) headers[i] = mm->scene->treelets[first + i].h;
    if (slots) slots[i] = mm->treelet_slot[first + i];
  }
  return SVO_OK;
}

extern "C" svo_status svo_mm_get_incore_buffer(const svo_mm* mm, const svo_node** nodes, uint64_t* count) {
  if (!mm || !nodes) { set_error("svo_mm_get_incore_buffer: null argument"); return SVO_ERR_ARG; }
  *nodes = mm->incore.data();
  if (count) *count = mm->incore.size();
  return SVO_OK;
}

extern "C" svo_status svo_mm_get_attribute_incore_buffer(const svo_mm* mm, const uint32_t** normals,
                                                         const uint32_t** colors, uint64_t* count) {
  if (!mm || (!normals && !colors)) { set_error("svo_mm_get_attribute_incore_buffer: null argument"); return SVO_ERR_ARG; }
  if (normals) *normals = mm->normals.data();
  if (colors) *colors = mm->colors.data();
  if (count) *count = mm->colors.size();
  return SVO_OK;
}

extern "C" svo_status svo_mm_free_incore_position(svo_mm* mm, int32_t* freed_gid) {
  if (!mm || !freed_gid) { set_error("svo_mm_free_incore_position: null argument"); return SVO_ERR_ARG; }
  const uint32_t g = mm->free_position();
  *freed_gid = g == 0 ? -1 : (int32_t)g;
  return SVO_OK;
}

extern "C" svo_status svo_mm_get_free_slot_stack(const svo_mm* mm, uint32_t* slots, uint32_t cap, uint32_t* n) {
  if (!mm || !n || (cap && !slots)) { set_error("svo_mm_get_free_slot_stack: null argument"); return SVO_ERR_ARG; }
  *n = (uint32_t)mm->free_slots.size();
  for (uint32_t i = 0; i < *n && i < cap; ++i) slots[i] = mm->free_slots[i];
  return SVO_OK;
}

extern "C" svo_status svo_mm_get_slots_to_upload(const svo_mm* mm, uint32_t* slots, uint32_t cap, uint32_t* n) {
  if (!mm || !n || (cap && !slots)) { set_error("svo_mm_get_slots_to_upload: null argument"); return SVO_ERR_ARG; }
  *n = (uint32_t)mm->dirty.size();
  uint32_t i = 0;
  for (const uint32_t s : mm->dirty) {
    if (i >= cap) break;
    slots[i++] = s;
  }
  return SVO_OK;
}

extern "C" svo_status svo_mm_get_slots(const svo_mm* mm, uint32_t first, uint32_t count, svo_slot_info* out) {
  if (!mm || (count && !out)) { set_error("svo_mm_get_slots: null argument"); return SVO_ERR_ARG; }
  if (first > mm->num_slots || count > mm->num_slots - first) { set_error("svo_mm_get_slots: slots [%u, %u+%u) out of range (%u)", first, first, count, mm->num_slots); return SVO_ERR_ARG; }
  for (uint32_t i = 0; i < count; ++i) {
    const SlotInfo& si = mm->slots[first + i];
    out[i] = svo_slot_info{si.gid, si.visibility, si.error, 0u};
  }
  return SVO_OK;
}

extern "C" svo_status svo_mm_get_device_pools(svo_mm* mm, svo_device_pools* out) {
  if (!mm || !out) { set_error("svo_mm_get_device_pools: null argument"); return SVO_ERR_ARG; }
  if (mm->host_only()) { set_error("svo_mm_get_device_pools: handle has no device pools"); return SVO_ERR_STATE; }
  HIP_TRY(hipSetDevice(mm->device));
  // kernels outside the library may read the normals: bring them up to date
  // now and send them with every later update
  mm->eager_normals = true;
  const svo_status r = mm->sync_normals(nullptr);
  if (r != SVO_OK) return r;
  *out = svo_device_pools{mm->d_nodes, mm->d_colors, mm->d_normals, mm->d_slot_gid, (uint64_t)mm->num_slots * mm->N,
                          mm->num_slots, mm->N, mm->device, 0u};
  return SVO_OK;
}

extern "C" svo_status svo_mm_init_device_buffer(svo_mm* mm, void* stream) {
  if (!mm) { set_error("svo_mm_init_device_buffer: null handle"); return SVO_ERR_ARG; }
  if (mm->host_only()) { set_error("svo_mm_init_device_buffer: handle has no device pools"); return SVO_ERR_STATE; }
  HIP_TRY(hipSetDevice(mm->device));
  // the whole buffer, as the reference's initClBuffer: the pools are
  // cleared, then every slot goes up within its extent (beyond it the host
  // buffer is zero)
  const hipStream_t st = (hipStream_t)stream;
  const size_t pn = (size_t)mm->num_slots * mm->N;
  HIP_TRY(hipMemsetAsync(mm->d_nodes, 0, pn * 8, st));
  HIP_TRY(hipMemsetAsync(mm->d_colors, 0, pn * 4, st));
  HIP_TRY(hipMemsetAsync(mm->d_normals, 0, pn * 4, st));
  for (uint32_t s = 0; s < mm->num_slots; ++s) {
    mm->refilled[s] = 1u;
    mm->dev_extent[s] = 0u;
    mm->dev_nrm_extent[s] = 0u;
    if (!mm->nrm_stale[s]) {
      mm->nrm_stale[s] = 1u;
      mm->nrm_slots.push_back(s);
    }
    mm->mark(s);
  }
  svo_upload_stats us;
  svo_status r = mm->update_device(st, &us);
  if (r != SVO_OK) return r;
  return mm->sync_normals(st);
}

// ------------------------------------------------------------ render hooks

namespace svo_impl {

int mm_device(const svo_mm* mm) { return mm->device; }

uint32_t mm_dirty_slots(const svo_mm* mm) { return (uint32_t)mm->dirty.size(); }

svo_status mm_sync_normals(svo_mm* mm, void* stream) { return mm->sync_normals((hipStream_t)stream); }

svo_renderer* mm_default_renderer(svo_mm* mm) {
  if (!mm->renderer && svo_renderer_create(mm->device, &mm->renderer) != SVO_OK) return nullptr;
  return mm->renderer;
}

svo_status mm_trace_pools(const svo_mm* mm, int policy, TracePools* P) {
  if (mm->host_only()) { set_error("residency handle has no device pools"); return SVO_ERR_STATE; }
  if (!mm->dirty.empty()) {
    set_error("%zu incore slots not uploaded; call svo_mm_update_device first", mm->dirty.size());
    return SVO_ERR_STATE;
  }
  const uint32_t D = mm->scene->max_depth;
  *P = TracePools{};
  P->nodes = mm->d_nodes;
  P->colors = mm->d_colors;
  P->normals = mm->d_normals;
  P->slot_gid = mm->d_slot_gid;
  P->nodes_per_treelet = mm->N;
  P->slot_shift = ((mm->N & (mm->N - 1)) == 0) ? (uint32_t)__builtin_ctz(mm->N) : 32u;
  P->num_slots = mm->num_slots;
  P->pool_nodes = mm->num_slots * mm->N;
  P->num_gids = (uint32_t)mm->scene->treelets.size();
  P->fb_slot = mm->d_fb_slot;
  P->fb_gid = mm->d_fb_gid;
  P->fb_queue = mm->d_fb_queue;
  P->fb_state = mm->d_fb_state;
  P->fb_cap = P->num_gids;
  if (mm->scene->has_package0) {
    P->package0_normal = mm->scene->package0[0];
    P->package0_color = mm->scene->package0[1];
  } else if (!mm->scene->treelets.empty() && mm->scene->treelets[0].attribs.size() >= 2) {
    P->package0_normal = mm->scene->treelets[0].attribs[0];
    P->package0_color = mm->scene->treelets[0].attribs[1];
  }
  if (policy == SVO_POLICY_TREELET) {
    // MAX_STACK_SIZE 13 / MAX_SVO_RAYCAST_DEPTH 12 (treeletRenderer_rgba.cl:7-8),
    // widened to max(13, D+1) so that leaves deeper than 12 are reachable.
    P->scale_max = std::max(13u, D + 1);
    P->lod_cap = P->scale_max - 1;
    P->max_loops = (P->scale_max + 1) * (P->scale_max + 1) * 5;
  } else if (policy == SVO_POLICY_CPU3) {
    // scaleMax = SVO depth, maxLoops = (D+1)^2 * 5
    // (src/incore/CpuRaycasterSingleRay3.cpp:167,231)
    P->scale_max = std::max(1u, D);
    P->lod_cap = 0;
    P->max_loops = (D + 1) * (D + 1) * 5;
  } else if (policy == SVO_POLICY_SINGLE) {
    // scaleMax 15, maxLoops (15+1)^2 * 5 (fillFrameBuffer_rgba.cl:256,290)
    if (D > 15) { set_error("the single-SVO policy reaches depth 15 at most (scene depth %u)", D); return SVO_ERR_ARG; }
    if (mm->scene->treelets.size() != 1 || mm->treelet_slot[0] != 0) {
      set_error("the single-SVO policy renders scenes of one treelet (in slot 0)");
      return SVO_ERR_ARG;
    }
    P->scale_max = 15;
    P->lod_cap = 0;
    P->max_loops = 16 * 16 * 5;
  } else {
    set_error("unknown traversal policy %d", policy);
    return SVO_ERR_ARG;
  }
  if (P->scale_max > 16) { set_error("scene depth %u exceeds the 16-level traversal stack", D); return SVO_ERR_ARG; }
  // A push leaves scale s for s-1 and makes the child at depth scale_max - s
  // the current node.  Policy A never pushes from scale scale_max - lod_cap
  // (the depth cap stops there), policy C never from 0; and no policy pushes
  // a node deeper than the scene's deepest pushable node (inner_depth), i.e.
  // from a scale below scale_max - inner_depth (a well-formed tree of depth
  // D pushes nodes down to depth D - 1: one stack level less than D).
  uint32_t mps = P->lod_cap != 0 ? P->scale_max - P->lod_cap + 1 : 1;
  const uint32_t inner = mm->scene->inner_depth;
  if (inner < P->scale_max && P->scale_max - inner > mps) mps = P->scale_max - inner;
  P->min_push_scale = mps;
  P->stack_levels = P->scale_max > P->min_push_scale ? P->scale_max - P->min_push_scale : 1;
  return SVO_OK;
}

}  // namespace svo_impl

namespace {

std::vector<svo_view> make_views(svo_mm* mm, const svo_camera* cams, uint32_t n, const svo_frame_out* outs) {
  std::vector<svo_view> v(n);
  for (uint32_t i = 0; i < n; ++i) v[i] = svo_view{mm, cams[i], outs[i]};
  return v;
}

}  // namespace

extern "C" svo_status svo_render(svo_mm* mm, const svo_camera* cams, uint32_t n_cams, const svo_frame_out* outs,
                                 int policy, void* stream) {
  if (!mm || !cams || !outs || n_cams == 0) { set_error("svo_render: null argument or no camera"); return SVO_ERR_ARG; }
  if (mm->host_only()) { set_error("svo_render: handle has no device pools"); return SVO_ERR_STATE; }
  svo_renderer* r = mm_default_renderer(mm);
  if (!r) return SVO_ERR_HIP;
  const std::vector<svo_view> v = make_views(mm, cams, n_cams, outs);
  return svo_render_views(r, v.data(), n_cams, policy, stream);
}

extern "C" svo_status svo_render_timed(svo_mm* mm, const svo_camera* cams, uint32_t n_cams, const svo_frame_out* outs,
                                       int policy, void* stream, uint32_t iters, float* mean_ms) {
  if (!mm || !cams || !outs || n_cams == 0) { set_error("svo_render_timed: null argument or no camera"); return SVO_ERR_ARG; }
  if (mm->host_only()) { set_error("svo_render_timed: handle has no device pools"); return SVO_ERR_STATE; }
  svo_renderer* r = mm_default_renderer(mm);
  if (!r) return SVO_ERR_HIP;
  const std::vector<svo_view> v = make_views(mm, cams, n_cams, outs);
  return svo_render_views_timed(r, v.data(), n_cams, policy, stream, iters, mean_ms);
}

// ------------------------------------------------------------ analysis feedback

namespace {
float key_error(unsigned long long key) {
  const uint32_t k = (uint32_t)key;
  const uint32_t b = (k & 0x80000000u) ? (k & 0x7FFFFFFFu) : ~k;
  float f;
  std::memcpy(&f, &b, 4);
  return f;
}
}  // namespace

extern "C" svo_status svo_analysis_camera(const svo_camera* cam, uint32_t divide, float jitter_x, float jitter_y,
                                          svo_camera* out) {
  if (!cam || !out) { set_error("svo_analysis_camera: null argument"); return SVO_ERR_ARG; }
  if (divide == 0 || cam->width < divide || cam->height < divide) {
    set_error("svo_analysis_camera: divide %u does not fit a %ux%u frame", divide, cam->width, cam->height);
    return SVO_ERR_ARG;
  }
  if (!(jitter_x >= -1.0f && jitter_x <= 1.0f && jitter_y >= -1.0f && jitter_y <= 1.0f)) {
    set_error("svo_analysis_camera: jitter must lie in [-1, 1]");
    return SVO_ERR_ARG;
  }
  svo_camera c = *cam;
  c.width = cam->width / divide;
  c.height = cam->height / divide;
  const float sx = (float)cam->width / (float)c.width, sy = (float)cam->height / (float)c.height;
  for (int a = 0; a < 3; ++a) {
    c.pixel_h[a] = cam->pixel_h[a] * sx;
    c.pixel_v[a] = cam->pixel_v[a] * sy;
    c.far_lower_left[a] = (cam->far_lower_left[a] + ((jitter_x * (float)divide) * cam->pixel_h[a]) * 0.5f) +
                          ((jitter_y * (float)divide) * cam->pixel_v[a]) * 0.5f;
  }
  *out = c;
  return SVO_OK;
}

extern "C" svo_status svo_mm_analyse(svo_mm* mm, const svo_camera* cam, uint32_t divide, float jitter_x,
                                     float jitter_y, int policy, void* stream) {
  if (!mm || !cam) { set_error("svo_mm_analyse: null argument"); return SVO_ERR_ARG; }
  if (mm->host_only()) { set_error("svo_mm_analyse: handle has no device pools"); return SVO_ERR_STATE; }
  svo_view v{};
  v.mm = mm;
  svo_status r = svo_analysis_camera(cam, divide, jitter_x, jitter_y, &v.camera);
  if (r != SVO_OK) return r;
  v.out = svo_frame_out{nullptr, nullptr, nullptr, nullptr, nullptr};
  v.feedback_stride = 1;
  svo_renderer* rd = mm_default_renderer(mm);
  if (!rd) return SVO_ERR_HIP;
  return svo_render_views(rd, &v, 1, policy, stream);
}

extern "C" svo_status svo_mm_apply_records(svo_mm* mm, const uint32_t* slots, const float* slot_errors,
                                           uint32_t n_slots, const uint32_t* gids, const float* gid_errors,
                                           uint32_t n_gids, svo_feedback_stats* stats) {
  if (!mm || (n_slots && (!slots || !slot_errors)) || (n_gids && (!gids || !gid_errors))) {
    set_error("svo_mm_apply_records: null argument");
    return SVO_ERR_ARG;
  }
  std::vector<std::pair<uint32_t, float>> seen(n_slots), asked;
  for (uint32_t i = 0; i < n_slots; ++i) seen[i] = {slots[i], slot_errors[i]};
  std::set<uint32_t> dup;
  for (uint32_t i = 0; i < n_gids; ++i) {
    if (gids[i] == 0 || gids[i] >= mm->scene->treelets.size()) { set_error("svo_mm_apply_records: gid %u out of range", gids[i]); return SVO_ERR_ARG; }
    if (!dup.insert(gids[i]).second) { set_error("svo_mm_apply_records: gid %u listed twice", gids[i]); return SVO_ERR_ARG; }
    asked.push_back({gids[i], gid_errors[i]});
  }
  mm->apply_records(seen, asked, stats);
  return SVO_OK;
}

namespace {

// Readback of a recording cycle, in two halves.  enqueue_export: one
// export kernel, behind the recording launches on `st`, writes the slot
// keys, the count and the asked gids with their keys into the pinned
// buffers (only the queued entries of the per-gid table cross PCIe), the
// count is reset from a pinned zero, and the cycle ends: later recordings
// carry the next stamp.  finish_export waits for that work and parses the
// buffers into mm->fb_seen / fb_asked.
svo_status enqueue_export(svo_mm* mm, hipStream_t st) {
  HIP_TRY(hipSetDevice(mm->device));
  const uint32_t ng = (uint32_t)mm->scene->treelets.size();
  // the kernel reads this cycle and moves the device stamp on: records of
  // launches ordered after it carry the next stamp
  const FeedbackExport x{mm->d_fb_slot, mm->d_fb_state, mm->d_fb_queue, mm->d_fb_gid, mm->num_slots, ng,
                         mm->h_fb_slot, mm->h_fb_count, mm->h_fb_queue, mm->h_fb_gid, mm->fb_stamp + 1u, 0u};
  const int e = launch_feedback_export(x, st);
  if (e != 0) { set_error("feedback readback launch failed: %s", hipGetErrorString((hipError_t)e)); return SVO_ERR_HIP; }
  HIP_TRY(hipMemcpyAsync(mm->d_fb_state, mm->h_zero, 4, hipMemcpyHostToDevice, st));
  HIP_TRY(hipEventRecord(mm->fb_event, st));
  mm->fb_export_stamp = mm->fb_stamp;
  ++mm->fb_stamp;
  mm->fb_pending = true;
  return SVO_OK;
}

svo_status finish_export(svo_mm* mm) {
  HIP_TRY(hipEventSynchronize(mm->fb_event));
  mm->fb_pending = false;
  const uint32_t ng = (uint32_t)mm->scene->treelets.size();
  const uint32_t n = std::min(*mm->h_fb_count, ng);
  const uint32_t stamp = mm->fb_export_stamp;
  mm->fb_seen.clear();
  mm->fb_asked.clear();
  for (uint32_t s = 0; s < mm->num_slots; ++s) {
    const unsigned long long k = mm->h_fb_slot[s];
    if ((uint32_t)(k >> 32) == stamp) mm->fb_seen.push_back({s, key_error(k)});
  }
  for (uint32_t i = 0; i < n; ++i) {
    const uint32_t g = mm->h_fb_queue[i];
    if (g < ng) mm->fb_asked.push_back({g, key_error(mm->h_fb_gid[i])});
  }
  mm->fb_parsed = true;
  return SVO_OK;
}

// The cycle's tables in mm->fb_seen / fb_asked: parsed already, or the
// pending export finished, or a readback now.  A readback of its own waits
// for the recording launches before it enqueues the export: work queued
// behind a running kernel was measured to wait for a render on another
// stream as well (fly-through with overlapped renders: 614 -> 800
// frames/s); callers that overlap renders use svo_mm_fetch_feedback before
// they enqueue the render.
svo_status parsed_tables(svo_mm* mm, hipStream_t st) {
  if (mm->fb_parsed) return SVO_OK;
  if (!mm->fb_pending) {
    HIP_TRY(hipSetDevice(mm->device));
    HIP_TRY(hipStreamSynchronize(st));
    svo_status r = enqueue_export(mm, st);
    if (r != SVO_OK) return r;
  }
  return finish_export(mm);
}

}  // namespace

extern "C" svo_status svo_mm_read_feedback(svo_mm* mm, void* stream, uint32_t* old_gids, float* old_errors,
                                           uint32_t old_cap, uint32_t* n_old, uint32_t* new_gids, float* new_errors,
                                           uint32_t new_cap, uint32_t* n_new) {
  if (!mm || !n_old || !n_new || (old_cap && (!old_gids || !old_errors)) || (new_cap && (!new_gids || !new_errors))) {
    set_error("svo_mm_read_feedback: null argument");
    return SVO_ERR_ARG;
  }
  if (mm->host_only()) { set_error("svo_mm_read_feedback: handle has no device pools"); return SVO_ERR_STATE; }
  if (!mm->fb_cached) {
    svo_status r = parsed_tables(mm, (hipStream_t)stream);
    if (r != SVO_OK) return r;
    mm->fb_parsed = false;
    mm->fb_sets = mm->make_sets(mm->fb_seen, mm->fb_asked);
    mm->fb_cached = true;
  }
  const svo_mm::FeedbackSets& F = mm->fb_sets;
  uint32_t i = 0;
  for (const auto& e : F.old) {
    if (i < old_cap) { old_gids[i] = e.first; old_errors[i] = e.second; }
    ++i;
  }
  *n_old = i;
  for (i = 0; i < F.asked.size(); ++i)
    if (i < new_cap) { new_gids[i] = F.asked[i].first; new_errors[i] = F.asked[i].second; }
  *n_new = (uint32_t)F.asked.size();
  return SVO_OK;
}

extern "C" svo_status svo_mm_apply_feedback(svo_mm* mm, void* stream, svo_feedback_stats* stats) {
  if (!mm) { set_error("svo_mm_apply_feedback: null handle"); return SVO_ERR_ARG; }
  if (mm->host_only()) { set_error("svo_mm_apply_feedback: handle has no device pools"); return SVO_ERR_STATE; }
  if (!mm->fb_cached) {
    svo_status r = parsed_tables(mm, (hipStream_t)stream);
    if (r != SVO_OK) return r;
    mm->fb_sets = mm->make_sets(mm->fb_seen, mm->fb_asked);
  }
  mm->fb_cached = false;
  mm->fb_parsed = false;
  mm->apply_sets(mm->fb_sets, stats);
  return SVO_OK;
}

extern "C" svo_status svo_mm_fetch_feedback(svo_mm* mm, void* stream) {
  if (!mm) { set_error("svo_mm_fetch_feedback: null handle"); return SVO_ERR_ARG; }
  if (mm->host_only()) { set_error("svo_mm_fetch_feedback: handle has no device pools"); return SVO_ERR_STATE; }
  if (mm->fb_cached || mm->fb_parsed || mm->fb_pending) return SVO_OK;  // this cycle is read already
  return enqueue_export(mm, (hipStream_t)stream);
}

// ------------------------------------------------------------ device utilities

extern "C" svo_status svo_device_count(int* count) {
  if (!count) { set_error("svo_device_count: null"); return SVO_ERR_ARG; }
  hipError_t e = hipGetDeviceCount(count);
  if (e != hipSuccess) { *count = 0; set_error("hipGetDeviceCount: %s", hipGetErrorString(e)); return SVO_ERR_HIP; }
  return SVO_OK;
}
extern "C" svo_status svo_set_device(int device) { HIP_TRY(hipSetDevice(device)); return SVO_OK; }
extern "C" svo_status svo_device_malloc(void** p, size_t bytes) {
  if (!p) { set_error("svo_device_malloc: null"); return SVO_ERR_ARG; }
  if (hipMalloc(p, bytes) != hipSuccess) { set_error("hipMalloc(%zu) failed", bytes); return SVO_ERR_NOMEM; }
  return SVO_OK;
}
extern "C" svo_status svo_device_free(void* p) { HIP_TRY(hipFree(p)); return SVO_OK; }
extern "C" svo_status svo_memcpy_h2d(void* d, const void* s, size_t n) { HIP_TRY(hipMemcpy(d, s, n, hipMemcpyHostToDevice)); return SVO_OK; }
extern "C" svo_status svo_memcpy_d2h(void* d, const void* s, size_t n) { HIP_TRY(hipMemcpy(d, s, n, hipMemcpyDeviceToHost)); return SVO_OK; }
extern "C" svo_status svo_memset_device(void* d, int v, size_t n) { HIP_TRY(hipMemset(d, v, n)); return SVO_OK; }
extern "C" svo_status svo_stream_create(void** s) {
  if (!s) { set_error("svo_stream_create: null"); return SVO_ERR_ARG; }
  HIP_TRY(hipStreamCreateWithFlags((hipStream_t*)s, hipStreamNonBlocking));
  return SVO_OK;
}
extern "C" svo_status svo_stream_destroy(void* s) { HIP_TRY(hipStreamDestroy((hipStream_t)s)); return SVO_OK; }
extern "C" svo_status svo_stream_synchronize(void* s) { HIP_TRY(hipStreamSynchronize((hipStream_t)s)); return SVO_OK; }
extern "C" svo_status svo_device_synchronize(void) { HIP_TRY(hipDeviceSynchronize()); return SVO_OK; }

extern "C" svo_status svo_debug_ray_dirs(const svo_camera* cam, float* dirs_dev, void* stream) {
  if (!cam || !dirs_dev) { set_error("svo_debug_ray_dirs: null argument"); return SVO_ERR_ARG; }
  int e = launch_ray_dirs(*cam, dirs_dev, stream);
  if (e != 0) { set_error("ray-direction launch failed: %s", hipGetErrorString((hipError_t)e)); return SVO_ERR_HIP; }
  return SVO_OK;
}

extern "C" svo_status svo_debug_valu(uint32_t* out, uint32_t blocks, uint32_t iters, void* stream) {
  if (!out || blocks == 0 || blocks > (1u << 20)) { set_error("svo_debug_valu: null output or blocks not in 1..2^20"); return SVO_ERR_ARG; }
  int e = launch_valu_probe(out, blocks, iters, stream);
  if (e != 0) { set_error("VALU probe launch failed: %s", hipGetErrorString((hipError_t)e)); return SVO_ERR_HIP; }
  return SVO_OK;
}

extern "C" svo_status svo_debug_copy(void* dst, const void* src, size_t bytes, void* stream) {
  if (!dst || !src || (bytes % 16) != 0) { set_error("svo_debug_copy: null pointer or size not a multiple of 16"); return SVO_ERR_ARG; }
  int e = launch_copy(dst, src, bytes, stream);
  if (e != 0) { set_error("copy launch failed: %s", hipGetErrorString((hipError_t)e)); return SVO_ERR_HIP; }
  return SVO_OK;
}
