// Launch parameters shared by the host side of libsvo_hip.so and the HIP
// traversal kernels in trace.hip.  Plain C++ (no HIP types).
#pragma once

#include <cstdint>
#include <vector>

#include "../../include/svo_hip.h"

namespace svo_impl {

// Octree pools of one scene in HBM.  nodes[i] = {first_child, masks},
// colors[i] = packed_color of node i, slot_gid[s] = gid held by slot s
// (-1 empty).  Traversal parameters follow the policy.
struct TracePools {
  const uint64_t* nodes;
  const uint32_t* colors;
  const uint32_t* normals;
  const int32_t* slot_gid;
  uint32_t nodes_per_treelet;
  uint32_t slot_shift;      // log2(nodes_per_treelet) when a power of two, else 32
  uint32_t num_slots;
  uint32_t pool_nodes;      // num_slots * nodes_per_treelet (< 2^32)
  uint32_t scale_max;       // policy scaleMax
  uint32_t max_loops;       // policy maxLoops
  uint32_t lod_cap;         // policy-A depth cap (MAX_SVO_RAYCAST_DEPTH), 0 = none
  uint32_t min_push_scale;  // lowest scale a push can start from
  uint32_t stack_levels;    // scale_max - min_push_scale (stack depth)
  uint32_t num_gids;        // treelets in the scene (size of fb_gid)
  // Analysis feedback tables of the handle (renderToFeedbackBuffer fused
  // into the traversal).  Keys are (stamp << 32) | order-preserving bits of
  // the error, combined with atomicMax, so a table needs no clearing between
  // analysis cycles: entries of older stamps are stale.  The stamp of the
  // current cycle lives in device memory (fb_state[1]) and is advanced by the
  // readback kernel on the stream, so descriptors staged once (a graph
  // batch) record into whatever cycle is current when they run.
  unsigned long long* fb_slot;   // per slot: the largest key of a record whose hit node lies in it
  unsigned long long* fb_gid;    // per treelet: the largest key of a record asking for it
  uint32_t* fb_queue;            // treelets asked for in this cycle, first-come order
  uint32_t* fb_state;            // [0] entries appended to fb_queue (may exceed fb_cap), [1] stamp
  uint32_t fb_cap;
  // attribute package 0 {packed_normal, packed_color}: what the single-SVO
  // renderer's reflection mode shades every bounce with
  uint32_t package0_normal;
  uint32_t package0_color;
  uint32_t pad;
};

// One view of a launch: its scene, camera uniforms, outputs and first tile.
// Views are kept in device memory sorted by tile_begin.
struct TraceFrame {
  TracePools pools;
  svo_camera cam;
  svo_frame_out out;
  uint32_t tile_begin;    // first global tile id of this view
  uint32_t tiles_x;       // tiles per row
  uint32_t fb_stride;     // record analysis feedback for pixels with x % s == y % s == 0 (0 = off)
  uint32_t pad;
};

// Shape of a launch: pixel tile of one wavefront (tile_w x 64/tile_w),
// wavefronts per workgroup and where the per-ray stack lives.
struct LaunchShape {
  uint32_t tile_w;  // 8, 16 or 32
  uint32_t waves;   // 1, 2 or 4
  uint32_t stack;   // SVO_STACK_LDS, SVO_STACK_PRIVATE or SVO_STACK_LDS_NODE
};

bool shape_supported(const LaunchShape& s);

// Enqueues the traversal of `num_tiles` tiles spread over `num_frames` views
// (frames_dev in device memory) on `stream`; `stack_levels` is the largest
// stack depth over the views; `shade` adds the view-mode pass for views
// with a shade output (with policy SVO_POLICY_SINGLE that pass can trace the
// single-SVO renderer's shadow and reflection rays, with an LDS stack of
// `stack_levels`).  Returns a hipError_t value (0 = success).
int launch_trace(const LaunchShape& shape, const TraceFrame* frames_dev, uint32_t num_frames, uint32_t num_tiles,
                 uint32_t stack_levels, int policy, bool shade, void* stream);

// Host image of one launch's view descriptors (render.cpp): views in order,
// tiles numbered consecutively for the given shape.
struct StagedLaunch {
  std::vector<TraceFrame> frames;
  uint32_t num_tiles = 0;
  uint32_t stack_levels = 1;
  bool shade = false;  // some view has a view-mode output
  // residency handles of the views that shade with normals, each once
  std::vector<svo_mm*> normal_handles;
};

// Tiles, stack depth and shading of a set of staged views.
struct StageTotals {
  uint32_t num_tiles;
  uint32_t stack_levels;
  bool shade;
};
// Writes the descriptors of views[0..n) into frames[0..n) (no allocation).
svo_status stage_views_into(const svo_view* views, uint32_t n, int policy, const LaunchShape& shape, int device,
                            TraceFrame* frames, StageTotals* totals);
// The view shades with the normals pool (normal, Phong, and the single-SVO
// renderer's shadow and reflection modes).
bool view_needs_normals(const svo_view& v);
bool mode_needs_normals(uint32_t mode);
// A view mode the library knows, and the policy it needs (the shadow and
// reflection modes trace single-SVO secondary rays).
svo_status check_view_mode(uint32_t mode, int policy, uint32_t view_index);

// Device scratch of the allocation-free launch: descriptors of at most
// kNoAllocMaxViews views, in a fixed array that each GPU holds for the life
// of the process.  noalloc_scratch returns this device's address of it (the
// current device must be `device`), looked up once per device.  Returns a
// hipError_t value.
constexpr uint32_t kNoAllocMaxViews = SVO_NOALLOC_MAX_VIEWS;
int noalloc_scratch(int device, TraceFrame** ptr);

// Sends the stale normals of a staged launch's handles (see
// mm_sync_normals) before it runs on `stream`.
svo_status sync_staged_normals(const StagedLaunch& sg, void* stream);
svo_status stage_views(const svo_view* views, uint32_t n, int policy, const LaunchShape& shape, int device,
                       StagedLaunch* out);
LaunchShape renderer_shape(const svo_renderer* r);
int renderer_device(const svo_renderer* r);

// Writes the normalised FP32 ray direction of every pixel of `cam`
// (3 floats per pixel, row-major) -- used to pin the ray generation.
int launch_ray_dirs(const svo_camera& cam, float* dirs_dev, void* stream);

// Copies bytes/16 16-byte words with a plain streaming kernel.
int launch_copy(void* dst, const void* src, size_t bytes, void* stream);

// VALU issue probe: `blocks` workgroups of 256 lanes, `iters` iterations of
// 16 VALU instructions each; out holds blocks * 256 words.
int launch_valu_probe(uint32_t* out, uint32_t blocks, uint32_t iters, void* stream);

// One span of an upload: `bytes` (a multiple of 4) from `src` (pinned host
// memory or device memory) to `dst` (device memory); first_chunk is the
// running sum of ceil(bytes / 16) over the spans before it.
struct UploadSpan {
  uint64_t src;
  uint64_t dst;
  uint64_t bytes;
  uint64_t first_chunk;
};

// Copies every span of `spans` (device memory, ascending first_chunk,
// total_chunks 16-byte chunks in all) in one launch.  Reading pinned host
// memory from a kernel streams over PCIe at the rate of one large DMA copy
// whatever the number of spans, where one DMA command per span pays a
// fixed cost per command.
int launch_upload(const UploadSpan* spans, uint32_t num_spans, uint64_t total_chunks, void* stream);

// Device feedback tables of one residency handle and the pinned host buffers
// a readback fills (h_keys[i] = key of the gid in h_queue[i]).
struct FeedbackExport {
  const unsigned long long* fb_slot;
  uint32_t* fb_state;  // [0] queue count, [1] stamp (set to next_stamp at the end)
  const uint32_t* fb_queue;
  const unsigned long long* fb_gid;
  uint32_t num_slots;
  uint32_t num_gids;
  unsigned long long* h_slot;
  uint32_t* h_count;
  uint32_t* h_queue;
  unsigned long long* h_keys;
  uint32_t next_stamp;
  uint32_t pad;
};

// Enqueues the readback of one analysis cycle into the pinned buffers; the
// kernel then starts the next cycle (device stamp = next_stamp).
int launch_feedback_export(const FeedbackExport& x, void* stream);

}  // namespace svo_impl
