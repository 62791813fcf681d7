// Launch parameters shared by the host side of libsvo_hip.so and the HIP
// traversal kernels in trace.hip.  Plain C++ (no HIP types).
#pragma once

#include <cstdint>

#include "../../include/svo_hip.h"

namespace svo_impl {

// Octree pools in HBM.  nodes[i] = {first_child, masks}, colors[i] =
// packed_color of node i, slot_gid[s] = gid held by slot s (-1 empty).
struct TracePools {
  const uint64_t* nodes;
  const uint32_t* colors;
  const int32_t* slot_gid;
  uint32_t nodes_per_treelet;
  uint32_t slot_shift;    // log2(nodes_per_treelet) when a power of two, else 32
  uint32_t num_slots;
  uint32_t pool_nodes;    // num_slots * nodes_per_treelet (< 2^32)
  uint32_t scale_max;     // policy scaleMax
  uint32_t max_loops;     // policy maxLoops
  uint32_t lod_cap;       // policy-A depth cap (MAX_SVO_RAYCAST_DEPTH), 0 = none
  uint32_t min_push_scale;  // lowest scale a push can start from
  uint32_t stack_levels;    // scale_max - min_push_scale (LDS stack depth)
};

// Pixel tile of one wavefront (64 lanes).
constexpr uint32_t kTileW = 16;
constexpr uint32_t kTileH = 4;

// One camera of a launch: its uniforms, its outputs and its first tile.
struct TraceFrame {
  svo_camera cam;
  svo_frame_out out;
  uint32_t tile_begin;    // first global tile id of this frame
  uint32_t tiles_x;       // tiles per row (kTileW x kTileH pixels per tile)
  uint32_t pad[2];
};

// Enqueues the traversal of `num_tiles` tiles spread over `num_frames`
// frames (frames[] in device memory, sorted by tile_begin) on `stream`.
// Returns a hipError_t value (0 = success).
int launch_trace(const TracePools& pools, const TraceFrame* frames_dev, uint32_t num_frames,
                 uint32_t num_tiles, int policy, void* stream);

// Writes the normalised FP32 ray direction of every pixel of `cam`
// (3 floats per pixel, row-major) -- used to pin the ray generation.
int launch_ray_dirs(const svo_camera& cam, float* dirs_dev, void* stream);

// Copies bytes/16 16-byte words with a plain streaming kernel.
int launch_copy(void* dst, const void* src, size_t bytes, void* stream);

}  // namespace svo_impl
