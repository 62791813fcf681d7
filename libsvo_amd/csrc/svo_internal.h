// Internal C++ declarations shared by the host translation units of
// libsvo_hip.so.  Not installed; the public surface is include/svo_hip.h and
// the C++ wrappers in libsvo_amd/csrc/svo/.
#pragma once

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/svo_hip.h"

namespace svo_impl {

// ------------------------------------------------------------ errors

void set_error(const char* fmt, ...) __attribute__((format(printf, 1, 2)));
void clear_error();

// ------------------------------------------------------------ node masks
// Bit layout of CpuSvoNode::_masks (src/CpuSvoNode.h:46-48): leaf mask in
// bits 0-7, valid mask in bits 8-15.
constexpr uint32_t kLeafShift = 0;
constexpr uint32_t kValidShift = 8;

inline bool valid_flag(uint32_t masks, uint32_t c) { return (masks >> (kValidShift + c)) & 1u; }
inline bool leaf_flag(uint32_t masks, uint32_t c) { return (masks >> (kLeafShift + c)) & 1u; }
// getNthChildIndex (src/CpuSvoNode.cpp:145-152): first + number of valid
// children with a smaller child index, modulo 2^32.
inline uint32_t nth_child(uint32_t masks, uint32_t first, uint32_t c) {
  return first + (uint32_t)__builtin_popcount((masks >> kValidShift) & ((1u << c) - 1u));
}

// ------------------------------------------------------------ attributes
// Packed colour r<<24|g<<16|b<<8|a, unpacked by unpackRgbaFromUnsigned
// (opencl/treeletRenderer_rgba.cl:183-195).
inline uint32_t pack_rgba(uint32_t r, uint32_t g, uint32_t b, uint32_t a) {
  return (r << 24) | (g << 16) | (b << 8) | a;
}

// ------------------------------------------------------------ hashing
inline uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
inline uint64_t hash_xyz(uint64_t seed, uint32_t i, uint32_t j, uint32_t k) {
  return mix64(mix64(seed) ^ (((uint64_t)i << 42) | ((uint64_t)j << 21) | (uint64_t)k));
}

// ------------------------------------------------------------ morton
// Child index of a node inside its parent is 4x+2y+z (x = 1 for the
// positive half), opencl/treeletRenderer_rgba.cl:289-291.  A Morton code of
// depth D holds D such triplets, the root's child first.
uint64_t morton_encode(uint32_t i, uint32_t j, uint32_t k, uint32_t depth);
void morton_decode(uint64_t code, uint32_t depth, uint32_t* i, uint32_t* j, uint32_t* k);

// In-place LSD radix sort of 64-bit keys whose set bits lie below `bits`.
void radix_sort_u64(std::vector<uint64_t>& keys, uint32_t bits);

// ------------------------------------------------------------ scene
struct TreeletData {
  svo_treelet_header h{};
  std::vector<svo_node> nodes;     // h.num_nodes
  std::vector<uint32_t> attribs;   // 2 * h.num_nodes: {normal, color}
};

}  // namespace svo_impl

struct svo_scene {
  uint32_t treelet_bytes = 0;
  uint32_t nodes_per_treelet = 0;
  uint32_t max_depth = 0;
  // depth of the deepest node a traversal can push: a valid non-leaf node
  // or a treelet root (an anchor leaf becomes one when its sub-treelet is
  // stitched in); max_depth - 1 in a well-formed tree
  uint32_t inner_depth = 0;
  // attribute package 0 of a single-SVO file {packed_normal, packed_color}
  // (the package the single-SVO renderer's reflection mode shades with);
  // scenes from treelet bundles or builders use the root node's attributes
  bool has_package0 = false;
  uint32_t package0[2] = {0, 0};
  uint64_t tree_nodes = 0;
  uint64_t tree_leaves = 0;
  std::vector<svo_impl::TreeletData> treelets;
};

namespace svo_impl {

// Depth of the deepest valid node of a scene (root = 0; the deepest leaf in a
// well-formed tree), checking that every child offset stays inside its
// treelet; -1 (and *why) on a bad offset.  scene.cpp.
int scene_depth(const svo_scene& sc, std::string* why, uint32_t* inner_depth = nullptr);

struct TracePools;

// Residency handles as seen by the renderer (memory_manager.cpp).
int mm_device(const svo_mm* mm);
// Slots waiting for svo_mm_update_device.
uint32_t mm_dirty_slots(const svo_mm* mm);
svo_renderer* mm_default_renderer(svo_mm* mm);
// Pools and traversal parameters of `mm` for `policy`; fails while slots
// wait for upload or when the handle has no device pools.
svo_status mm_trace_pools(const svo_mm* mm, int policy, TracePools* out);
// Sends the normals that uploads left stale (host-synchronous on `stream`);
// launches whose views shade with normals call it first.
svo_status mm_sync_normals(svo_mm* mm, void* stream);

// Builds the scene from sorted, unique leaf codes at `depth` and two
// attribute words per leaf.  Consumes `codes` (cleared on return).
svo_status build_scene(uint32_t depth, std::vector<uint64_t>& codes,
                       std::vector<uint32_t>& leaf_attribs,
                       uint32_t treelet_bytes, uint32_t flags, svo_scene** out);

}  // namespace svo_impl
