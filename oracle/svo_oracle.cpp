// CPU oracle for the serialized-SVO treelet ray caster.
//
// TEST INFRASTRUCTURE ONLY.  This file restates the reference's algorithms on
// the CPU so that the HIP path can be checked against them; it is loaded by
// the tests, by __graft_entry__.smoke() to check the GPU result and by
// bench.py's CPU-baseline leg, and never by the product library.
//
// It is written independently of libsvo_amd/csrc: its own .svo/.ia reader,
// its own incore-buffer stitching and a traversal that keeps the reference's
// full explicit stack (index, tmin, tmax, centre per level) instead of the
// GPU kernel's recomputed stack.
//
// Restated reference functions (paths relative to the libsvo tree):
//   read_bundle      TreeletMemoryManager::loadFromFile + Treelet::loadFromFile
//                    (src/TreeletMemoryManager.cpp:229-291,
//                    src/Treelet.cpp:459-478); .ia is this project's layout
//   make_incore      resetIncoreBuffer + insertTreeletIntoIncoreBuffer for
//                    every treelet in gid order
//                    (src/TreeletMemoryManager.cpp:303-380,455-548)
//   intersect_aabb   opencl/treeletRenderer_rgba.cl:100-178
//   pixel_ray        renderToBuffer ray set-up, treeletRenderer_rgba.cl:627-633
//                    (normalize(v) taken as v / sqrt(v.v))
//   sample_treelet   treeletRenderer_rgba.cl:201-382 (policy A)
//   travers_cpu3     src/incore/CpuRaycasterSingleRay3.cpp:158-387 (policy C),
//                    FP32 arithmetic, relative child indexing on the incore
//                    buffer, the same root-box routine as policy A
//   get_color        treeletRenderer_rgba.cl:183-195,399-404,645-655
//
// Pinning status of each function is recorded in tests/test_oracle_cpu.py;
// the boundary-ray flag is this project's definition (see kBoundary below).
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include <omp.h>

namespace {

struct Node {
  uint32_t first;
  uint32_t masks;
};

struct Header {
  uint32_t num_nodes, num_leaves, mem_size, gid, parent_gid, leaf_pos, leaf_idx, leaf_parent_pos;
};

struct Treelet {
  Header h;
  std::vector<Node> nodes;      // padded to mem_size / 8
  std::vector<uint32_t> attr;   // normal, colour per node (padded)
};

// Same memory layout as the library's svo_camera (80 bytes).
struct Camera {
  float far_lower_left[4];
  float pixel_h[4];
  float pixel_v[4];
  float position[4];
  float t_scale_ratio;
  uint32_t width, height, view_mode;
};
static_assert(sizeof(Camera) == 80, "camera layout");

}  // namespace

struct oracle_scene {
  uint32_t treelet_bytes = 0, N = 0, max_depth = 0;
  std::vector<Treelet> treelets;
  // incore buffer
  uint32_t num_slots = 0;
  std::vector<Node> incore;
  std::vector<uint32_t> colors, normals;
  std::vector<int32_t> slot_gid;
};

namespace {

bool read_u32(FILE* f, uint32_t* v) { return fread(v, 4, 1, f) == 1; }

int read_bundle(const std::string& base, oracle_scene* s) {
  FILE* f = fopen((base + ".svo").c_str(), "rb");
  if (!f) return 2;
  uint32_t nt = 0;
  if (!read_u32(f, &nt) || !read_u32(f, &s->treelet_bytes) || s->treelet_bytes < 128) { fclose(f); return 3; }
  s->N = s->treelet_bytes / 8;
  s->treelets.resize(nt);
  for (uint32_t t = 0; t < nt; ++t) {
    Treelet& T = s->treelets[t];
    if (fread(&T.h, sizeof(Header), 1, f) != 1 || T.h.num_nodes > s->N) { fclose(f); return 3; }
    T.nodes.assign(s->N, Node{0, 0});
    if (fread(T.nodes.data(), 8, T.h.num_nodes, f) != T.h.num_nodes) { fclose(f); return 3; }
  }
  fclose(f);
  f = fopen((base + ".ia").c_str(), "rb");
  if (!f) return 2;
  uint32_t na = 0;
  if (!read_u32(f, &na) || na < nt) { fclose(f); return 3; }
  for (uint32_t t = 0; t < nt; ++t) {
    Treelet& T = s->treelets[t];
    uint32_t np = 0, w = 0;
    if (!read_u32(f, &np) || !read_u32(f, &w) || w != 2 || np != T.h.num_nodes) { fclose(f); return 3; }
    T.attr.assign(2 * (size_t)s->N, 0u);
    if (fread(T.attr.data(), 4, 2 * (size_t)np, f) != 2 * (size_t)np) { fclose(f); return 3; }
  }
  fclose(f);
  return 0;
}

// Depth of the deepest leaf, walking each treelet from its root.
uint32_t bundle_depth(const oracle_scene& s) {
  std::vector<uint32_t> root_depth(s.treelets.size(), 0);
  std::vector<std::vector<uint32_t>> kids(s.treelets.size());
  for (size_t u = 1; u < s.treelets.size(); ++u)
    if (s.treelets[u].h.parent_gid < u) kids[s.treelets[u].h.parent_gid].push_back((uint32_t)u);
  uint32_t maxd = 0;
  for (size_t t = 0; t < s.treelets.size(); ++t) {
    const Treelet& T = s.treelets[t];
    std::vector<uint32_t> d(T.h.num_nodes, 0);
    d[0] = root_depth[t];
    for (uint32_t i = 0; i < T.h.num_nodes; ++i) {
      uint32_t idx = T.nodes[i].first;
      for (uint32_t c = 0; c < 8; ++c) {
        if (!(T.nodes[i].masks & (1u << (8 + c)))) continue;
        const uint32_t ch = i + idx;
        ++idx;
        if (ch < T.h.num_nodes) {
          d[ch] = d[i] + 1;
          if ((T.nodes[i].masks & (1u << c)) && d[ch] > maxd) maxd = d[ch];
        }
      }
    }
    for (uint32_t u : kids[t])
      if (s.treelets[u].h.leaf_pos < d.size()) root_depth[u] = d[s.treelets[u].h.leaf_pos];
  }
  return maxd;
}

int make_incore(oracle_scene* s, uint64_t incore_bytes) {
  const uint64_t slots = incore_bytes / s->treelet_bytes;
  if (slots < s->treelets.size()) return 7;
  s->num_slots = (uint32_t)slots;
  const uint64_t n = slots * s->N;
  s->incore.assign(n, Node{0, 0});
  s->colors.assign(n, 0u);
  s->normals.assign(n, 0u);
  s->slot_gid.assign(slots, -1);
  std::vector<int32_t> slot_of(s->treelets.size(), -1);
  // the free-slot stack yields 0, 1, 2, ... ; inserting in gid order puts
  // treelet g into slot g
  for (uint32_t g = 0; g < s->treelets.size(); ++g) {
    const Treelet& T = s->treelets[g];
    const uint64_t pos = (uint64_t)g * s->N;
    slot_of[g] = (int32_t)g;
    s->slot_gid[g] = (int32_t)g;
    for (uint32_t i = 0; i < s->N; ++i) {
      s->incore[pos + i] = T.nodes[i];
      s->normals[pos + i] = T.attr[2 * (size_t)i];
      s->colors[pos + i] = T.attr[2 * (size_t)i + 1];
    }
    if (g == 0) continue;
    const uint64_t pslot = (uint64_t)slot_of[T.h.parent_gid];
    const uint64_t leaf = pslot * s->N + T.h.leaf_pos;
    const uint64_t leaf_parent = pslot * s->N + T.h.leaf_parent_pos;
    s->incore[leaf] = s->incore[pos];
    s->incore[leaf_parent].masks &= ~(1u << T.h.leaf_idx);
    s->incore[leaf].first = (uint32_t)((int32_t)(pos + 1) - (int32_t)leaf);
  }
  return 0;
}

// ------------------------------------------------------------ traversal

struct Buffers {
  const Node* nodes;
  const uint32_t* colors;
  const int32_t* slot_gid;
  uint64_t pool;
  uint32_t N;
  uint32_t num_slots;
  uint32_t depth;
};

struct Result {
  bool hit;
  uint32_t node;
  uint32_t depth;
  float t;
  uint32_t loops;
  float center[3];
  uint32_t pushes;
  bool boundary;
  float quality;  // SampleResult.quality (treeletRenderer_rgba.cl:275,329,343)
  // analysis record (sampleAnalyse, opencl/treeletRenderer_rgba.cl:848-875):
  // kind 1 leaf hit, 2 inner node at the LOD / depth limit, 0 none
  uint32_t rec_kind;
  float rec_err;
  uint32_t rec_sub;  // leaf's first-child field (sub-treelet gid), 0 otherwise
};

// A leaf record's sub-treelet gid is the leaf node's own first-child field.
inline void leaf_record(const Buffers& B, uint32_t leaf, float scale_exp2, float ratio, float tc_min, Result* r);

inline float cl_max(float x, float y) { return (x < y) ? y : x; }
inline float cl_min(float x, float y) { return (y < x) ? y : x; }
inline float cl_sign(float x) { return x > 0.0f ? 1.0f : (x < 0.0f ? -1.0f : (x == x ? x : 0.0f)); }
inline void swap_to_increasing(float& a, float& b) {
  if (a > b) { float t = a; a = b; b = t; }
}
inline bool valid_flag(uint32_t masks, uint32_t i) { return (masks & (1u << (i + 8))) != 0; }
inline bool leaf_flag(uint32_t masks, uint32_t i) { return (masks & (1u << i)) != 0; }
inline uint32_t nth_child_idx(uint32_t masks, uint32_t first, uint32_t child) {
  uint32_t index = first;
  for (uint32_t i = 0; i != child; ++i) index += valid_flag(masks, i);
  return index;
}

bool intersect_aabb(const float o[3], const float d[3], float* t_min, float* t_max) {
  float t_near = -10000000.0f, t_far = 10000000.0f;
  const float r = 0.5f;
  for (int i = 0; i != 3; ++i) {
    if (d[i] == 0 && (-r > o[i] || r < o[i])) return false;
    float t1 = (-r - o[i]) / d[i];
    float t2 = (r - o[i]) / d[i];
    if (t1 > t2) { float tmp = t1; t1 = t2; t2 = tmp; }
    if (t1 > t_near) t_near = t1;
    if (t2 < t_far) t_far = t2;
    if (t_near > t_far) return false;
    if (t_far < 0.0f) return false;
  }
  *t_min = t_near;
  *t_max = t_far;
  return true;
}

void pixel_ray(const Camera& c, uint32_t x, uint32_t y, float d[3]) {
  float v[3];
  for (int a = 0; a < 3; ++a) {
    float p = c.far_lower_left[a] + c.pixel_h[a] * (float)x;
    p = p + c.pixel_v[a] * (float)(c.height - y);
    v[a] = p - c.position[a];
  }
  float len = std::sqrt((v[0] * v[0] + v[1] * v[1]) + v[2] * v[2]);
  for (int a = 0; a < 3; ++a) d[a] = v[a] / len;
}

// Boundary rays (this project's definition; the reference has none).  A ray
// is flagged when its result can depend on rounding or on the policy:
//   * a direction component below 1e-4 (the policies clamp differently);
//   * a step enters or leaves a child within 1e-6 (in t) of an edge or corner
//     (the two largest entry-plane t, or the two smallest exit-plane t, tie),
//     or clips a child over less than 1e-6;
//   * a pop test is within 1e-6 of the epsilon threshold, or sees an inverted
//     interval tmin > tmax + epsilon (policy A processes it, policy C pops);
//   * the ray exhausts maxLoops.
constexpr float kBoundary = 1e-6f;

inline bool near_edge(const float t0[3], const float t1[3], float tc_min, float tc_max) {
  if (tc_max - tc_min < kBoundary) return true;
  float a0 = t0[0], a1 = t0[1], a2 = t0[2];
  if (a0 < a1) std::swap(a0, a1);
  if (a1 < a2) std::swap(a1, a2);
  if (a0 < a1) std::swap(a0, a1);
  if (std::fabs(a0 - a1) < kBoundary) return true;
  float c0 = t1[0], c1 = t1[1], c2 = t1[2];
  if (c0 > c1) std::swap(c0, c1);
  if (c1 > c2) std::swap(c1, c2);
  if (c0 > c1) std::swap(c0, c1);
  return std::fabs(c0 - c1) < kBoundary;
}

inline bool near_pop_threshold(float tmin, float tmax, float eps) {
  const float diff = std::fabs(tmin - tmax);
  return std::fabs(diff - eps) < kBoundary || (tmin > tmax && diff > eps);
}

inline bool small_component(const float d[3]) {
  return std::fabs(d[0]) < 1e-4f || std::fabs(d[1]) < 1e-4f || std::fabs(d[2]) < 1e-4f;
}

struct StackElement {
  uint32_t parent_node_index;
  float parent_tmin;
  float parent_tmax;
  float center[3];
};

constexpr int kMaxStack = 24;

// Policy A: treeletRenderer_rgba.cl::sample, scaleMax = max(13, D+1).
inline void leaf_record(const Buffers& B, uint32_t leaf, float scale_exp2, float ratio, float tc_min, Result* r) {
  r->rec_kind = 1;
  r->rec_err = scale_exp2 - (ratio * tc_min);
  r->quality = r->rec_err;
  r->rec_sub = leaf < B.pool ? B.nodes[leaf].first : 0u;
}
inline void stop_record(float scale_exp2, float ratio, float tc_min, Result* r) {
  r->rec_kind = 2;
  r->rec_err = (ratio * tc_min > scale_exp2) ? 1.0f : 0.0f;
  r->rec_sub = 0;
}

void sample_treelet(const Buffers& B, const float o[3], float d[3], float ratio, Result* r) {
  r->rec_kind = 0;
  r->rec_err = 0.0f;
  r->rec_sub = 0;
  r->quality = 0.0f;
  r->hit = false;
  r->loops = 0;
  r->pushes = 0;
  r->boundary = false;
  r->node = 0xFFFFFFFFu;
  r->depth = 0;
  r->t = INFINITY;
  r->center[0] = r->center[1] = r->center[2] = 0.0f;
  float t_min, t_max;
  if (!intersect_aabb(o, d, &t_min, &t_max)) return;

  const int scale_max = (int)std::max(13u, B.depth + 1);
  const int max_depth_cap = scale_max - 1;  // MAX_SVO_RAYCAST_DEPTH
  int scale = scale_max - 1;
  float scale_exp2 = 0.5f;
  const float min_normal = 0.0001f;
  const float epsilon = 0.00001f;
  StackElement stack[kMaxStack];
  stack[scale] = StackElement{0, t_min, t_max, {0.0f, 0.0f, 0.0f}};
  if (small_component(d)) r->boundary = true;
  for (int a = 0; a < 3; ++a)
    if (std::fabs(d[a]) < min_normal) d[a] = min_normal * cl_sign(d[a]);
  const float dx_r = 1.0f / d[0], mpx = -o[0] * dx_r;
  const float dy_r = 1.0f / d[1], mpy = -o[1] * dy_r;
  const float dz_r = 1.0f / d[2], mpz = -o[2] * dz_r;
  unsigned while_counter = 0;
  const unsigned max_loops = (scale_max + 1) * (scale_max + 1) * 5;

  while (scale < scale_max) {
    StackElement* parent = &stack[scale];
    scale_exp2 = std::ldexp(1.0f, scale - scale_max);
    const float child_size_half = scale_exp2 * 0.5f;
    ++while_counter;
    if (while_counter == max_loops) {
      r->boundary = true;
      r->hit = true;
      r->node = parent->parent_node_index;
      r->depth = scale_max - scale;
      r->t = parent->parent_tmin;
      r->loops = while_counter;
      std::memcpy(r->center, parent->center, sizeof r->center);
      r->quality = scale_exp2 - (ratio * parent->parent_tmin);
      return;
    }
    if (near_pop_threshold(parent->parent_tmin, parent->parent_tmax, epsilon)) r->boundary = true;
    if (parent->parent_tmax < 0.0f) { ++scale; continue; }
    if (std::fabs(parent->parent_tmin - parent->parent_tmax) > epsilon) {
      float e[3];
      for (int a = 0; a < 3; ++a) e[a] = (o[a] + (parent->parent_tmin + epsilon) * d[a]) - parent->center[a];
      const int child_idx = 4 * (e[0] > 0.0f) + 2 * (e[1] > 0.0f) + (e[2] > 0.0f);
      float cc[3];
      cc[0] = (-0.5f + (bool)(child_idx & 4)) * scale_exp2 + parent->center[0];
      cc[1] = (-0.5f + (bool)(child_idx & 2)) * scale_exp2 + parent->center[1];
      cc[2] = (-0.5f + (bool)(child_idx & 1)) * scale_exp2 + parent->center[2];
      float tx0 = dx_r * (cc[0] - child_size_half) + mpx;
      float tx1 = dx_r * (cc[0] + child_size_half) + mpx;
      swap_to_increasing(tx0, tx1);
      float ty0 = dy_r * (cc[1] - child_size_half) + mpy;
      float ty1 = dy_r * (cc[1] + child_size_half) + mpy;
      swap_to_increasing(ty0, ty1);
      float tz0 = dz_r * (cc[2] - child_size_half) + mpz;
      float tz1 = dz_r * (cc[2] + child_size_half) + mpz;
      swap_to_increasing(tz0, tz1);
      const float tc_min = cl_max(tx0, cl_max(ty0, tz0));
      const float tc_max = cl_min(tx1, cl_min(ty1, tz1));
      {
        const float t0[3] = {tx0, ty0, tz0}, t1[3] = {tx1, ty1, tz1};
        if (near_edge(t0, t1, tc_min, tc_max)) r->boundary = true;
      }
      const uint64_t pidx = parent->parent_node_index;
      if (pidx >= B.pool) { r->t = t_max; r->loops = while_counter; return; }  // corrupt index: miss
      const Node pn = B.nodes[pidx];
      if (valid_flag(pn.masks, child_idx)) {
        const uint32_t child = parent->parent_node_index + nth_child_idx(pn.masks, pn.first, child_idx);
        if (leaf_flag(pn.masks, child_idx)) {
          r->hit = true;
          r->node = child;
          r->depth = scale_max - scale;
          r->t = parent->parent_tmin;
          r->loops = while_counter;
          std::memcpy(r->center, cc, sizeof cc);
          leaf_record(B, child, scale_exp2, ratio, tc_min, r);
          return;
        }
        if (ratio * tc_min > scale_exp2 || scale_max - scale == max_depth_cap) {
          stop_record(scale_exp2, ratio, tc_min, r);
          r->hit = true;
          r->node = child;
          r->depth = scale_max - scale + 1;
          r->t = parent->parent_tmin;
          r->loops = while_counter;
          std::memcpy(r->center, cc, sizeof cc);
          return;
        }
        parent->parent_tmin = tc_max;
        --scale;
        stack[scale] = StackElement{child, tc_min, tc_max, {cc[0], cc[1], cc[2]}};
        ++r->pushes;
        continue;
      } else {
        parent->parent_tmin = tc_max;
      }
    } else {
      ++scale;
      continue;
    }
  }
  r->hit = false;
  r->node = 0xFFFFFFFFu;
  r->t = t_max;
  r->loops = while_counter;
}

// Policy C: CpuRaycasterSingleRay3::traversSvo, scaleMax = SVO depth.
void travers_cpu3(const Buffers& B, const float o[3], float d[3], float ratio, Result* r) {
  r->rec_kind = 0;
  r->rec_err = 0.0f;
  r->rec_sub = 0;
  r->quality = 0.0f;
  r->hit = false;
  r->loops = 0;
  r->pushes = 0;
  r->boundary = false;
  r->node = 0xFFFFFFFFu;
  r->depth = 0;
  r->t = INFINITY;
  r->center[0] = r->center[1] = r->center[2] = 0.0f;
  float t_min, t_max;
  if (!intersect_aabb(o, d, &t_min, &t_max)) return;
  r->t = t_max;

  const int scale_max = (int)std::max(1u, B.depth);
  int scale = scale_max - 1;
  float scale_exp2 = 0.5f;
  const float epsilon = 0.00001f;
  StackElement stack[kMaxStack];
  stack[scale] = StackElement{0, t_min, t_max, {0.0f, 0.0f, 0.0f}};
  if (small_component(d)) r->boundary = true;
  for (int a = 0; a < 3; ++a)
    if (std::fabs(d[a]) < epsilon) d[a] = epsilon * cl_sign(d[a]) * 10.0f;
  const float dx_r = 1.0f / d[0], mpx = -o[0] * dx_r;
  const float dy_r = 1.0f / d[1], mpy = -o[1] * dy_r;
  const float dz_r = 1.0f / d[2], mpz = -o[2] * dz_r;
  float child_size_half = 0.25f;
  bool refetch_parent = true;
  uint32_t parent_node_index = 0;
  float parent_tmin = 0, parent_tmax = 0, parent_center[3] = {0, 0, 0};
  unsigned while_counter = 0;
  const unsigned max_loops = (unsigned)((B.depth + 1) * (B.depth + 1) * 5.0);

  while (scale < scale_max) {
    ++while_counter;
    if (while_counter == max_loops) {  // "MAX LOOPS REACHED": no hit
      r->boundary = true;
      r->loops = while_counter;
      return;
    }
    if (refetch_parent) {
      parent_node_index = stack[scale].parent_node_index;
      parent_tmin = stack[scale].parent_tmin;
      parent_tmax = stack[scale].parent_tmax;
      std::memcpy(parent_center, stack[scale].center, sizeof parent_center);
      scale_exp2 = (float)std::pow(2.0, scale - scale_max);
      child_size_half = scale_exp2 * 0.5f;
      refetch_parent = false;
    }
    if (near_pop_threshold(parent_tmin, parent_tmax, epsilon)) r->boundary = true;
    if (parent_tmax < 0.0) { ++scale; refetch_parent = true; continue; }
    if (std::fabs(parent_tmin - parent_tmax) < epsilon) { ++scale; refetch_parent = true; continue; }
    if (parent_tmin < parent_tmax) {
      float e[3];
      for (int a = 0; a < 3; ++a) e[a] = (o[a] + (parent_tmin + epsilon) * d[a]) - parent_center[a];
      const int child_index = 4 * (e[0] > 0) + 2 * (e[1] > 0) + (e[2] > 0);
      float cc[3];
      cc[0] = (-0.5f + (bool)(child_index & 4)) * scale_exp2 + parent_center[0];
      cc[1] = (-0.5f + (bool)(child_index & 2)) * scale_exp2 + parent_center[1];
      cc[2] = (-0.5f + (bool)(child_index & 1)) * scale_exp2 + parent_center[2];
      const float x0 = cc[0] - child_size_half, x1 = cc[0] + child_size_half;
      const float y0 = cc[1] - child_size_half, y1 = cc[1] + child_size_half;
      const float z0 = cc[2] - child_size_half, z1 = cc[2] + child_size_half;
      float tx0 = dx_r * x0 + mpx, tx1 = dx_r * x1 + mpx;
      swap_to_increasing(tx0, tx1);
      float ty0 = dy_r * y0 + mpy, ty1 = dy_r * y1 + mpy;
      swap_to_increasing(ty0, ty1);
      float tz0 = dz_r * z0 + mpz, tz1 = dz_r * z1 + mpz;
      swap_to_increasing(tz0, tz1);
      const float tc_min = cl_max(tx0, cl_max(ty0, tz0));
      const float tc_max = cl_min(tx1, cl_min(ty1, tz1));
      {
        const float t0[3] = {tx0, ty0, tz0}, t1[3] = {tx1, ty1, tz1};
        if (near_edge(t0, t1, tc_min, tc_max)) r->boundary = true;
      }
      if (parent_node_index >= B.pool) { r->t = t_max; r->loops = while_counter; return; }
      const Node pn = B.nodes[parent_node_index];
      if (valid_flag(pn.masks, child_index)) {
        const uint32_t child = parent_node_index + nth_child_idx(pn.masks, pn.first, child_index);
        if (leaf_flag(pn.masks, child_index) || ratio * tc_max > scale_exp2 || scale == 0) {
          if (leaf_flag(pn.masks, child_index)) leaf_record(B, child, scale_exp2, ratio, tc_min, r);
          else stop_record(scale_exp2, ratio, tc_min, r);
          r->hit = true;
          r->node = child;
          r->depth = scale_max - scale;
          r->t = parent_tmin;
          r->loops = while_counter;
          std::memcpy(r->center, cc, sizeof cc);
          return;
        }
        stack[scale].parent_tmin = tc_max;
        --scale;
        stack[scale] = StackElement{child, tc_min, tc_max, {cc[0], cc[1], cc[2]}};
        ++r->pushes;
        refetch_parent = true;
        continue;
      } else {
        parent_tmin = tc_max;
      }
    } else {
      ++scale;
      refetch_parent = true;
      continue;
    }
  }
  r->loops = while_counter;  // left the root: miss
}

// Policy B: fillFrameBuffer_rgba.cl::sample (:239-425), the single-SVO
// renderer, over a node array with ABSOLUTE first-child indices:
// scaleMax 15, epsilon pow(2, -16), |d| < epsilon -> epsilon*sign(d)*10,
// getNthChildIndex = first + number of valid children before idx, LOD stop
// on tcMax, maxLoops (16^2*5) ends the loop as a miss.
void sample_fill(const Node* nodes, uint64_t n, const float o[3], float d[3], float ratio, Result* r) {
  r->rec_kind = 0;
  r->rec_err = 0.0f;
  r->rec_sub = 0;
  r->quality = 0.0f;
  r->hit = false;
  r->loops = 0;
  r->pushes = 0;
  r->boundary = false;
  r->node = 0xFFFFFFFFu;
  r->depth = 0;
  r->t = INFINITY;
  r->center[0] = r->center[1] = r->center[2] = 0.0f;
  float t_min, t_max;
  if (!intersect_aabb(o, d, &t_min, &t_max)) return;
  const int scale_max = 15;
  int scale = scale_max - 1;
  float scale_exp2 = 0.5f;
  const float epsilon = std::ldexp(1.0f, -scale_max - 1);
  StackElement stack[kMaxStack];
  stack[scale] = StackElement{0, t_min, t_max, {0.0f, 0.0f, 0.0f}};
  if (small_component(d)) r->boundary = true;
  for (int a = 0; a < 3; ++a)
    if (std::fabs(d[a]) < epsilon) d[a] = epsilon * cl_sign(d[a]) * 10.0f;
  const float dx_r = 1.0f / d[0], mpx = -o[0] * dx_r;
  const float dy_r = 1.0f / d[1], mpy = -o[1] * dy_r;
  const float dz_r = 1.0f / d[2], mpz = -o[2] * dz_r;
  unsigned while_counter = 0;
  const unsigned max_loops = (unsigned)((scale_max + 1) * (scale_max + 1) * 5.0f);
  while (scale < scale_max) {
    ++while_counter;
    if (while_counter == max_loops) { r->boundary = true; break; }
    StackElement* parent = &stack[scale];
    scale_exp2 = std::ldexp(1.0f, scale - scale_max);
    const float child_size_half = scale_exp2 * 0.5f;
    if (near_pop_threshold(parent->parent_tmin, parent->parent_tmax, epsilon)) r->boundary = true;
    if (parent->parent_tmax < 0.0f) { ++scale; continue; }
    if (std::fabs(parent->parent_tmin - parent->parent_tmax) > epsilon) {
      float e[3];
      for (int a = 0; a < 3; ++a) e[a] = (o[a] + (parent->parent_tmin + epsilon) * d[a]) - parent->center[a];
      const int idx = 4 * (e[0] > 0.0f) + 2 * (e[1] > 0.0f) + (e[2] > 0.0f);
      float cc[3];
      cc[0] = (-0.5f + (bool)(idx & 4)) * scale_exp2 + parent->center[0];
      cc[1] = (-0.5f + (bool)(idx & 2)) * scale_exp2 + parent->center[1];
      cc[2] = (-0.5f + (bool)(idx & 1)) * scale_exp2 + parent->center[2];
      float tx0 = dx_r * (cc[0] - child_size_half) + mpx, tx1 = dx_r * (cc[0] + child_size_half) + mpx;
      swap_to_increasing(tx0, tx1);
      float ty0 = dy_r * (cc[1] - child_size_half) + mpy, ty1 = dy_r * (cc[1] + child_size_half) + mpy;
      swap_to_increasing(ty0, ty1);
      float tz0 = dz_r * (cc[2] - child_size_half) + mpz, tz1 = dz_r * (cc[2] + child_size_half) + mpz;
      swap_to_increasing(tz0, tz1);
      const float tc_min = cl_max(tx0, cl_max(ty0, tz0));
      const float tc_max = cl_min(tx1, cl_min(ty1, tz1));
      {
        const float t0[3] = {tx0, ty0, tz0}, t1[3] = {tx1, ty1, tz1};
        if (near_edge(t0, t1, tc_min, tc_max)) r->boundary = true;
      }
      if (parent->parent_node_index >= n) { r->t = t_max; r->loops = while_counter; return; }
      const Node pn = nodes[parent->parent_node_index];
      if (valid_flag(pn.masks, idx)) {
        const uint32_t child = nth_child_idx(pn.masks, pn.first, idx);  // absolute
        if (leaf_flag(pn.masks, idx) || ratio * tc_max > scale_exp2) {
          if (leaf_flag(pn.masks, idx)) r->quality = scale_exp2 - (ratio * tc_min);
          r->hit = true;
          r->node = child;
          r->depth = scale_max - scale;
          r->t = parent->parent_tmin;
          r->loops = while_counter;
          std::memcpy(r->center, cc, sizeof cc);
          return;
        }
        // A push from scale 0 would write stack[-1] (the reference's stack
        // has levels 0..14 and no depth stop): only an inner node below
        // depth 15, i.e. a malformed tree, gets here.  The ray ends as a
        // miss, as for a child index outside the buffer.
        if (scale == 0) { r->t = t_max; r->loops = while_counter; return; }
        parent->parent_tmin = tc_max;
        --scale;
        stack[scale] = StackElement{child, tc_min, tc_max, {cc[0], cc[1], cc[2]}};
        ++r->pushes;
        continue;
      } else {
        parent->parent_tmin = tc_max;
      }
    } else {
      ++scale;
      continue;
    }
  }
  r->hit = false;
  r->node = 0xFFFFFFFFu;
  r->t = t_max;
  r->loops = while_counter;
}

// ------------------------------------------------------------ view modes
// renderToBuffer's modes (treeletRenderer_rgba.cl:644-732), FP32 RGBA,
// restated: per-component vector arithmetic, dot = (x*x + y*y) + z*z,
// normalize = v / sqrt(dot(v, v)), OpenCL max(0, x), smoothstep(0, 1, x),
// the FP64 literals of the phong and colour-map code taken as FP32.
struct V3 { float x, y, z; };
inline void unpack_rgba(uint32_t v, float o[4]) {
  o[0] = (float)((v >> 24) & 255u); o[1] = (float)((v >> 16) & 255u); o[2] = (float)((v >> 8) & 255u); o[3] = (float)(v & 255u);
}
inline V3 get_normal(uint32_t packed) {
  float u[4];
  unpack_rgba(packed, u);
  const float k = 0.003921569f;
  return V3{((u[0] * k) - 0.5f) * 2.0f, ((u[1] * k) - 0.5f) * 2.0f, ((u[2] * k) - 0.5f) * 2.0f};
}
inline float dot3(V3 a, V3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
inline V3 normalize3(V3 v) {
  const float l = std::sqrt(dot3(v, v));
  return V3{v.x / l, v.y / l, v.z / l};
}
inline float smoothstep01(float x) {
  float t = (x - 0.0f) / (1.0f - 0.0f);
  t = t < 0.0f ? 0.0f : (t > 1.0f ? 1.0f : t);
  return t * t * (3.0f - 2.0f * t);
}
inline float adjust(float c, float f) {
  if (c == 0.0f) return 0.0f;
  int r = (int)std::round(255.0f * std::pow(c * f, 0.8f));
  return (float)std::min(255, std::max(0, r));
}
inline V3 data_point_color(float value, float vmin, float vmax) {
  const float w = (value - vmin) / (vmax - vmin) * (780.0f - 380.0f) + 380.0f;
  float r, g, b, f;
  if (380.0f <= w && w <= 440.0f) { r = -(w - 440.0f) / (440.0f - 380.0f); g = 0.0f; b = 1.0f; }
  else if (440.0f < w && w <= 490.0f) { r = 0.0f; g = (w - 440.0f) / (490.0f - 440.0f); b = 1.0f; }
  else if (490.0f < w && w <= 510.0f) { r = 0.0f; g = 1.0f; b = -(w - 510.0f) / (510.0f - 490.0f); }
  else if (510.0f < w && w <= 580.0f) { r = (w - 510.0f) / (580.0f - 510.0f); g = 1.0f; b = 0.0f; }
  else if (580.0f < w && w <= 645.0f) { r = 1.0f; g = -(w - 645.0f) / (645.0f - 580.0f); b = 0.0f; }
  else if (645.0f < w && w <= 780.0f) { r = 1.0f; g = 0.0f; b = 0.0f; }
  else { r = 0.0f; g = 0.0f; b = 0.0f; }
  if (380.0f <= w && w <= 420.0f) f = 0.3f + 0.7f * (w - 380.0f) / (420.0f - 380.0f);
  else if (420.0f < w && w <= 701.0f) f = 1.0f;
  else if (701.0f < w && w <= 780.0f) f = 0.3f + 0.7f * (780.0f - w) / (780.0f - 701.0f);
  else f = 0.0f;
  const float fac = 1.0f / 255.0f;
  return V3{adjust(r, f) * fac, adjust(g, f) * fac, adjust(b, f) * fac};
}

void shade(uint32_t mode, const Result& r, bool hit, const float d[3], uint32_t color, uint32_t normal, float out[4]) {
  out[3] = 1.0f;
  const float k = 0.003921569f;
  switch (mode) {
    case 1:
      if (hit) { const V3 n = get_normal(normal); out[0] = n.x; out[1] = n.y; out[2] = n.z; }
      else { out[0] = 0.2f; out[1] = 0.3f; out[2] = 0.2f; }
      return;
    case 2: {
      const V3 c = data_point_color((float)r.loops, 0.0f, (float)(13 * 30));
      out[0] = c.x; out[1] = c.y; out[2] = c.z;
      return;
    }
    case 3: {
      const float v = (float)r.depth / (float)13, h = hit ? 1.0f : 0.0f;
      out[0] = out[1] = out[2] = v * h;
      return;
    }
    case 4:
      out[0] = out[1] = out[2] = 1.0f - r.t * 0.1f;
      return;
    case 5:
      out[0] = out[1] = out[2] = 1.0f - (float)r.loops / 200.0f;
      return;
    case 6: {
      if (!hit) { out[0] = 0.4f; out[1] = 0.6f; out[2] = 0.4f; return; }
      const V3 n = normalize3(get_normal(normal));
      float u[4];
      unpack_rgba(color, u);
      const float col[3] = {u[0] * k, u[1] * k, u[2] * k};
      const V3 l = normalize3(V3{-120.0f - r.center[0], 0.0f - r.center[1], 0.0f - r.center[2]});
      const float ndl = cl_max(0.0f, dot3(n, l));
      const V3 rf = normalize3(V3{((2.0f * n.x) * ndl) - l.x, ((2.0f * n.y) * ndl) - l.y, ((2.0f * n.z) * ndl) - l.z});
      const float rdv = cl_max(0.0f, dot3(rf, V3{-d[0], -d[1], -d[2]}));
      const float spec = cl_max(0.0f, std::pow(rdv, 10.0f)) * (ndl > 0.0f ? 1.0f : 0.0f);
      const float shadow = 1.0f - 0.0f;
      for (int a = 0; a < 3; ++a)
        out[a] = smoothstep01(col[a] * 0.5f + ((col[a] * ndl) * shadow + (spec * shadow) * 0.2f) * 0.5f);
      return;
    }
    case 7:
      if (hit) out[0] = out[1] = out[2] = std::sqrt(std::sqrt(r.quality));
      else { out[0] = 0.6f; out[1] = 0.3f; out[2] = 0.3f; }
      return;
    case 8: {  // fillFrameBuffer_rgba.cl::shade_svoDepth (:482-489): depth/12 * hit
      const float v = (float)r.depth / 12.0f, h = hit ? 1.0f : 0.0f;
      out[0] = out[1] = out[2] = v * h;
      return;
    }
    case 9:  // fillFrameBuffer_rgba.cl::shade_iterationDepth (:454-461)
      out[0] = out[1] = out[2] = 1.0f - (float)r.loops / 300.0f;
      return;
    default: {
      if (hit) { float u[4]; unpack_rgba(color, u); out[0] = u[0] * k; out[1] = u[1] * k; out[2] = u[2] * k; }
      else { out[0] = 0.2f; out[1] = 0.3f; out[2] = 0.2f; }
      return;
    }
  }
}

// Single-SVO files, read independently of the library (own parser).
struct oracle_single {
  std::vector<Node> nodes;        // absolute first-child indices, as stored
  std::vector<uint32_t> aidx;     // attribute index per node
  std::vector<uint32_t> packages; // {normal, color} per package
  uint32_t depth = 0;
};

int read_single(const std::string& base, oracle_single* s) {
  FILE* f = fopen((base + ".svo").c_str(), "rb");
  if (!f) return 2;
  std::vector<std::string> words;
  bool v1 = false;
  for (int k = 0; k < 3; ++k) {
    std::string w;
    int c;
    while ((c = fgetc(f)) == ' ') {}
    while (c != EOF && c != ' ') { w.push_back((char)c); c = fgetc(f); }
    if (c == EOF) { fclose(f); return 3; }
    words.push_back(w);
    if (k == 0) v1 = (w == "v1");
    if (k == 1 && !v1) break;
  }
  const uint32_t n = (uint32_t)std::strtoul(words.back().c_str(), nullptr, 10);
  s->depth = (uint32_t)std::strtoul(words[words.size() - 2].c_str(), nullptr, 10);
  s->nodes.resize(n);
  s->aidx.resize(n);
  bool ok = fread(s->nodes.data(), 8, n, f) == n;
  if (v1) ok = ok && fread(s->aidx.data(), 4, n, f) == n;
  else for (uint32_t i = 0; i < n; ++i) s->aidx[i] = i;
  fclose(f);
  if (!ok) return 3;
  f = fopen((base + ".ia").c_str(), "rb");
  if (!f) return 2;
  uint32_t blocks = 0, np = 0, w = 0;
  ok = read_u32(f, &blocks) && read_u32(f, &np) && read_u32(f, &w) && blocks == 1 && w == 2;
  if (ok) {
    s->packages.resize(2 * (size_t)np);
    ok = fread(s->packages.data(), 4, s->packages.size(), f) == s->packages.size();
  }
  fclose(f);
  if (!ok) return 3;
  for (uint32_t i = 0; i < n; ++i)
    if (s->aidx[i] >= np) return 3;
  return 0;
}

}  // namespace

// ------------------------------------------------------------ C interface

extern "C" {

int oracle_open(const char* base, oracle_scene** out) {
  if (!base || !out) return 1;
  std::unique_ptr<oracle_scene> s(new oracle_scene());
  int r = read_bundle(base, s.get());
  if (r) return r;
  s->max_depth = bundle_depth(*s);
  *out = s.release();
  return 0;
}

void oracle_free(oracle_scene* s) { delete s; }

int oracle_info(const oracle_scene* s, uint32_t* num_treelets, uint32_t* nodes_per_treelet, uint32_t* max_depth) {
  if (!s) return 1;
  if (num_treelets) *num_treelets = (uint32_t)s->treelets.size();
  if (nodes_per_treelet) *nodes_per_treelet = s->N;
  if (max_depth) *max_depth = s->max_depth;
  return 0;
}

int oracle_make_incore(oracle_scene* s, uint64_t incore_bytes) { return s ? make_incore(s, incore_bytes) : 1; }

// Borrowed pointers to the stitched incore buffer: nodes (2 words each),
// colours, normals, slot table.
int oracle_incore(const oracle_scene* s, const uint32_t** nodes, const uint32_t** colors, const uint32_t** normals,
                  const int32_t** slot_gid, uint64_t* pool_nodes, uint32_t* num_slots) {
  if (!s || s->incore.empty()) return 6;
  if (nodes) *nodes = reinterpret_cast<const uint32_t*>(s->incore.data());
  if (colors) *colors = s->colors.data();
  if (normals) *normals = s->normals.data();
  if (slot_gid) *slot_gid = s->slot_gid.data();
  if (pool_nodes) *pool_nodes = s->incore.size();
  if (num_slots) *num_slots = s->num_slots;
  return 0;
}

// Renders one frame over arbitrary incore buffers.  Any output may be null.
//   t, node, treelet, rgba : as the library's svo_frame_out
//   boundary               : 1 if the ray is a boundary ray
//   pushes                 : number of stack pushes of the ray
//   depth, loops           : SampleResult depth / numWhileLoops
//   center                 : 3 floats, nodeCenter of the hit
//   rec_kind, rec_err,     : analysis record of the ray (sampleAnalyse):
//   rec_sub                  kind 1 leaf / 2 LOD-depth stop / 0 none, its
//                            error, the leaf's sub-treelet gid
int oracle_render_buffers(const uint32_t* nodes, uint64_t pool_nodes, const uint32_t* colors, const int32_t* slot_gid,
                          uint32_t nodes_per_treelet, uint32_t num_slots, uint32_t depth, const void* camera,
                          int policy, float* t, uint32_t* node, uint32_t* treelet, uint32_t* rgba, uint8_t* boundary,
                          uint32_t* pushes, uint32_t* depth_out, uint32_t* loops, float* center, uint8_t* rec_kind,
                          float* rec_err, uint32_t* rec_sub, const uint32_t* normals, float* shade_out,
                          int threads) {
  if (!nodes || !colors || !slot_gid || !camera || nodes_per_treelet == 0) return 1;
  if (shade_out && !normals) return 1;
  if (policy != 0 && policy != 2) return 1;
  Camera cam;
  std::memcpy(&cam, camera, sizeof cam);
  const Buffers B{reinterpret_cast<const Node*>(nodes), colors, slot_gid, pool_nodes, nodes_per_treelet, num_slots, depth};
  if (std::max(13u, depth + 1) >= (uint32_t)kMaxStack) return 1;
  const int W = (int)cam.width, H = (int)cam.height;
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads > 0 ? threads : 1)
  for (int y = 0; y < H; ++y) {
    for (int x = 0; x < W; ++x) {
      float d[3];
      pixel_ray(cam, (uint32_t)x, (uint32_t)y, d);
      const float d0[3] = {d[0], d[1], d[2]};  // the traversal clamps d
      const float o[3] = {cam.position[0], cam.position[1], cam.position[2]};
      Result r;
      if (policy == 0) sample_treelet(B, o, d, cam.t_scale_ratio, &r);
      else travers_cpu3(B, o, d, cam.t_scale_ratio, &r);
      const size_t p = (size_t)y * W + x;
      uint32_t tr = 0xFFFFFFFFu, col = 0x334D33FFu;
      if (r.hit && r.node < pool_nodes) {
        tr = (uint32_t)slot_gid[r.node / nodes_per_treelet];
        col = colors[r.node] | 0xFFu;
      }
      if (t) t[p] = r.t;
      if (node) node[p] = r.hit ? r.node : 0xFFFFFFFFu;
      if (treelet) treelet[p] = tr;
      if (rgba) rgba[p] = col;
      if (boundary) boundary[p] = r.boundary ? 1 : 0;
      if (pushes) pushes[p] = r.pushes;
      if (depth_out) depth_out[p] = r.depth;
      if (loops) loops[p] = r.loops;
      if (center) { center[3 * p] = r.center[0]; center[3 * p + 1] = r.center[1]; center[3 * p + 2] = r.center[2]; }
      if (rec_kind) rec_kind[p] = (uint8_t)r.rec_kind;
      if (rec_err) rec_err[p] = r.rec_err;
      if (rec_sub) rec_sub[p] = r.rec_sub;
      if (shade_out) {
        const bool h = r.hit && r.node < pool_nodes;
        shade(cam.view_mode, r, h, d0, h ? colors[r.node] : 0u, h ? normals[r.node] : 0u, shade_out + 4 * p);
      }
    }
  }
  return 0;
}

int oracle_render(const oracle_scene* s, const void* camera, int policy, float* t, uint32_t* node, uint32_t* treelet,
                  uint32_t* rgba, uint8_t* boundary, uint32_t* pushes, uint32_t* depth_out, uint32_t* loops,
                  float* center, float* shade_out, int threads) {
  if (!s || s->incore.empty()) return 6;
  return oracle_render_buffers(reinterpret_cast<const uint32_t*>(s->incore.data()), s->incore.size(), s->colors.data(),
                               s->slot_gid.data(), s->N, s->num_slots, s->max_depth, camera, policy, t, node, treelet,
                               rgba, boundary, pushes, depth_out, loops, center, nullptr, nullptr, nullptr,
                               s->normals.data(), shade_out, threads);
}

// One ray with explicit origin / direction (direction used as given).
int oracle_trace_ray(const oracle_scene* s, const float origin[3], const float direction[3], float ratio, int policy,
                     uint32_t* node, float* t, uint32_t* depth_out, float center[3], uint8_t* boundary) {
  if (!s || s->incore.empty() || !origin || !direction) return 1;
  const Buffers B{s->incore.data(), s->colors.data(), s->slot_gid.data(), s->incore.size(), s->N, s->num_slots,
                  s->max_depth};
  float d[3] = {direction[0], direction[1], direction[2]};
  Result r;
  if (policy == 0) sample_treelet(B, origin, d, ratio, &r);
  else travers_cpu3(B, origin, d, ratio, &r);
  if (node) *node = r.hit ? r.node : 0xFFFFFFFFu;
  if (t) *t = r.t;
  if (depth_out) *depth_out = r.depth;
  if (center) std::memcpy(center, r.center, sizeof r.center);
  if (boundary) *boundary = r.boundary;
  return 0;
}

// Ray directions of every pixel (3 floats per pixel, row-major).
int oracle_ray_dirs(const void* camera, float* dirs) {
  if (!camera || !dirs) return 1;
  Camera cam;
  std::memcpy(&cam, camera, sizeof cam);
  for (uint32_t y = 0; y < cam.height; ++y)
    for (uint32_t x = 0; x < cam.width; ++x) pixel_ray(cam, x, y, dirs + 3 * ((size_t)y * cam.width + x));
  return 0;
}

int oracle_max_threads(void) { return omp_get_max_threads(); }

// The iteration view mode's colour map (DataPointToColor over
// WavelengthToRGB, treeletRenderer_rgba.cl:518-596) at one value, for tests.
void oracle_data_point_color(float value, float vmin, float vmax, float* rgb) {
  const V3 c = data_point_color(value, vmin, vmax);
  rgb[0] = c.x; rgb[1] = c.y; rgb[2] = c.z;
}

int oracle_open_single(const char* base, oracle_single** out) {
  if (!base || !out) return 1;
  oracle_single* s = new oracle_single();
  const int r = read_single(base, s);
  if (r) { delete s; return r; }
  *out = s;
  return 0;
}

void oracle_free_single(oracle_single* s) { delete s; }

// The single-SVO renderer's shadow (11) and reflection (10) modes,
// fillFrameBuffer_rgba.cl:494-594, written as the reference's two nested
// functions: shade_shadow is shade_diffuse_color_shadow (a policy-B sample
// from nodeCenter + normal*0.01 towards normalize(-0.5, 1, 0.6) with
// tScaleRatio 0.001; shadow = its hit as 0/1), shade_reflection is
// shade_diffuse_color_reflection with numReflections 2, whose samples never
// set attribIndex, so every bounce reads attribute package 0.
void shade_shadow(const oracle_single* s, const Result& r, uint32_t attrib_pos, float out[4]) {
  if (!r.hit) { out[0] = 0.2f; out[1] = 0.3f; out[2] = 0.2f; out[3] = 1.0f; return; }
  const V3 n = get_normal(s->packages[2 * (size_t)attrib_pos]);
  float u[4];
  unpack_rgba(s->packages[2 * (size_t)attrib_pos + 1], u);
  const float k = 0.003921569f;
  const float color[4] = {u[0] * k, u[1] * k, u[2] * k, u[3] * k};
  const V3 light = normalize3(V3{-0.5f, 1.0f, 0.6f});
  const float ndl = cl_max(0.0f, dot3(n, light));
  const float o[3] = {r.center[0] + n.x * 0.01f, r.center[1] + n.y * 0.01f, r.center[2] + n.z * 0.01f};
  float d[3] = {light.x, light.y, light.z};
  Result rs;
  sample_fill(s->nodes.data(), s->nodes.size(), o, d, 0.001f, &rs);
  const float shadow = rs.hit ? 1.0f : 0.0f;
  for (int a = 0; a < 4; ++a) out[a] = color[a] * ndl * (1.0f - shadow) + color[a] * 0.2f;
  out[3] = 1.0f;
}

void shade_reflection(const oracle_single* s, const Result& primary, const float ray_dir[3], float out[4]) {
  if (!primary.hit) { out[0] = 0.2f; out[1] = 0.3f; out[2] = 0.2f; out[3] = 1.0f; return; }
  Result cur = primary;
  float result_color[4] = {0.0f, 0.0f, 0.0f, 0.0f};
  V3 view{ray_dir[0], ray_dir[1], ray_dir[2]};
  int reflect_counter = 2;
  float strength = 0.75f;
  while (true) {
    const V3 normal = get_normal(s->packages[0]);
    float sh[4];
    shade_shadow(s, cur, 0, sh);
    for (int a = 0; a < 4; ++a) result_color[a] = result_color[a] * (1.0f - strength) + sh[a] * strength;
    strength *= 0.75f;
    if (reflect_counter == 0 || !cur.hit) break;
    const float dn = dot3(normal, view);
    const V3 rd = normalize3(V3{view.x + -2.0f * normal.x * dn, view.y + -2.0f * normal.y * dn,
                                view.z + -2.0f * normal.z * dn});
    const float o[3] = {cur.center[0] + normal.x * 0.01f, cur.center[1] + normal.y * 0.01f,
                        cur.center[2] + normal.z * 0.01f};
    float d[3] = {rd.x, rd.y, rd.z};
    sample_fill(s->nodes.data(), s->nodes.size(), o, d, 0.001f, &cur);
    view = rd;
    --reflect_counter;
  }
  for (int a = 0; a < 4; ++a) out[a] = result_color[a];
}

// Single-SVO frame with policy B (fillFrameBuffer_rgba.cl::renderToBuffer
// mode 0): colour = packages[attribIndices[node]] (:640), alpha forced,
// treelet id 0 for a hit (the whole SVO is one treelet).
int oracle_render_single(const oracle_single* s, const void* camera, float* t, uint32_t* node, uint32_t* treelet,
                         uint32_t* rgba, uint8_t* boundary, uint32_t* pushes, float* shade_out, int threads) {
  if (!s || !camera) return 1;
  Camera cam;
  std::memcpy(&cam, camera, sizeof cam);
  const int W = (int)cam.width, H = (int)cam.height;
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads > 0 ? threads : 1)
  for (int y = 0; y < H; ++y) {
    for (int x = 0; x < W; ++x) {
      float d[3];
      pixel_ray(cam, (uint32_t)x, (uint32_t)y, d);
      const float d0[3] = {d[0], d[1], d[2]};
      const float o[3] = {cam.position[0], cam.position[1], cam.position[2]};
      Result r;
      sample_fill(s->nodes.data(), s->nodes.size(), o, d, cam.t_scale_ratio, &r);
      const size_t p = (size_t)y * W + x;
      const bool hit = r.hit && r.node < s->nodes.size();
      if (t) t[p] = r.t;
      if (node) node[p] = hit ? r.node : 0xFFFFFFFFu;
      if (treelet) treelet[p] = hit ? 0u : 0xFFFFFFFFu;
      if (rgba) rgba[p] = hit ? (s->packages[2 * (size_t)s->aidx[r.node] + 1] | 0xFFu) : 0x334D33FFu;
      if (boundary) boundary[p] = r.boundary ? 1 : 0;
      if (pushes) pushes[p] = r.pushes;
      if (shade_out) {
        const size_t a = hit ? 2 * (size_t)s->aidx[r.node] : 0;
        if (cam.view_mode == 11) shade_shadow(s, r, hit ? s->aidx[r.node] : 0u, shade_out + 4 * p);
        else if (cam.view_mode == 10) shade_reflection(s, r, d0, shade_out + 4 * p);
        else shade(cam.view_mode, r, hit, d0, hit ? s->packages[a + 1] : 0u, hit ? s->packages[a] : 0u, shade_out + 4 * p);
      }
    }
  }
  return 0;
}

}  // extern "C"
