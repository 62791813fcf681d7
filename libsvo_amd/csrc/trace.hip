// Serialized-SVO treelet traversal for gfx950.
//
// One ray per lane, one pixel tile per wavefront (tile_w x 64/tile_w), a few
// wavefronts per workgroup; every view of a launch may use its own scene.
// The per-ray stack holds, per level, the node index of the parent left by a
// push and that parent's tmax (8 bytes), in LDS ([level][lane],
// conflict-free) or in private memory, or the parent's child base, tmax and
// 16-bit masks (10 bytes, LDS) so that a pop reads no global memory (the
// launch shape picks; the autotuner measures all three).  The child index
// taken from each parent goes to a 64-bit path register (3 bits per level);
// with the 8-byte layouts the parent node itself is re-read on a pop (a
// cache hit: it was fetched a few iterations earlier).  The stack has a
// level for every scale a push can leave, down to the scene's deepest
// pushable node.
// Everything else about a parent is recomputed bit-exactly instead of stored:
//   * centres are dyadic, so parent = child - (-0.5 + bit) * 2^(s - scaleMax)
//     is exact;
//   * after a pop the parent's tmin equals the popped child's tmax, because
//     the reference sets parentTMin = tcMax right before the push
//     (opencl/treeletRenderer_rgba.cl:347-357,
//     src/incore/CpuRaycasterSingleRay3.cpp:355-363) and nothing changes it
//     until the pop; its tmax is not changed by its own steps either.
// Every floating-point expression follows the reference's operation order
// and is compiled with -ffp-contract=off, so the kernel reproduces the CPU
// restatement bit for bit.
//
// The loop is bound by VALU issue (several waves per SIMD take turns on one
// VALU), so its cost is its dynamic instruction count.  Rays whose
// reciprocal direction is finite on every axis (all but exactly axis-aligned
// ones) and whose origin lies off the planes through 0 run traverse_fast:
// two planes per axis per child step ordered by a min and a max instead of
// compare-and-swap, scales held as the bits of their half-edge, the parent's
// tmax restored from the stack on a pop, cycle checks scheduled on the
// wave's (scalar) trip count, and a single loop exit (exits inside the
// loop make the compiler copy the loop state at every exit).  One trip of
// the fast loop is up to two of the reference's iterations: a pop, if the
// ray pops, then a child step, if the ray does not pop again.  In a wave
// some lanes pop while others step; in one trip, a lane that pops also takes
// its next step instead of idling through the others' step code.  The
// iteration count (loops, maxLoops) still counts every pop and every step.
// A wave with any other ray runs the generic loop.
//
// Tile rows are dispatched bottom-up.  The last waves of a launch run almost
// alone on their SIMDs, where node-fetch latency, not VALU issue, sets their
// pace; in upright views the rows below the horizon hold the long traversals
// and the rows above it the short ones, so bottom-up ends a launch on short
// traversals (1080p terrain frame: 0.350 -> 0.291 ms).  The order changes
// no pixel.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>

#include "trace.h"

namespace svo_impl {
namespace {

constexpr int kMaxScale = 16;  // stack levels; scaleMax <= 16

// OpenCL max/min as specified for floats: max -> y if x < y else x,
// min -> y if y < x else x.
__device__ __forceinline__ float cl_max(float x, float y) { return (x < y) ? y : x; }
__device__ __forceinline__ float cl_min(float x, float y) { return (y < x) ? y : x; }
// OpenCL sign(): 1, -1, signed zero preserved, 0 for NaN.
__device__ __forceinline__ float cl_sign(float x) {
  return x > 0.0f ? 1.0f : (x < 0.0f ? -1.0f : (x == x ? x : 0.0f));
}
// 2^k for k in [-126, 127], exact.
__device__ __forceinline__ float exp2i(int k) { return __int_as_float((127 + k) << 23); }

// Traversal policies (the template parameter POLICY):
//   0  treeletRenderer_rgba.cl::sample (:201-382): epsilon 1e-5, |d| < 1e-4
//      clamped to 1e-4*sign(d), LOD stop on tcMin or the depth cap, maxLoops
//      returns the parent as a hit;
//   1  fillFrameBuffer_rgba.cl::sample (:239-425), the single-SVO renderer:
//      epsilon 2^-16, |d| < epsilon clamped to epsilon*sign(d)*10, LOD stop
//      on tcMax, maxLoops is a miss, absolute child indexing (on a scene
//      held as one treelet in slot 0 this is the same node as relative
//      indexing, which is how the buffer is stored);
//   2  CpuRaycasterSingleRay3::traversSvo (src/incore/CpuRaycasterSingleRay3.
//      cpp:158-387): epsilon 1e-5, |d| < 1e-5 clamped to 1e-5*sign(d)*10, an
//      extra pop unless tmin < tmax, LOD stop on tcMax or scale 0, maxLoops
//      is a miss.
template <int POLICY>
__device__ __forceinline__ float policy_eps() { return POLICY == 1 ? 1.52587890625e-05f : 0.00001f; }
template <int POLICY>
__device__ __forceinline__ void clamp_direction(float d[3]) {
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    if (POLICY == 0) {
      if (fabsf(d[a]) < 0.0001f) d[a] = 0.0001f * cl_sign(d[a]);
    } else {
      const float e = policy_eps<POLICY>();
      if (fabsf(d[a]) < e) d[a] = e * cl_sign(d[a]) * 10.0f;
    }
  }
}
template <int POLICY>
__device__ __forceinline__ bool policy_pop(float tmin, float tmax) {
  const float eps = policy_eps<POLICY>();
  if (POLICY == 2) return (tmax < 0.0f) || (fabsf(tmin - tmax) < eps) || !(tmin < tmax);
  return (tmax < 0.0f) || !(fabsf(tmin - tmax) > eps);
}

// intersectAABB against [-0.5,0.5]^3 (opencl/treeletRenderer_rgba.cl:100-178).
__device__ __forceinline__ bool root_box(const float o[3], const float d[3], float& t_min, float& t_max) {
  float t_near = -10000000.0f, t_far = 10000000.0f;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    if (d[i] == 0.0f && (-0.5f > o[i] || 0.5f < o[i])) return false;
    float t1 = (-0.5f - o[i]) / d[i];
    float t2 = (0.5f - o[i]) / d[i];
    if (t1 > t2) { float tmp = t1; t1 = t2; t2 = tmp; }
    if (t1 > t_near) t_near = t1;
    if (t2 < t_far) t_far = t2;
    if (t_near > t_far) return false;
    if (t_far < 0.0f) return false;
  }
  t_min = t_near;
  t_max = t_far;
  return true;
}

// Ray of pixel (x,y) as in renderToBuffer (treeletRenderer_rgba.cl:627-633):
// no half-pixel offset, y flipped; normalize(v) defined as v / sqrt(v.v).
__device__ __forceinline__ void pixel_ray(const svo_camera& c, uint32_t x, uint32_t y, float d[3]) {
  const float fx = (float)x;
  const float fy = (float)(c.height - y);
  float v[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const float p = (c.far_lower_left[a] + c.pixel_h[a] * fx) + c.pixel_v[a] * fy;
    v[a] = p - c.position[a];
  }
  const float len = sqrtf((v[0] * v[0] + v[1] * v[1]) + v[2] * v[2]);
#pragma unroll
  for (int a = 0; a < 3; ++a) d[a] = v[a] / len;
}

struct Hit {
  uint32_t node;  // 0xFFFFFFFF = miss
  float t;
  uint32_t pushes;
  // Analysis record of sampleAnalyse (treeletRenderer_rgba.cl:848-875):
  // kind 1 = leaf hit, error = scale_exp2 - tScaleRatio*tcMin; kind 2 = inner
  // node at the LOD / depth limit, error = (tScaleRatio*tcMin > scale_exp2);
  // kind 0 = no record (miss or maxLoops).
  uint32_t kind;
  float err;
  // SampleResult fields of the view modes (treeletRenderer_rgba.cl:39-48)
  uint32_t depth;    // scaleMax - scale of the hit (+1 at a policy-A LOD stop), 0 on a miss
  uint32_t loops;    // numWhileLoops
  float center[3];   // nodeCenter: centre of the hit node (the parent's at maxLoops)
  float quality;     // scale_exp2 - tScaleRatio*tcMin at a leaf, 0 at an LOD stop,
                     // scale_exp2 - tScaleRatio*parentTMin at a policy-A maxLoops
};

__device__ __forceinline__ float leaf_error(float se, float ratio, float tc_min) { return se - (ratio * tc_min); }
__device__ __forceinline__ float stop_error(float se, float ratio, float tc_min) { return (ratio * tc_min > se) ? 1.0f : 0.0f; }

// tcMax of the node centred at c with half-edge h (the value computed when
// the node was pushed).
__device__ __forceinline__ float node_tmax(const float c[3], float h, const float rc[3], const float mo[3]) {
  float t1[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const float lo = rc[a] * (c[a] - h) + mo[a];
    const float hi = rc[a] * (c[a] + h) + mo[a];
    t1[a] = (lo > hi) ? lo : hi;  // swapToIncreasing keeps the larger in t1
  }
  return cl_min(t1[0], cl_min(t1[1], t1[2]));
}

// Per-ray stack: per level the node index of the parent left by a push and
// that parent's tmax (the fast loop restores it on a pop instead of
// recomputing it; the generic loop uses the index only).  The generic loop
// addresses a level by its number; the fast loop, which holds a scale as the
// bits of its half-edge (exponent field only, 2^23 apart from one scale to
// the next), by those bits: in the LDS layouts ([level][lane] of 8-byte
// entries, 8 * BLOCK bytes a level) the entry of scale bits `sb` sits at
// byte address sb_base + (sb >> kShift), where bind_scales sets sb_base once
// per ray so that the deepest level (min_push_scale) lands on this lane's
// first entry; the private layout indexes its arrays by the level
// (sb - mps_sb) >> 23, from the deepest level's bits passed in (a scalar
// register: a member would live in scratch beside the arrays).
typedef __attribute__((address_space(3))) uint64_t lds_u64;  // an entry: low word first
typedef __attribute__((address_space(3))) uint16_t lds_u16;
__device__ __forceinline__ uint64_t entry(uint32_t lo, uint32_t hi) { return (uint64_t)lo | ((uint64_t)hi << 32); }
template <int BLOCK>
__host__ __device__ constexpr int lds_level_shift() { return 23 - 3 - (BLOCK == 64 ? 6 : BLOCK == 128 ? 7 : 8); }
template <typename T>
__device__ __forceinline__ uint32_t lds_address(const T* p) {
  return (uint32_t)(size_t)(const __attribute__((address_space(3))) void*)p;
}
template <int STACK, int BLOCK>
struct RayStack;
template <int BLOCK>
struct RayStack<SVO_STACK_LDS, BLOCK> {  // LDS, [level][lane] of 8-byte entries
  static constexpr bool kHoldsNodes = false;
  static constexpr int kShift = lds_level_shift<BLOCK>();
  uint2* base;
  uint32_t tid;
  uint32_t sb_base;
  __device__ __forceinline__ void put(int l, uint32_t v) { base[l * BLOCK + tid].x = v; }
  __device__ __forceinline__ uint32_t get(int l) const { return base[l * BLOCK + tid].x; }
  __device__ __forceinline__ void bind_scales(uint32_t mps_sb) { sb_base = lds_address(base + tid) - (mps_sb >> kShift); }
  __device__ __forceinline__ lds_u64* at(uint32_t sb) const { return (lds_u64*)(size_t)(sb_base + (sb >> kShift)); }
  __device__ __forceinline__ void put2_sb(uint32_t sb, uint32_t, uint32_t v, float t) { *at(sb) = entry(v, __float_as_uint(t)); }
  __device__ __forceinline__ void get2_sb(uint32_t sb, uint32_t, uint32_t& v, float& t) const {
    const uint64_t e = *at(sb);
    v = (uint32_t)e;
    t = __uint_as_float((uint32_t)(e >> 32));
  }
};
// LDS, [level][lane]: 8-byte entries {node word, tmax} and, in a second
// array behind the first, the 16-bit child masks of the node left by a push.
// The fast loop keeps the parent's node itself here (its child base, i.e.
// node index + first-child offset, and its masks), so that a pop reads LDS
// instead of re-reading the parent from global memory; the generic loop
// keeps the node index in the first word, as with SVO_STACK_LDS.
template <int BLOCK>
struct RayStack<SVO_STACK_LDS_NODE, BLOCK> {
  static constexpr bool kHoldsNodes = true;
  static constexpr int kShift = lds_level_shift<BLOCK>();
  uint2* base;
  uint16_t* masks;
  uint32_t tid;
  __device__ __forceinline__ void put(int l, uint32_t v) { base[l * BLOCK + tid].x = v; }
  __device__ __forceinline__ uint32_t get(int l) const { return base[l * BLOCK + tid].x; }
  uint32_t sb_base, sb_mbase;  // the mask array has 2-byte entries: 2 * BLOCK bytes a level
  __device__ __forceinline__ void bind_scales(uint32_t mps_sb) {
    sb_base = lds_address(base + tid) - (mps_sb >> kShift);
    sb_mbase = lds_address(masks + tid) - (mps_sb >> (kShift + 2));
  }
  __device__ __forceinline__ void put3_sb(uint32_t sb, uint32_t cb, uint32_t m, float t) {
    *(lds_u64*)(size_t)(sb_base + (sb >> kShift)) = entry(cb, __float_as_uint(t));
    *(lds_u16*)(size_t)(sb_mbase + (sb >> (kShift + 2))) = (uint16_t)m;
  }
  __device__ __forceinline__ void get3_sb(uint32_t sb, uint32_t& cb, uint32_t& m, float& t) const {
    const uint64_t e = *(const lds_u64*)(size_t)(sb_base + (sb >> kShift));
    cb = (uint32_t)e;
    t = __uint_as_float((uint32_t)(e >> 32));
    m = *(const lds_u16*)(size_t)(sb_mbase + (sb >> (kShift + 2)));
  }
};
template <int BLOCK>
struct RayStack<SVO_STACK_PRIVATE, BLOCK> {  // private (scratch-backed) arrays
  static constexpr bool kHoldsNodes = false;
  uint32_t v[kMaxScale];
  float tm[kMaxScale];
  __device__ __forceinline__ void put(int l, uint32_t x) { v[l] = x; }
  __device__ __forceinline__ uint32_t get(int l) const { return v[l]; }
  __device__ __forceinline__ void bind_scales(uint32_t) {}
  __device__ __forceinline__ void put2_sb(uint32_t sb, uint32_t mps_sb, uint32_t x, float t) {
    const uint32_t l = (sb - mps_sb) >> 23;
    v[l] = x;
    tm[l] = t;
  }
  __device__ __forceinline__ void get2_sb(uint32_t sb, uint32_t mps_sb, uint32_t& x, float& t) const {
    const uint32_t l = (sb - mps_sb) >> 23;
    x = v[l];
    t = tm[l];
  }
};

// POLICY 0: treeletRenderer_rgba.cl::sample (:201-382).
// POLICY 2: CpuRaycasterSingleRay3::traversSvo
//           (src/incore/CpuRaycasterSingleRay3.cpp:158-387).
//
// Stack: level lvl = scale - min_push_scale holds the node index of the
// parent left by a push; the child index taken at each push is kept in
// `path`, 3 bits per level.  A pop re-reads the parent node (a cache hit: it
// was fetched a few iterations earlier).  Every node value therefore comes
// from the one conditional load at the end of an iteration, and its first
// use is the mask test of the next iteration, after the entry-point and plane
// arithmetic that overlaps the fetch.  (Running a lane's pops back to back
// before its next child step was measured slower: it serialises the pop
// chains of a wave.)
template <int POLICY, typename Stack>
__device__ Hit traverse(const TracePools& S, const float o[3], float d[3], float ratio, Stack& stk) {
  Hit r{0xFFFFFFFFu, __int_as_float(0x7F800000), 0u, 0u, 0.0f, 0u, 0u, {0.0f, 0.0f, 0.0f}, 0.0f};
  float t_min, t_max;
  if (!root_box(o, d, t_min, t_max)) return r;
  r.t = t_max;

  const float eps = policy_eps<POLICY>();
  clamp_direction<POLICY>(d);
  float rc[3], mo[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    rc[a] = 1.0f / d[a];
    mo[a] = -o[a] * rc[a];
  }

  const int smax = (int)S.scale_max;
  const int mps = (int)S.min_push_scale;
  int scale = smax - 1;
  float c[3] = {0.0f, 0.0f, 0.0f};
  float tmin = t_min, tmax = t_max;
  uint32_t nidx = 0;  // index of the current node
  uint64_t nd = S.nodes[0];
  uint64_t path = 0;
  uint32_t iter = 0;
  uint32_t pushes = 0;
  // Cycle detection.  The traversal is deterministic and its whole state is
  // (scale, path, tmin, tmax): the path fixes the node and its ancestors, and
  // every other quantity is recomputed from them.  Some rays fall into an
  // exact cycle (an advance that leaves tmin unchanged, or a push/pop loop
  // around a child that is re-entered at the same t) and spin until maxLoops,
  // which in the reference returns the parent as a hit (policy A) or a miss
  // (policies B, C).  A snapshot taken at every power-of-two iteration >= 128
  // is compared with later states; on a repeat with period p the counter
  // jumps whole periods towards maxLoops, so the ray reaches maxLoops in the
  // same state as the reference (pushes are credited per skipped period).
  uint64_t snap_path = 0;
  uint32_t snap_iter = 0, snap_scale = 0, snap_tmin = 0, snap_tmax = 0, snap_pushes = 0;
  // outcome: 0 running, 1 hit (leaf / LOD stop), 2 miss, 3 maxLoops, 4 corrupt
  // offset or push below the stack (reported as a miss)
  uint32_t outcome = 0, hit_node = 0, hit_kind = 0, hit_ci = 0;
  float hit_tcmin = 0.0f;

  do {
    const float se = exp2i(scale - smax);  // scale_exp2
    const float h = se * 0.5f;             // childSizeHalf
    ++iter;
    if (iter >= 128u) {
      if ((iter & (iter - 1u)) == 0u) {
        snap_iter = iter; snap_scale = (uint32_t)scale; snap_path = path;
        snap_tmin = __float_as_uint(tmin); snap_tmax = __float_as_uint(tmax); snap_pushes = pushes;
      } else if (snap_iter != 0u && (uint32_t)scale == snap_scale && path == snap_path &&
                 __float_as_uint(tmin) == snap_tmin && __float_as_uint(tmax) == snap_tmax &&
                 iter < S.max_loops) {
        // the state repeats with period p: jump whole periods towards maxLoops
        const uint32_t p = iter - snap_iter;
        const uint32_t skip = (S.max_loops - iter) / p;
        pushes += skip * (pushes - snap_pushes);
        iter += skip * p;
        snap_iter = 0u;
      }
    }
    if (iter == S.max_loops) {
      outcome = 3;
      break;
    }
    const bool pop = policy_pop<POLICY>(tmin, tmax);

    bool fetch = false;
    uint32_t fidx = 0;
    if (!pop) {
      const float te = tmin + eps;
      const float ex = (o[0] + te * d[0]) - c[0];
      const float ey = (o[1] + te * d[1]) - c[1];
      const float ez = (o[2] + te * d[2]) - c[2];
      const uint32_t ci = (ex > 0.0f ? 4u : 0u) | (ey > 0.0f ? 2u : 0u) | (ez > 0.0f ? 1u : 0u);
      float cc[3];
      cc[0] = (-0.5f + (float)((ci >> 2) & 1u)) * se + c[0];
      cc[1] = (-0.5f + (float)((ci >> 1) & 1u)) * se + c[1];
      cc[2] = (-0.5f + (float)(ci & 1u)) * se + c[2];
      float t0[3], t1[3];
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        const float lo = rc[a] * (cc[a] - h) + mo[a];
        const float hi = rc[a] * (cc[a] + h) + mo[a];
        const bool sw = lo > hi;
        t0[a] = sw ? hi : lo;
        t1[a] = sw ? lo : hi;
      }
      const float tc_min = cl_max(t0[0], cl_max(t0[1], t0[2]));
      const float tc_max = cl_min(t1[0], cl_min(t1[1], t1[2]));

      const uint32_t masks = (uint32_t)(nd >> 32);
      if ((masks >> (8u + ci)) & 1u) {
        const uint32_t child = nidx + (uint32_t)nd + (uint32_t)__popc((masks >> 8) & ((1u << ci) - 1u));
        const bool is_leaf = (masks >> ci) & 1u;
        bool stop;
        if (POLICY == 0) stop = (ratio * tc_min > se) || (S.lod_cap != 0 && (uint32_t)(smax - scale) == S.lod_cap);
        else if (POLICY == 1) stop = ratio * tc_max > se;
        else stop = (ratio * tc_max > se) || scale == 0;
        if (is_leaf || stop) {
          outcome = 1;
          hit_node = child;
          hit_kind = is_leaf ? 1u : 2u;
          hit_ci = ci;
          hit_tcmin = tc_min;
        } else if (child >= S.pool_nodes || scale < mps) {
          // corrupt offset (never read outside the pool), or a push below
          // the stack: an inner node deeper than the scene's deepest
          // pushable node, or (policy 1, which has no depth stop) deeper
          // than its 15 levels, can only come from a malformed tree; the ray
          // ends here, and the stack is never written below level 0
          outcome = 4;
        } else {  // push
          ++pushes;
          stk.put(scale - mps, nidx);
          path = (path << 3) | ci;
          --scale;
          c[0] = cc[0]; c[1] = cc[1]; c[2] = cc[2];
          tmin = tc_min;
          tmax = tc_max;
          fidx = child;
          fetch = true;
        }
      } else {
        tmin = tc_max;  // advance
      }
    } else if (scale + 1 >= smax) {
      outcome = 2;  // left the root: miss, t = root tMax
    } else {
      ++scale;
      const float t_child = tmax;
      fidx = stk.get(scale - mps);
      fetch = true;
      const uint32_t ci = (uint32_t)(path & 7u);
      path >>= 3;
      const float sp = exp2i(scale - smax);
      c[0] = c[0] - (-0.5f + (float)((ci >> 2) & 1u)) * sp;
      c[1] = c[1] - (-0.5f + (float)((ci >> 1) & 1u)) * sp;
      c[2] = c[2] - (-0.5f + (float)(ci & 1u)) * sp;
      tmin = t_child;
      tmax = (scale == smax - 1) ? t_max : node_tmax(c, sp, rc, mo);
    }
    if (fetch) {
      nidx = fidx;
      nd = S.nodes[fidx];
    }
  } while (outcome == 0);

  r.pushes = pushes;
  r.loops = iter;
  const float se = exp2i(scale - smax);
  if (outcome == 1) {
    r.node = hit_node;
    r.t = tmin;
    r.kind = hit_kind;
    r.err = hit_kind == 1u ? leaf_error(se, ratio, hit_tcmin) : stop_error(se, ratio, hit_tcmin);
    r.depth = (uint32_t)(smax - scale) + ((POLICY == 0 && hit_kind == 2u) ? 1u : 0u);
    r.center[0] = (-0.5f + (float)((hit_ci >> 2) & 1u)) * se + c[0];
    r.center[1] = (-0.5f + (float)((hit_ci >> 1) & 1u)) * se + c[1];
    r.center[2] = (-0.5f + (float)(hit_ci & 1u)) * se + c[2];
    r.quality = hit_kind == 1u ? r.err : 0.0f;
  } else if (outcome == 3 && POLICY == 0) {  // maxLoops returns the current parent as a hit
    r.node = nidx;
    r.t = tmin;
    r.depth = (uint32_t)(smax - scale);
    r.center[0] = c[0]; r.center[1] = c[1]; r.center[2] = c[2];
    r.quality = se - (ratio * tmin);
  }
  return r;
}

// Wave-uniform traversal parameters, held in scalar registers.
// Scales are held as the bits of the half-edge 2^(scale - scaleMax) of a
// node at that scale (exponent field 127 + scale - scaleMax, mantissa 0):
// positive floats order as their bits do, so scales compare as these bits.
struct Uniform {
  const __attribute__((address_space(1))) uint64_t* nodes;
  int smax;
  int mps;
  uint32_t max_loops;
  uint32_t pool_nodes;
  uint32_t mps_sb;   // min_push_scale: the deepest level of the stack
  uint32_t lod_sb;   // policy A: the scale whose inner nodes stop the descent (0 = none)
  uint32_t zero_sb;  // scale 0 (policy C stops there)
};
__device__ __forceinline__ uint32_t scale_bits(int scale, int smax) { return (uint32_t)(127 + scale - smax) << 23; }
constexpr uint32_t kRootScaleBits = 0x3F000000u;  // scale scaleMax - 1: the root, half-edge 0.5
constexpr uint32_t kScaleStep = 0x00800000u;      // one scale

__device__ __forceinline__ uint32_t sgpr(uint32_t v) { return (uint32_t)__builtin_amdgcn_readfirstlane((int)v); }

__device__ __forceinline__ Uniform make_uniform(const TracePools& S) {
  Uniform U;
  const uint64_t p = (uint64_t)S.nodes;
  const uint64_t q = ((uint64_t)sgpr((uint32_t)(p >> 32)) << 32) | sgpr((uint32_t)p);
  U.nodes = (const __attribute__((address_space(1))) uint64_t*)q;
  U.smax = (int)sgpr(S.scale_max);
  U.mps = (int)sgpr(S.min_push_scale);
  U.max_loops = sgpr(S.max_loops);
  const uint32_t cap = sgpr(S.lod_cap);
  U.pool_nodes = sgpr(S.pool_nodes);
  U.mps_sb = scale_bits(U.mps, U.smax);
  U.lod_sb = (cap != 0u && cap < 127u) ? scale_bits(U.smax - (int)cap, U.smax) : 0u;
  U.zero_sb = scale_bits(0, U.smax);
  return U;
}

// Traversal for rays whose reciprocal direction is finite on every axis
// (the common case; a direction component that is exactly zero stays zero
// after the reference's clamp, since sign(0) = 0, and takes the generic
// loop).  Same states and FP32 operations as the generic loop, with less
// work per step:
//   * every plane coordinate (node centre plus or minus a power of two) is
//     exact, so t(p) = rc*p + mo is the value the reference computes for
//     that plane however the coordinate was formed;
//   * with rc finite, t(p) is monotone in p (rounding is monotone), so the
//     compare-and-swap of the reference (swapToIncreasing) puts the smaller
//     of the two plane values into t0 and the larger into t1, and a min and
//     a max compute them; where the two rounded values are equal they are
//     the same number: t = -0 needs rc*p = -0 and mo = -o*rc = -0, so an
//     origin coordinate of +-0 (|rc| >= 1 keeps o*rc from underflowing),
//     and rays from such an origin take the generic loop;
//   * a child box on one axis is bounded by the node's midplane and one face
//     of the node, so a step evaluates two planes per axis; a pop evaluates
//     none (the parent's tmax comes back from the stack, its centre from the
//     child's, exactly).
template <int POLICY, typename Stack>
__device__ Hit traverse_fast(const Uniform& U, const float o[3], const float d[3], const float rc[3], const float mo[3],
                             float t_min, float t_max, float ratio, Stack& stk, Hit r) {
  const float eps = policy_eps<POLICY>();

  // the current node's scale as the bits of its half-edge (see Uniform)
  uint32_t sb = kRootScaleBits;
  stk.bind_scales(U.mps_sb);
  float c[3] = {0.0f, 0.0f, 0.0f};
  float tmin = t_min, tmax = t_max;
  uint32_t nidx = 0;
  uint64_t nd = U.nodes[0];
  uint64_t path = 0;
  uint32_t iter = 0;
  uint32_t trip = 0;  // loop trips of the wave (uniform); schedules the cycle checks
  uint32_t next_check = 128u;
  uint32_t pushes = 0;
  uint64_t snap_path = 0;
  uint32_t snap_iter = 0, snap_scale = 0, snap_tmin = 0, snap_tmax = 0, snap_pushes = 0;
  // outcome: 0 running, 1 hit at `hit_node` with t = tmin, 2 miss (t = root tMax),
  // 3 maxLoops, 4 corrupt offset
  uint32_t outcome = 0;
  uint32_t hit_node = 0;
  uint32_t hit_kind = 0;
  uint32_t hit_ci = 0;
  float hit_err = 0.0f;

  do {
    ++trip;
    // snapshots at power-of-two trips >= 128, comparisons every 8th trip: a
    // repeat found at any multiple of the period is a period too
    if (trip == next_check) {
      next_check += 8u;
      if ((trip & (trip - 1u)) == 0u) {
        snap_iter = iter; snap_scale = sb; snap_path = path;
        snap_tmin = __float_as_uint(tmin); snap_tmax = __float_as_uint(tmax); snap_pushes = pushes;
      } else if (snap_iter != 0u && sb == snap_scale && path == snap_path &&
                 __float_as_uint(tmin) == snap_tmin && __float_as_uint(tmax) == snap_tmax &&
                 iter + 2u < U.max_loops) {
        const uint32_t p = iter - snap_iter;
        const uint32_t skip = (U.max_loops - 2u - iter) / p;
        pushes += skip * (pushes - snap_pushes);
        iter += skip * p;
        snap_iter = 0u;
      }
    }
    // first half of a trip: one reference iteration if it pops
    bool pops = policy_pop<POLICY>(tmin, tmax);
    if (pops) {
      ++iter;
      if (iter >= U.max_loops) {
        outcome = 3;
      } else if (sb >= kRootScaleBits) {
        outcome = 2;  // left the root
      } else {
        const uint32_t seb = sb;  // half-edge of the node left
        sb += kScaleStep;
        float tparent;
        uint32_t fidx;
        uint32_t pmask = 0;
        if constexpr (Stack::kHoldsNodes) stk.get3_sb(sb, fidx, pmask, tparent);
        else stk.get2_sb(sb, U.mps_sb, fidx, tparent);
        const uint32_t pl = (uint32_t)path;
        path >>= 3;
        // parent centre = child centre -/+ se (bit set: the child was the upper half)
        c[0] = c[0] + __uint_as_float(seb | ((pl << 29) & 0x80000000u));
        c[1] = c[1] + __uint_as_float(seb | ((pl << 30) & 0x80000000u));
        c[2] = c[2] + __uint_as_float(seb | (pl << 31));
        tmin = tmax;
        tmax = tparent;
        nidx = fidx;
        if constexpr (Stack::kHoldsNodes) {
          // the parent's child base and masks come back from the stack: no
          // global load on a pop (nidx now holds the child base, nd the
          // masks over a zero offset)
          nd = (uint64_t)pmask << 32;
        } else {
          nd = U.nodes[fidx];
        }
        // only a lane that popped has a new interval to test again (the
        // other lanes' test above still holds): the step half's test costs
        // nothing on trips where no lane of the wave pops
        pops = policy_pop<POLICY>(tmin, tmax);
      }
    }
    // second half: the next reference iteration if it steps into a child
    if (outcome == 0 && !pops) {
      ++iter;
      if (iter >= U.max_loops) {
        outcome = 3;
      } else {
        const float se = __uint_as_float(sb);  // half-edge of the current node
        const float te = tmin + eps;
        bool up[3];
        up[0] = ((o[0] + te * d[0]) - c[0]) > 0.0f;
        up[1] = ((o[1] + te * d[1]) - c[1]) > 0.0f;
        up[2] = ((o[2] + te * d[2]) - c[2]) > 0.0f;
        const uint32_t ci = (up[0] ? 4u : 0u) | (up[1] ? 2u : 0u) | (up[2] ? 1u : 0u);
        float t0[3], t1[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
          const float tmid = rc[a] * c[a] + mo[a];
          const float tface = rc[a] * (c[a] + (up[a] ? se : -se)) + mo[a];
          t0[a] = fminf(tface, tmid);  // the near plane has the smaller t
          t1[a] = fmaxf(tface, tmid);
        }
        // IEEE max / min equal OpenCL's max / min here: no NaN (finite rc
        // and planes), and no -0 (that needs the origin on a plane through
        // 0, which takes the generic loop), so ties are equal bit patterns
        float tc_min = fmaxf(t0[0], fmaxf(t0[1], t0[2]));
        const float tc_max = fminf(t1[0], fminf(t1[1], t1[2]));
        // keep tc_min next to the plane arithmetic: computed after the mask
        // test its inputs would come from another block, and the compiler
        // quiets such inputs before a max
        asm("" : "+v"(tc_min));
        const uint32_t masks = (uint32_t)(nd >> 32);
        if ((masks >> (8u + ci)) & 1u) {
          const uint32_t child = nidx + (uint32_t)nd + (uint32_t)__popc((masks >> 8) & ((1u << ci) - 1u));  // relative
          bool stop;
          if (POLICY == 0) stop = (ratio * tc_min > se) || sb == U.lod_sb;
          else if (POLICY == 1) stop = ratio * tc_max > se;
          else stop = (ratio * tc_max > se) || sb == U.zero_sb;
          if (((masks >> ci) & 1u) || stop) {  // leaf, or inner node at the LOD / depth limit
            outcome = 1;
            hit_node = child;
            const bool is_leaf = (masks >> ci) & 1u;
            hit_kind = is_leaf ? 1u : 2u;
            hit_ci = ci;
            hit_err = is_leaf ? leaf_error(se, ratio, tc_min) : stop_error(se, ratio, tc_min);
          } else if (child >= U.pool_nodes || sb < U.mps_sb) {
            outcome = 4;  // corrupt offset, or a push below the stack
          } else {
            ++pushes;
            if constexpr (Stack::kHoldsNodes) stk.put3_sb(sb, nidx + (uint32_t)nd, masks, tmax);
            else stk.put2_sb(sb, U.mps_sb, nidx, tmax);
            path = (path << 3) | ci;
            sb -= kScaleStep;
            const float h = __uint_as_float(sb);  // the child's half-edge, se / 2
#pragma unroll
            for (int a = 0; a < 3; ++a) c[a] = c[a] + (up[a] ? h : -h);
            tmin = tc_min;
            tmax = tc_max;
            nidx = child;
            nd = U.nodes[child];
          }
        } else {
          tmin = tc_max;  // advance
        }
      }
    }
  } while (outcome == 0);

  r.pushes = pushes;
  r.loops = iter;
  const int scale = (int)(sb >> 23) - 127 + U.smax;
  const float se = __uint_as_float(sb);
  if (outcome == 1) {
    r.node = hit_node;
    r.t = tmin;
    r.kind = hit_kind;
    r.err = hit_err;
    r.depth = (uint32_t)(U.smax - scale) + ((POLICY == 0 && hit_kind == 2u) ? 1u : 0u);
    const float h = se * 0.5f;
#pragma unroll
    for (int a = 0; a < 3; ++a) r.center[a] = c[a] + (((hit_ci >> (2 - a)) & 1u) ? h : -h);
    r.quality = hit_kind == 1u ? hit_err : 0.0f;
  } else if (outcome == 3) {
    if (POLICY == 0) {  // maxLoops returns the current parent as a hit
      if constexpr (Stack::kHoldsNodes) {
        // after a pop nidx holds the node's child base, not its index: walk
        // the path from the root to find the index (maxLoops rays only)
        uint32_t n = 0;
        for (int l = U.smax - 2 - scale; l >= 0; --l) {
          const uint64_t v = U.nodes[n];
          const uint32_t ci = (uint32_t)(path >> (3 * l)) & 7u;
          n = n + (uint32_t)v + (uint32_t)__popc(((uint32_t)(v >> 32) >> 8) & ((1u << ci) - 1u));
        }
        nidx = n;
      }
      r.node = nidx;
      r.t = tmin;
      r.depth = (uint32_t)(U.smax - scale);
      r.center[0] = c[0]; r.center[1] = c[1]; r.center[2] = c[2];
      r.quality = se - (ratio * tmin);
    }
  }
  return r;
}

// Root-box test, direction clamp and reciprocals, then the fast loop when
// every active lane of the wave has a finite reciprocal direction and an
// origin off the planes through 0, else the generic loop.
template <int POLICY, typename Stack>
__device__ Hit trace_ray(const TracePools& S, const float o[3], float d[3], float ratio, Stack& stk) {
  Hit r{0xFFFFFFFFu, __int_as_float(0x7F800000), 0u, 0u, 0.0f, 0u, 0u, {0.0f, 0.0f, 0.0f}, 0.0f};
  float t_min, t_max;
  if (!root_box(o, d, t_min, t_max)) return r;
  float dc[3] = {d[0], d[1], d[2]};
  clamp_direction<POLICY>(dc);
  float rc[3], mo[3];
  bool finite = true;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    rc[a] = 1.0f / dc[a];
    mo[a] = -o[a] * rc[a];
    finite = finite && (fabsf(rc[a]) < __int_as_float(0x7F800000)) && (fabsf(mo[a]) < __int_as_float(0x7F800000)) &&
             o[a] != 0.0f;
  }
  if (__builtin_amdgcn_ballot_w64(!finite) == 0) {
    const Uniform U = make_uniform(S);
    r.t = t_max;
    return traverse_fast<POLICY>(U, o, dc, rc, mo, t_min, t_max, ratio, stk, r);
  }
  return traverse<POLICY>(S, o, d, ratio, stk);
}

// Order-preserving map of a float onto unsigned (larger float -> larger key).
__device__ __forceinline__ uint32_t float_key(float f) {
  const uint32_t b = __float_as_uint(f);
  return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}

// RenderPassAnalyse::performAnalysePass (src/RenderPassAnalyse.cpp:248-304),
// first half, on the device: a record marks the treelet holding its node as
// visible with the largest error seen, and a leaf whose first-child field
// holds a sub-treelet gid asks for that treelet (largest error wins; the
// first record of a cycle appends the gid to the queue).
__device__ __forceinline__ void record_feedback(const TracePools& S, const Hit& hit) {
  const uint32_t stamp = S.fb_state[1];  // the current cycle (advanced by feedback_export_kernel)
  const unsigned long long key = ((unsigned long long)stamp << 32) | float_key(hit.err);
  const uint32_t slot = (S.slot_shift < 32) ? (hit.node >> S.slot_shift) : (hit.node / S.nodes_per_treelet);
  atomicMax(&S.fb_slot[slot], key);
  if (hit.kind == 1u) {
    const uint32_t sub = (uint32_t)S.nodes[hit.node];  // a leaf's first child: its sub-treelet gid, 0 if none
    if (sub != 0u && sub < S.num_gids) {
      const unsigned long long old = atomicMax(&S.fb_gid[sub], key);
      if ((uint32_t)(old >> 32) != stamp) {
        const uint32_t i = atomicAdd(&S.fb_state[0], 1u);
        if (i < S.fb_cap) S.fb_queue[i] = sub;
      }
    }
  }
}

// ------------------------------------------------------------ view modes
// renderToBuffer's modes (treeletRenderer_rgba.cl:644-732) as FP32 RGBA.
// OpenCL vector arithmetic is per component with the same rounding; dot()
// is (x*x + y*y) + z*z and normalize(v) is v / sqrt(dot(v, v)), as in the
// ray generation; max(0, x) is OpenCL's (x if 0 < x, else 0).
__device__ __forceinline__ float4 unpack_rgba(uint32_t v) {  // unpackRgbaFromUnsigned (:183-195)
  return make_float4((float)((v >> 24) & 255u), (float)((v >> 16) & 255u), (float)((v >> 8) & 255u), (float)(v & 255u));
}
__device__ __forceinline__ float3 get_normal(uint32_t packed) {  // getNormal (:389-393)
  const float4 u = unpack_rgba(packed);
  const float k = 0.003921569f;
  return make_float3(((u.x * k) - 0.5f) * 2.0f, ((u.y * k) - 0.5f) * 2.0f, ((u.z * k) - 0.5f) * 2.0f);
}
__device__ __forceinline__ float4 get_color(uint32_t packed) {  // getColor (:399-404)
  const float4 u = unpack_rgba(packed);
  const float k = 0.003921569f;
  return make_float4(u.x * k, u.y * k, u.z * k, u.w * k);
}
__device__ __forceinline__ float dot3(float3 a, float3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
__device__ __forceinline__ float3 normalize3(float3 v) {
  const float l = sqrtf(dot3(v, v));
  return make_float3(v.x / l, v.y / l, v.z / l);
}
__device__ __forceinline__ float smoothstep01(float x) {  // OpenCL smoothstep(0, 1, x)
  const float t = fminf(fmaxf((x - 0.0f) / (1.0f - 0.0f), 0.0f), 1.0f);
  return t * t * (3.0f - 2.0f * t);
}
// Adjust (:518-529): round(255 * pow(c*f, 0.8)) clamped to [0, 255]
__device__ __forceinline__ float adjust(float c, float f) {
  if (c == 0.0f) return 0.0f;
  const int r = (int)roundf(255.0f * powf(c * f, 0.8f));
  return (float)min(255, max(0, r));
}
// DataPointToColor (:580-596) over WavelengthToRGB (:533-576)
__device__ float3 data_point_color(float value, float vmin, float vmax) {
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
  return make_float3(adjust(r, f) * fac, adjust(g, f) * fac, adjust(b, f) * fac);
}

__device__ float4 shade_pixel(uint32_t mode, const Hit& hit, const float d[3], const TracePools& S) {
  const bool hit_ok = hit.node != 0xFFFFFFFFu && hit.node < S.pool_nodes;
  switch (mode) {
    case SVO_VIEW_NORMAL:
      if (hit_ok) {
        const float3 n = get_normal(S.normals[hit.node]);
        return make_float4(n.x, n.y, n.z, 1.0f);
      }
      return make_float4(0.2f, 0.3f, 0.2f, 1.0f);
    case SVO_VIEW_ITERATIONS: {
      const float3 c = data_point_color((float)hit.loops, 0.0f, (float)(13 * 30));
      return make_float4(c.x, c.y, c.z, 1.0f);
    }
    case SVO_VIEW_SVO_DEPTH: {
      const float v = (float)hit.depth / (float)13;
      const float h = hit_ok ? 1.0f : 0.0f;
      return make_float4(v * h, v * h, v * h, 1.0f);
    }
    case SVO_VIEW_T: {
      const float v = 1.0f - hit.t * 0.1f;
      return make_float4(v, v, v, 1.0f);
    }
    case SVO_VIEW_LOOPS: {
      const float v = 1.0f - (float)hit.loops / 200.0f;
      return make_float4(v, v, v, 1.0f);
    }
    case SVO_VIEW_PHONG: {
      if (!hit_ok) return make_float4(0.4f, 0.6f, 0.4f, 1.0f);
      const float3 n = normalize3(get_normal(S.normals[hit.node]));
      float4 col = get_color(S.colors[hit.node]);
      const float3 l = normalize3(make_float3(-120.0f - hit.center[0], 0.0f - hit.center[1], 0.0f - hit.center[2]));
      const float ndl = cl_max(0.0f, dot3(n, l));
      const float3 rf = normalize3(make_float3(((2.0f * n.x) * ndl) - l.x, ((2.0f * n.y) * ndl) - l.y,
                                               ((2.0f * n.z) * ndl) - l.z));
      const float rdv = cl_max(0.0f, dot3(rf, make_float3(-d[0], -d[1], -d[2])));
      const float spec = cl_max(0.0f, powf(rdv, 10.0f)) * (ndl > 0.0f ? 1.0f : 0.0f);
      const float shadow = 1.0f - 0.0f;
      float4 o;
      o.x = col.x * 0.5f + ((col.x * ndl) * shadow + (spec * shadow) * 0.2f) * 0.5f;
      o.y = col.y * 0.5f + ((col.y * ndl) * shadow + (spec * shadow) * 0.2f) * 0.5f;
      o.z = col.z * 0.5f + ((col.z * ndl) * shadow + (spec * shadow) * 0.2f) * 0.5f;
      return make_float4(smoothstep01(o.x), smoothstep01(o.y), smoothstep01(o.z), 1.0f);
    }
    case SVO_VIEW_QUALITY:
      if (hit_ok) {
        const float q = sqrtf(sqrtf(hit.quality));
        return make_float4(q, q, q, 1.0f);
      }
      return make_float4(0.6f, 0.3f, 0.3f, 1.0f);
    case SVO_VIEW_FILL_SVO_DEPTH: {  // fillFrameBuffer_rgba.cl::shade_svoDepth (:482-489)
      const float v = (float)hit.depth / 12.0f;
      const float h = hit_ok ? 1.0f : 0.0f;
      return make_float4(v * h, v * h, v * h, 1.0f);
    }
    case SVO_VIEW_FILL_LOOPS: {  // fillFrameBuffer_rgba.cl::shade_iterationDepth (:454-461)
      const float v = 1.0f - (float)hit.loops / 300.0f;
      return make_float4(v, v, v, 1.0f);
    }
    default:  // SVO_VIEW_COLOR
      if (hit_ok) {
        const float4 c = get_color(S.colors[hit.node]);
        return make_float4(c.x, c.y, c.z, 1.0f);
      }
      return make_float4(0.2f, 0.3f, 0.2f, 1.0f);
  }
}

// The single-SVO renderer's shadow and reflection modes
// (fillFrameBuffer_rgba.cl:494-594), from the primary hit's node and centre:
// every secondary ray is a policy-B sample() with tScaleRatio 0.001 from
// nodeCenter + 0.01 * normal.  The shadow shading of a hit is
// colour * nDotL * (1 - shadow) + colour * 0.2 (alpha 1); reflection blends
// the shadow shadings of the hit and of up to two reflection hits with
// strengths 0.75, 0.75^2, 0.75^3 and shades every bounce with attribute
// package 0 (the reference's reflection mode never looks the attribute
// index up).  One loop, one traversal per trip: the shadow ray of the
// current hit, then (reflection) the ray to the next hit.
template <typename Stack>
__device__ float4 fill_secondary(uint32_t mode, const Hit& prim, const float d[3], const TracePools& S, Stack& stk) {
  const float4 bg = make_float4(0.2f, 0.3f, 0.2f, 1.0f);
  bool hit = prim.node != 0xFFFFFFFFu && prim.node < S.pool_nodes;
  if (!hit) return bg;
  const bool reflect = mode == SVO_VIEW_FILL_REFLECTION;
  const float3 n = get_normal(reflect ? S.package0_normal : S.normals[prim.node]);
  const float4 col = get_color(reflect ? S.package0_color : S.colors[prim.node]);
  const float3 light = normalize3(make_float3(-0.5f, 1.0f, 0.6f));
  float c[3] = {prim.center[0], prim.center[1], prim.center[2]};
  float3 view = make_float3(d[0], d[1], d[2]);
  float4 acc = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  float strength = 0.75f;
  uint32_t bounces_left = 2;
  bool shadow_ray = true;
  while (true) {
    // origin nodeCenter + normal * 0.01f, direction towards the light or
    // reflected about the normal
    const float o[3] = {c[0] + (n.x * 0.01f), c[1] + (n.y * 0.01f), c[2] + (n.z * 0.01f)};
    float3 dir = light;
    if (!shadow_ray) {
      const float k = dot3(n, view);
      dir = normalize3(make_float3(view.x + ((-2.0f * n.x) * k), view.y + ((-2.0f * n.y) * k),
                                   view.z + ((-2.0f * n.z) * k)));
    }
    float dd[3] = {dir.x, dir.y, dir.z};
    const Hit h = trace_ray<1>(S, o, dd, 0.001f, stk);
    const bool h_hit = h.node != 0xFFFFFFFFu && h.node < S.pool_nodes;
    if (shadow_ray) {
      const float ndl = cl_max(0.0f, dot3(n, light));
      const float shadow = h_hit ? 1.0f : 0.0f;
      const float4 sh = make_float4(((col.x * ndl) * (1.0f - shadow)) + (col.x * 0.2f),
                                    ((col.y * ndl) * (1.0f - shadow)) + (col.y * 0.2f),
                                    ((col.z * ndl) * (1.0f - shadow)) + (col.z * 0.2f), 1.0f);
      if (!reflect) return sh;
      acc = make_float4((acc.x * (1.0f - strength)) + (sh.x * strength), (acc.y * (1.0f - strength)) + (sh.y * strength),
                        (acc.z * (1.0f - strength)) + (sh.z * strength), (acc.w * (1.0f - strength)) + (sh.w * strength));
      strength *= 0.75f;
      if (bounces_left == 0) return acc;
      shadow_ray = false;
    } else {
      if (!h_hit) {  // the reflection ray left the scene: the background's turn
        return make_float4((acc.x * (1.0f - strength)) + (bg.x * strength),
                           (acc.y * (1.0f - strength)) + (bg.y * strength),
                           (acc.z * (1.0f - strength)) + (bg.z * strength),
                           (acc.w * (1.0f - strength)) + (bg.w * strength));
      }
      c[0] = h.center[0]; c[1] = h.center[1]; c[2] = h.center[2];
      view = dir;
      --bounces_left;
      shadow_ray = true;
    }
  }
}

template <int POLICY, int TW, int WAVES, int STACK>
// Eight waves per SIMD: the register budget is held to 64 VGPRs (a few
// spills outside the inner loop), and a 10-level LDS stack (5 KiB a wave)
// lets 32 waves share a CU's 160 KiB.
__global__ __launch_bounds__(64 * WAVES) __attribute__((amdgpu_waves_per_eu(8, 8)))
void trace_kernel(const TraceFrame* __restrict__ frames, uint32_t num_frames, uint32_t num_tiles,
                  uint32_t stack_levels) {
  constexpr int kBlock = 64 * WAVES;
  constexpr uint32_t TH = 64u / TW;
  extern __shared__ uint32_t s_stack[];
  const uint32_t tid = threadIdx.x;
  const uint32_t lane = tid & 63u;
  const uint32_t wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  // Workgroups are dispatched round-robin over the 8 XCDs (b -> XCD b % 8),
  // each with its own L2.  Renumber them so that every XCD gets runs of
  // kXcdRun neighbouring tiles, whose rays share the upper levels of the
  // tree, instead of every 8th tile (a bijection on each full group of
  // 8 * kXcdRun workgroups; the last partial group keeps its order).
  constexpr uint32_t kXcdRun = 32u;
  uint32_t wg = blockIdx.x;
  if (wg < gridDim.x / (8u * kXcdRun) * (8u * kXcdRun)) {
    const uint32_t g = wg / (8u * kXcdRun), r = wg % (8u * kXcdRun);
    wg = g * (8u * kXcdRun) + (r % 8u) * kXcdRun + r / 8u;
  }
  const uint32_t tile = wg * WAVES + wave;
  if (tile >= num_tiles) return;

  uint32_t lo = 0, hi = num_frames;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (frames[mid].tile_begin <= tile) lo = mid;
    else hi = mid;
  }
  const TraceFrame& F = frames[lo];
  const uint32_t lt = tile - F.tile_begin;
  const uint32_t tiles_y = (F.cam.height + TH - 1u) / TH;
  const uint32_t x = (lt % F.tiles_x) * TW + (lane % TW);
  const uint32_t y = (tiles_y - 1u - lt / F.tiles_x) * TH + (lane / TW);
  if (x >= F.cam.width || y >= F.cam.height) return;
  const TracePools& S = F.pools;

  const float o[3] = {F.cam.position[0], F.cam.position[1], F.cam.position[2]};
  float d[3];
  pixel_ray(F.cam, x, y, d);
  RayStack<STACK, kBlock> stk;
  if constexpr (STACK == SVO_STACK_LDS) {
    stk.base = reinterpret_cast<uint2*>(s_stack);
    stk.tid = tid;
  } else if constexpr (STACK == SVO_STACK_LDS_NODE) {
    stk.base = reinterpret_cast<uint2*>(s_stack);
    stk.masks = reinterpret_cast<uint16_t*>(s_stack + 2u * stack_levels * kBlock);
    stk.tid = tid;
  }
  float dt[3] = {d[0], d[1], d[2]};  // the traversal may clamp its copy
  const Hit hit = trace_ray<POLICY>(S, o, dt, F.cam.t_scale_ratio, stk);

  const size_t pix = (size_t)y * F.cam.width + x;
  uint32_t tree = 0xFFFFFFFFu, rgba = 0x334D33FFu;
  if (hit.node != 0xFFFFFFFFu && hit.node < S.pool_nodes) {
    const uint32_t slot = (S.slot_shift < 32) ? (hit.node >> S.slot_shift) : (hit.node / S.nodes_per_treelet);
    tree = (uint32_t)S.slot_gid[slot];
    rgba = S.colors[hit.node] | 0xFFu;
  }
  if (F.fb_stride != 0u && hit.kind != 0u && hit.node < S.pool_nodes && x % F.fb_stride == 0u &&
      y % F.fb_stride == 0u)
    record_feedback(S, hit);
  if (F.out.t) F.out.t[pix] = hit.t;
  if (F.out.node) F.out.node[pix] = hit.node;
  if (F.out.treelet) F.out.treelet[pix] = tree;
  if (F.out.rgba) F.out.rgba[pix] = rgba;
  if (F.out.pushes) F.out.pushes[pix] = hit.pushes;
  if (F.out.shade) {
    // raw hit data for shade_kernel, which turns it into the view mode's
    // colour in place (keeps the shading out of this kernel's registers)
    uint4 raw;
    raw.x = hit.node;
    if (F.cam.view_mode == SVO_VIEW_PHONG || F.cam.view_mode == SVO_VIEW_FILL_SHADOW ||
        F.cam.view_mode == SVO_VIEW_FILL_REFLECTION) {
      raw.y = __float_as_uint(hit.center[0]);
      raw.z = __float_as_uint(hit.center[1]);
      raw.w = __float_as_uint(hit.center[2]);
    } else {
      raw.y = __float_as_uint(hit.t);
      raw.z = hit.depth | (hit.loops << 8);
      raw.w = __float_as_uint(hit.quality);
    }
    reinterpret_cast<uint4*>(F.out.shade)[pix] = raw;
  }
}

// Second pass of a launch with view-mode outputs: one lane per pixel of each
// view whose shade output is set, tiles numbered as in trace_kernel.
template <int TW, bool SECONDARY>
__global__ __launch_bounds__(64) void shade_kernel(const TraceFrame* __restrict__ frames, uint32_t num_frames,
                                                   uint32_t num_tiles, uint32_t stack_levels) {
  constexpr uint32_t TH = 64u / TW;
  const uint32_t tile = blockIdx.x;
  const uint32_t lane = threadIdx.x;
  if (tile >= num_tiles) return;
  uint32_t lo = 0, hi = num_frames;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (frames[mid].tile_begin <= tile) lo = mid;
    else hi = mid;
  }
  const TraceFrame& F = frames[lo];
  if (!F.out.shade) return;
  const uint32_t lt = tile - F.tile_begin;
  const uint32_t x = (lt % F.tiles_x) * TW + (lane % TW);
  const uint32_t y = (lt / F.tiles_x) * TH + (lane / TW);
  if (x >= F.cam.width || y >= F.cam.height) return;
  const size_t pix = (size_t)y * F.cam.width + x;
  const uint4 raw = reinterpret_cast<const uint4*>(F.out.shade)[pix];
  Hit hit{};
  hit.node = raw.x;
  const uint32_t mode = F.cam.view_mode;
  const bool secondary = mode == SVO_VIEW_FILL_SHADOW || mode == SVO_VIEW_FILL_REFLECTION;
  if (mode == SVO_VIEW_PHONG || secondary) {
    hit.center[0] = __uint_as_float(raw.y);
    hit.center[1] = __uint_as_float(raw.z);
    hit.center[2] = __uint_as_float(raw.w);
  } else {
    hit.t = __uint_as_float(raw.y);
    hit.depth = raw.z & 255u;
    hit.loops = raw.z >> 8;
    hit.quality = __uint_as_float(raw.w);
  }
  float d[3];
  pixel_ray(F.cam, x, y, d);
  float4 out;
  if constexpr (SECONDARY) {
    if (secondary) {
      extern __shared__ uint32_t s_stack[];
      RayStack<SVO_STACK_LDS, 64> stk;
      stk.base = reinterpret_cast<uint2*>(s_stack);
      stk.tid = lane;
      out = fill_secondary(mode, hit, d, F.pools, stk);
    } else {
      out = shade_pixel(mode, hit, d, F.pools);
    }
  } else {
    out = shade_pixel(mode, hit, d, F.pools);
  }
  reinterpret_cast<float4*>(F.out.shade)[pix] = out;
}

__global__ void ray_dir_kernel(svo_camera cam, float* dirs) {
  const uint32_t x = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t y = blockIdx.y;
  if (x >= cam.width || y >= cam.height) return;
  float d[3];
  pixel_ray(cam, x, y, d);
  const size_t p = (size_t)y * cam.width + x;
  dirs[3 * p + 0] = d[0];
  dirs[3 * p + 1] = d[1];
  dirs[3 * p + 2] = d[2];
}

// Plain streaming copy, 16 bytes per lane per iteration (counter calibration).
__global__ __launch_bounds__(256) void copy_kernel(uint4* __restrict__ dst, const uint4* __restrict__ src, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

// VALU issue probe: eight independent integer chains per lane (no memory
// traffic but one store per lane at the end), 16 VALU instructions per
// iteration, eight iterations per loop trip so that the loop's own scalar
// instructions are few.  Launched with enough waves to fill every SIMD, its
// counters give the VALU-busy reading of a kernel that does nothing but
// issue VALU instructions, against which the traversal's reading is
// compared.
__global__ __launch_bounds__(256) void valu_probe_kernel(uint32_t* __restrict__ out, uint32_t iters) {
  const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t a[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) a[k] = g * (uint32_t)(2 * k + 1);
  for (uint32_t i = 0; i < iters; i += 8u) {
#pragma unroll
    for (int j = 0; j < 8; ++j) {
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        a[k] = a[k] + (0x9E3779B9u + (uint32_t)j);
        a[k] = a[k] ^ (0x85EBCA6Bu + (uint32_t)(8 * j + k));
      }
    }
  }
  uint32_t r = 0;
#pragma unroll
  for (int k = 0; k < 8; ++k) r ^= a[k];
  out[g] = r;
}

// One lane per 16-byte chunk of the concatenated spans; the span of a chunk
// is found by binary search over first_chunk.  A chunk moves as one 16-byte
// word when source and destination are 16-byte aligned, else as 4-byte words.
__global__ __launch_bounds__(256) void upload_kernel(const UploadSpan* __restrict__ spans, uint32_t num_spans,
                                                     uint64_t total_chunks) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t c = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; c < total_chunks; c += stride) {
    uint32_t lo = 0, hi = num_spans;
    while (hi - lo > 1u) {
      const uint32_t mid = (lo + hi) >> 1;
      if (spans[mid].first_chunk <= c) lo = mid;
      else hi = mid;
    }
    const UploadSpan sp = spans[lo];
    const uint64_t off = (c - sp.first_chunk) * 16u;
    if (off >= sp.bytes) continue;
    const uint64_t n = sp.bytes - off < 16u ? sp.bytes - off : 16u;
    const uint64_t s = sp.src + off, d = sp.dst + off;
    if (n == 16u && ((s | d) & 15u) == 0u) {
      *reinterpret_cast<uint4*>(d) = *reinterpret_cast<const uint4*>(s);
    } else {
      for (uint64_t k = 0; k + 4u <= n; k += 4u)
        *reinterpret_cast<uint32_t*>(d + k) = *reinterpret_cast<const uint32_t*>(s + k);
    }
  }
}

// Readback of one analysis cycle into pinned host memory: the slot keys
// whole, the count of asked gids as recorded, and the first min(count,
// num_gids) queue entries with the key of each (the per-gid table stays on
// the device).  Stores go straight over PCIe; the host reads the buffers
// after a stream sync.  Lane 0 then advances the device stamp: launches
// ordered after this one record into the next cycle.  (Only the count word
// is read by the other lanes; the queue count itself is reset by a copy
// behind this kernel.)
__global__ __launch_bounds__(256) void feedback_export_kernel(FeedbackExport x) {
  const uint32_t stride = gridDim.x * blockDim.x;
  const uint32_t t0 = blockIdx.x * blockDim.x + threadIdx.x;
  for (uint32_t s = t0; s < x.num_slots; s += stride) x.h_slot[s] = x.fb_slot[s];
  const uint32_t count = x.fb_state[0];
  const uint32_t n = count < x.num_gids ? count : x.num_gids;
  for (uint32_t i = t0; i < n; i += stride) {
    const uint32_t g = x.fb_queue[i];
    x.h_queue[i] = g;
    x.h_keys[i] = g < x.num_gids ? x.fb_gid[g] : 0ull;
  }
  if (t0 == 0) {
    *x.h_count = count;
    x.fb_state[1] = x.next_stamp;
  }
}

// View descriptors of the allocation-free launch (svo_render_views_noalloc):
// a fixed array in device memory, one copy per GPU for the life of the
// process, shared by every such launch on that GPU.  Host code never names
// it: it uses the device address noalloc_scratch looks up.
static __device__ __align__(256) TraceFrame scratch[kNoAllocMaxViews];

using TraceFn = void (*)(const TraceFrame*, uint32_t, uint32_t, uint32_t);

template <int POLICY, int TW, int WAVES>
TraceFn pick_stack(uint32_t stack) {
  return stack == SVO_STACK_PRIVATE    ? trace_kernel<POLICY, TW, WAVES, SVO_STACK_PRIVATE>
         : stack == SVO_STACK_LDS_NODE ? trace_kernel<POLICY, TW, WAVES, SVO_STACK_LDS_NODE>
                                       : trace_kernel<POLICY, TW, WAVES, SVO_STACK_LDS>;
}

template <int POLICY, int TW>
TraceFn pick_waves(uint32_t waves, uint32_t stack) {
  switch (waves) {
    case 1: return pick_stack<POLICY, TW, 1>(stack);
    case 2: return pick_stack<POLICY, TW, 2>(stack);
    default: return pick_stack<POLICY, TW, 4>(stack);
  }
}

template <int POLICY>
TraceFn pick_tile(const LaunchShape& s) {
  switch (s.tile_w) {
    case 8: return pick_waves<POLICY, 8>(s.waves, s.stack);
    case 32: return pick_waves<POLICY, 32>(s.waves, s.stack);
    default: return pick_waves<POLICY, 16>(s.waves, s.stack);
  }
}

}  // namespace

int noalloc_scratch(int device, TraceFrame** ptr) {
  constexpr int kMaxDevices = 64;
  static std::mutex lock;
  static void* addr[kMaxDevices] = {};
  if (device < 0 || device >= kMaxDevices) return (int)hipErrorInvalidDevice;
  std::lock_guard<std::mutex> g(lock);
  if (!addr[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return (int)e;
    if (!p) return (int)hipErrorInvalidSymbol;
    addr[device] = p;
  }
  *ptr = static_cast<TraceFrame*>(addr[device]);
  return 0;
}

bool shape_supported(const LaunchShape& s) {
  return (s.tile_w == 8 || s.tile_w == 16 || s.tile_w == 32) && (s.waves == 1 || s.waves == 2 || s.waves == 4) &&
         (s.stack == SVO_STACK_LDS || s.stack == SVO_STACK_PRIVATE || s.stack == SVO_STACK_LDS_NODE);
}

int launch_copy(void* dst, const void* src, size_t bytes, void* stream) {
  const size_t n = bytes / 16;
  if (n == 0) return 0;
  const unsigned blocks = (unsigned)std::min<size_t>((n + 255) / 256, 256 * 64);
  (void)hipGetLastError();  // see launch_trace
  hipLaunchKernelGGL(copy_kernel, dim3(blocks), dim3(256), 0, (hipStream_t)stream, (uint4*)dst, (const uint4*)src, n);
  return (int)hipGetLastError();
}

int launch_valu_probe(uint32_t* out, uint32_t blocks, uint32_t iters, void* stream) {
  if (blocks == 0) return 0;
  (void)hipGetLastError();  // see launch_trace
  hipLaunchKernelGGL(valu_probe_kernel, dim3(blocks), dim3(256), 0, (hipStream_t)stream, out, iters);
  return (int)hipGetLastError();
}

int launch_upload(const UploadSpan* spans, uint32_t num_spans, uint64_t total_chunks, void* stream) {
  if (num_spans == 0 || total_chunks == 0) return 0;
  const unsigned blocks = (unsigned)std::min<uint64_t>((total_chunks + 255) / 256, 2048);
  (void)hipGetLastError();  // see launch_trace
  hipLaunchKernelGGL(upload_kernel, dim3(blocks), dim3(256), 0, (hipStream_t)stream, spans, num_spans, total_chunks);
  return (int)hipGetLastError();
}

int launch_feedback_export(const FeedbackExport& x, void* stream) {
  const unsigned blocks = std::max(1u, std::min((x.num_slots + 255u) / 256u, 64u));
  (void)hipGetLastError();  // see launch_trace
  hipLaunchKernelGGL(feedback_export_kernel, dim3(blocks), dim3(256), 0, (hipStream_t)stream, x);
  return (int)hipGetLastError();
}

int launch_trace(const LaunchShape& shape, const TraceFrame* frames_dev, uint32_t num_frames, uint32_t num_tiles,
                 uint32_t stack_levels, int policy, bool shade, void* stream) {
  if (num_tiles == 0) return 0;
  if (!shape_supported(shape) || stack_levels > (uint32_t)kMaxScale) return (int)hipErrorInvalidValue;
  const TraceFn fn = policy == SVO_POLICY_CPU3     ? pick_tile<2>(shape)
                     : policy == SVO_POLICY_SINGLE ? pick_tile<1>(shape)
                                                   : pick_tile<0>(shape);
  const uint32_t block = 64u * shape.waves;
  const dim3 grid((num_tiles + shape.waves - 1) / shape.waves);
  // LDS stack: 8 bytes per level and lane, plus 2 bytes of masks for the
  // node-holding layout
  const size_t lds = shape.stack == SVO_STACK_LDS        ? (size_t)stack_levels * block * sizeof(uint2)
                     : shape.stack == SVO_STACK_LDS_NODE ? (size_t)stack_levels * block * (sizeof(uint2) + 2)
                                                         : 0;
  // hipGetLastError reports the thread's last error from any earlier call:
  // clear it so that the check below sees this launch only
  (void)hipGetLastError();
  hipLaunchKernelGGL(fn, grid, dim3(block), lds, (hipStream_t)stream, frames_dev, num_frames, num_tiles, stack_levels);
  int e = (int)hipGetLastError();
  if (e != 0 || !shade) return e;
  const dim3 sgrid(num_tiles), sblock(64);
  hipStream_t st = (hipStream_t)stream;
  if (policy == SVO_POLICY_SINGLE) {
    // the pass that can also trace the shadow and reflection rays of the
    // single-SVO renderer's view modes (a 64-lane LDS stack)
    const size_t slds = (size_t)stack_levels * 64u * sizeof(uint2);
    if (shape.tile_w == 8) hipLaunchKernelGGL((shade_kernel<8, true>), sgrid, sblock, slds, st, frames_dev, num_frames, num_tiles, stack_levels);
    else if (shape.tile_w == 32) hipLaunchKernelGGL((shade_kernel<32, true>), sgrid, sblock, slds, st, frames_dev, num_frames, num_tiles, stack_levels);
    else hipLaunchKernelGGL((shade_kernel<16, true>), sgrid, sblock, slds, st, frames_dev, num_frames, num_tiles, stack_levels);
  } else {
    if (shape.tile_w == 8) hipLaunchKernelGGL((shade_kernel<8, false>), sgrid, sblock, 0, st, frames_dev, num_frames, num_tiles, stack_levels);
    else if (shape.tile_w == 32) hipLaunchKernelGGL((shade_kernel<32, false>), sgrid, sblock, 0, st, frames_dev, num_frames, num_tiles, stack_levels);
    else hipLaunchKernelGGL((shade_kernel<16, false>), sgrid, sblock, 0, st, frames_dev, num_frames, num_tiles, stack_levels);
  }
  return (int)hipGetLastError();
}

int launch_ray_dirs(const svo_camera& cam, float* dirs_dev, void* stream) {
  const dim3 grid((cam.width + 63) / 64, cam.height), block(64);
  (void)hipGetLastError();  // see launch_trace
  hipLaunchKernelGGL(ray_dir_kernel, grid, block, 0, (hipStream_t)stream, cam, dirs_dev);
  return (int)hipGetLastError();
}

}  // namespace svo_impl
