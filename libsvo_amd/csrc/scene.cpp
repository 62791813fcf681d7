// Treelet scenes: octree construction from leaf voxels, treelet packing, and
// the .svo / .ia bundle files.
//
// Every scene has the reference's format invariants:
//   * a treelet is stored breadth-first from its root (local index 0); the
//     children of a node are contiguous, in child-index order, and the node
//     stores the offset of its first child relative to itself
//     (src/TreeletBuilderFromFaces.cpp:188-331);
//   * nodes a treelet does not expand are its leaves and get their leaf bit
//     set in their parent (:339-358); each such leaf above the scene depth
//     anchors a sub-treelet, whose gid it stores in its first-child field;
//     gids are handed out in FIFO order over treelets
//     (src/TreeletBuildManager.cpp:174-258);
//   * inner-node attributes are the mean of the valid children's attributes,
//     normal renormalised, colour clamped
//     (src/attribute_generators/Ag_colorAndNormalsTriangles.cpp:249-300), and
//     a sub-treelet's root carries the same attribute as the parent's leaf
//     (:185-199).
// Which nodes a treelet expands is the packing policy:
//   * SVO_BUILD_PACK_BFS is the reference rule: expand in queue order while at
//     least 9 slots remain (`currentNodeIndex < size-8`, :200);
//   * the default dense policy cuts the octree into the fewest connected
//     subtrees of at most N nodes, so that slots are mostly full (the BFS rule
//     leaves most slots nearly empty on large scenes: a 4096-node depth-11
//     terrain becomes millions of treelets of about ten nodes).
// One deliberate difference from the reference's triangle builder, which keeps
// subdividing while a treelet has room: here every octree leaf sits exactly at
// the scene depth.
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstring>
#include <deque>
#include <memory>

#include <omp.h>

#include "svo_internal.h"

namespace svo_impl {

static thread_local std::string g_error;

void set_error(const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_error = buf;
}
void clear_error() { g_error.clear(); }

}  // namespace svo_impl

extern "C" const char* svo_last_error(void) { return svo_impl::g_error.c_str(); }

namespace svo_impl {

// ------------------------------------------------------------ morton

// Bit b of i, j, k goes to bit 3b+2, 3b+1, 3b of the code (bits >= depth
// dropped).  The spread and its inverse move bit groups of halving width
// with shifts and masks (21 bits per axis).
static inline uint64_t spread3(uint64_t x) {
  x &= 0x1FFFFFull;
  x = (x | x << 32) & 0x001F00000000FFFFull;
  x = (x | x << 16) & 0x001F0000FF0000FFull;
  x = (x | x << 8) & 0x100F00F00F00F00Full;
  x = (x | x << 4) & 0x10C30C30C30C30C3ull;
  x = (x | x << 2) & 0x1249249249249249ull;
  return x;
}

static inline uint32_t compact3(uint64_t x) {
  x &= 0x1249249249249249ull;
  x = (x | x >> 2) & 0x10C30C30C30C30C3ull;
  x = (x | x >> 4) & 0x100F00F00F00F00Full;
  x = (x | x >> 8) & 0x001F0000FF0000FFull;
  x = (x | x >> 16) & 0x001F00000000FFFFull;
  x = (x | x >> 32) & 0x1FFFFFull;
  return (uint32_t)x;
}

uint64_t morton_encode(uint32_t i, uint32_t j, uint32_t k, uint32_t depth) {
  const uint32_t m = depth >= 21 ? 0x1FFFFFu : (1u << depth) - 1u;
  return spread3(i & m) << 2 | spread3(j & m) << 1 | spread3(k & m);
}

void morton_decode(uint64_t c, uint32_t depth, uint32_t* i, uint32_t* j, uint32_t* k) {
  const uint32_t m = depth >= 21 ? 0x1FFFFFu : (1u << depth) - 1u;
  *i = compact3(c >> 2) & m;
  *j = compact3(c >> 1) & m;
  *k = compact3(c) & m;
}

void radix_sort_u64(std::vector<uint64_t>& keys, uint32_t bits) {
  const size_t n = keys.size();
  if (n < 2) return;
  std::vector<uint64_t> tmp(n);
  const uint32_t kRadix = 11, kBuckets = 1u << kRadix;
  // LSD passes; each thread histograms and scatters its own contiguous chunk,
  // the chunks' offsets per bucket following in thread order (stable)
  const int nt = n < (1u << 16) ? 1 : std::max(1, omp_get_max_threads());
  std::vector<size_t> count((size_t)nt * kBuckets);
  for (uint32_t shift = 0; shift < bits; shift += kRadix) {
#pragma omp parallel num_threads(nt)
    {
      const int t = omp_get_thread_num();
      const size_t lo = n * t / nt, hi = n * (t + 1) / nt;
      size_t* c = &count[(size_t)t * kBuckets];
      std::fill(c, c + kBuckets, 0);
      for (size_t i = lo; i < hi; ++i) ++c[(keys[i] >> shift) & (kBuckets - 1)];
#pragma omp barrier
#pragma omp single
      {
        size_t sum = 0;
        for (uint32_t b = 0; b < kBuckets; ++b)
          for (int u = 0; u < nt; ++u) {
            size_t& x = count[(size_t)u * kBuckets + b];
            const size_t v = x;
            x = sum;
            sum += v;
          }
      }
      for (size_t i = lo; i < hi; ++i) tmp[c[(keys[i] >> shift) & (kBuckets - 1)]++] = keys[i];
    }
    keys.swap(tmp);
  }
}

// ------------------------------------------------------------ attributes

static void unpack_color(uint32_t c, float v[3]) {
  v[0] = (float)((c >> 24) & 255u) / 255.0f;
  v[1] = (float)((c >> 16) & 255u) / 255.0f;
  v[2] = (float)((c >> 8) & 255u) / 255.0f;
}
static uint32_t pack_unit(float x) {  // [0,1] -> byte, round half up
  float y = x * 255.0f + 0.5f;
  if (!(y > 0.0f)) return 0;
  if (y >= 255.0f) return 255;
  return (uint32_t)y;
}
static void unpack_normal(uint32_t c, float v[3]) {
  unpack_color(c, v);
  for (int a = 0; a < 3; ++a) v[a] = v[a] * 2.0f - 1.0f;
}
static uint32_t pack_normal(const float n[3]) {
  return pack_rgba(pack_unit(n[0] * 0.5f + 0.5f), pack_unit(n[1] * 0.5f + 0.5f),
                   pack_unit(n[2] * 0.5f + 0.5f), 0u);
}

// Mean of `count` child attribute packages (normal, colour).
static void average_attribs(const uint32_t* child_attr, uint32_t count, uint32_t out[2]) {
  float n[3] = {0, 0, 0}, c[3] = {0, 0, 0};
  for (uint32_t i = 0; i < count; ++i) {
    float a[3], b[3];
    unpack_normal(child_attr[2 * i], a);
    unpack_color(child_attr[2 * i + 1], b);
    for (int k = 0; k < 3; ++k) { n[k] += a[k]; c[k] += b[k]; }
  }
  for (int k = 0; k < 3; ++k) { n[k] /= (float)count; c[k] /= (float)count; }
  float len = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
  if (len > 0.0f) for (int k = 0; k < 3; ++k) n[k] /= len;
  for (int k = 0; k < 3; ++k) c[k] = std::min(1.0f, std::max(0.0f, c[k]));
  out[0] = pack_normal(n);
  out[1] = pack_rgba(pack_unit(c[0]), pack_unit(c[1]), pack_unit(c[2]), 255u);
}

// ------------------------------------------------------------ octree levels

namespace {

struct Level {
  std::vector<uint8_t> mask;     // valid-child mask (inner levels)
  std::vector<uint32_t> first;   // index of the first child in the next level
  std::vector<uint32_t> attr;    // 2 words per node
};

struct QueueItem {
  uint32_t level;
  uint32_t index;        // node index within its level
  uint32_t local;        // node index within the treelet
  uint32_t child_idx;    // 0..7 inside its parent
  uint32_t parent_local;
};

struct Incomplete {
  uint32_t level, index, local, child_idx, parent_local;
};

}  // namespace

svo_status build_scene(uint32_t depth, std::vector<uint64_t>& codes,
                       std::vector<uint32_t>& leaf_attribs, uint32_t treelet_bytes,
                       uint32_t flags, svo_scene** out) {
  if (!out) { set_error("build_scene: out is null"); return SVO_ERR_ARG; }
  *out = nullptr;
  if (depth < 1 || depth > 20) { set_error("scene depth %u outside [1,20]", depth); return SVO_ERR_ARG; }
  if (codes.empty()) { set_error("scene has no leaf voxels"); return SVO_ERR_ARG; }
  if (leaf_attribs.size() != 2 * codes.size()) { set_error("attribute count mismatch"); return SVO_ERR_ARG; }
  if (treelet_bytes != 0 && (treelet_bytes % 8 != 0 || treelet_bytes < 8 * 16)) {
    set_error("treelet_bytes %u must be a multiple of 8 and >= 128", treelet_bytes);
    return SVO_ERR_ARG;
  }
  if (flags & ~(uint32_t)SVO_BUILD_PACK_BFS) { set_error("unknown build flags 0x%x", flags); return SVO_ERR_ARG; }
  const bool bfs = (flags & SVO_BUILD_PACK_BFS) != 0;
  for (size_t i = 1; i < codes.size(); ++i)
    if (codes[i] <= codes[i - 1]) { set_error("leaf codes not strictly increasing at %zu", i); return SVO_ERR_ARG; }
  if ((codes.back() >> (3 * depth)) != 0) { set_error("leaf code beyond depth %u", depth); return SVO_ERR_ARG; }

  // Bottom-up: level d+1 -> level d.
  std::vector<Level> lv(depth + 1);
  lv[depth].attr.swap(leaf_attribs);
  std::vector<uint64_t> cur;
  cur.swap(codes);
  uint64_t tree_nodes = cur.size();
  const uint64_t leaves = cur.size();
  // Each level is grouped in parallel: thread chunks start at a parent
  // boundary, count their parents, then write them at their prefix offsets.
  const int nt = std::max(1, omp_get_max_threads());
  std::vector<size_t> chunk(nt + 1), nparents(nt + 1);
  for (int d = (int)depth - 1; d >= 0; --d) {
    Level& P = lv[d];
    const Level& C = lv[d + 1];
    const size_t nc = cur.size();
    const int use = nc < (1u << 16) ? 1 : nt;
    for (int t = 0; t <= use; ++t) {
      size_t b = nc * t / use;
      while (b > 0 && b < nc && (cur[b] >> 3) == (cur[b - 1] >> 3)) ++b;
      chunk[t] = b;
    }
    nparents[0] = 0;
#pragma omp parallel for num_threads(use) schedule(static, 1)
    for (int t = 0; t < use; ++t) {
      size_t c = 0;
      for (size_t i = chunk[t]; i < chunk[t + 1]; ++i) c += (i == chunk[t] || (cur[i] >> 3) != (cur[i - 1] >> 3));
      nparents[t + 1] = c;
    }
    for (int t = 0; t < use; ++t) nparents[t + 1] += nparents[t];
    const size_t np = nparents[use];
    std::vector<uint64_t> parent_codes(np);
    P.mask.assign(np, 0);
    P.first.resize(np);
    P.attr.resize(2 * np);
#pragma omp parallel for num_threads(use) schedule(static, 1)
    for (int t = 0; t < use; ++t) {
      size_t q = nparents[t];
      for (size_t i = chunk[t]; i < chunk[t + 1]; ++i) {
        const uint64_t pc = cur[i] >> 3;
        if (i == chunk[t] || pc != (cur[i - 1] >> 3)) {
          parent_codes[q] = pc;
          P.first[q] = (uint32_t)i;
          ++q;
        }
        P.mask[q - 1] |= (uint8_t)(1u << (cur[i] & 7u));
      }
    }
#pragma omp parallel for schedule(static) if (np >= 4096)
    for (int64_t p = 0; p < (int64_t)np; ++p) {
      uint32_t cnt = (uint32_t)__builtin_popcount(P.mask[p]);
      average_attribs(&C.attr[2 * (size_t)P.first[p]], cnt, &P.attr[2 * p]);
    }
    cur.swap(parent_codes);
    tree_nodes += np;
  }
  cur.clear();
  cur.shrink_to_fit();

  uint32_t tb = treelet_bytes;
  if (tb == 0) {
    uint64_t need = std::max<uint64_t>((tree_nodes + 9) * 8, 128);  // the .svo reader's minimum
    if (need > 0xFFFFFFF8ull) { set_error("tree of %llu nodes does not fit one treelet", (unsigned long long)tree_nodes); return SVO_ERR_ARG; }
    tb = (uint32_t)need;
  }
  const uint32_t N = tb / 8;

  // Dense packing: cut[d][i] marks the roots of sub-treelets.  Bottom-up,
  // each node carries w = the size of the part of its subtree that stays in
  // its treelet (a cut node costs one anchor node); while a node's w exceeds
  // N the heaviest cuttable nodes below it are cut (greedy partition of a
  // tree into connected parts of bounded size).
  // Cuts are made only at even levels >= 2.  The anchor of a sub-treelet is
  // therefore never a child of a treelet root: the reference's stitch step
  // clears the anchor's leaf bit in its parent inside the parent treelet's
  // slot (src/TreeletMemoryManager.cpp:534-535), which would miss the copy of
  // a treelet root stitched into the level above.  A node whose children sit
  // at an odd level cuts among its grandchildren instead.
  std::vector<std::vector<uint8_t>> cut(depth + 1);
  if (!bfs) {
    // (a tree that fits one treelet needs no cut, whatever N)
    if (N < 80 && tree_nodes > N) { set_error("dense packing needs treelets of at least 80 nodes (got %u)", N); return SVO_ERR_ARG; }
    std::vector<std::vector<uint32_t>> w(depth + 1);
    w[depth].assign(lv[depth].attr.size() / 2, 1u);
    cut[depth].assign(w[depth].size(), 0);
    struct Cand { uint32_t w; uint32_t level; uint32_t index; uint32_t parent_slot; };
    for (int d = (int)depth - 1; d >= 0; --d) {
      const Level& P = lv[d];
      const size_t np = P.mask.size();
      w[d].assign(np, 0);
      cut[d].assign(np, 0);
      const bool child_even = ((d + 1) % 2) == 0 && d + 1 >= 2;
      // A node reads and writes only its own children and grandchildren, so
      // the nodes of a level are independent; the lowest failing node names
      // the error.
      int64_t bad = -1;
      uint64_t bad_sum = 0;
#pragma omp parallel for schedule(dynamic, 1024) if (np >= 4096)
      for (int64_t p = 0; p < (int64_t)np; ++p) {
        const uint32_t f = P.first[p];
        const uint32_t nc = (uint32_t)__builtin_popcount(P.mask[p]);
        uint64_t sum = 1;
        for (uint32_t k = 0; k < nc; ++k) sum += w[d + 1][f + k];
        if (sum > N) {
          Cand cand[64];
          uint32_t n = 0;
          if (child_even) {
            for (uint32_t k = 0; k < nc; ++k)
              if ((uint32_t)(d + 1) < depth) cand[n++] = Cand{w[d + 1][f + k], (uint32_t)d + 1, f + k, k};
          } else if ((uint32_t)(d + 2) <= depth) {
            for (uint32_t k = 0; k < nc; ++k) {
              const uint32_t c = f + k;
              if ((uint32_t)(d + 1) >= depth) continue;
              const uint32_t gf = lv[d + 1].first[c];
              const uint32_t gn = (uint32_t)__builtin_popcount(lv[d + 1].mask[c]);
              for (uint32_t q = 0; q < gn; ++q)
                if ((uint32_t)(d + 2) < depth && !cut[d + 2][gf + q]) cand[n++] = Cand{w[d + 2][gf + q], (uint32_t)d + 2, gf + q, c};
            }
          }
          std::sort(cand, cand + n, [](const Cand& a, const Cand& b) {
            return a.w != b.w ? a.w > b.w : (a.level != b.level ? a.level < b.level : a.index < b.index);
          });
          for (uint32_t k = 0; k < n && sum > N; ++k) {
            if (cand[k].w <= 1) break;
            cut[cand[k].level][cand[k].index] = 1;
            const uint64_t dec = (uint64_t)cand[k].w - 1;
            sum -= dec;
            if (cand[k].level == (uint32_t)d + 2) w[d + 1][cand[k].parent_slot] -= (uint32_t)dec;
          }
          if (sum > N) {
#pragma omp critical(svo_dense_cut_error)
            if (bad < 0 || p < bad) { bad = p; bad_sum = sum; }
            continue;
          }
        }
        w[d][p] = (uint32_t)sum;
      }
      if (bad >= 0) { set_error("dense packing: node at level %d keeps %llu > %u nodes", d, (unsigned long long)bad_sum, N); return SVO_ERR_ARG; }
      if (d + 2 <= (int)depth) std::vector<uint32_t>().swap(w[d + 2]);
    }
  }

  std::unique_ptr<svo_scene> sc(new svo_scene());
  sc->treelet_bytes = tb;
  sc->nodes_per_treelet = N;
  sc->max_depth = depth;
  sc->inner_depth = depth > 0 ? depth - 1 : 0;  // every leaf lies at `depth`
  sc->tree_nodes = tree_nodes;
  sc->tree_leaves = leaves;

  // Treelet contents depend only on their root node and the cuts, so each
  // generation (the treelets anchored in the previous one) is expanded in
  // parallel.  Gids follow the reference's FIFO order over treelets: the
  // sub-treelets of treelet t are numbered after those of t-1, in the order
  // of t's anchor leaves.
  struct Root { uint32_t level, index, parent_gid, leaf_pos, leaf_idx, leaf_parent_pos; };
  auto build_treelet = [&](uint32_t gid, const Root& r, std::vector<Incomplete>& inc) {
    TreeletData& T = sc->treelets[gid];
    T.h.mem_size = tb;
    T.h.gid = gid;
    T.h.parent_gid = r.parent_gid;
    T.h.parent_leaf_pos = r.leaf_pos;
    T.h.parent_leaf_idx = r.leaf_idx;
    T.h.parent_leaf_parent_pos = r.leaf_parent_pos;
    T.nodes.push_back(svo_node{0, 0});
    T.attribs.push_back(lv[r.level].attr[2 * (size_t)r.index]);
    T.attribs.push_back(lv[r.level].attr[2 * (size_t)r.index + 1]);
    std::vector<QueueItem> q;  // FIFO: q[head..] are waiting
    q.reserve(std::min<uint64_t>(N, tree_nodes) + 8);
    size_t head = 0;
    std::vector<QueueItem> frontier;  // treelet leaves in local-index order
    q.push_back(QueueItem{r.level, r.index, 0, 0, 0});
    uint32_t cur_idx = 1;
    while (head < q.size()) {
      const QueueItem e = q[head];
      bool expand;
      if (bfs) {
        // reference rule: stop while fewer than 9 slots remain
        if (!(cur_idx < N - 8) || e.level == depth) break;
        expand = true;
      } else {
        expand = e.level < depth && (e.local == 0 || !cut[e.level][e.index]);
      }
      ++head;
      if (!expand) { frontier.push_back(e); continue; }
      const uint32_t m = lv[e.level].mask[e.index];
      uint32_t child = lv[e.level].first[e.index];
      bool first_valid = true;
      for (uint32_t c = 0; c < 8; ++c) {
        if (!(m & (1u << c))) continue;
        if (first_valid) {
          T.nodes[e.local].first_child = cur_idx - e.local;
          first_valid = false;
        }
        T.nodes[e.local].masks |= 1u << (kValidShift + c);
        T.nodes.push_back(svo_node{0, 0});
        const uint32_t* a = &lv[e.level + 1].attr[2 * (size_t)child];
        T.attribs.push_back(a[0]);
        T.attribs.push_back(a[1]);
        q.push_back(QueueItem{e.level + 1, child, cur_idx, c, e.local});
        ++child;
        ++cur_idx;
      }
    }
    for (size_t i = head; i < q.size(); ++i) frontier.push_back(q[i]);
    T.h.num_nodes = cur_idx;
    T.h.num_leaves = (uint32_t)frontier.size();
    inc.clear();
    for (const QueueItem& e : frontier) {
      T.nodes[e.parent_local].masks |= 1u << (kLeafShift + e.child_idx);
      if (e.level < depth) inc.push_back(Incomplete{e.level, e.index, e.local, e.child_idx, e.parent_local});
    }
  };

  std::vector<Root> gen{Root{0, 0, 0, 0, 0, 0}};
  while (!gen.empty()) {
    const size_t base = sc->treelets.size();
    sc->treelets.resize(base + gen.size());
    std::vector<std::vector<Incomplete>> inc(gen.size());
#pragma omp parallel for schedule(dynamic, 4) if (gen.size() >= 8)
    for (int64_t g = 0; g < (int64_t)gen.size(); ++g) build_treelet((uint32_t)(base + g), gen[g], inc[g]);
    std::vector<Root> next;
    for (size_t g = 0; g < gen.size(); ++g) {
      const uint32_t t = (uint32_t)(base + g);
      for (const Incomplete& e : inc[g]) {
        const uint32_t gid = (uint32_t)(base + gen.size() + next.size());
        sc->treelets[t].nodes[e.local].first_child = gid;
        next.push_back(Root{e.level, e.index, t, e.local, e.child_idx, e.parent_local});
      }
    }
    gen.swap(next);
  }
  *out = sc.release();
  return SVO_OK;
}

// ------------------------------------------------------------ file IO

namespace {

struct FileCloser { void operator()(FILE* f) const { if (f) fclose(f); } };
using FilePtr = std::unique_ptr<FILE, FileCloser>;

bool write_u32(FILE* f, uint32_t v) { return fwrite(&v, 4, 1, f) == 1; }
bool read_u32(FILE* f, uint32_t* v) { return fread(v, 4, 1, f) == 1; }

}  // namespace

// Depth of every node of every treelet, checking that child offsets stay
// inside the treelet.  Returns the depth of the deepest valid node (the
// deepest leaf in a well-formed tree), or -1 on a bad offset.
int scene_depth(const svo_scene& sc, std::string* why, uint32_t* inner_depth) {
  const size_t nt = sc.treelets.size();
  std::vector<uint32_t> root_depth(nt, 0);
  // children of each treelet in CSR form (parent gid < child gid)
  std::vector<uint32_t> cstart(nt + 1, 0), clist(nt ? nt - 1 : 0);
  for (size_t u = 1; u < nt; ++u) ++cstart[sc.treelets[u].h.parent_gid + 1];
  for (size_t t = 0; t < nt; ++t) cstart[t + 1] += cstart[t];
  {
    std::vector<uint32_t> fill(cstart.begin(), cstart.end() - 1);
    for (size_t u = 1; u < nt; ++u) clist[fill[sc.treelets[u].h.parent_gid]++] = (uint32_t)u;
  }
  int maxd = 0;
  uint32_t inner = 0;
  std::vector<uint32_t> nd;
  for (size_t t = 0; t < nt; ++t) {
    const TreeletData& T = sc.treelets[t];
    nd.assign(T.h.num_nodes, 0);
    nd[0] = root_depth[t];
    if (T.h.num_nodes != 0 && nd[0] > inner) inner = nd[0];  // a (sub-)treelet root
    // BFS order: a node's children have larger local indices.
    for (uint32_t i = 0; i < T.h.num_nodes; ++i) {
      const svo_node& n = T.nodes[i];
      const uint32_t vm = (n.masks >> kValidShift) & 255u;
            if (vm == 0) continue;
      for (uint32_t c = 0; c < 8; ++c) {
        if (!(vm & (1u << c))) continue;
        const uint32_t ci = i + nth_child(n.masks, n.first_child, c);
        if (ci <= i || ci >= T.h.num_nodes) {
          if (why) {
            char b[160];
            snprintf(b, sizeof b, "treelet %zu node %u child %u offset out of range", t, i, c);
            *why = b;
          }
          return -1;
        }
        nd[ci] = nd[i] + 1;
        // every valid node counts, not only leaves: an inner node below the
        // deepest leaf (a malformed tree) must still fit the traversal stack
        if (nd[ci] > (uint32_t)maxd) maxd = (int)nd[ci];
        if (!((n.masks >> c) & 1u) && nd[ci] > inner) inner = nd[ci];  // a valid non-leaf child
      }
    }
    // children treelets of t take their root depth from t's leaf
    for (uint32_t k = cstart[t]; k < cstart[t + 1]; ++k) {
      const uint32_t u = clist[k];
      root_depth[u] = nd[sc.treelets[u].h.parent_leaf_pos];
    }
  }
  if (inner_depth) *inner_depth = inner;
  return maxd;
}

}  // namespace svo_impl

using namespace svo_impl;

extern "C" svo_status svo_scene_write(const svo_scene* sc, const char* base) {
  if (!sc || !base) { set_error("svo_scene_write: null argument"); return SVO_ERR_ARG; }
  std::string p = std::string(base) + ".svo";
  FilePtr f(fopen(p.c_str(), "wb"));
  if (!f) { set_error("cannot open %s for writing", p.c_str()); return SVO_ERR_IO; }
  bool ok = write_u32(f.get(), (uint32_t)sc->treelets.size()) && write_u32(f.get(), sc->treelet_bytes);
  for (const TreeletData& T : sc->treelets) {
    ok = ok && fwrite(&T.h, sizeof(T.h), 1, f.get()) == 1;
    ok = ok && fwrite(T.nodes.data(), sizeof(svo_node), T.h.num_nodes, f.get()) == T.h.num_nodes;
  }
  if (!ok) { set_error("write failed: %s", p.c_str()); return SVO_ERR_IO; }
  f.reset();
  std::string pa = std::string(base) + ".ia";
  FilePtr g(fopen(pa.c_str(), "wb"));
  if (!g) { set_error("cannot open %s for writing", pa.c_str()); return SVO_ERR_IO; }
  ok = write_u32(g.get(), (uint32_t)sc->treelets.size());
  for (const TreeletData& T : sc->treelets) {
    ok = ok && write_u32(g.get(), T.h.num_nodes) && write_u32(g.get(), 2u);
    ok = ok && fwrite(T.attribs.data(), 4, 2 * (size_t)T.h.num_nodes, g.get()) == 2 * (size_t)T.h.num_nodes;
  }
  if (!ok) { set_error("write failed: %s", pa.c_str()); return SVO_ERR_IO; }
  return SVO_OK;
}

extern "C" svo_status svo_scene_open(const char* base, svo_scene** out) {
  if (!base || !out) { set_error("svo_scene_open: null argument"); return SVO_ERR_ARG; }
  *out = nullptr;
  std::string p = std::string(base) + ".svo";
  FilePtr f(fopen(p.c_str(), "rb"));
  if (!f) { set_error("could not open %s", p.c_str()); return SVO_ERR_IO; }
  uint32_t nt = 0, tb = 0;
  if (!read_u32(f.get(), &nt) || !read_u32(f.get(), &tb)) { set_error("%s: truncated header", p.c_str()); return SVO_ERR_FORMAT; }
  if (nt == 0 || tb < 128 || tb % 8) { set_error("%s: bad header (treelets %u, bytes %u)", p.c_str(), nt, tb); return SVO_ERR_FORMAT; }
  std::unique_ptr<svo_scene> sc(new svo_scene());
  sc->treelet_bytes = tb;
  sc->nodes_per_treelet = tb / 8;
  sc->treelets.resize(nt);
  for (uint32_t t = 0; t < nt; ++t) {
    TreeletData& T = sc->treelets[t];
    if (fread(&T.h, sizeof(T.h), 1, f.get()) != 1) { set_error("%s: truncated at treelet %u", p.c_str(), t); return SVO_ERR_FORMAT; }
    if (T.h.mem_size != tb || T.h.num_nodes == 0 || T.h.num_nodes > tb / 8 || T.h.num_leaves > T.h.num_nodes ||
        T.h.gid != t || (t > 0 && T.h.parent_gid >= t)) {
      set_error("%s: treelet %u header inconsistent", p.c_str(), t);
      return SVO_ERR_FORMAT;
    }
    if (t > 0) {
      const svo_treelet_header& ph = sc->treelets[T.h.parent_gid].h;
      if (T.h.parent_leaf_pos >= ph.num_nodes || T.h.parent_leaf_parent_pos >= ph.num_nodes || T.h.parent_leaf_idx > 7) {
        set_error("%s: treelet %u parent link out of range", p.c_str(), t);
        return SVO_ERR_FORMAT;
      }
    }
    T.nodes.resize(T.h.num_nodes);
    if (fread(T.nodes.data(), sizeof(svo_node), T.h.num_nodes, f.get()) != T.h.num_nodes) {
      set_error("%s: truncated nodes of treelet %u", p.c_str(), t);
      return SVO_ERR_FORMAT;
    }
  }
  f.reset();
  std::string pa = std::string(base) + ".ia";
  FilePtr g(fopen(pa.c_str(), "rb"));
  if (!g) { set_error("could not open %s", pa.c_str()); return SVO_ERR_IO; }
  uint32_t na = 0;
  if (!read_u32(g.get(), &na) || na < nt) { set_error("%s: attribute count %u < treelets %u", pa.c_str(), na, nt); return SVO_ERR_FORMAT; }
  for (uint32_t t = 0; t < nt; ++t) {
    TreeletData& T = sc->treelets[t];
    uint32_t np = 0, wpp = 0;
    if (!read_u32(g.get(), &np) || !read_u32(g.get(), &wpp) || wpp != 2 || np != T.h.num_nodes) {
      set_error("%s: attribute block %u malformed", pa.c_str(), t);
      return SVO_ERR_FORMAT;
    }
    T.attribs.resize(2 * (size_t)np);
    if (fread(T.attribs.data(), 4, 2 * (size_t)np, g.get()) != 2 * (size_t)np) {
      set_error("%s: truncated attributes of treelet %u", pa.c_str(), t);
      return SVO_ERR_FORMAT;
    }
  }
  std::string why;
  int d = scene_depth(*sc, &why, &sc->inner_depth);
  if (d < 0) { set_error("%s: %s", p.c_str(), why.c_str()); return SVO_ERR_FORMAT; }
  sc->max_depth = (uint32_t)d;
  uint64_t nodes = 0, leaves = 0;
  for (const TreeletData& T : sc->treelets) {
    nodes += T.h.num_nodes;
    for (uint32_t i = 0; i < T.h.num_nodes; ++i) leaves += __builtin_popcount(T.nodes[i].masks & 255u);
  }
  sc->tree_nodes = nodes - (nt - 1);  // each sub-treelet root duplicates its parent's leaf
  sc->tree_leaves = leaves - (nt - 1);
  *out = sc.release();
  return SVO_OK;
}

extern "C" void svo_scene_destroy(svo_scene* sc) { delete sc; }

extern "C" svo_status svo_scene_get_info(const svo_scene* sc, svo_scene_info* out) {
  if (!sc || !out) { set_error("svo_scene_get_info: null argument"); return SVO_ERR_ARG; }
  out->num_treelets = (uint32_t)sc->treelets.size();
  out->treelet_bytes = sc->treelet_bytes;
  out->nodes_per_treelet = sc->nodes_per_treelet;
  out->max_depth = sc->max_depth;
  uint64_t tot = 0;
  for (const TreeletData& T : sc->treelets) tot += T.h.num_nodes;
  out->total_nodes = tot;
  out->tree_nodes = sc->tree_nodes;
  out->tree_leaves = sc->tree_leaves;
  return SVO_OK;
}

extern "C" svo_status svo_scene_get_treelet(const svo_scene* sc, uint32_t gid, svo_treelet_header* h,
                                            const svo_node** nodes, const uint32_t** attribs) {
  if (!sc) { set_error("svo_scene_get_treelet: null scene"); return SVO_ERR_ARG; }
  if (gid >= sc->treelets.size()) { set_error("treelet gid %u out of range (%zu treelets)", gid, sc->treelets.size()); return SVO_ERR_ARG; }
  const TreeletData& T = sc->treelets[gid];
  if (h) *h = T.h;
  if (nodes) *nodes = T.nodes.data();
  if (attribs) *attribs = T.attribs.data();
  return SVO_OK;
}

extern "C" svo_status svo_scene_build_from_leaves(uint32_t depth, const uint64_t* morton, const uint32_t* attribs,
                                                  size_t count, uint32_t treelet_bytes, uint32_t flags,
                                                  svo_scene** out) {
  if (!morton || !attribs || !out) { set_error("svo_scene_build_from_leaves: null argument"); return SVO_ERR_ARG; }
  std::vector<uint64_t> codes(morton, morton + count);
  std::vector<uint32_t> at(attribs, attribs + 2 * count);
  return build_scene(depth, codes, at, treelet_bytes, flags, out);
}
