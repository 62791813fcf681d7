// Legacy single-SVO files: the serialized octree of the in-core path, read
// into a scene of one treelet so that every traversal policy (and the
// residency manager) handles it like any treelet bundle.
//
//   SvoLoader (src/SvoLoader.cpp:80-107): "<depth> <numNodes> " in ASCII,
//       then numNodes 8-byte nodes with ABSOLUTE first-child indices;
//       attribute package i belongs to node i.
//   SvoBranch v1 (src/incore/SvoBranch.cpp:320-395): "v1 <depth>
//       <numNodes> " in ASCII, the nodes, then numNodes u32 attribute
//       indices (fillFrameBuffer_rgba.cl:640 reads attribs[attribIndices[n]]).
//
// The words are separated by one space and the binary part starts right
// after the space that ends the last word (how SvoBranch writes them).  The
// attribute file <base>.ia holds one block of packages in this library's
// .ia layout (u32 1, u32 packages, u32 2, packages x {normal, color}); the
// reference reads gloost InterleavedAttributes there, whose layout is not
// part of the reference's sources.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "svo_internal.h"

using namespace svo_impl;

namespace {

struct FileCloser {
  void operator()(FILE* f) const {
    if (f) fclose(f);
  }
};
using File = std::unique_ptr<FILE, FileCloser>;

bool read_word(FILE* f, std::string* w) {
  w->clear();
  int c;
  while ((c = fgetc(f)) != EOF && (c == ' ' || c == '\n' || c == '\t' || c == '\r')) {
  }
  while (c != EOF && c != ' ' && c != '\n' && c != '\t' && c != '\r') {
    w->push_back((char)c);
    if (w->size() > 32) return false;
    c = fgetc(f);
  }
  return !w->empty() && c != EOF;  // the separator after the word was consumed
}

bool parse_u32(const std::string& w, uint32_t* v) {
  if (w.empty() || w.size() > 10) return false;
  uint64_t x = 0;
  for (char ch : w) {
    if (ch < '0' || ch > '9') return false;
    x = x * 10 + (uint64_t)(ch - '0');
  }
  if (x > 0xFFFFFFFFull) return false;
  *v = (uint32_t)x;
  return true;
}

bool read_u32(FILE* f, uint32_t* v) { return fread(v, 4, 1, f) == 1; }
bool write_u32(FILE* f, uint32_t v) { return fwrite(&v, 4, 1, f) == 1; }

}  // namespace

extern "C" svo_status svo_scene_open_single(const char* base, svo_scene** out) {
  if (!base || !out) { set_error("svo_scene_open_single: null argument"); return SVO_ERR_ARG; }
  *out = nullptr;
  const std::string p = std::string(base) + ".svo";
  File f(fopen(p.c_str(), "rb"));
  if (!f) { set_error("could not open %s", p.c_str()); return SVO_ERR_IO; }
  std::string w;
  if (!read_word(f.get(), &w)) { set_error("%s: no header", p.c_str()); return SVO_ERR_FORMAT; }
  const bool v1 = (w == "v1");
  if (v1 && !read_word(f.get(), &w)) { set_error("%s: truncated header", p.c_str()); return SVO_ERR_FORMAT; }
  uint32_t depth = 0, n = 0;
  if (!parse_u32(w, &depth)) { set_error("%s: bad depth word '%s'", p.c_str(), w.c_str()); return SVO_ERR_FORMAT; }
  if (!read_word(f.get(), &w) || !parse_u32(w, &n)) { set_error("%s: bad node count", p.c_str()); return SVO_ERR_FORMAT; }
  if (n == 0 || n > (1u << 29)) { set_error("%s: node count %u out of range", p.c_str(), n); return SVO_ERR_FORMAT; }
  std::vector<svo_node> nodes(n);
  if (fread(nodes.data(), sizeof(svo_node), n, f.get()) != n) { set_error("%s: truncated nodes", p.c_str()); return SVO_ERR_FORMAT; }
  std::vector<uint32_t> aidx(n);
  if (v1) {
    if (fread(aidx.data(), 4, n, f.get()) != n) { set_error("%s: truncated attribute indices", p.c_str()); return SVO_ERR_FORMAT; }
  } else {
    for (uint32_t i = 0; i < n; ++i) aidx[i] = i;
  }
  f.reset();

  const std::string pa = std::string(base) + ".ia";
  File g(fopen(pa.c_str(), "rb"));
  if (!g) { set_error("could not open %s", pa.c_str()); return SVO_ERR_IO; }
  uint32_t blocks = 0, np = 0, wpp = 0;
  if (!read_u32(g.get(), &blocks) || blocks != 1 || !read_u32(g.get(), &np) || !read_u32(g.get(), &wpp) || wpp != 2) {
    set_error("%s: expected one block of 2-word packages", pa.c_str());
    return SVO_ERR_FORMAT;
  }
  std::vector<uint32_t> packages(2 * (size_t)np);
  if (fread(packages.data(), 4, packages.size(), g.get()) != packages.size()) { set_error("%s: truncated", pa.c_str()); return SVO_ERR_FORMAT; }

  std::unique_ptr<svo_scene> sc(new svo_scene());
  TreeletData T;
  T.nodes.resize(n);
  T.attribs.resize(2 * (size_t)n);
  uint32_t leaves = 0;
  for (uint32_t i = 0; i < n; ++i) {
    svo_node nd = nodes[i];
    const uint32_t vm = (nd.masks >> kValidShift) & 255u;
    if ((nd.masks & 255u) & ~vm) { set_error("%s: node %u has leaf bits outside its valid mask", p.c_str(), i); return SVO_ERR_FORMAT; }
    leaves += (uint32_t)__builtin_popcount(nd.masks & 255u);
    if (vm != 0) {
      const uint64_t last = (uint64_t)nd.first_child + (uint64_t)__builtin_popcount(vm) - 1;
      if (nd.first_child <= i || last >= n) { set_error("%s: node %u children out of range", p.c_str(), i); return SVO_ERR_FORMAT; }
      nd.first_child -= i;  // absolute -> relative
    }
    T.nodes[i] = nd;
    if (aidx[i] >= np) { set_error("%s: node %u attribute index %u >= %u packages", p.c_str(), i, aidx[i], np); return SVO_ERR_FORMAT; }
    T.attribs[2 * (size_t)i] = packages[2 * (size_t)aidx[i]];
    T.attribs[2 * (size_t)i + 1] = packages[2 * (size_t)aidx[i] + 1];
  }
  const uint64_t tb = std::max<uint64_t>(128, ((uint64_t)n * 8 + 7) / 8 * 8);
  if (tb > 0xFFFFFFF8ull) { set_error("%s: too many nodes for one treelet", p.c_str()); return SVO_ERR_FORMAT; }
  T.h = svo_treelet_header{n, leaves, (uint32_t)tb, 0, 0, 0, 0, 0};
  sc->treelet_bytes = (uint32_t)tb;
  sc->nodes_per_treelet = (uint32_t)(tb / 8);
  sc->treelets.push_back(std::move(T));
  if (np > 0) {
    sc->has_package0 = true;
    sc->package0[0] = packages[0];
    sc->package0[1] = packages[1];
  }
  std::string why;
  const int d = scene_depth(*sc, &why, &sc->inner_depth);
  if (d < 0) { set_error("%s: %s", p.c_str(), why.c_str()); return SVO_ERR_FORMAT; }
  if ((uint32_t)d != depth) { set_error("%s: header depth %u, tree depth %d", p.c_str(), depth, d); return SVO_ERR_FORMAT; }
  sc->max_depth = (uint32_t)d;
  sc->tree_nodes = n;
  sc->tree_leaves = leaves;
  *out = sc.release();
  return SVO_OK;
}

extern "C" svo_status svo_scene_write_single(const svo_scene* sc, const char* base, uint32_t format,
                                             const uint32_t* attrib_index) {
  if (!sc || !base) { set_error("svo_scene_write_single: null argument"); return SVO_ERR_ARG; }
  if (format != SVO_SINGLE_SVOLOADER && format != SVO_SINGLE_V1) { set_error("unknown single-SVO format %u", format); return SVO_ERR_ARG; }
  if (sc->treelets.size() != 1) { set_error("a single-SVO file holds one treelet, the scene has %zu", sc->treelets.size()); return SVO_ERR_ARG; }
  if (attrib_index && format != SVO_SINGLE_V1) { set_error("attribute indices need the v1 format"); return SVO_ERR_ARG; }
  const TreeletData& T = sc->treelets[0];
  const uint32_t n = T.h.num_nodes;
  std::vector<uint32_t> idx(n);
  std::vector<uint32_t> packages(2 * (size_t)n, 0);
  std::vector<uint8_t> used(n, 0);
  for (uint32_t i = 0; i < n; ++i) {
    idx[i] = attrib_index ? attrib_index[i] : i;
    if (idx[i] >= n || used[idx[i]]) { set_error("attribute indices must be a permutation of 0..%u", n - 1); return SVO_ERR_ARG; }
    used[idx[i]] = 1;
    packages[2 * (size_t)idx[i]] = T.attribs[2 * (size_t)i];
    packages[2 * (size_t)idx[i] + 1] = T.attribs[2 * (size_t)i + 1];
  }
  const std::string p = std::string(base) + ".svo";
  File f(fopen(p.c_str(), "wb"));
  if (!f) { set_error("cannot open %s for writing", p.c_str()); return SVO_ERR_IO; }
  char head[64];
  const int hl = format == SVO_SINGLE_V1 ? snprintf(head, sizeof head, "v1 %u %u ", sc->max_depth, n)
                                         : snprintf(head, sizeof head, "%u %u ", sc->max_depth, n);
  bool ok = fwrite(head, 1, (size_t)hl, f.get()) == (size_t)hl;
  for (uint32_t i = 0; i < n && ok; ++i) {
    svo_node nd = T.nodes[i];
    if ((nd.masks >> kValidShift) & 255u) nd.first_child += i;  // relative -> absolute
    ok = fwrite(&nd, sizeof nd, 1, f.get()) == 1;
  }
  if (ok && format == SVO_SINGLE_V1) ok = fwrite(idx.data(), 4, n, f.get()) == n;
  if (!ok) { set_error("write failed: %s", p.c_str()); return SVO_ERR_IO; }
  f.reset();
  const std::string pa = std::string(base) + ".ia";
  File g(fopen(pa.c_str(), "wb"));
  if (!g) { set_error("cannot open %s for writing", pa.c_str()); return SVO_ERR_IO; }
  ok = write_u32(g.get(), 1) && write_u32(g.get(), n) && write_u32(g.get(), 2) &&
       fwrite(packages.data(), 4, packages.size(), g.get()) == packages.size();
  if (!ok) { set_error("write failed: %s", pa.c_str()); return SVO_ERR_IO; }
  return SVO_OK;
}
