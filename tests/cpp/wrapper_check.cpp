// Exercises the C++ wrappers (include/svo/svo_hip.hpp) the way libsvo code
// would, on a host-only manager (device -1): argv[1] = scene base path,
// argv[2] = incore slots.  Prints one "key value..." line per check; the
// Python test compares them with the Python API on the same scene.
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "svo/svo_hip.hpp"

int main(int argc, char** argv) {
  if (argc != 3) return 2;
  // a pool that holds no treelet: the constructor throws with a status
  try {
    svo::TreeletMemoryManagerHip none(argv[1], 8ull, -1);
    std::printf("error_status none\n");
  } catch (const svo::HipError& e) {
    std::printf("error_status %d\n", e.status());
  }
  try {
    svo::TreeletMemoryManagerHip missing(std::string(argv[1]) + "_missing", 1ull << 20, -1);
    std::printf("missing_status none\n");
  } catch (const svo::HipError& e) {
    std::printf("missing_status %d\n", e.status());
  }
  const unsigned slots = (unsigned)std::atoi(argv[2]);
  // the byte size of the incore buffer, as TreeletMemoryManager takes it
  svo_scene* probe = nullptr;
  svo::check(svo_scene_open(argv[1], &probe), "svo_scene_open");
  svo_scene_info si;
  svo::check(svo_scene_get_info(probe, &si), "svo_scene_get_info");
  svo_scene_destroy(probe);
  const unsigned long long tb = si.treelet_bytes;
  svo::TreeletMemoryManagerHip mm(argv[1], tb * slots, -1);
  std::printf("treelets %u\n", mm.getNumTreelets());
  std::printf("nodes_per_treelet %u\n", mm.getNumNodesPerTreelet());
  // insert the root's children while slots last
  unsigned inserted = 0;
  for (unsigned g = 1; g < mm.getNumTreelets(); ++g)
    if (mm.getTreelet(g).parent_gid == 0 && mm.insertTreeletIntoIncoreBuffer(g, 1.0f)) ++inserted;
  std::printf("inserted %u\n", inserted);
  std::printf("slots");
  for (int g : mm.getSlots()) std::printf(" %d", g);
  std::printf("\n");
  // inserting an already resident treelet is refused, not fatal
  std::printf("reinsert %d\n", mm.insertTreeletIntoIncoreBuffer((unsigned)mm.getSlots()[1]) ? 1 : 0);
  const int removed = mm.getSlots()[1];
  std::printf("remove %d %d\n", removed, mm.removeTreeletFromIncoreBuffer((unsigned)removed) ? 1 : 0);
  std::vector<uint32_t> attr;
  const std::vector<svo_node> nodes = mm.getIncoreBuffer(&attr);
  unsigned long long sum = 0;
  for (const svo_node& n : nodes) sum += n.first_child * 3ull + n.masks;
  std::printf("incore_checksum %llu %zu\n", sum, nodes.size());
  // device-side calls on a host-only manager throw
  try {
    mm.updateClientSideIncoreBuffer(nullptr);
    std::printf("feedback_status none\n");
  } catch (const svo::HipError& e) {
    std::printf("feedback_status %d\n", e.status());
  }
  mm.resetIncoreBuffer();
  std::printf("after_reset %u\n", mm.info().resident_treelets);
  return 0;
}
