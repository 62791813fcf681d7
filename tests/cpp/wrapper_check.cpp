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

  // The rest of the manager surface on a second manager: two children of the
  // root, then their children while slots last.
  svo::TreeletMemoryManagerHip m2(argv[1], tb * slots, -1);
  const std::vector<svo::TreeletMemoryManagerHip::TreeletState> tl = m2.getTreelets();
  std::vector<unsigned> kids;
  for (unsigned g = 1; g < tl.size() && kids.size() < 2; ++g)
    if (tl[g].header.parent_gid == 0 && m2.insertTreeletIntoIncoreBuffer(g, 1.0f)) kids.push_back(g);
  for (unsigned g = 1; g < tl.size(); ++g)
    if (tl[g].header.parent_gid == kids.at(0) || tl[g].header.parent_gid == kids.at(1))
      m2.insertTreeletIntoIncoreBuffer(g, 2.0f);  // refused once the pool is full
  auto print_list = [](const char* key, const std::vector<unsigned>& v) {
    std::printf("%s", key);
    for (unsigned x : v) std::printf(" %u", x);
    std::printf("\n");
  };
  print_list("free_stack", m2.getFreeIncoreSlotStack());
  print_list("to_upload", m2.getIncoreSlotsToUpload());
  std::printf("treelet_slots");
  for (const auto& t : m2.getTreelets())
    if (t.slot >= 0) std::printf(" %u:%d", t.header.gid, t.slot);
  std::printf("\n");
  // nine updates without records: every slot but the root's ends at
  // visibility -1, so grandchildren of the root become freeable
  for (int i = 0; i < 9; ++i) {
    svo_feedback_stats st;
    svo::check(svo_mm_apply_records(m2.handle(), nullptr, nullptr, 0, nullptr, nullptr, 0, &st), "apply_records");
  }
  std::printf("slot_infos");
  for (const svo_slot_info& si : m2.getSlotInfos()) std::printf(" %d/%d/%g", si.treelet_gid, si.visibility, si.error);
  std::printf("\n");
  int freed = -1;
  const bool ok = m2.freeIncorePosition(&freed);
  std::printf("freed %d %d\n", ok ? 1 : 0, freed);
  std::printf("slots2");
  for (int g : m2.getSlots()) std::printf(" %d", g);
  std::printf("\n");
  print_list("free_stack2", m2.getFreeIncoreSlotStack());
  print_list("to_upload2", m2.getIncoreSlotsToUpload());
  // the borrowed buffers are the ones getIncoreBuffer copies
  unsigned long long n = 0, na = 0;
  const svo_node* live = m2.getIncoreBufferData(&n);
  const uint32_t *nrm = nullptr, *col = nullptr;
  m2.getAttributeIncoreBuffer(&nrm, &col, &na);
  std::vector<uint32_t> attr2;
  const std::vector<svo_node> copy = m2.getIncoreBuffer(&attr2);
  bool same = n == copy.size() && na == n;
  for (size_t i = 0; same && i < copy.size(); ++i)
    same = live[i].first_child == copy[i].first_child && live[i].masks == copy[i].masks && nrm[i] == attr2[2 * i] &&
           col[i] == attr2[2 * i + 1];
  std::printf("borrowed_equal %d %llu\n", same ? 1 : 0, n);
  // device-side calls on a host-only manager throw
  try {
    m2.getDevicePools();
    std::printf("pools_status none\n");
  } catch (const svo::HipError& e) {
    std::printf("pools_status %d\n", e.status());
  }
  return 0;
}
