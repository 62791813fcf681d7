// Drives the GPU path through the C++ classes of include/svo/svo_hip.hpp the
// way the OpenCL app's frame loop does (analyse pass, visible sets, client
// update, device update, renderToBuffer), on device 0:
//   argv[1] scene base path, argv[2] incore slots, argv[3] frames,
//   argv[4] output file for the last frame (t, node, rgba of W x H pixels).
// Prints one "key value..." line per step for the Python test to compare.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "svo/svo_hip.hpp"

namespace {
constexpr unsigned kW = 160, kH = 100, kDivide = 4;

svo_camera orbit_camera(int f) {
  const double a = 0.4 * f;
  const double eye[3] = {1.6 * std::cos(a), 0.9, 1.6 * std::sin(a)};
  const double target[3] = {0.0, -0.1, 0.0};
  const double up[3] = {0.0, 1.0, 0.0};
  svo_camera cam;
  svo::check(svo_camera_look_at(eye, target, up, 55.0, kW, kH, 20.0, 0.25, &cam), "svo_camera_look_at");
  return cam;
}
}  // namespace

int main(int argc, char** argv) {
  if (argc != 5) return 2;
  svo_scene* probe = nullptr;
  svo::check(svo_scene_open(argv[1], &probe), "svo_scene_open");
  svo_scene_info si;
  svo::check(svo_scene_get_info(probe, &si), "svo_scene_get_info");
  svo_scene_destroy(probe);
  const unsigned slots = (unsigned)std::atoi(argv[2]);
  const int frames = std::atoi(argv[3]);

  svo::TreeletMemoryManagerHip mm(argv[1], (unsigned long long)si.treelet_bytes * slots, 0);
  svo::RenderPassAnalyseHip analyse(&mm, kW / kDivide, kH / kDivide);
  svo_camera cam{};
  for (int f = 0; f < frames; ++f) {
    cam = orbit_camera(f);
    const float jx = 0.25f * (float)(f % 3) - 0.25f, jy = 0.5f - 0.25f * (float)(f % 5);
    analyse.performAnalysePass(cam, kDivide, jx, jy);
    const auto vold = analyse.getVisibleOldTreeletsGids();
    const auto vnew = analyse.getVisibleNewTreeletsGids();
    double eo = 0.0, en = 0.0;
    for (const auto& e : vold) eo += e._treeletGid * (double)e._error;
    for (const auto& e : vnew) en += e._treeletGid * (double)e._error;
    std::printf("visible %d %zu %zu %.9g %.9g\n", f, vold.size(), vnew.size(), eo, en);
    const svo_feedback_stats st = mm.updateClientSideIncoreBuffer(&analyse);
    const svo_upload_stats up = mm.updateDeviceMemory();
    std::printf("update %d %u %u %u %llu\n", f, st.inserted, st.evicted, st.resident_treelets,
                (unsigned long long)up.bytes);
  }
  std::printf("slots");
  for (int g : mm.getSlots()) std::printf(" %d", g);
  std::printf("\n");

  const size_t n = (size_t)kW * kH;
  void *t = nullptr, *node = nullptr, *tree = nullptr, *rgba = nullptr;
  svo::check(svo_device_malloc(&t, n * 4), "svo_device_malloc");
  svo::check(svo_device_malloc(&node, n * 4), "svo_device_malloc");
  svo::check(svo_device_malloc(&tree, n * 4), "svo_device_malloc");
  svo::check(svo_device_malloc(&rgba, n * 4), "svo_device_malloc");
  const svo_frame_out out = {(float*)t, (uint32_t*)node, (uint32_t*)tree, (uint32_t*)rgba, nullptr, nullptr};
  svo::renderToBuffer(mm, cam, out);
  svo::check(svo_device_synchronize(), "svo_device_synchronize");
  std::vector<uint32_t> host(3 * n);
  svo::check(svo_memcpy_d2h(host.data(), t, n * 4), "svo_memcpy_d2h");
  svo::check(svo_memcpy_d2h(host.data() + n, node, n * 4), "svo_memcpy_d2h");
  svo::check(svo_memcpy_d2h(host.data() + 2 * n, rgba, n * 4), "svo_memcpy_d2h");
  for (void* p : {t, node, tree, rgba}) svo::check(svo_device_free(p), "svo_device_free");
  FILE* fo = std::fopen(argv[4], "wb");
  if (!fo || std::fwrite(host.data(), 4, host.size(), fo) != host.size()) return 3;
  std::fclose(fo);
  std::printf("frame %u %u\n", kW, kH);

  // The device pools (getClIncoreBuffer / getClAttributeIncoreBuffers) equal
  // the host incore buffers; after one more frame's update they still do
  // (normals included); initClBuffer restores pools that were overwritten.
  auto pools_equal = [&mm]() {
    const svo_device_pools p = mm.getDevicePools();
    unsigned long long cnt = 0;
    const svo_node* hn = mm.getIncoreBufferData(&cnt);
    const uint32_t *hnrm = nullptr, *hcol = nullptr;
    mm.getAttributeIncoreBuffer(&hnrm, &hcol);
    std::vector<svo_node> dn(cnt);
    std::vector<uint32_t> dnrm(cnt), dcol(cnt);
    std::vector<int32_t> dslot(p.num_slots);
    svo::check(svo_memcpy_d2h(dn.data(), p.nodes, cnt * 8), "svo_memcpy_d2h");
    svo::check(svo_memcpy_d2h(dnrm.data(), p.normals, cnt * 4), "svo_memcpy_d2h");
    svo::check(svo_memcpy_d2h(dcol.data(), p.colors, cnt * 4), "svo_memcpy_d2h");
    svo::check(svo_memcpy_d2h(dslot.data(), p.slot_gid, p.num_slots * 4ull), "svo_memcpy_d2h");
    bool same = cnt == p.pool_nodes && p.device == 0 && p.nodes_per_treelet == mm.getNumNodesPerTreelet();
    for (unsigned long long i = 0; same && i < cnt; ++i)
      same = dn[i].first_child == hn[i].first_child && dn[i].masks == hn[i].masks && dnrm[i] == hnrm[i] &&
             dcol[i] == hcol[i];
    const std::vector<int> slots_now = mm.getSlots();
    for (unsigned s = 0; same && s < p.num_slots; ++s) same = dslot[s] == slots_now[s];
    return same;
  };
  std::printf("pools_equal %d\n", pools_equal() ? 1 : 0);
  cam = orbit_camera(frames);
  analyse.performAnalysePass(cam, kDivide, 0.0f, 0.0f);
  const svo_feedback_stats st2 = mm.updateClientSideIncoreBuffer(&analyse);
  mm.updateDeviceMemory();
  std::printf("pools_after_update %u %d\n", st2.inserted, pools_equal() ? 1 : 0);
  const svo_device_pools p = mm.getDevicePools();
  svo::check(svo_memset_device(p.nodes, 0xAB, p.pool_nodes * 8), "svo_memset_device");
  svo::check(svo_memset_device(p.normals, 0xCD, p.pool_nodes * 4), "svo_memset_device");
  std::printf("pools_clobbered %d\n", pools_equal() ? 1 : 0);
  mm.initClBuffer();
  std::printf("pools_after_init %d\n", pools_equal() ? 1 : 0);
  return 0;
}
