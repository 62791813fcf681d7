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
  return 0;
}
