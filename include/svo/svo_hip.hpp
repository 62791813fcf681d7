// C++ face of libsvo_hip.so for libsvo code: the class names and methods of
// the reference's OpenCL path, header-only over the C-ABI of svo_hip.h.
//
//   TreeletMemoryManagerHip  <- TreeletMemoryManager + TreeletMemoryManagerCl
//                               (src/TreeletMemoryManager.h:97-157,
//                               src/TreeletMemoryManagerCl.h:69-93)
//   RenderPassAnalyseHip     <- RenderPassAnalyse (src/RenderPassAnalyse.h:72-103)
//   renderToBuffer()         <- the renderToBuffer enqueue of
//                               raycast_serialized_opencl.cpp:729-749
//
// Differences a caller sees: errors throw svo::HipError (status + message)
// instead of printing and exit(); device buffers are plain device pointers
// (svo_frame_out) instead of libbencl gids; cameras are the FP32 uniforms of
// svo_camera instead of gloost cameras: the far-plane lower-left corner, the
// one-pixel steps along the far plane and the eye position, all in model
// space, plus tScaleRatio and the frame size (the kernel arguments that
// raycast_serialized_opencl.cpp:729-749 sets).
#ifndef SVO_HIP_HPP
#define SVO_HIP_HPP

#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../svo_hip.h"

namespace svo {

class HipError : public std::runtime_error {
 public:
  HipError(svo_status status, const std::string& where)
      : std::runtime_error(where + ": " + svo_last_error()), _status(status) {}
  svo_status status() const { return _status; }

 private:
  svo_status _status;
};

inline void check(svo_status s, const char* where) {
  if (s != SVO_OK) throw HipError(s, where);
}

class RenderPassAnalyseHip;

// The treelet bundle on the host plus its HBM incore pools.  device = -1
// keeps residency bookkeeping only (no pools, no rendering).
class TreeletMemoryManagerHip {
 public:
  TreeletMemoryManagerHip(const std::string& svoBasePath, unsigned long long incoreBufferSizeInByte, int device = 0) {
    check(svo_scene_open(svoBasePath.c_str(), &_scene), "TreeletMemoryManagerHip: loadFromFile");
    const svo_status s = svo_mm_create(_scene, incoreBufferSizeInByte, device, &_mm);
    if (s != SVO_OK) {
      svo_scene_destroy(_scene);
      throw HipError(s, "TreeletMemoryManagerHip: initClBuffer");
    }
  }
  ~TreeletMemoryManagerHip() {
    svo_mm_destroy(_mm);
    svo_scene_destroy(_scene);
  }
  TreeletMemoryManagerHip(const TreeletMemoryManagerHip&) = delete;
  TreeletMemoryManagerHip& operator=(const TreeletMemoryManagerHip&) = delete;

  unsigned getTreeletSizeInByte() const { return sceneInfo().treelet_bytes; }
  unsigned getNumNodesPerTreelet() const { return sceneInfo().nodes_per_treelet; }
  unsigned getNumTreelets() const { return sceneInfo().num_treelets; }

  svo_treelet_header getTreelet(unsigned gid) const {
    svo_treelet_header h;
    check(svo_scene_get_treelet(_scene, gid, &h, nullptr, nullptr), "getTreelet");
    return h;
  }

  bool insertTreeletIntoIncoreBuffer(unsigned gid, float error = 0.0f) {
    return svo_mm_insert(_mm, gid, error) == SVO_OK;
  }
  bool removeTreeletFromIncoreBuffer(unsigned gid) { return svo_mm_remove(_mm, gid) == SVO_OK; }
  void resetIncoreBuffer() { check(svo_mm_reset(_mm), "resetIncoreBuffer"); }
  void makeAllResident() { check(svo_mm_make_resident_all(_mm), "makeAllResident"); }

  // updateClientSideIncoreBuffer(RenderPassAnalyse*): applies the analyse
  // pass's feedback (visibility, page-ins, evictions).
  svo_feedback_stats updateClientSideIncoreBuffer(RenderPassAnalyseHip* analyse, void* stream = nullptr);

  svo_upload_stats updateDeviceMemory(void* stream = nullptr) {
    svo_upload_stats st;
    check(svo_mm_update_device(_mm, stream, &st), "updateDeviceMemory");
    return st;
  }

  // getSlots()[i]._treeletGid
  std::vector<int> getSlots() const {
    const svo_mm_info i = info();
    std::vector<int> out(i.num_slots);
    for (unsigned s = 0; s < i.num_slots; ++s) {
      int32_t g = -1;
      check(svo_mm_slot_treelet(_mm, s, &g), "getSlots");
      out[s] = g;
    }
    return out;
  }

  // Copy of the host incore buffer (getIncoreBuffer) and its colours and
  // normals (getAttributeIncoreBuffer).
  std::vector<svo_node> getIncoreBuffer(std::vector<uint32_t>* attribs = nullptr) const {
    const svo_mm_info i = info();
    std::vector<svo_node> nodes(i.pool_nodes);
    std::vector<uint32_t> a(2 * i.pool_nodes);
    check(svo_mm_read_host_incore(_mm, 0, i.pool_nodes, nodes.data(), a.data()), "getIncoreBuffer");
    if (attribs) *attribs = std::move(a);
    return nodes;
  }

  svo_mm_info info() const {
    svo_mm_info i;
    check(svo_mm_get_info(_mm, &i), "svo_mm_get_info");
    return i;
  }
  svo_mm* handle() { return _mm; }
  const svo_scene* scene() const { return _scene; }

 private:
  svo_scene_info sceneInfo() const {
    svo_scene_info i;
    check(svo_scene_get_info(_scene, &i), "svo_scene_get_info");
    return i;
  }
  svo_scene* _scene = nullptr;
  svo_mm* _mm = nullptr;
};

// RenderPassAnalyse(memoryManager, bufferWidth, bufferHeight): the analyse
// pass traces a frame of bufferWidth x bufferHeight rays over the camera's
// frustum; its records stay on the device until the memory manager's
// updateClientSideIncoreBuffer.
class RenderPassAnalyseHip {
 public:
  RenderPassAnalyseHip(TreeletMemoryManagerHip* memoryManager, unsigned bufferWidth, unsigned bufferHeight)
      : _mm(memoryManager), _w(bufferWidth), _h(bufferHeight) {}

  // performAnalysePass(..., frameBufferToFeedbackBufferRatio, drawOnly):
  // `camera` is the full frame's; the feedback frame is its 1/ratio per axis,
  // moved by (jitterX, jitterY) half feedback pixels when rumble is on
  // (RenderPassAnalyse.cpp:172-184).
  void performAnalysePass(const svo_camera& camera, unsigned frameBufferToFeedbackBufferRatio, float jitterX = 0.0f,
                          float jitterY = 0.0f, int policy = SVO_POLICY_TREELET, void* stream = nullptr) {
    if (camera.width / frameBufferToFeedbackBufferRatio != _w || camera.height / frameBufferToFeedbackBufferRatio != _h)
      throw std::invalid_argument("RenderPassAnalyseHip: camera / ratio does not match the feedback buffer size");
    const float jx = _rumble ? jitterX : 0.0f, jy = _rumble ? jitterY : 0.0f;
    check(svo_mm_analyse(_mm->handle(), &camera, frameBufferToFeedbackBufferRatio, jx, jy, policy, stream),
          "performAnalysePass");
  }
  // getVisibleOldTreeletsGids / getVisibleNewTreeletsGids: the visible sets
  // of the records since the last update (svo_mm_read_feedback); the next
  // updateClientSideIncoreBuffer applies these records.
  struct TreeletGidAndError {
    unsigned _treeletGid;
    float _error;
  };
  std::vector<TreeletGidAndError> getVisibleOldTreeletsGids(void* stream = nullptr) { return visible(stream, true); }
  std::vector<TreeletGidAndError> getVisibleNewTreeletsGids(void* stream = nullptr) { return visible(stream, false); }

  void setEnableRumble(bool onOrOff) { _rumble = onOrOff; }
  bool getEnableRumble() const { return _rumble; }
  TreeletMemoryManagerHip* memoryManager() { return _mm; }

 private:
  std::vector<TreeletGidAndError> visible(void* stream, bool old) {
    uint32_t n_old = 0, n_new = 0;
    check(svo_mm_read_feedback(_mm->handle(), stream, nullptr, nullptr, 0, &n_old, nullptr, nullptr, 0, &n_new),
          "getVisibleTreeletsGids");
    std::vector<uint32_t> og(n_old), ng(n_new);
    std::vector<float> oe(n_old), ne(n_new);
    check(svo_mm_read_feedback(_mm->handle(), stream, og.data(), oe.data(), n_old, &n_old, ng.data(), ne.data(), n_new,
                               &n_new),
          "getVisibleTreeletsGids");
    std::vector<TreeletGidAndError> out;
    const std::vector<uint32_t>& g = old ? og : ng;
    const std::vector<float>& e = old ? oe : ne;
    for (size_t i = 0; i < g.size(); ++i) out.push_back(TreeletGidAndError{g[i], e[i]});
    return out;
  }

  TreeletMemoryManagerHip* _mm;
  unsigned _w, _h;
  bool _rumble = true;
};

inline svo_feedback_stats TreeletMemoryManagerHip::updateClientSideIncoreBuffer(RenderPassAnalyseHip* analyse,
                                                                                void* stream) {
  if (analyse && analyse->memoryManager() != this)
    throw std::invalid_argument("updateClientSideIncoreBuffer: the analyse pass belongs to another manager");
  svo_feedback_stats st;
  check(svo_mm_apply_feedback(_mm, stream, &st), "updateClientSideIncoreBuffer");
  return st;
}

// The renderToBuffer enqueue: one frame (or many, see svo_render) into
// caller-owned device outputs, asynchronously on `stream`.
inline void renderToBuffer(TreeletMemoryManagerHip& mm, const svo_camera& camera, const svo_frame_out& out,
                           int policy = SVO_POLICY_TREELET, void* stream = nullptr) {
  check(svo_render(mm.handle(), &camera, 1, &out, policy, stream), "renderToBuffer");
}

}  // namespace svo

#endif  // SVO_HIP_HPP
