// C++ face of libsvo_hip.so for libsvo code: the class names and methods of
// the reference's OpenCL path, header-only over the C-ABI of svo_hip.h.
//
//   TreeletMemoryManagerHip  <- TreeletMemoryManager + TreeletMemoryManagerCl
//                               (src/TreeletMemoryManager.h:97-157,
//                               src/TreeletMemoryManagerCl.h:69-93): getTreelet(s),
//                               getTreeletSizeInByte, getNumNodesPerTreelet,
//                               getIncoreBuffer, getAttributeIncoreBuffer,
//                               updateClientSideIncoreBuffer, insert/remove,
//                               freeIncorePosition, updateDeviceMemory,
//                               resetIncoreBuffer, getFreeIncoreSlotStack,
//                               getIncoreSlotsToUpload, getSlots, initClBuffer,
//                               the device pools and device
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

  // getTreelets(): every treelet's header and its slot (Treelet::getSlotGid,
  // -1 = not resident).
  struct TreeletState {
    svo_treelet_header header;
    int slot;
  };
  std::vector<TreeletState> getTreelets() const {
    const unsigned n = getNumTreelets();
    std::vector<svo_treelet_header> h(n);
    std::vector<int32_t> s(n);
    check(svo_mm_get_treelets(_mm, 0, n, h.data(), s.data()), "getTreelets");
    std::vector<TreeletState> out(n);
    for (unsigned i = 0; i < n; ++i) out[i] = TreeletState{h[i], s[i]};
    return out;
  }

  bool insertTreeletIntoIncoreBuffer(unsigned gid, float error = 0.0f) {
    return svo_mm_insert(_mm, gid, error) == SVO_OK;
  }
  bool removeTreeletFromIncoreBuffer(unsigned gid) { return svo_mm_remove(_mm, gid) == SVO_OK; }
  void resetIncoreBuffer() { check(svo_mm_reset(_mm), "resetIncoreBuffer"); }
  // freeIncorePosition(): frees an invisible treelet found by the probe rule
  // and everything resident below it; false when none was found.
  bool freeIncorePosition(int* freedGid = nullptr) {
    int32_t g = -1;
    check(svo_mm_free_incore_position(_mm, &g), "freeIncorePosition");
    if (freedGid) *freedGid = g;
    return g >= 0;
  }
  // getFreeIncoreSlotStack(): bottom of the stack first (back() is the top).
  std::vector<unsigned> getFreeIncoreSlotStack() const {
    return listOf(svo_mm_get_free_slot_stack, "getFreeIncoreSlotStack");
  }
  // getIncoreSlotsToUpload(): slots changed since the last upload, ascending.
  std::vector<unsigned> getIncoreSlotsToUpload() const {
    return listOf(svo_mm_get_slots_to_upload, "getIncoreSlotsToUpload");
  }
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

  // getSlots() with the whole SlotInfo (gid, visibility, error).
  std::vector<svo_slot_info> getSlotInfos() const {
    std::vector<svo_slot_info> out(info().num_slots);
    check(svo_mm_get_slots(_mm, 0, (uint32_t)out.size(), out.data()), "getSlotInfos");
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
  // The host incore buffer itself (the reference returns a reference to its
  // vector): pool_nodes nodes, valid while the manager lives.
  const svo_node* getIncoreBufferData(unsigned long long* count = nullptr) const {
    const svo_node* p = nullptr;
    uint64_t n = 0;
    check(svo_mm_get_incore_buffer(_mm, &p, &n), "getIncoreBuffer");
    if (count) *count = n;
    return p;
  }
  // The attribute incore buffer itself, as its normal and colour arrays.
  void getAttributeIncoreBuffer(const uint32_t** normals, const uint32_t** colors,
                                unsigned long long* count = nullptr) const {
    uint64_t n = 0;
    check(svo_mm_get_attribute_incore_buffer(_mm, normals, colors, &n), "getAttributeIncoreBuffer");
    if (count) *count = n;
  }

  // TreeletMemoryManagerCl's device side: initClBuffer() uploads the whole
  // incore buffer; getClIncoreBuffer() / getClAttributeIncoreBuffers() give
  // the HBM pools (device pointers in place of libbencl gids) and
  // getDevice() the device (in place of getContext()).  Once the pools were
  // taken, every update also sends the normals.
  void initClBuffer(void* stream = nullptr) { check(svo_mm_init_device_buffer(_mm, stream), "initClBuffer"); }
  svo_device_pools getDevicePools() {
    svo_device_pools p;
    check(svo_mm_get_device_pools(_mm, &p), "getDevicePools");
    return p;
  }
  const svo_node* getClIncoreBuffer() { return static_cast<const svo_node*>(getDevicePools().nodes); }
  void getClAttributeIncoreBuffers(const uint32_t** normals, const uint32_t** colors) {
    const svo_device_pools p = getDevicePools();
    if (normals) *normals = static_cast<const uint32_t*>(p.normals);
    if (colors) *colors = static_cast<const uint32_t*>(p.colors);
  }
  int getDevice() const { return (int)info().device; }

  svo_mm_info info() const {
    svo_mm_info i;
    check(svo_mm_get_info(_mm, &i), "svo_mm_get_info");
    return i;
  }
  svo_mm* handle() { return _mm; }
  const svo_scene* scene() const { return _scene; }

 private:
  template <typename Fn>
  std::vector<unsigned> listOf(Fn fn, const char* where) const {
    uint32_t n = 0;
    check(fn(_mm, nullptr, 0, &n), where);
    std::vector<uint32_t> v(n);
    check(fn(_mm, v.data(), n, &n), where);
    return std::vector<unsigned>(v.begin(), v.end());
  }
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
