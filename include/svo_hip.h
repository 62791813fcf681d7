/*
 * svo_hip.h -- C-ABI of the MI355X serialized-SVO treelet ray caster.
 *
 * Plain C, plain pointers and sizes.  Every function returns an svo_status
 * (0 = SVO_OK) and leaves a thread-local message for svo_last_error() on
 * failure; nothing in the library calls exit().
 *
 * Handles are opaque and not thread-safe; device work is ordered on the
 * hipStream_t passed in (as void*, NULL = the default stream).  Device pools
 * are owned by the handles, frame outputs are owned by the caller.
 *
 * Reference interfaces replaced (paths relative to the libsvo tree):
 *   svo_scene_*      Treelet / TreeletBuildManager / TreeletMemoryManager
 *                    loadFromFile (src/Treelet.cpp:403-478,
 *                    src/TreeletBuildManager.cpp:286-342,
 *                    src/TreeletMemoryManager.cpp:229-291)
 *   svo_mm_*         TreeletMemoryManager + TreeletMemoryManagerCl
 *                    (src/TreeletMemoryManager.h:97-157,
 *                    src/TreeletMemoryManagerCl.h:69-93)
 *   svo_render       kernel renderToBuffer
 *                    (opencl/treeletRenderer_rgba.cl:610-655) as enqueued by
 *                    projects/raycast_serialized_opencl/src/
 *                    raycast_serialized_opencl.cpp:729-749
 */
#ifndef SVO_HIP_H
#define SVO_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ status */

typedef int svo_status;
enum {
  SVO_OK = 0,
  SVO_ERR_ARG = 1,      /* bad argument (null, out of range)               */
  SVO_ERR_IO = 2,       /* file could not be opened / read / written       */
  SVO_ERR_FORMAT = 3,   /* malformed .svo / .ia content                    */
  SVO_ERR_HIP = 4,      /* a HIP runtime call failed                       */
  SVO_ERR_NOMEM = 5,    /* host or device allocation failed                */
  SVO_ERR_STATE = 6,    /* call not valid in the handle's current state    */
  SVO_ERR_CAPACITY = 7  /* incore pool has no free / freeable slot         */
};

/* Message of the last failing call on this thread ("" if none). */
const char* svo_last_error(void);

/* ------------------------------------------------------------ data model */

/* One serialized octree node, 8 bytes, identical bit layout to
 * CpuSvoNode (src/CpuSvoNode.h:46-48,100-102) and the OpenCL SvoNode
 * (opencl/treeletRenderer_rgba.cl:15-19):
 *   first_child : inner node -> offset of its first child RELATIVE to the
 *                 node's own index (may be negative after stitching, kept
 *                 modulo 2^32); treelet leaf -> gid of the sub-treelet that
 *                 continues below it (0 = none).
 *   masks       : bits 0-7 leaf mask, bits 8-15 valid mask, 16-31 zero.   */
typedef struct svo_node {
  uint32_t first_child;
  uint32_t masks;
} svo_node;

/* Per-treelet header of the .svo bundle, in file order
 * (src/Treelet.cpp:406-418). */
typedef struct svo_treelet_header {
  uint32_t num_nodes;
  uint32_t num_leaves;
  uint32_t mem_size;               /* bytes of one treelet slot          */
  uint32_t gid;
  uint32_t parent_gid;
  uint32_t parent_leaf_pos;        /* leaf's local index in the parent   */
  uint32_t parent_leaf_idx;        /* 0..7 child slot of that leaf       */
  uint32_t parent_leaf_parent_pos; /* local index of the leaf's parent   */
} svo_treelet_header;

typedef struct svo_scene svo_scene;

typedef struct svo_scene_info {
  uint32_t num_treelets;
  uint32_t treelet_bytes;       /* mem_size of every treelet            */
  uint32_t nodes_per_treelet;   /* treelet_bytes / 8                    */
  uint32_t max_depth;           /* depth of the deepest valid node (root =
                                   0): the deepest leaf of a well-formed
                                   tree; an inner node below every leaf
                                   counts too, so that the traversal stack
                                   holds every push                      */
  uint64_t total_nodes;         /* sum of num_nodes over treelets       */
  uint64_t tree_nodes;          /* distinct octree nodes                */
  uint64_t tree_leaves;         /* octree leaves at max_depth           */
} svo_scene_info;

/* --------------------------------------------------------- scene handles */

/* Reads <base_path>.svo and <base_path>.ia (TreeletMemoryManager::
 * loadFromFile, src/TreeletMemoryManager.cpp:229-291).  The .ia layout is
 * this library's own (u32 count, then per treelet u32 packages, u32 words
 * per package = 2, then packages x {u32 packed_normal, u32 packed_color}).*/
svo_status svo_scene_open(const char* base_path, svo_scene** out);

/* Writes <base_path>.svo and <base_path>.ia
 * (TreeletBuildManager::writeToFile, src/TreeletBuildManager.cpp:286-342).*/
svo_status svo_scene_write(const svo_scene* scene, const char* base_path);

void svo_scene_destroy(svo_scene* scene);

svo_status svo_scene_get_info(const svo_scene* scene, svo_scene_info* out);

/* Borrowed view of one treelet: header, num_nodes nodes and num_nodes
 * attribute packages {packed_normal, packed_color}.  Valid until the scene
 * is destroyed. */
svo_status svo_scene_get_treelet(const svo_scene* scene, uint32_t gid,
                                 svo_treelet_header* header,
                                 const svo_node** nodes,
                                 const uint32_t** attribs);

/* Treelet packing of the scene builders (flags argument).
 *   0                  dense packing (default): the octree is cut into the
 *                      fewest connected subtrees of at most nodes_per_treelet
 *                      nodes (greedy bottom-up, heaviest child cut first);
 *                      each is stored breadth-first.
 *   SVO_BUILD_PACK_BFS the reference builder's rule: fill each treelet
 *                      breadth-first until fewer than 9 slots remain and start
 *                      a sub-treelet at every unexpanded node above the scene
 *                      depth (src/TreeletBuilderFromFaces.cpp:188-365,
 *                      src/TreeletBuildManager.cpp:101-274).
 * Both emit the reference's .svo invariants: root at local index 0,
 * contiguous children with a relative first-child offset, sub-treelet anchor
 * leaves holding the sub-treelet gid, gids in FIFO order over treelets,
 * inner attributes = mean of the children. */
enum { SVO_BUILD_PACK_BFS = 1u };

/* Legacy single-SVO files (one octree, absolute first-child indices), read
 * into a scene of one treelet (relative offsets, gid 0), which every
 * traversal policy renders; node indices are unchanged:
 *   SVO_SINGLE_SVOLOADER  SvoLoader (src/SvoLoader.cpp:80-107): ASCII
 *                         "<depth> <numNodes> ", then the nodes; attribute
 *                         package i belongs to node i.
 *   SVO_SINGLE_V1         SvoBranch v1 (src/incore/SvoBranch.cpp:320-395):
 *                         ASCII "v1 <depth> <numNodes> ", the nodes, then
 *                         numNodes u32 attribute indices (the indirection of
 *                         fillFrameBuffer_rgba.cl:640, resolved at load).
 * <base>.ia holds one block of packages in this library's .ia layout.
 * svo_scene_open_single detects the format from the first word. */
enum { SVO_SINGLE_SVOLOADER = 1u, SVO_SINGLE_V1 = 2u };
svo_status svo_scene_open_single(const char* base_path, svo_scene** out);
/* Writes a one-treelet scene as a single-SVO file; attrib_index (v1 only,
 * may be NULL = identity) is a permutation: node i's package is written at
 * attrib_index[i]. */
svo_status svo_scene_write_single(const svo_scene* scene, const char* base_path,
                                  uint32_t format, const uint32_t* attrib_index);

/* Builds a treelet scene from leaf voxels at depth `depth`: `morton` holds
 * strictly increasing Morton codes (3 bits per level, x most significant,
 * i.e. child index 4x+2y+z), `attribs` two words per leaf
 * {packed_normal, packed_color}.  treelet_bytes == 0 sizes one treelet to
 * hold the whole tree. */
svo_status svo_scene_build_from_leaves(uint32_t depth, const uint64_t* morton,
                                       const uint32_t* attribs, size_t count,
                                       uint32_t treelet_bytes, uint32_t flags,
                                       svo_scene** out);

/* Procedural scenes in the unit cube [-0.5,0.5]^3, stored as the surface
 * shell of the solid (filled voxels with a face on empty space or on the
 * cube boundary).
 *   menger : level-`level` Menger sponge on a 2^depth grid, leaf colour =
 *            hash(seed, i, j, k);
 *   terrain: voxels whose centre lies below an 8-octave value-noise
 *            heightfield h(x,z) in [0.2,0.8), colour by height band plus
 *            hash noise;
 *   spheres: `count` solid spheres, radius uniform in [r_min, r_max],
 *            centres uniform in the cube. */
svo_status svo_scene_build_menger(uint32_t level, uint32_t depth,
                                  uint64_t seed, uint32_t treelet_bytes,
                                  uint32_t flags, svo_scene** out);
svo_status svo_scene_build_terrain(uint32_t depth, uint64_t seed,
                                   uint32_t treelet_bytes, uint32_t flags,
                                   svo_scene** out);
svo_status svo_scene_build_spheres(uint32_t depth, uint32_t count,
                                   float r_min, float r_max, uint64_t seed,
                                   uint32_t treelet_bytes, uint32_t flags,
                                   svo_scene** out);

/* Height of the terrain surface h(x,z) in [0.2,0.8] (unit-cube units,
 * y measured from the cube's bottom face) at x,z in [0,1]. */
svo_status svo_terrain_height(uint64_t seed, double x, double z, double* h);

/* ----------------------------------------------------------- camera */

/* Per-frame ray-generation uniforms (renderToBuffer args 3-7,
 * raycast_serialized_opencl.cpp:670-675,729-733), all FP32:
 *   ray(x,y) dir = normalize(far_lower_left + pixel_h*x
 *                            + pixel_v*(height - y) - position). */
typedef struct svo_camera {
  float far_lower_left[4];
  float pixel_h[4];
  float pixel_v[4];
  float position[4];
  float t_scale_ratio; /* LOD: frameBufferSize.z; 0 disables the LOD stop */
  uint32_t width;
  uint32_t height;
  uint32_t view_mode;  /* renderToBuffer's otherParams.x, SVO_VIEW_*: what
                          svo_frame_out.shade holds (other outputs are the
                          same in every mode) */
} svo_camera;

/* View modes of renderToBuffer (opencl/treeletRenderer_rgba.cl:644-732),
 * written as FP32 RGBA to svo_frame_out.shade:
 *   COLOR      colour of the hit node (alpha 1), miss 0.2,0.3,0.2,1
 *   NORMAL     unpacked normal of the hit node, miss 0.2,0.3,0.2,1
 *   ITERATIONS loop count through the wavelength colour map
 *              (DataPointToColor(loops, 0, 390), :488-596, pow in FP32)
 *   SVO_DEPTH  depth/13 on a hit, 0 on a miss (shade_svoDepth)
 *   T          1 - 0.1 t (shade_pixelDepth)
 *   LOOPS      1 - loops/200 (shade_iterationDepth)
 *   PHONG      shade_phong (:452-510): diffuse + specular towards a light
 *              at (-120,0,0), smoothstep, no shadow ray; miss 0.4,0.6,0.4,1
 *   QUALITY    sqrt(sqrt(quality)) with quality = scale_exp2 -
 *              tScaleRatio*tcMin at a leaf, 0 at an LOD stop; miss
 *              0.6,0.3,0.3,1
 * The reference's FP64 literals in the phong and colour-map code are taken
 * as FP32 (that code does not build as OpenCL C without double support).
 *
 * The single-SVO renderer's modes (opencl/fillFrameBuffer_rgba.cl:600-704,
 * its otherParams.x 0..6 = COLOR, NORMAL, FILL_SVO_DEPTH, T, FILL_LOOPS,
 * FILL_REFLECTION, FILL_SHADOW):
 *   FILL_SVO_DEPTH   depth/12 on a hit, 0 on a miss (its shade_svoDepth)
 *   FILL_LOOPS       1 - loops/300 (its shade_iterationDepth)
 *   FILL_SHADOW      shade_diffuse_color_shadow (:494-530): colour * nDotL *
 *                    (1 - shadow) + colour * 0.2, alpha 1, light direction
 *                    normalize(-0.5, 1, 0.6); shadow = 1 when a policy-B ray
 *                    from nodeCenter + 0.01 * normal towards the light hits
 *                    (tScaleRatio 0.001; like every sample() it is traced
 *                    from where its line enters the root box); miss
 *                    0.2,0.3,0.2,1
 *   FILL_REFLECTION  shade_diffuse_color_reflection (:536-594) with two
 *                    reflections: the shadow shading of the hit, then of up
 *                    to two reflection hits (origin nodeCenter + 0.01 *
 *                    normal, direction reflected about the normal), blended
 *                    with strengths 0.75, 0.5625, 0.421875.  The reference
 *                    never looks the attribute index up in this mode (its
 *                    case 5 leaves SampleResult.attribIndex at 0), so every
 *                    bounce shades with attribute package 0: for a scene
 *                    from a single-SVO file the file's package 0, otherwise
 *                    the root node's attributes.
 * FILL_SHADOW and FILL_REFLECTION trace their secondary rays with the
 * single-SVO policy and need SVO_POLICY_SINGLE for the primary rays. */
enum {
  SVO_VIEW_COLOR = 0, SVO_VIEW_NORMAL = 1, SVO_VIEW_ITERATIONS = 2, SVO_VIEW_SVO_DEPTH = 3,
  SVO_VIEW_T = 4, SVO_VIEW_LOOPS = 5, SVO_VIEW_PHONG = 6, SVO_VIEW_QUALITY = 7,
  SVO_VIEW_FILL_SVO_DEPTH = 8, SVO_VIEW_FILL_LOOPS = 9, SVO_VIEW_FILL_REFLECTION = 10,
  SVO_VIEW_FILL_SHADOW = 11
};

/* Pinhole camera at `eye` looking at `target` with `up`, vertical field of
 * view fov_y_deg, far plane `far_dist`; fills the FP32 uniforms (computed in
 * double, rounded once).  t_scale_ratio = lod_multiplier * tan(fov/2) *
 * aspect / (width/2) (raycast_serialized_opencl.cpp:303-304). */
svo_status svo_camera_look_at(const double eye[3], const double target[3],
                              const double up[3], double fov_y_deg,
                              uint32_t width, uint32_t height,
                              double far_dist, double lod_multiplier,
                              svo_camera* out);

/* ------------------------------------------------ residency (HBM pools) */

typedef struct svo_mm svo_mm;

typedef struct svo_mm_info {
  uint32_t num_slots;
  uint32_t nodes_per_treelet;
  uint32_t free_slots;
  uint32_t dirty_slots;       /* slots waiting for svo_mm_update_device  */
  uint32_t resident_treelets;
  uint32_t device;
  uint64_t pool_nodes;        /* num_slots * nodes_per_treelet           */
  uint64_t device_bytes;      /* HBM held by the handle                  */
} svo_mm_info;

typedef struct svo_upload_stats {
  uint32_t slots_uploaded;    /* zz_memory_manager-cl_slots_to_upload   */
  uint32_t copy_calls;        /* zz_memory_manager-num_copy_calls       */
  uint64_t bytes;             /* the merged ranges' bytes (what the
                                 reference's enqueueWrite calls move)   */
  uint64_t bytes_moved;       /* bytes actually sent: within each range
                                 only the part of a slot that is, or on
                                 the device was, not zero padding; nodes
                                 and colours (normals follow with the
                                 first launch whose view shades with
                                 them, or a device readback)            */
} svo_upload_stats;

/* TreeletMemoryManagerCl(path, incoreBufferSizeInByte, ctx) +
 * initClBuffer(): slots = floor(incore_bytes / treelet_bytes), the root
 * treelet inserted into slot 0 (resetIncoreBuffer,
 * src/TreeletMemoryManager.cpp:303-380), HBM pools allocated on `device`
 * and the whole incore buffer uploaded.  device = -1 creates a host-only
 * handle (no HBM pools: residency bookkeeping only, svo_render refuses it).
 * The scene must outlive the handle. */
svo_status svo_mm_create(const svo_scene* scene, uint64_t incore_bytes,
                         int device, svo_mm** out);
void svo_mm_destroy(svo_mm* mm);

/* resetIncoreBuffer(): empties every slot and re-inserts the root. */
svo_status svo_mm_reset(svo_mm* mm);

/* insertTreeletIntoIncoreBuffer(SlotInfo(gid, MAX_VISIBILITY, error))
 * (src/TreeletMemoryManager.cpp:455-548), incl. eviction through
 * freeIncorePosition when no slot is free (:560-653). */
svo_status svo_mm_insert(svo_mm* mm, uint32_t gid, float error);

/* removeTreeletFromIncoreBuffer(gid) (src/TreeletMemoryManager.cpp:665-700)*/
svo_status svo_mm_remove(svo_mm* mm, uint32_t gid);

/* Inserts every treelet in gid order (parents precede children). */
svo_status svo_mm_make_resident_all(svo_mm* mm);

/* updateDeviceMemory(): dirty slots merged into ranges with the reference
 * ratio rule (src/TreeletMemoryManagerCl.cpp:155-251) and copied to HBM in
 * one launch on `stream` that reads the pinned host buffers, followed by a
 * stream synchronize.  The device ends equal to the host incore buffer. */
svo_status svo_mm_update_device(svo_mm* mm, void* stream,
                                svo_upload_stats* stats);

svo_status svo_mm_get_info(const svo_mm* mm, svo_mm_info* out);

/* gid held by `slot`, -1 if empty (getSlots()[slot]._treeletGid). */
svo_status svo_mm_slot_treelet(const svo_mm* mm, uint32_t slot, int32_t* gid);

/* slot holding `gid`, -1 if not resident (Treelet::getSlotGid). */
svo_status svo_mm_treelet_slot(const svo_mm* mm, uint32_t gid, int32_t* slot);

/* Copies nodes [first, first+count) of the HOST incore buffer
 * (getIncoreBuffer) and their attribute packages (2 words each). */
svo_status svo_mm_read_host_incore(const svo_mm* mm, uint64_t first,
                                   uint64_t count, svo_node* nodes,
                                   uint32_t* attribs);

/* Same range read back from HBM (packed_normal/packed_color arrays). */
svo_status svo_mm_read_device_incore(const svo_mm* mm, uint64_t first,
                                     uint64_t count, svo_node* nodes,
                                     uint32_t* attribs);

/* getTreelets() (src/TreeletMemoryManager.h:107, the manager's
 * std::vector<Treelet*>): header and slot (Treelet::getSlotGid, -1 = not
 * resident) of treelets [first, first+count); either array may be NULL. */
svo_status svo_mm_get_treelets(const svo_mm* mm, uint32_t first, uint32_t count,
                               svo_treelet_header* headers, int32_t* slots);

/* getIncoreBuffer() (src/TreeletMemoryManager.h:117): the host incore buffer
 * itself, *count = pool_nodes nodes (node n of slot s at s*N + n), borrowed:
 * valid until the handle is destroyed, changed by inserts, removes, resets and
 * feedback updates.  Page-locked when the handle has a device. */
svo_status svo_mm_get_incore_buffer(const svo_mm* mm, const svo_node** nodes,
                                    uint64_t* count);

/* getAttributeIncoreBuffer() (src/TreeletMemoryManager.h:121): the attribute
 * incore buffer, one {packed_normal, packed_color} package per incore node,
 * held as two borrowed arrays of *count words (same lifetime as
 * svo_mm_get_incore_buffer); either pointer may be NULL. */
svo_status svo_mm_get_attribute_incore_buffer(const svo_mm* mm,
                                              const uint32_t** normals,
                                              const uint32_t** colors,
                                              uint64_t* count);

/* freeIncorePosition() (src/TreeletMemoryManager.h:135,
 * src/TreeletMemoryManager.cpp:560-653): probes up to 100 slots from the
 * manager's cursor for a treelet with visibility < 0 whose parent also has
 * visibility < 0 and frees it and every resident treelet below it.
 * *freed_gid = the treelet freed, -1 if none was found (the reference
 * returns false). */
svo_status svo_mm_free_incore_position(svo_mm* mm, int32_t* freed_gid);

/* getFreeIncoreSlotStack() (src/TreeletMemoryManager.h:150): the free slots,
 * bottom of the stack first; slots[n-1] is the top, the slot the next insert
 * takes.  Up to `cap` entries are written, *n receives the full count. */
svo_status svo_mm_get_free_slot_stack(const svo_mm* mm, uint32_t* slots,
                                      uint32_t cap, uint32_t* n);

/* getIncoreSlotsToUpload() (src/TreeletMemoryManager.h:153): the slots
 * changed since the last svo_mm_update_device, ascending.  Up to `cap`
 * entries are written, *n receives the full count. */
svo_status svo_mm_get_slots_to_upload(const svo_mm* mm, uint32_t* slots,
                                      uint32_t cap, uint32_t* n);

/* getSlots() (src/TreeletMemoryManager.h:157) with the whole SlotInfo
 * (:72-89): treelet gid (-1 empty), visibility, error of slots
 * [first, first+count). */
typedef struct svo_slot_info {
  int32_t treelet_gid;
  int32_t visibility;
  float error;
  uint32_t reserved;
} svo_slot_info;
svo_status svo_mm_get_slots(const svo_mm* mm, uint32_t first, uint32_t count,
                            svo_slot_info* out);

/* The HBM pools of the handle, for kernels of the caller's own that read the
 * incore buffer, as the OpenCL app binds getClIncoreBufferGid() and
 * getClAttributeIncoreBufferGid() (src/TreeletMemoryManagerCl.h:76-90,
 * raycast_serialized_opencl.cpp:323-324); `device` stands for getContext().
 *   nodes    : pool_nodes x 8 B {first_child, masks}
 *   colors   : pool_nodes x u32 packed_color   } the reference's interleaved
 *   normals  : pool_nodes x u32 packed_normal  } Attribs, as two arrays
 *   slot_gid : num_slots x i32, gid per slot (-1 empty)
 * Owned by the handle.  The pools equal the host incore buffer once
 * svo_mm_update_device has run after the last change; this call sends any
 * normals still pending and, from then on, every update sends normals with
 * nodes and colours. */
typedef struct svo_device_pools {
  void* nodes;
  void* colors;
  void* normals;
  void* slot_gid;
  uint64_t pool_nodes;
  uint32_t num_slots;
  uint32_t nodes_per_treelet;
  int32_t device;
  uint32_t reserved;
} svo_device_pools;
svo_status svo_mm_get_device_pools(svo_mm* mm, svo_device_pools* out);

/* initClBuffer() (src/TreeletMemoryManagerCl.cpp:101-113): uploads the whole
 * incore buffer (nodes and attributes) to the HBM pools on `stream` and waits
 * for it.  svo_mm_create already did this once. */
svo_status svo_mm_init_device_buffer(svo_mm* mm, void* stream);

/* ------------------------------------------------------------ rendering */

/* Traversal policy.  SVO_POLICY_TREELET is the OpenCL renderer's
 * sample() (opencl/treeletRenderer_rgba.cl:201-382); SVO_POLICY_SINGLE is
 * the single-SVO OpenCL renderer's sample()
 * (opencl/fillFrameBuffer_rgba.cl:239-425: scaleMax 15, epsilon 2^-16, LOD
 * on tcMax, maxLoops = miss; its absolute child indexing addresses the same
 * nodes as relative indexing on a scene held as one treelet, which is how
 * svo_scene_open_single stores a single SVO); SVO_POLICY_CPU3 is
 * CpuRaycasterSingleRay3::traversSvo
 * (src/incore/CpuRaycasterSingleRay3.cpp:158-387).  Every policy runs on the
 * relative-indexed incore buffer. */
enum { SVO_POLICY_TREELET = 0, SVO_POLICY_SINGLE = 1, SVO_POLICY_CPU3 = 2 };

/* Per-camera outputs, device pointers, width*height elements each, row y
 * at offset y*width (y as in renderToBuffer's get_global_id(1)).  Any
 * pointer may be NULL to skip that output.
 *   t       : FP32 ray parameter of the hit (parentTMin), the box exit
 *             tMax on a miss inside the box, +inf if the box is missed
 *   node    : incore node index of the hit, 0xFFFFFFFF on a miss
 *   treelet : gid of the treelet whose slot holds `node`, 0xFFFFFFFF miss
 *   rgba    : packed_color of the hit node with alpha forced to 255
 *             (r<<24|g<<16|b<<8|a); miss = background 0.2,0.3,0.2,1
 *             -> 0x334D33FF
 *   pushes  : number of stack pushes of the ray (traversal statistics;
 *             8 bytes of node data are fetched per push)
 *   shade   : 4 floats per pixel, the camera's view mode (SVO_VIEW_*)    */
typedef struct svo_frame_out {
  float* t;
  uint32_t* node;
  uint32_t* treelet;
  uint32_t* rgba;
  uint32_t* pushes;
  float* shade;
} svo_frame_out;

/* Renders n_cams frames in one launch on `stream` (asynchronous). */
svo_status svo_render(svo_mm* mm, const svo_camera* cams, uint32_t n_cams,
                      const svo_frame_out* outs, int policy, void* stream);

/* ------------------------------------------ multi-view launches, graphs */

/* One view of a launch: the scene (its residency handle), the camera and
 * the outputs.  A launch may mix views of different scenes and sizes;
 * all handles must live on the renderer's device. */
typedef struct svo_view {
  svo_mm* mm;
  svo_camera camera;
  svo_frame_out out;
  /* > 0: pixels with x % feedback_stride == 0 and y % feedback_stride == 0
   * also record analysis feedback into mm's tables (see svo_mm_analyse) */
  uint32_t feedback_stride;
  uint32_t reserved;
} svo_view;

/* Launch shape chosen by svo_autotune (or set by hand):
 *   tile_width      : pixels per wavefront tile row, 8, 16 or 32 (the tile
 *                     is tile_width x 64/tile_width)
 *   waves_per_group : wavefronts per workgroup, 1, 2 or 4
 *   stack_layout    : SVO_STACK_LDS (shared memory, [level][lane], the
 *                     parent's node index and tmax: a pop re-reads the
 *                     parent node), SVO_STACK_PRIVATE (per-lane private
 *                     array) or SVO_STACK_LDS_NODE (shared memory holding the
 *                     parent node's child base and masks as well: a pop
 *                     reads no global memory; 10 instead of 8 bytes per
 *                     level and lane) */
enum { SVO_STACK_LDS = 0, SVO_STACK_PRIVATE = 1, SVO_STACK_LDS_NODE = 2 };
typedef struct svo_launch_config {
  uint32_t tile_width;
  uint32_t waves_per_group;
  uint32_t stack_layout;
  uint32_t reserved;
} svo_launch_config;

/* A renderer owns the device-side view descriptors of its launches (a ring,
 * so that launches on different streams can be in flight together) and the
 * launch shape.  Default shape: 8-wide tiles, 1 wave, LDS stack. */
typedef struct svo_renderer svo_renderer;
svo_status svo_renderer_create(int device, svo_renderer** out);
void svo_renderer_destroy(svo_renderer* r);
svo_status svo_renderer_set_config(svo_renderer* r, const svo_launch_config* c);
svo_status svo_renderer_get_config(const svo_renderer* r, svo_launch_config* c);

/* One launch rendering every view (the renderToBuffer launch, generalised
 * to many cameras and scenes).  Asynchronous on `stream`. */
svo_status svo_render_views(svo_renderer* r, const svo_view* views, uint32_t n,
                            int policy, void* stream);

/* Mean ms per launch over `iters` back-to-back launches (HIP events on
 * `stream`). */
svo_status svo_render_views_timed(svo_renderer* r, const svo_view* views,
                                  uint32_t n, int policy, void* stream,
                                  uint32_t iters, float* mean_ms);

/* Allocation-free variant of svo_render_views, for callers that cannot
 * allocate device or pinned memory (tight serving loops, harnesses that
 * forbid hipMalloc).  No counterpart in the reference.  The view descriptors
 * (which svo_render_views keeps in its renderer's ring, allocated on first
 * use) go to a fixed device scratch array that each GPU holds for the life
 * of the process: at most SVO_NOALLOC_MAX_VIEWS views per call (larger calls
 * fail with SVO_ERR_ARG); config = NULL takes the default launch shape.  The
 * residency handles were created beforehand (they own the pools); outputs
 * are the caller's.  Every call on a GPU shares that GPU's scratch: calls
 * ordered on one stream are safe, calls on different streams of one GPU that
 * overlap race.  Results are those of svo_render_views.  Asynchronous on
 * `stream` (the descriptors are copied from host memory before it returns). */
enum { SVO_NOALLOC_MAX_VIEWS = 256 };
svo_status svo_render_views_noalloc(const svo_view* views, uint32_t n, int policy,
                                    const svo_launch_config* config, void* stream);

/* Times every supported launch shape on these views (`iters` launches each,
 * after one warm-up) and installs the fastest in the renderer.  Results are
 * identical for every shape; only the speed differs. */
svo_status svo_autotune(svo_renderer* r, const svo_view* views, uint32_t n,
                        int policy, void* stream, uint32_t iters,
                        svo_launch_config* best, float* best_ms);

/* Autotune cache: the shapes svo_autotune picked, kept per workload class so
 * that a process, or the next one through a file, does not time the 27
 * shapes again.  No counterpart in the reference, whose launch shape is the
 * fixed work size of the renderToBuffer enqueue
 * (raycast_serialized_opencl.cpp:729-749) and whose stack depth is a
 * compile-time constant (opencl/treeletRenderer_rgba.cl:7-8).
 * The key (svo_tune_cache_key) names the device architecture and CU count,
 * the policy, the number of views and the total rays (each rounded up to a
 * power of two), and every distinct scene class of the views: scaleMax,
 * stack levels, nodes per treelet, LOD on or off.
 * File: text, one entry per line, `key<TAB>tile_width<TAB>waves<TAB>stack
 * <TAB>ms` with stack lds / private / lds_node; lines starting with '#' are
 * comments.  svo_tune_cache_load merges a file into the cache (a missing file
 * is an empty cache, a malformed line SVO_ERR_FORMAT and nothing merged);
 * svo_tune_cache_save writes the whole cache through a temporary file and a
 * rename.  A cache may be shared by renderers and threads.
 * svo_autotune_cached sets the renderer's shape from the cache when the key
 * is there (*cache_hit = 1, *best_ms = the time recorded with it), otherwise
 * runs svo_autotune and records its pick (*cache_hit = 0). */
typedef struct svo_tune_cache svo_tune_cache;
svo_status svo_tune_cache_create(svo_tune_cache** out);
void svo_tune_cache_destroy(svo_tune_cache* c);
uint32_t svo_tune_cache_size(const svo_tune_cache* c);
svo_status svo_tune_cache_load(svo_tune_cache* c, const char* path);
svo_status svo_tune_cache_save(svo_tune_cache* c, const char* path);
svo_status svo_tune_cache_key(svo_renderer* r, const svo_view* views, uint32_t n,
                              int policy, char* key, size_t cap);
svo_status svo_autotune_cached(svo_renderer* r, svo_tune_cache* c,
                               const svo_view* views, uint32_t n, int policy,
                               void* stream, uint32_t iters,
                               svo_launch_config* best, float* best_ms,
                               int* cache_hit);

/* A batch is a multi-view launch captured once into a hipGraph and replayed
 * with svo_batch_launch.  Its view descriptors live in device memory owned by
 * the batch; svo_batch_set_cameras rewrites the camera uniforms (same frame
 * sizes; view modes are checked, and the normals of handles whose views now
 * shade with them are sent before the next replay) with an asynchronous copy
 * on `stream`, ordered before the next launch on that stream.  Views with
 * feedback_stride > 0 record into the analysis cycle that is current on the
 * device when the replay runs (the readback of a cycle advances it on the
 * stream), so a batch keeps feeding out-of-core updates across cycles.  A
 * replay is refused while a view's handle has slots waiting for
 * svo_mm_update_device, like svo_render. */
typedef struct svo_batch svo_batch;
svo_status svo_batch_create(svo_renderer* r, const svo_view* views, uint32_t n,
                            int policy, void* stream, svo_batch** out);
svo_status svo_batch_set_cameras(svo_batch* b, const svo_camera* cams,
                                 uint32_t n, void* stream);
svo_status svo_batch_launch(svo_batch* b, void* stream);
void svo_batch_destroy(svo_batch* b);

/* ---------------------------------------------- analysis feedback (OOC) */

/* The analyse pass of the OpenCL app (RenderPassAnalyse::performAnalysePass,
 * src/RenderPassAnalyse.cpp:149-369, kernel renderToFeedbackBuffer +
 * sampleAnalyse, opencl/treeletRenderer_rgba.cl:773-1011), fused into the
 * traversal: every ray of a W/divide x H/divide frame over the same frustum
 * (far lower-left corner moved by jitter_x, jitter_y in [-1, 1] half analysis
 * pixels, the reference's "rumble") traces with the given policy and records
 * its hit on the device:
 *   - leaf hit: error = scale_exp2 - tScaleRatio*tcMin; when the leaf's
 *     first-child field holds a sub-treelet gid, that treelet is asked for;
 *   - inner node at the LOD / depth limit: error = (tScaleRatio*tcMin >
 *     scale_exp2);
 *   - miss or maxLoops: no record.
 * The treelet holding each hit node is marked visible with the largest
 * error, each asked treelet keeps its largest error (device tables, no
 * per-pixel readback).  Views with feedback_stride > 0 record the same way
 * during an ordinary render. */
svo_status svo_mm_analyse(svo_mm* mm, const svo_camera* cam, uint32_t divide,
                          float jitter_x, float jitter_y, int policy,
                          void* stream);

/* The analysis frame svo_mm_analyse traces: width/divide x height/divide
 * pixels spanning the same far plane, the far lower-left corner moved by
 * (jitter_x * pixel_h + jitter_y * pixel_v) * divide / 2
 * (RenderPassAnalyse.cpp:172-184; FP32, in that operation order). */
svo_status svo_analysis_camera(const svo_camera* cam, uint32_t divide,
                               float jitter_x, float jitter_y,
                               svo_camera* out);

typedef struct svo_feedback_stats {
  uint32_t visible_slots;      /* slots holding a recorded hit             */
  uint32_t visible_treelets;   /* those treelets and their ancestors       */
  uint32_t requests;           /* distinct treelets asked for              */
  uint32_t requests_kept;      /* after the 12 MiB cap, largest error first */
  uint32_t inserted;           /* page-ins                                 */
  uint32_t insert_failed;      /* no free or freeable slot                 */
  uint32_t evicted;            /* treelets freed by freeIncorePosition     */
  uint32_t resident_treelets;
} svo_feedback_stats;

/* updateClientSideIncoreBuffer(RenderPassAnalyse*)
 * (src/TreeletMemoryManager.cpp:392-443) from the records gathered since the
 * previous call: reads the tables back on `stream` (the recording launches
 * must be ordered before it on that stream), marks visible treelets and their
 * ancestors, ages every other slot, and inserts the asked treelets in
 * descending error (at most 12 MiB of treelets), evicting invisible ones as
 * needed.  Follow with svo_mm_update_device to upload the changed slots. */
svo_status svo_mm_apply_feedback(svo_mm* mm, void* stream,
                                 svo_feedback_stats* stats);

/* The visible sets of the records gathered since the previous update, as
 * RenderPassAnalyse::getVisibleOldTreeletsGids / getVisibleNewTreeletsGids
 * (src/RenderPassAnalyse.h:89-93) give them after performAnalysePass
 * (src/RenderPassAnalyse.cpp:248-356):
 *   old: every treelet holding a recorded hit and all its ancestors, with
 *        the largest error below it (the root: 99), in ascending gid;
 *   new: the treelets asked for, by descending error (ties: ascending gid),
 *        capped at 12 MiB of treelets.
 * Reads the tables back on `stream` and ends the recording cycle (later
 * recordings start a new one).  The sets are computed against the slot
 * table of this moment and kept; the next svo_mm_apply_feedback applies
 * them instead of reading the tables (as performAnalysePass computes them
 * before updateClientSideIncoreBuffer applies them).  Up to
 * old_cap / new_cap entries are written; *n_old / *n_new receive the full
 * counts (arrays may be null when their cap is 0). */
svo_status svo_mm_read_feedback(svo_mm* mm, void* stream,
                                uint32_t* old_gids, float* old_errors,
                                uint32_t old_cap, uint32_t* n_old,
                                uint32_t* new_gids, float* new_errors,
                                uint32_t new_cap, uint32_t* n_new);

/* Enqueues the readback of the records gathered since the previous update
 * on `stream`, behind the launches that recorded them, and ends the
 * recording cycle (later recordings start a new one); returns at once.  The
 * next svo_mm_apply_feedback (or svo_mm_read_feedback) waits for it instead
 * of reading, and builds the visible sets then, against the slot table of
 * that moment.  Lets a caller enqueue the readback and its render on one
 * stream and do the host update while the render runs.  The reference reads
 * its feedback buffer back with a blocking read inside performAnalysePass
 * (src/RenderPassAnalyse.cpp:228-232); this splits that read off. */
svo_status svo_mm_fetch_feedback(svo_mm* mm, void* stream);

/* The same update from caller-supplied aggregated records (no device
 * access; works on host-only handles): slots[i] was hit with largest error
 * slot_errors[i]; gids[i] (distinct, non-root) asked for with largest error
 * gid_errors[i]. */
svo_status svo_mm_apply_records(svo_mm* mm, const uint32_t* slots,
                                const float* slot_errors, uint32_t n_slots,
                                const uint32_t* gids, const float* gid_errors,
                                uint32_t n_gids, svo_feedback_stats* stats);

/* ------------------------------------------------------------ serving */

/* A render server keeps several residency handles (scenes) on one device and
 * turns a stream of render requests into multi-view launches as they
 * arrive (continuous batching).  It has no counterpart in the reference,
 * whose app renders one camera per frame (raycast_serialized_opencl.cpp:
 * 684-749); each launch it makes is the svo_render_views launch.
 *
 *  - Scenes are admitted by HBM budget: svo_server_add_scene fails with
 *    SVO_ERR_CAPACITY when the scene's device bytes plus the frame pool
 *    would exceed hbm_budget_bytes.
 *  - Requests whose outputs are all null get their outputs (t, node,
 *    treelet, rgba: 16 bytes per pixel) from the server's frame pool of
 *    frame_pool_bytes; a request waits (in arrival order) until its frame
 *    fits.  Results stay in HBM; the frame is returned to the pool when the
 *    request completes.  Requests with any output set render into the
 *    caller's buffers.
 *  - Up to max_inflight launches run at once, each on its own stream; a
 *    launch takes waiting requests in arrival order up to max_batch_views
 *    views and max_batch_rays rays (at least one request).
 *    max_inflight = max_batch_views = 1 is the plain one-frame-per-launch
 *    loop.
 *  - Overload: with max_waiting > 0 at most that many requests wait for a
 *    launch; svo_server_submit refuses more with SVO_ERR_CAPACITY (replay
 *    completes such a request as SVO_REQ_REJECTED).  A request with
 *    timeout_us > 0 that is still waiting that long after its arrival is
 *    dropped unrendered and completes as SVO_REQ_EXPIRED.
 * A server is driven from one host thread. */
typedef struct svo_server svo_server;

typedef struct svo_server_config {
  uint64_t hbm_budget_bytes;   /* scenes + frame pool; 0 = no limit */
  uint64_t frame_pool_bytes;   /* output pool for requests without outputs */
  uint64_t max_batch_rays;     /* 0 = no limit */
  uint32_t max_batch_views;    /* 1..65536 */
  uint32_t max_inflight;       /* 1..16 */
  int32_t policy;              /* SVO_POLICY_* */
  uint32_t max_waiting;        /* requests waiting for a launch; 0 = no limit */
} svo_server_config;

typedef struct svo_request {
  uint32_t scene;              /* id from svo_server_add_scene */
  uint32_t timeout_us;         /* longest wait for a launch; 0 = none */
  double arrival_s;            /* svo_server_replay: offset from the start */
  svo_camera camera;
  svo_frame_out out;           /* all null: outputs from the frame pool */
} svo_request;

/* Times in seconds since the server was created (online use) or since the
 * start of the replay. */
enum { SVO_REQ_DONE = 0, SVO_REQ_EXPIRED = 1, SVO_REQ_REJECTED = 2 };

typedef struct svo_completion {
  uint64_t id;                 /* svo_server_submit id / replay index */
  double arrival_s, launch_s, done_s;   /* not rendered: launch_s = done_s */
  uint32_t batch;              /* launch sequence number */
  uint32_t batch_views;        /* views in that launch (0: not rendered) */
  uint32_t status;             /* SVO_REQ_* */
  uint32_t reserved;
} svo_completion;

typedef struct svo_server_stats {
  uint64_t requests, rays, batches;
  uint64_t peak_pool_bytes;
  uint32_t max_batch_views_seen, max_inflight_seen;
  double wall_s;               /* first arrival to last completion */
  double frames_per_s, mrays_per_s;
  double p50_ms, p99_ms, mean_ms, max_ms;   /* done - arrival, rendered */
  double mean_batch_views;
  uint64_t expired, rejected;  /* not rendered (frames_per_s counts the rest) */
} svo_server_stats;

svo_status svo_server_create(svo_renderer* r, const svo_server_config* cfg,
                             svo_server** out);
void svo_server_destroy(svo_server* s);
svo_status svo_server_add_scene(svo_server* s, svo_mm* mm, uint32_t* scene_id);

/* Online use: queue a request now (its arrival time), then call
 * svo_server_poll repeatedly; it retires finished launches, starts new ones
 * and returns up to `cap` completions. */
svo_status svo_server_submit(svo_server* s, const svo_request* req, uint64_t* id);
svo_status svo_server_poll(svo_server* s, svo_completion* done, uint32_t cap,
                           uint32_t* n_done);
/* Requests queued or in flight. */
svo_status svo_server_pending(const svo_server* s, uint32_t* n);

/* Benchmark driver: submits reqs[i] at its arrival_s offset in real time
 * (arrival_s ascending), serves them until all complete and fills one
 * completion per request (may be null) and the summary. */
svo_status svo_server_replay(svo_server* s, const svo_request* reqs, uint32_t n,
                             svo_completion* done, svo_server_stats* stats);

/* ------------------------------------------- device memory and streams */

svo_status svo_device_count(int* count);
svo_status svo_set_device(int device);
svo_status svo_device_malloc(void** ptr, size_t bytes);
svo_status svo_device_free(void* ptr);
svo_status svo_memcpy_h2d(void* dst, const void* src, size_t bytes);
svo_status svo_memcpy_d2h(void* dst, const void* src, size_t bytes);
svo_status svo_memset_device(void* dst, int value, size_t bytes);
svo_status svo_stream_create(void** stream);
svo_status svo_stream_destroy(void* stream);
svo_status svo_stream_synchronize(void* stream);
svo_status svo_device_synchronize(void);

/* Times `iters` back-to-back svo_render launches on `stream` with HIP
 * events recorded on that stream; returns the mean ms per launch. */
svo_status svo_render_timed(svo_mm* mm, const svo_camera* cams,
                            uint32_t n_cams, const svo_frame_out* outs,
                            int policy, void* stream, uint32_t iters,
                            float* mean_ms);

/* Diagnostic: writes the FP32 ray direction of every pixel of *cam
 * (3 floats per pixel, row-major) into dirs_dev, computed exactly as the
 * traversal kernel computes it (renderToBuffer's ray generation,
 * opencl/treeletRenderer_rgba.cl:627-633). */
svo_status svo_debug_ray_dirs(const svo_camera* cam, float* dirs_dev,
                              void* stream);

/* Diagnostic: device-to-device copy of `bytes` (a multiple of 16) with a
 * plain streaming kernel; a dispatch of known traffic for calibrating the
 * FETCH_SIZE / WRITE_SIZE hardware counters. */
svo_status svo_debug_copy(void* dst, const void* src, size_t bytes,
                          void* stream);

/* Diagnostic: a kernel that only issues VALU instructions (eight
 * independent integer chains per lane, 16 VALU instructions per iteration,
 * iterations rounded up to a multiple of 8, one store per lane), `blocks`
 * workgroups of 256 lanes; `out` holds blocks * 256 words.  Its hardware counters give the VALU-busy reading of a
 * kernel bound by VALU issue alone, the reference for the traversal's. */
svo_status svo_debug_valu(uint32_t* out, uint32_t blocks, uint32_t iters,
                          void* stream);

#ifdef __cplusplus
}
#endif

#endif /* SVO_HIP_H */
