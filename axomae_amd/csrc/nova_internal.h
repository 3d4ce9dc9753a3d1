// Internal layouts of the HIP backend (host + device).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/nova_hip.h"

namespace nova_hip {

// Traversal record, 48 B, in BVH leaf order: (v0, prim id) (e1, leaf size) (e2, mesh tag);
// the leaf size (int bits) is set in the first record of each leaf.
// e1 = v1 - v0, e2 = v2 - v0 are the edges Triangle::hit forms (Triangle.h:158-160); a
// triangle with repeated indices gets zero edges so that det == 0 rejects it, as
// Triangle::isDegenerate does. Mesh tag (int bits) 0: the mesh's transform is the identity
// and v0, e1, e2 are world (= object) coordinates; k > 0: v0, e1, e2 are object
// coordinates of mesh k - 1, whose rays are taken to object space first.
struct alignas(16) TriMT {
  float4 a, b, c;
};

// Shading record, 192 B, in global primitive order: object-space corners and the
// per-vertex attributes the Embree-path hit record interpolates (embree.cpp:287-390).
struct alignas(16) ShadeTri {
  float p0[3], p1[3], p2[3];
  float uv0[2], uv1[2], uv2[2];
  float n0[3], n1[3], n2[3];
  float t0[3], t1[3], t2[3];
  float b0[3], b1[3], b2[3];
  int32_t material, mesh;
  int32_t pad[2];
};
static_assert(sizeof(ShadeTri) == 192, "ShadeTri layout");

// Per-mesh 3x3 blocks: mat3(transform.m) and the normal matrix transpose(inverse(mat3(m))),
// column-major.
struct MeshXf {
  float m3[9];
  float n3[9];
};

struct MatRec {
  int32_t albedo, metallic, normal, emissive, albedo_f32, roughness_f32;
  float eta[3], k[3];
  float anisotropy, emission_scale;
};

struct TexRec {
  int32_t width, height, channels, invert_u, invert_v, format;
  uint64_t offset;  // element offset into the u32 or f32 texel pool
};

// Binary SAH node with the reference layout (Bvhnl, BvhBuilder.h:10-15).
struct BNode {
  float mn[3], mx[3];
  int32_t left, count;
};

// The binary tree arrays the traversal reads only to settle exact equal-t ties, gathered in
// one device record so a kernel carries a single pointer to them.
struct TieTables {
  const BNode *bnodes;
  const int32_t *bparent, *bdepth, *leaf_of_pos;
  // Rows 0-2 of each mesh's inverse model matrix (three float4 per mesh): the triangle test
  // of a mesh with a non-identity transform takes the ray to object space, as
  // Triangle::hit does (Triangle.h:150-153). Read only for such meshes' triangles.
  const float4 *xf_inv;
};

// 4-wide node, 64 B, four 16-byte loads per visit. The child boxes are on a per-node grid:
// on axis a, plane value q (one byte) stands for the float fmaf(q, step_a, org_a), and the
// build picks each min plane's q so that this value is <= the child's box minimum and each
// max plane's q so that it is >= the maximum (checked in that same float arithmetic), so a
// decoded box holds the child's exact box. Byte k of qmin*/qmax* belongs to child k.
// child[k] >= 0: an inner node index; child[k] <= -2: a leaf whose triangles start at
// -(child[k] + 2) in the TriMT array (the size is in that first record's b.w);
// child[k] == -1: empty slot.
struct alignas(16) Node4 {
  float orgx, orgy, orgz, stepx;
  float stepy, stepz;
  uint32_t qminx, qmaxx;
  uint32_t qminy, qmaxy, qminz, qmaxz;
  int32_t child[4];
};
__host__ __device__ inline int32_t leaf_ref(int32_t first) { return -(first + 2); }
__host__ __device__ inline int32_t leaf_first(int32_t ref) { return -(ref + 2); }
static_assert(sizeof(Node4) == 64, "Node4 layout");

// The same node with its planes decoded (each exactly the float fmaf(q, step, org) of the
// Node4), 128 B, for the wave-coherent walk, which reads nodes through the scalar unit:
// plane[axis][0][k] / plane[axis][1][k] are child k's min / max planes on that axis, so the
// near and far planes of an axis are a choice of row, made once per walk when the wave's
// rays agree on the direction's sign.
struct alignas(16) Node4F {
  float plane[3][2][4];
  int32_t child[4];
  int32_t pad[4];
};
static_assert(sizeof(Node4F) == 128, "Node4F layout");

struct DeviceBuffer {
  void *ptr = nullptr;
  size_t bytes = 0;
};

}  // namespace nova_hip

struct nova_hip_ctx {
  int device = 0;
  std::string last_error;

  // scene
  uint32_t prim_count = 0;
  uint32_t mesh_count = 0;
  nova_hip::TriMT *tri_world = nullptr;      // world-space MT records, primitive order
  nova_hip::ShadeTri *shade = nullptr;       // shading records, primitive order
  float *prim_aabb = nullptr;                // 6 floats per primitive (min3, max3)
  float *prim_centroid = nullptr;            // 3 floats per primitive
  nova_hip::MeshXf *mesh_xf = nullptr;
  float4 *xf_inv = nullptr;                  // rows 0-2 of each mesh's inverse model matrix
  nova_hip::MatRec *materials = nullptr;
  nova_hip::TexRec *textures = nullptr;
  uint32_t *tex_u32 = nullptr;
  float *tex_f32 = nullptr;
  uint64_t tex_f32_floats = 0;               // FP32 texel pool size
  int32_t envmap = -1;
  float *acos_lut = nullptr;
  uint32_t *sobol_table = nullptr;
  std::vector<nova_hip::TexRec> host_textures;
  bool scene_ready = false;

  // BVH
  int32_t *prim_idx = nullptr;
  nova_hip::BNode *bnodes = nullptr;
  uint32_t bnode_count = 0;
  nova_hip::Node4 *nodes4 = nullptr;
  nova_hip::Node4F *nodes4f = nullptr;        // the nodes with decoded planes (wave-coherent walk)
  nova_hip::TriMT *tri_leaf = nullptr;       // MT records in leaf order (prim id in a.w)
  int32_t *bparent = nullptr;                // binary tree parent (-1 at the root)
  int32_t *bdepth = nullptr;                 // binary tree depth
  int32_t *leaf_of_pos = nullptr;            // binary leaf containing each prim_idx position
  nova_hip::TieTables *tie = nullptr;        // device record of bnodes, bparent, bdepth, leaf_of_pos
  nova_hip_bvh_info bvh_info{};
  bool bvh_ready = false;
  // bounce at which the last render's tail_kernel started (0: none), for batch size,
  // max_depth and rr_start tail_hint_key; cleared by a new scene or BVH
  int tail_hint = 0;
  uint64_t tail_hint_key[3] = {0, 0, 0};
  // bumped by every change of the scene, its BVH or its textures (keys the frames below)
  uint64_t scene_gen = 0;
  // frames rendered ahead by nova_hip_render_frame_lookahead (images in scratch slot 12):
  // frames ahead_first .. ahead_first + ahead_count - 1 of the parameters ahead_key
  // (frame_start 0) on scene generation ahead_gen
  nova_hip_render_params ahead_key{};
  int32_t ahead_first = 0, ahead_count = 0;
  uint64_t ahead_gen = 0;
  bool ahead_valid = false;
  // timing events reused by every render call (created on first use, destroyed with the
  // context): the call's start and end, then two per timed launch
  hipEvent_t ev_begin = nullptr, ev_end = nullptr;
  std::vector<hipEvent_t> ev_pool;

  // integrator scratch (grown on demand)
  nova_hip::DeviceBuffer scratch[16];
  uint32_t *gstack = nullptr;                // overflow traversal stacks
  size_t gstack_entries = 0;
  nova_hip_render_stats stats{};
  hipStream_t batch_stream = nullptr;        // batched renders capture their wave graph on it
};

namespace nova_hip {
// error helpers shared by the translation units
int set_error(nova_hip_ctx *ctx, int code, const std::string &msg);
int check_hip(nova_hip_ctx *ctx, hipError_t e, const char *what);
void *ensure_scratch(nova_hip_ctx *ctx, int slot, size_t bytes, hipError_t *err);
uint32_t traversal_stack_bound(const nova_hip_ctx *ctx);
const uint32_t *device_sobol_table();
// Device address of the no-allocation entry point's fixed scratch array on `device`, looked
// up once per device (the device must be current); NOVA_HIP_DEVICE_ERROR when unavailable.
int noalloc_scratch_setup(int device);
}  // namespace nova_hip

#define NOVA_HIP_CHECK(ctx, expr)                                   \
  do {                                                              \
    hipError_t _e = (expr);                                         \
    if (_e != hipSuccess) return nova_hip::check_hip(ctx, _e, #expr); \
  } while (0)
