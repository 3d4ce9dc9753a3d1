// C++ host facade over the C ABI, shaped like the reference's device-accelerator slot:
//   nova::aggregate::DeviceAcceleratorInterface / DeviceIntersectorInterface
//   (corelib/nova/aggregate/device_acceleration_interface.h:11-38).
// The data types mirror the reference's field names and order so that the engine's glue
// reads the same: device_traversal_param_s (corelib/nova/gpu/nova_gpu.h:20-29) with
// HdrBufferStruct (api_datastructures.h:7-19) and CameraResourcesHolder
// (camera/nova_camera.h:6-29); primitive_aggregate_data_s (aggregate_datastructures.h:17-20)
// with MeshBundleViews (shape/MeshContext.h:9-34), Object3D's attribute spans
// (internal/geometry/Object3D.h:387-393) and the 57-float transform records
// (shape/shape_datastructures.h:41-62). glm::mat4 / glm::vec3 are stored as glm keeps them
// (column-major float arrays). The reference compiles these types from its own headers;
// here they are restated so that the facade builds on its own.
//
// build() flattens the scene (flatten(), from the engine's holders: engine.cpp:159-220,
// material.cpp:181-226) and uploads it, then builds the BVH; copyParamsToDevice() keeps the
// frame's parameters; traverse() renders one frame of 1 spp into
// render_buffers.partial_buffer, as gpu_draw -> device_start_integrator does per frame
// (engine.cpp:461-494). The engine then accumulates (writeFramebuffer, engine.cpp:351-363).
#pragma once
#include <cstddef>
#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/nova_hip.h"

namespace nova {

// axstd::span: a pointer and a count.
template <class T>
struct span {
  T *ptr = nullptr;
  size_t count = 0;
  size_t size() const { return count; }
  bool empty() const { return count == 0; }
  T *data() const { return ptr; }
  T &operator[](size_t i) const { return ptr[i]; }
};

struct glm_mat4 {  // glm::mat4, column-major (glm::value_ptr order)
  float m[16];
};
struct glm_vec3 {
  float x, y, z;
};

// RenderBuffers<float> (api_datastructures.h:7-19)
template <class T>
struct RenderBuffers {
  T *partial_buffer = nullptr;  // RGBA32F, (y * width + x) * 4, device memory on this path
  T *depth_buffer = nullptr;
  T *normal_buffer = nullptr;
  size_t color_buffers_pitch{};
  size_t depth_buffers_pitch{};
  size_t normal_buffers_pitch{};
  size_t byte_size_color_buffers{};
  size_t byte_size_depth_buffers{};
  int channels{};
};
using HdrBufferStruct = RenderBuffers<float>;

namespace camera {
// CameraResourcesHolder (camera/nova_camera.h:6-29)
struct CameraResourcesHolder {
  glm_mat4 P, inv_P, V, inv_V, PV, inv_PV;
  glm_vec3 position, up_vector, direction;
  unsigned int screen_width, screen_height;
  float far, near, fov;
};
}  // namespace camera

// Object3D attribute spans (internal/geometry/Object3D.h:387-393): xyz per vertex for
// vertices, colors, normals, bitangents, tangents; uv two per vertex; three indices a face.
struct Object3D {
  span<float> vertices{}, uv{}, colors{}, normals{}, bitangents{}, tangents{};
  span<unsigned> indices{};
  static constexpr int face_stride = 3;
};

namespace shape {
namespace triangle {
struct mesh_vertex_attrib_views_t {  // shape_datastructures.h:12-15
  span<const Object3D> device_geometry_view;
  span<const Object3D> host_geometry_view;
};
}  // namespace triangle
namespace transform {
// 57 floats per record: m, inv, t (16 each), then n (9) (shape_datastructures.h:41-62)
constexpr size_t kTransformPadding = 57;
struct mesh_transform_views_t {  // shape_datastructures.h:21-27
  span<const float> matrix_components_view{};
  span<const uint64_t> mesh_offsets_to_matrix{};
};
}  // namespace transform
struct MeshBundleViews {  // shape/MeshContext.h:9-34
  transform::mesh_transform_views_t transforms{};
  triangle::mesh_vertex_attrib_views_t triangle_mesh_geometry{};
};
}  // namespace shape

namespace primitive {
// NovaGeoPrimitive: a triangle shape (mesh index, index of its first index) and the
// material it was added with (engine.cpp:168-174).
struct primitive_record_s {
  uint32_t mesh_index;
  uint32_t triangle_index;
  int32_t material_index;
};
}  // namespace primitive

namespace texturing {
// TextureRawData (texturing/nova_texturing.h:98-112): u32 texels packed
// (a<<24)|(r<<16)|(g<<8)|b, or `channels` floats per texel.
struct TextureRawData {
  const uint32_t *raw_data = nullptr;
  const float *f_data = nullptr;
  int width = 0, height = 0, channels = 4;
  bool invert_x = false, invert_y = false;
};
}  // namespace texturing

namespace material {
// texture_pack of PrincipledMaterial (material.cpp:208-216), as texture indices (-1: none),
// the refractive indices and anisotropy passed to addMaterial (material.cpp:218-225), and
// the FP32-texture / emission extensions (defaults: the reference's behaviour).
struct principled_material_s {
  int32_t albedo = -1, normalmap = -1, metallic = -1, roughness = -1, emissive = -1, opacity = -1, specular = -1, ao = -1;
  float eta[3] = {0.f, 0.f, 0.f}, k[3] = {0.f, 0.f, 0.f};
  float anisotropy = 0.f;
  int32_t albedo_f32 = -1, roughness_f32 = -1;
  float emission_scale = 1.f;
};
}  // namespace material

// Shading resources of NovaResourceManager that the HIP build uploads with the geometry
// (the OptiX path reads them through the per-frame views of device_traversal_param_s).
struct shading_resources_s {
  span<const material::principled_material_s> materials{};
  span<const texturing::TextureRawData> textures{};
  int32_t current_envmap_index = -1;  // TextureResourcesHolder::getEnvmapId
};

struct device_random_generators_s {};  // the library keeps its own Sobol direction table

// device_traversal_param_s (gpu/nova_gpu.h:20-29), same fields in the same order. The
// mesh/texture/primitive/material views are the OptiX pipeline's device views; the HIP
// path uploads the scene once in build() and does not read them per frame.
struct device_traversal_param_s {
  HdrBufferStruct render_buffers;
  device_random_generators_s device_random_generators;
  shape::MeshBundleViews mesh_bundle_views;
  const void *texture_bundle_views = nullptr;
  const void *primitives_view = nullptr;
  const void *material_view = nullptr;
  camera::CameraResourcesHolder camera;
  uint32_t width, height, depth, current_envmap_index, sample_index, sample_max;
};

namespace aggregate {

// primitive_aggregate_data_s (aggregate_datastructures.h:17-20)
struct primitive_aggregate_data_s {
  span<const primitive::primitive_record_s> primitive_list_view;
  shape::MeshBundleViews mesh_geometry;
};

// Host storage behind a flattened scene descriptor: the descriptor points into it.
struct flat_scene_s {
  std::vector<nova_hip_mesh_desc> meshes;
  std::vector<nova_hip_material_desc> materials;
  std::vector<nova_hip_texture_desc> textures;
  std::vector<std::vector<float>> completed;  // tangents or bitangents the meshes lack
  nova_hip_scene_desc desc{};
};

// Engine::buildScene's holders -> the C-ABI scene descriptor: per mesh its Object3D spans,
// the model matrix `m` of its transform record, and the material of its primitives; the
// texture list and the envmap index. A mesh without tangents or bitangents gets them per
// vertex as Object3D::compute_missing_vectors derives them (Object3D.h:208-238, 369-384):
// bitangent = normalize(cross(n, t)), tangent = normalize(cross(n, b)). Throws
// std::invalid_argument on inconsistent input.
void flatten(const primitive_aggregate_data_s &primitives, const shading_resources_s &shading, flat_scene_s &out);

class DeviceIntersectorInterface {
 public:
  virtual ~DeviceIntersectorInterface() = default;
  virtual void traverse(const device_traversal_param_s &params) const = 0;
};

class DeviceAcceleratorInterface {
 public:
  virtual ~DeviceAcceleratorInterface() = default;
  virtual void build(primitive_aggregate_data_s primitive_data_list) = 0;
  virtual void cleanup() = 0;
  virtual unsigned getMaxRecursiveDepth() const = 0;
  virtual void copyParamsToDevice(const device_traversal_param_s &params) const = 0;
  virtual std::unique_ptr<DeviceIntersectorInterface> getIntersectorObject() const = 0;
  /* Replaces the factory's OptiX branch (aggregate/api/device_accelerator_factory.cpp:9-17). */
  static std::unique_ptr<DeviceAcceleratorInterface> make(int device = 0);
};

// Device errors surface as exceptions on the host side, as GenericException does in the
// reference's device layer (engine.cpp:478-486 maps them to RENDERER_EXCEPTION).
class HipDeviceError : public std::runtime_error {
 public:
  HipDeviceError(int code, const std::string &msg) : std::runtime_error(msg), code_(code) {}
  int code() const { return code_; }

 private:
  int code_;
};

class HipAccelerator final : public DeviceAcceleratorInterface {
 public:
  explicit HipAccelerator(int device = 0);
  ~HipAccelerator() override;
  // The shading resources build() uploads with the geometry (the engine glue passes
  // NovaResourceManager's material and texture holders before Engine::buildAcceleration).
  void setShadingResources(const shading_resources_s &shading) { shading_ = shading; }
  // Sample index policy of traverse(): NOVA_HIP_INDEX_CPU (the CPU renderer's
  // 4 * pixel * frame, the parity default; sample_max is not added, as the CPU engine's
  // renderer_max_samples is 0) or NOVA_HIP_INDEX_OPTIX (the GPU raygen's
  // (x * H + y) * sample_max + frame). Russian roulette: first scatter that rolls, -1 = off.
  void setIndexPolicy(int policy) { index_policy_ = policy; }
  void setRussianRoulette(int rr_start) { rr_start_ = rr_start; }
  void setVerticalInvert(bool v) { vertical_invert_ = v; }
  // Frames traverse() renders ahead while the camera and settings stay the same
  // (nova_hip_render_frame_lookahead): 1 (the default) renders each frame alone. Every
  // frame's radiance is the same either way.
  void setFrameLookahead(int frames) { lookahead_ = frames; }
  void build(primitive_aggregate_data_s scene) override;
  void cleanup() override;
  unsigned getMaxRecursiveDepth() const override { return 8; } /* MAX_REC_DEPTH (optix/host.cpp:25) */
  void copyParamsToDevice(const device_traversal_param_s &params) const override;
  std::unique_ptr<DeviceIntersectorInterface> getIntersectorObject() const override;
  nova_hip_ctx *context() const { return ctx_; }
  const nova_hip_bvh_info &bvhInfo() const { return info_; }

 private:
  int device_ = 0;
  nova_hip_ctx *ctx_ = nullptr;
  nova_hip_bvh_info info_{};
  shading_resources_s shading_{};
  int index_policy_ = NOVA_HIP_INDEX_CPU;
  int rr_start_ = -1;
  bool vertical_invert_ = false;
  int lookahead_ = 1;
  mutable device_traversal_param_s params_{};
};

}  // namespace aggregate
}  // namespace nova
