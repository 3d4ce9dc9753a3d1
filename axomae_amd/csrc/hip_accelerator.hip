// HipAccelerator: the C++ device-accelerator facade over the C ABI, and flatten(), which
// turns the engine's scene holders into the C-ABI descriptor.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>

#include "hip_accelerator.h"

namespace nova::aggregate {

static void throw_if(nova_hip_ctx *ctx, int code) {
  if (code != NOVA_HIP_SUCCESS) throw HipDeviceError(code, ctx ? nova_hip_last_error(ctx) : "no context");
}

// glm::normalize(glm::cross(a, b)) per vertex (Object3D.h:208-238)
static void cross_normalized(const span<float> &a, const span<float> &b, size_t vertex_count, std::vector<float> &out) {
  out.assign(vertex_count * 3, 0.f);
  for (size_t v = 0; v < vertex_count; v++) {
    float ax = 0.f, ay = 0.f, az = 0.f, bx = 0.f, by = 0.f, bz = 0.f;
    if (a.size() >= (v + 1) * 3) {
      ax = a[v * 3];
      ay = a[v * 3 + 1];
      az = a[v * 3 + 2];
    }
    if (b.size() >= (v + 1) * 3) {
      bx = b[v * 3];
      by = b[v * 3 + 1];
      bz = b[v * 3 + 2];
    }
    const float cx = ay * bz - by * az, cy = az * bx - bz * ax, cz = ax * by - bx * ay;
    const float inv = 1.f / std::sqrt((cx * cx + cy * cy) + cz * cz);
    out[v * 3] = cx * inv;
    out[v * 3 + 1] = cy * inv;
    out[v * 3 + 2] = cz * inv;
  }
}

void flatten(const primitive_aggregate_data_s &primitives, const shading_resources_s &shading, flat_scene_s &out) {
  const span<const Object3D> &geom = primitives.mesh_geometry.triangle_mesh_geometry.host_geometry_view;
  const shape::transform::mesh_transform_views_t &xf = primitives.mesh_geometry.transforms;
  out = flat_scene_s{};
  out.meshes.resize(geom.size());
  // a mesh's material: the one its primitives were added with (setup_trimesh adds every
  // triangle of a mesh with the same material, engine.cpp:168-174)
  std::vector<int32_t> mesh_material(geom.size(), -1);
  for (size_t i = 0; i < primitives.primitive_list_view.size(); i++) {
    const primitive::primitive_record_s &p = primitives.primitive_list_view[i];
    if (p.mesh_index >= geom.size()) throw std::invalid_argument("primitive refers to a missing mesh");
    int32_t &m = mesh_material[p.mesh_index];
    if (m >= 0 && m != p.material_index) throw std::invalid_argument("primitives of one mesh with different materials");
    m = p.material_index;
  }
  out.completed.reserve(geom.size() * 2);
  for (size_t m = 0; m < geom.size(); m++) {
    const Object3D &o = geom[m];
    nova_hip_mesh_desc &d = out.meshes[m];
    const size_t nv = o.vertices.size() / 3;
    if (o.vertices.size() % 3 != 0 || o.indices.size() % 3 != 0) throw std::invalid_argument("Object3D spans are not whole vertices / faces");
    d.vertices = o.vertices.data();
    d.normals = o.normals.data();
    d.uv = o.uv.data();
    d.tangents = o.tangents.data();
    d.bitangents = o.bitangents.data();
    d.indices = o.indices.data();
    d.vertex_count = nv;
    d.index_count = o.indices.size();
    d.material = mesh_material[m] >= 0 ? mesh_material[m] : 0;
    // attributes the mesh lacks: zeros (Object3D::getTri leaves them at zero), then the
    // tangent frame completed as compute_missing_vectors does
    auto zeros = [&](size_t n) -> const float * {
      out.completed.emplace_back(n, 0.f);
      return out.completed.back().data();
    };
    if (o.normals.size() < nv * 3) d.normals = zeros(nv * 3);
    if (o.uv.size() < nv * 2) d.uv = zeros(nv * 2);
    if (o.bitangents.empty()) {
      out.completed.emplace_back();
      cross_normalized(o.normals, o.tangents, nv, out.completed.back());
      d.bitangents = out.completed.back().data();
    } else if (o.tangents.empty()) {
      out.completed.emplace_back();
      cross_normalized(o.normals, o.bitangents, nv, out.completed.back());
      d.tangents = out.completed.back().data();
    }
    if (!d.tangents) d.tangents = zeros(nv * 3);
    // transform.m of the mesh's 57-float record
    if (m < xf.mesh_offsets_to_matrix.size()) {
      const uint64_t off = xf.mesh_offsets_to_matrix[m];
      if (off % shape::transform::kTransformPadding != 0 || off + 16 > xf.matrix_components_view.size())
        throw std::invalid_argument("transform offset outside the matrix storage");
      std::memcpy(d.transform, xf.matrix_components_view.data() + off, 16 * sizeof(float));
    } else {
      for (int k = 0; k < 16; k++) d.transform[k] = (k % 5 == 0) ? 1.f : 0.f;
    }
  }
  out.materials.resize(shading.materials.size());
  for (size_t i = 0; i < shading.materials.size(); i++) {
    const material::principled_material_s &s = shading.materials[i];
    nova_hip_material_desc &d = out.materials[i];
    d.albedo = s.albedo;
    d.metallic = s.metallic;
    d.normal = s.normalmap;
    d.emissive = s.emissive;
    d.albedo_f32 = s.albedo_f32;
    d.roughness_f32 = s.roughness_f32;
    for (int k = 0; k < 3; k++) {
      d.eta[k] = s.eta[k];
      d.k[k] = s.k[k];
    }
    d.anisotropy = s.anisotropy;
    d.emission_scale = s.emission_scale;
  }
  out.textures.resize(shading.textures.size());
  for (size_t i = 0; i < shading.textures.size(); i++) {
    const texturing::TextureRawData &t = shading.textures[i];
    nova_hip_texture_desc &d = out.textures[i];
    d.width = t.width;
    d.height = t.height;
    d.channels = t.channels;
    d.invert_u = t.invert_x ? 1 : 0;
    d.invert_v = t.invert_y ? 1 : 0;
    d.format = t.f_data ? NOVA_HIP_TEX_F32 : NOVA_HIP_TEX_U32;
    d.data = t.f_data ? static_cast<const void *>(t.f_data) : static_cast<const void *>(t.raw_data);
    if (!d.data) throw std::invalid_argument("texture without texels");
  }
  out.desc.meshes = out.meshes.data();
  out.desc.mesh_count = (uint32_t)out.meshes.size();
  out.desc.materials = out.materials.data();
  out.desc.material_count = (uint32_t)out.materials.size();
  out.desc.textures = out.textures.data();
  out.desc.texture_count = (uint32_t)out.textures.size();
  out.desc.envmap_texture = shading.current_envmap_index;
}

HipAccelerator::HipAccelerator(int device) : device_(device) {
  int code = nova_hip_create(device, &ctx_);
  if (code != NOVA_HIP_SUCCESS) throw HipDeviceError(code, "nova_hip_create failed");
}

HipAccelerator::~HipAccelerator() {
  if (ctx_) nova_hip_destroy(ctx_);
}

void HipAccelerator::build(primitive_aggregate_data_s scene) {
  flat_scene_s flat;
  try {
    flatten(scene, shading_, flat);
  } catch (const std::invalid_argument &e) {
    throw HipDeviceError(NOVA_HIP_INVALID_ARGUMENT, e.what());
  }
  throw_if(ctx_, nova_hip_scene_upload(ctx_, &flat.desc, nullptr));  // copies the host buffers
  throw_if(ctx_, nova_hip_bvh_build(ctx_, 4, &info_, nullptr));
}

void HipAccelerator::cleanup() {
  if (ctx_) nova_hip_destroy(ctx_);
  ctx_ = nullptr;
  throw_if(nullptr, nova_hip_create(device_, &ctx_));
}

void HipAccelerator::copyParamsToDevice(const device_traversal_param_s &params) const { params_ = params; }

namespace {
class HipIntersector final : public DeviceIntersectorInterface {
 public:
  HipIntersector(nova_hip_ctx *ctx, int policy, int rr_start, bool vertical_invert, int lookahead)
      : ctx_(ctx), policy_(policy), rr_start_(rr_start), vertical_invert_(vertical_invert), lookahead_(lookahead) {}
  // One frame (1 spp) of the path integrator into render_buffers.partial_buffer, stream
  // ordered on the null stream and finished on return, as the reference's synchronous
  // traverse is (aggregate/api/optix/device.cpp:18-21).
  void traverse(const device_traversal_param_s &p) const override {
    nova_hip_render_params rp{};
    rp.width = (int32_t)p.width;
    rp.height = (int32_t)p.height;
    rp.max_depth = (int32_t)p.depth;
    rp.frame_start = (int32_t)p.sample_index;
    rp.frame_count = 1;
    rp.index_policy = policy_;
    rp.max_samples_offset = policy_ == NOVA_HIP_INDEX_OPTIX ? p.sample_max : 0u;
    rp.vertical_invert = vertical_invert_ ? 1 : 0;
    rp.rr_start = rr_start_;
    rp.sampling_roughness_texture = -1;
    rp.z_near = p.camera.near;
    rp.z_far = p.camera.far;
    std::memcpy(rp.inv_P, p.camera.inv_P.m, sizeof(rp.inv_P));
    std::memcpy(rp.inv_V, p.camera.inv_V.m, sizeof(rp.inv_V));
    if (lookahead_ <= 1) {
      throw_if(ctx_, nova_hip_render_frames(ctx_, &rp, nullptr, nullptr, p.render_buffers.partial_buffer, nullptr));
      return;
    }
    throw_if(ctx_, nova_hip_render_frame_lookahead(ctx_, &rp, lookahead_, nullptr, nullptr, p.render_buffers.partial_buffer, nullptr));
    // a kept frame is copied by a stream-ordered kernel: wait for it, so the frame is in
    // partial_buffer on return as after the reference's synchronous traverse
    if (hipStreamSynchronize(nullptr) != hipSuccess) throw HipDeviceError(NOVA_HIP_DEVICE_ERROR, "traverse: stream synchronization failed");
  }

 private:
  nova_hip_ctx *ctx_;
  int policy_, rr_start_;
  bool vertical_invert_;
  int lookahead_;
};
}  // namespace

std::unique_ptr<DeviceIntersectorInterface> HipAccelerator::getIntersectorObject() const {
  return std::make_unique<HipIntersector>(ctx_, index_policy_, rr_start_, vertical_invert_, lookahead_);
}

std::unique_ptr<DeviceAcceleratorInterface> DeviceAcceleratorInterface::make(int device) { return std::make_unique<HipAccelerator>(device); }

}  // namespace nova::aggregate
