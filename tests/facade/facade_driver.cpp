// Test driver (test infrastructure, built by tests/facade/build_facade.py): the reference
// engine's GPU frame loop (corelib/nova/api/engine.cpp:461-494) through the C++ facade
// alone -- flatten the holders, build, then per frame copyParamsToDevice,
// getIntersectorObject()->traverse into HdrBufferStruct::partial_buffer, and the host-side
// writeFramebuffer (accumulator += partial; image = accumulator / (frame + 1),
// engine.cpp:351-363). Called from Python through ctypes with plain arrays.
#include <hip/hip_runtime.h>

#include <cstring>
#include <string>
#include <vector>

#include "hip_accelerator.h"

using namespace nova;

namespace {
struct Holders {
  std::vector<Object3D> objects;
  std::vector<float> matrices;
  std::vector<uint64_t> offsets;
  std::vector<primitive::primitive_record_s> prims;
  std::vector<material::principled_material_s> materials;
  std::vector<texturing::TextureRawData> textures;
};

// The arrays of one scene as the engine's holders: Object3D spans per mesh, 57-float
// transform records (m, then t = transpose(m); the inverse and normal blocks are not read by
// flatten and stay zero), one primitive record per triangle.
void make_holders(int mesh_count, float *const *vertices, float *const *normals, float *const *tangents, float *const *bitangents,
                  float *const *uv, unsigned *const *indices, const uint64_t *vertex_counts, const uint64_t *index_counts, const float *transforms,
                  const int32_t *mesh_material, int material_count, const int32_t *mat_tex, const float *mat_f, int texture_count,
                  const int32_t *tex_info, const void *const *tex_data, int omit_bitangents, Holders &h) {
  for (int m = 0; m < mesh_count; m++) {
    Object3D o;
    const size_t nv = vertex_counts[m];
    o.vertices = span<float>{vertices[m], nv * 3};
    o.normals = span<float>{normals[m], nv * 3};
    o.tangents = span<float>{tangents[m], nv * 3};
    if (!omit_bitangents) o.bitangents = span<float>{bitangents[m], nv * 3};
    o.uv = span<float>{uv[m], nv * 2};
    o.indices = span<unsigned>{indices[m], (size_t)index_counts[m]};
    h.objects.push_back(o);
    h.offsets.push_back(h.matrices.size());
    std::vector<float> rec(shape::transform::kTransformPadding, 0.f);
    for (int k = 0; k < 16; k++) rec[k] = transforms[m * 16 + k];
    for (int c = 0; c < 4; c++)
      for (int r = 0; r < 4; r++) rec[32 + c * 4 + r] = transforms[m * 16 + r * 4 + c];
    h.matrices.insert(h.matrices.end(), rec.begin(), rec.end());
    for (uint64_t t = 0; t < index_counts[m]; t += 3) h.prims.push_back({(uint32_t)m, (uint32_t)t, mesh_material[m]});
  }
  for (int i = 0; i < material_count; i++) {
    material::principled_material_s s;
    const int32_t *t = mat_tex + i * 6;
    s.albedo = t[0];
    s.metallic = t[1];
    s.normalmap = t[2];
    s.emissive = t[3];
    s.albedo_f32 = t[4];
    s.roughness_f32 = t[5];
    const float *f = mat_f + i * 8;
    for (int k = 0; k < 3; k++) {
      s.eta[k] = f[k];
      s.k[k] = f[3 + k];
    }
    s.anisotropy = f[6];
    s.emission_scale = f[7];
    h.materials.push_back(s);
  }
  for (int i = 0; i < texture_count; i++) {
    texturing::TextureRawData t;
    const int32_t *ti = tex_info + i * 6;
    t.width = ti[0];
    t.height = ti[1];
    t.channels = ti[2];
    t.invert_x = ti[3] != 0;
    t.invert_y = ti[4] != 0;
    if (ti[5] == NOVA_HIP_TEX_F32)
      t.f_data = static_cast<const float *>(tex_data[i]);
    else
      t.raw_data = static_cast<const uint32_t *>(tex_data[i]);
    h.textures.push_back(t);
  }
}

aggregate::primitive_aggregate_data_s aggregate_of(const Holders &h) {
  aggregate::primitive_aggregate_data_s a;
  a.primitive_list_view = span<const primitive::primitive_record_s>{h.prims.data(), h.prims.size()};
  a.mesh_geometry.triangle_mesh_geometry.host_geometry_view = span<const Object3D>{h.objects.data(), h.objects.size()};
  a.mesh_geometry.transforms.matrix_components_view = span<const float>{h.matrices.data(), h.matrices.size()};
  a.mesh_geometry.transforms.mesh_offsets_to_matrix = span<const uint64_t>{h.offsets.data(), h.offsets.size()};
  return a;
}

void set_err(char *err, int err_len, const std::string &m) {
  if (err && err_len > 0) {
    std::strncpy(err, m.c_str(), (size_t)err_len - 1);
    err[err_len - 1] = 0;
  }
}
}  // namespace

// flatten() alone (no device): the descriptor's counts, each mesh's material and transform,
// and mesh 0's bitangents (as given or completed) copied to bitangents0_out.
extern "C" int facade_flatten(int mesh_count, float *const *vertices, float *const *normals, float *const *tangents, float *const *bitangents,
                              float *const *uv, unsigned *const *indices, const uint64_t *vertex_counts, const uint64_t *index_counts,
                              const float *transforms, const int32_t *mesh_material, int material_count, const int32_t *mat_tex, const float *mat_f,
                              int texture_count, const int32_t *tex_info, const void *const *tex_data, int envmap, int omit_bitangents,
                              int32_t *counts_out, int32_t *mesh_material_out, float *transforms_out, float *bitangents0_out, char *err,
                              int err_len) {
  try {
    Holders h;
    make_holders(mesh_count, vertices, normals, tangents, bitangents, uv, indices, vertex_counts, index_counts, transforms, mesh_material, material_count,
                 mat_tex, mat_f, texture_count, tex_info, tex_data, omit_bitangents, h);
    shading_resources_s sh{span<const material::principled_material_s>{h.materials.data(), h.materials.size()},
                           span<const texturing::TextureRawData>{h.textures.data(), h.textures.size()}, envmap};
    aggregate::flat_scene_s flat;
    aggregate::flatten(aggregate_of(h), sh, flat);
    counts_out[0] = (int32_t)flat.desc.mesh_count;
    counts_out[1] = (int32_t)flat.desc.material_count;
    counts_out[2] = (int32_t)flat.desc.texture_count;
    counts_out[3] = flat.desc.envmap_texture;
    for (uint32_t m = 0; m < flat.desc.mesh_count; m++) {
      mesh_material_out[m] = flat.desc.meshes[m].material;
      std::memcpy(transforms_out + m * 16, flat.desc.meshes[m].transform, 64);
    }
    if (flat.desc.mesh_count > 0)
      std::memcpy(bitangents0_out, flat.desc.meshes[0].bitangents, flat.desc.meshes[0].vertex_count * 3 * sizeof(float));
    return 0;
  } catch (const std::exception &e) {
    set_err(err, err_len, e.what());
    return 1;
  }
}

// The engine's GPU frame loop through the facade: frames [0, frames) at 1 spp each into the
// host image and accumulator (width * height * 4 floats each).
extern "C" int facade_engine_loop(int mesh_count, float *const *vertices, float *const *normals, float *const *tangents, float *const *bitangents,
                                  float *const *uv, unsigned *const *indices, const uint64_t *vertex_counts, const uint64_t *index_counts,
                                  const float *transforms, const int32_t *mesh_material, int material_count, const int32_t *mat_tex,
                                  const float *mat_f, int texture_count, const int32_t *tex_info, const void *const *tex_data, int envmap,
                                  int omit_bitangents, const float *inv_P, const float *inv_V, int width, int height, int frames, int depth,
                                  int rr_start, int lookahead, float *image_out, float *accum_out, char *err, int err_len) {
  float *partial = nullptr;
  try {
    Holders h;
    make_holders(mesh_count, vertices, normals, tangents, bitangents, uv, indices, vertex_counts, index_counts, transforms, mesh_material, material_count,
                 mat_tex, mat_f, texture_count, tex_info, tex_data, omit_bitangents, h);
    std::unique_ptr<aggregate::DeviceAcceleratorInterface> accel = aggregate::DeviceAcceleratorInterface::make(0);
    auto *hip = static_cast<aggregate::HipAccelerator *>(accel.get());
    hip->setShadingResources(shading_resources_s{span<const material::principled_material_s>{h.materials.data(), h.materials.size()},
                                                 span<const texturing::TextureRawData>{h.textures.data(), h.textures.size()}, envmap});
    hip->setRussianRoulette(rr_start);
    hip->setFrameLookahead(lookahead);
    accel->build(aggregate_of(h));
    const size_t n = (size_t)width * height * 4;
    if (hipMalloc(&partial, n * sizeof(float)) != hipSuccess) throw std::runtime_error("partial buffer alloc");
    std::vector<float> host(n), accum(n, 0.f);
    for (int f = 0; f < frames; f++) {
      device_traversal_param_s params{};
      params.render_buffers.partial_buffer = partial;
      params.render_buffers.channels = 4;
      std::memcpy(params.camera.inv_P.m, inv_P, 64);
      std::memcpy(params.camera.inv_V.m, inv_V, 64);
      params.width = (uint32_t)width;
      params.height = (uint32_t)height;
      params.depth = (uint32_t)depth;
      params.current_envmap_index = (uint32_t)envmap;
      params.sample_index = (uint32_t)f;
      params.sample_max = (uint32_t)frames;
      accel->copyParamsToDevice(params);
      accel->getIntersectorObject()->traverse(params);
      if (hipMemcpy(host.data(), partial, n * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess) throw std::runtime_error("partial copy");
      for (size_t i = 0; i < n; i++) {  // Engine::writeFramebuffer
        accum[i] += host[i];
        image_out[i] = accum[i] / float(f + 1);
      }
    }
    std::memcpy(accum_out, accum.data(), n * sizeof(float));
    (void)hipFree(partial);
    return 0;
  } catch (const std::exception &e) {
    if (partial) (void)hipFree(partial);
    set_err(err, err_len, e.what());
    return 1;
  }
}
