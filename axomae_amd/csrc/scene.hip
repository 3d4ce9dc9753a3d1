// Scene upload: copies the mesh spans, textures and materials to the device and
// flattens them into per-triangle traversal/shading records with one kernel.
#include <cmath>
#include <cstring>
#include <vector>

#include "device_math.h"
#include "device_scene.h"
#include "nova_internal.h"

extern "C" void nova_hip_free_scene_internal(nova_hip_ctx *ctx);

namespace nova_hip {

// One thread per triangle (flatten_triangle in device_scene.h).
__global__ void flatten_kernel(uint32_t ntri, const TriRef *__restrict__ refs, const float *__restrict__ pos, const float *__restrict__ nrm,
                               const float *__restrict__ tan, const float *__restrict__ bit, const float *__restrict__ uv,
                               const float *__restrict__ xforms, const int32_t *__restrict__ mesh_mat, TriMT *__restrict__ tri,
                               ShadeTri *__restrict__ shade, float *__restrict__ aabb, float *__restrict__ centroid) {
  uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= ntri) return;
  flatten_triangle(p, refs[p], pos, nrm, tan, bit, uv, xforms, mesh_mat, tri, shade, aabb, centroid);
}


template <class T>
static hipError_t upload(T **dst, const void *src, size_t bytes, hipStream_t s) {
  hipError_t e = hipMalloc((void **)dst, bytes ? bytes : 16);
  if (e != hipSuccess) return e;
  if (bytes) e = hipMemcpyAsync(*dst, src, bytes, hipMemcpyHostToDevice, s);
  return e;
}

}  // namespace nova_hip

using namespace nova_hip;

extern "C" int nova_hip_scene_upload(nova_hip_ctx *ctx, const nova_hip_scene_desc *sd, void *stream) {
  if (!ctx) return NOVA_HIP_INVALID_ARGUMENT;
  if (!sd || (sd->mesh_count && !sd->meshes)) return set_error(ctx, NOVA_HIP_INVALID_ARGUMENT, "null scene descriptor");
  hipStream_t s = (hipStream_t)stream;
  NOVA_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  nova_hip_free_scene_internal(ctx);

  // ---- validate and gather on the host
  uint64_t nvert = 0, ntri = 0;
  for (uint32_t m = 0; m < sd->mesh_count; m++) {
    const nova_hip_mesh_desc &md = sd->meshes[m];
    if (md.index_count % 3) return set_error(ctx, NOVA_HIP_INVALID_SCENE_TYPE, "mesh " + std::to_string(m) + ": index count not a multiple of 3");
    if (md.vertex_count && (!md.vertices || !md.normals || !md.tangents || !md.bitangents || !md.uv))
      return set_error(ctx, NOVA_HIP_INVALID_SCENE_TYPE, "mesh " + std::to_string(m) + ": missing vertex attribute span");
    if (md.index_count && !md.indices) return set_error(ctx, NOVA_HIP_INVALID_SCENE_TYPE, "mesh " + std::to_string(m) + ": null indices");
    for (uint64_t i = 0; i < md.index_count; i++)
      if (md.indices[i] >= md.vertex_count)
        return set_error(ctx, NOVA_HIP_INVALID_SCENE_TYPE, "mesh " + std::to_string(m) + ": index out of range");
    if (md.material < 0 || (uint32_t)md.material >= sd->material_count)
      return set_error(ctx, NOVA_HIP_INVALID_ARGUMENT, "mesh " + std::to_string(m) + ": material out of range");
    nvert += md.vertex_count;
    ntri += md.index_count / 3;
  }
  if (ntri == 0) return set_error(ctx, NOVA_HIP_INVALID_SCENE_TYPE, "Primitives list is empty.");
  if (ntri >= (1ull << 31) || nvert >= (1ull << 32) / 3) return set_error(ctx, NOVA_HIP_INVALID_ARGUMENT, "scene too large");

  std::vector<TexRec> texrec(sd->texture_count);
  uint64_t nu32 = 0, nf32 = 0;
  for (uint32_t t = 0; t < sd->texture_count; t++) {
    const nova_hip_texture_desc &td = sd->textures[t];
    if (td.width < 2 || td.height < 2 || td.channels < 1 || td.channels > 4 || !td.data)
      return set_error(ctx, NOVA_HIP_INVALID_ARGUMENT, "texture " + std::to_string(t) + ": needs width, height >= 2, 1..4 channels and data");
    TexRec &r = texrec[t];
    r.width = td.width;
    r.height = td.height;
    r.channels = td.channels;
    r.invert_u = td.invert_u;
    r.invert_v = td.invert_v;
    r.format = td.format;
    if (td.format == NOVA_HIP_TEX_U32) {
      r.offset = nu32;
      nu32 += (uint64_t)td.width * td.height;
    } else if (td.format == NOVA_HIP_TEX_F32) {
      r.offset = nf32;
      nf32 += (uint64_t)td.width * td.height * td.channels;
    } else {
      return set_error(ctx, NOVA_HIP_INVALID_ARGUMENT, "texture " + std::to_string(t) + ": unknown format");
    }
  }
  std::vector<MatRec> mats(sd->material_count);
  for (uint32_t m = 0; m < sd->material_count; m++) {
    const nova_hip_material_desc &d = sd->materials[m];
    int32_t ids[6] = {d.albedo, d.metallic, d.normal, d.emissive, d.albedo_f32, d.roughness_f32};
    for (int k = 0; k < 6; k++) {
      if (ids[k] >= (int32_t)sd->texture_count) return set_error(ctx, NOVA_HIP_INVALID_ARGUMENT, "material texture out of range");
      int want = k < 4 ? NOVA_HIP_TEX_U32 : NOVA_HIP_TEX_F32;
      if (ids[k] >= 0 && texrec[ids[k]].format != want)
        return set_error(ctx, NOVA_HIP_INVALID_ARGUMENT, "material " + std::to_string(m) + ": texture format mismatch");
    }
    MatRec &r = mats[m];
    r.albedo = d.albedo;
    r.metallic = d.metallic;
    r.normal = d.normal;
    r.emissive = d.emissive;
    r.albedo_f32 = d.albedo_f32;
    r.roughness_f32 = d.roughness_f32;
    for (int k = 0; k < 3; k++) {
      r.eta[k] = d.eta[k];
      r.k[k] = d.k[k];
    }
    r.anisotropy = d.anisotropy;
    r.emission_scale = d.emission_scale;
  }
  if (sd->envmap_texture >= (int32_t)sd->texture_count ||
      (sd->envmap_texture >= 0 && texrec[sd->envmap_texture].format != NOVA_HIP_TEX_F32))
    return set_error(ctx, NOVA_HIP_INVALID_ARGUMENT, "envmap must be an f32 texture");

  std::vector<float> pos(nvert * 3), nrm(nvert * 3), tan(nvert * 3), bit(nvert * 3), uv(nvert * 2), xf(sd->mesh_count * 16);
  std::vector<int32_t> mesh_mat(sd->mesh_count);
  std::vector<MeshXf> mxf(sd->mesh_count);
  std::vector<float4> xinv(sd->mesh_count * 3 + 3);  // never empty
  std::vector<TriRef> refs(ntri);
  uint64_t vo = 0, to = 0;
  for (uint32_t m = 0; m < sd->mesh_count; m++) {
    const nova_hip_mesh_desc &md = sd->meshes[m];
    size_t n3 = md.vertex_count * 3 * sizeof(float);
    if (md.vertex_count) {
      std::memcpy(&pos[vo * 3], md.vertices, n3);
      std::memcpy(&nrm[vo * 3], md.normals, n3);
      std::memcpy(&tan[vo * 3], md.tangents, n3);
      std::memcpy(&bit[vo * 3], md.bitangents, n3);
      std::memcpy(&uv[vo * 2], md.uv, md.vertex_count * 2 * sizeof(float));
    }
    std::memcpy(&xf[m * 16], md.transform, 16 * sizeof(float));
    mesh_xf_blocks(md.transform, mxf[m]);
    mesh_inverse_rows(md.transform, &xinv[m * 3]);
    mesh_mat[m] = md.material;
    for (uint64_t i = 0; i < md.index_count; i += 3)
      refs[to++] = TriRef{(uint32_t)(md.indices[i] + vo), (uint32_t)(md.indices[i + 1] + vo), (uint32_t)(md.indices[i + 2] + vo), m};
    vo += md.vertex_count;
  }
  // acos table as the reference's build step writes it (generate_luts_trigonometrics.py:11-20)
  std::vector<float> lut(kAcosLut);
  for (int i = 0; i < kAcosLut; i++) lut[i] = (float)std::acos(((double)i / (double)kAcosLut) * 2.0 - 1.0);

  std::vector<uint32_t> u32pool(nu32 ? nu32 : 1);
  std::vector<float> f32pool(nf32 ? nf32 : 1);
  for (uint32_t t = 0; t < sd->texture_count; t++) {
    const nova_hip_texture_desc &td = sd->textures[t];
    if (td.format == NOVA_HIP_TEX_U32)
      std::memcpy(&u32pool[texrec[t].offset], td.data, (size_t)td.width * td.height * 4);
    else
      std::memcpy(&f32pool[texrec[t].offset], td.data, (size_t)td.width * td.height * td.channels * 4);
  }

  // ---- device copies
  float *dpos = nullptr, *dnrm = nullptr, *dtan = nullptr, *dbit = nullptr, *duv = nullptr, *dxf = nullptr;
  int32_t *dmm = nullptr;
  TriRef *drefs = nullptr;
  hipError_t e = hipSuccess;
#define UP(dst, vec) \
  if (e == hipSuccess) e = upload(&dst, vec.data(), vec.size() * sizeof(vec[0]), s);
  UP(dpos, pos) UP(dnrm, nrm) UP(dtan, tan) UP(dbit, bit) UP(duv, uv) UP(dxf, xf) UP(dmm, mesh_mat) UP(drefs, refs)
  UP(ctx->mesh_xf, mxf) UP(ctx->xf_inv, xinv) UP(ctx->materials, mats) UP(ctx->textures, texrec) UP(ctx->tex_u32, u32pool) UP(ctx->tex_f32, f32pool)
  UP(ctx->acos_lut, lut)
#undef UP
  if (e == hipSuccess) e = hipMalloc(&ctx->tri_world, ntri * sizeof(TriMT));
  if (e == hipSuccess) e = hipMalloc(&ctx->shade, ntri * sizeof(ShadeTri));
  if (e == hipSuccess) e = hipMalloc(&ctx->prim_aabb, ntri * 6 * sizeof(float));
  if (e == hipSuccess) e = hipMalloc(&ctx->prim_centroid, ntri * 3 * sizeof(float));
  if (e == hipSuccess) {
    const unsigned block = 256;
    hipLaunchKernelGGL(flatten_kernel, dim3((unsigned)((ntri + block - 1) / block)), dim3(block), 0, s, (uint32_t)ntri, drefs, dpos, dnrm, dtan, dbit,
                       duv, dxf, dmm, ctx->tri_world, ctx->shade, ctx->prim_aabb, ctx->prim_centroid);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  for (void *p : {(void *)dpos, (void *)dnrm, (void *)dtan, (void *)dbit, (void *)duv, (void *)dxf, (void *)dmm, (void *)drefs})
    if (p) hipFree(p);
  if (e != hipSuccess) {
    nova_hip_free_scene_internal(ctx);
    return check_hip(ctx, e, "scene upload");
  }
  ctx->prim_count = (uint32_t)ntri;
  ctx->mesh_count = sd->mesh_count;
  ctx->envmap = sd->envmap_texture;
  ctx->host_textures = texrec;
  ctx->tex_f32_floats = nf32;
  ctx->tail_hint = 0;
  ctx->scene_ready = true;
  ctx->bvh_ready = false;
  ctx->scene_gen++;
  return NOVA_HIP_SUCCESS;
}

extern "C" int nova_hip_texture_f32_info(nova_hip_ctx *ctx, int32_t texture, uint64_t *offset, uint64_t *pool_floats) {
  if (!ctx) return NOVA_HIP_INVALID_ARGUMENT;
  if (!ctx->scene_ready) return set_error(ctx, NOVA_HIP_SCENE_NOT_PROCESSED, "no scene uploaded");
  if (texture < 0 || (size_t)texture >= ctx->host_textures.size() || ctx->host_textures[texture].format != NOVA_HIP_TEX_F32)
    return set_error(ctx, NOVA_HIP_INVALID_ARGUMENT, "not an FP32 texture of the scene");
  if (offset) *offset = ctx->host_textures[texture].offset;
  if (pool_floats) *pool_floats = ctx->tex_f32_floats;
  return NOVA_HIP_SUCCESS;
}

extern "C" int nova_hip_texture_f32_update(nova_hip_ctx *ctx, int32_t texture, const float *texels_dev, void *stream) {
  if (!ctx) return NOVA_HIP_INVALID_ARGUMENT;
  if (!texels_dev) return set_error(ctx, NOVA_HIP_INVALID_BUFFER_STATE, "null texel buffer");
  uint64_t off = 0;
  int st = nova_hip_texture_f32_info(ctx, texture, &off, nullptr);
  if (st) return st;
  const TexRec &t = ctx->host_textures[texture];
  const size_t n = (size_t)t.width * t.height * t.channels;
  NOVA_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  NOVA_HIP_CHECK(ctx, hipMemcpyAsync(ctx->tex_f32 + off, texels_dev, n * sizeof(float), hipMemcpyDeviceToDevice, (hipStream_t)stream));
  ctx->scene_gen++;
  return NOVA_HIP_SUCCESS;
}
