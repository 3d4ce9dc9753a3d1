// Oracle (test infrastructure only): C entry points for ctypes. Loaded only by
// tests/, the smoke check and bench.py's CPU-baseline leg.
#include "oracle.hpp"
#include "fmath.hpp"
#include <atomic>
#include <cstring>
#include <thread>

using namespace oracle;

extern "C" {

void oracle_sobol_bits(const uint32_t *idx, const uint32_t *dims, size_t n, uint32_t *out) {
  for (size_t i = 0; i < n; i++) out[i] = sobol_sample_bits(idx[i], dims[i]);
}

void oracle_sobol_float(const uint32_t *idx, const uint32_t *dims, size_t n, float *out) {
  for (size_t i = 0; i < n; i++) out[i] = sobol_sample(idx[i], dims[i]);
}

void oracle_direction_table(uint32_t *out) { std::memcpy(out, kSobolTable, sizeof(kSobolTable)); }

void oracle_acos_lut(float *out) {
  const auto &l = acos_lut();
  std::memcpy(out, l.data(), l.size() * sizeof(float));
}

void oracle_acos_lt(const float *a, size_t n, float *out) {
  for (size_t i = 0; i < n; i++) out[i] = acos_lt(a[i]);
}

void oracle_atan2_approx(const float *y, const float *x, size_t n, float *out) {
  for (size_t i = 0; i < n; i++) out[i] = atan2_approx(y[i], x[i]);
}

void *oracle_scene_new() { return new Scene(); }
void oracle_scene_free(void *s) { delete static_cast<Scene *>(s); }

int oracle_scene_add_texture_u32(void *sp, int w, int h, int channels, int inv_u, int inv_v, const uint32_t *data) {
  Scene &s = *static_cast<Scene *>(sp);
  Texture t;
  t.width = w;
  t.height = h;
  t.channels = channels;
  t.invert_u = inv_u != 0;
  t.invert_v = inv_v != 0;
  t.u32.assign(data, data + (size_t)w * h);
  s.textures.push_back(std::move(t));
  return (int)s.textures.size() - 1;
}

int oracle_scene_add_texture_f32(void *sp, int w, int h, int channels, int inv_u, int inv_v, const float *data) {
  Scene &s = *static_cast<Scene *>(sp);
  Texture t;
  t.width = w;
  t.height = h;
  t.channels = channels;
  t.invert_u = inv_u != 0;
  t.invert_v = inv_v != 0;
  t.f32.assign(data, data + (size_t)w * h * channels);
  s.textures.push_back(std::move(t));
  return (int)s.textures.size() - 1;
}

// tex: albedo, metallic, normal, emissive (u32 ids), albedo_f32, roughness_f32 (f32 ids); -1 = absent
int oracle_scene_add_material(void *sp, const int32_t *tex, const float *eta, const float *k, float aniso, float emission_scale) {
  Scene &s = *static_cast<Scene *>(sp);
  Material m;
  m.albedo = tex[0];
  m.metallic = tex[1];
  m.normal = tex[2];
  m.emissive = tex[3];
  m.albedo_f32 = tex[4];
  m.roughness_f32 = tex[5];
  for (int i = 0; i < 3; i++) {
    m.eta[i] = eta[i];
    m.k[i] = k[i];
  }
  m.anisotropy = aniso;
  m.emission_scale = emission_scale;
  s.materials.push_back(m);
  return (int)s.materials.size() - 1;
}

int oracle_scene_add_mesh(void *sp, size_t nv, const float *vertices, const float *normals, const float *tangents, const float *bitangents,
                          const float *uv, size_t ni, const uint32_t *indices, const float *transform, int material) {
  Scene &s = *static_cast<Scene *>(sp);
  Mesh m;
  m.vertices.assign(vertices, vertices + nv * 3);
  m.normals.assign(normals, normals + nv * 3);
  m.tangents.assign(tangents, tangents + nv * 3);
  m.bitangents.assign(bitangents, bitangents + nv * 3);
  m.uv.assign(uv, uv + nv * 2);
  m.indices.assign(indices, indices + ni);
  if (transform)
    for (int i = 0; i < 16; i++) m.transform.m[i] = transform[i];
  m.material = material;
  s.meshes.push_back(std::move(m));
  return (int)s.meshes.size() - 1;
}

void oracle_scene_set_envmap(void *sp, int tex) { static_cast<Scene *>(sp)->envmap = tex; }
void oracle_scene_finalize(void *sp) { static_cast<Scene *>(sp)->finalize(); }
size_t oracle_scene_prim_count(void *sp) { return static_cast<Scene *>(sp)->prim_mesh.size(); }

void *oracle_bvh_build(void *sp) {
  Bvh *b = new Bvh();
  bvh_build(*static_cast<Scene *>(sp), *b);
  return b;
}
void oracle_bvh_free(void *b) { delete static_cast<Bvh *>(b); }
int oracle_bvh_nodes_used(void *b) { return static_cast<Bvh *>(b)->nodes_used; }
// nodes: nodes_used x 8 words (min3, max3 as float bits; left, count as int32)
void oracle_bvh_get(void *bp, int32_t *prim_idx, void *nodes) {
  Bvh &b = *static_cast<Bvh *>(bp);
  std::memcpy(prim_idx, b.prim_idx.data(), b.prim_idx.size() * sizeof(int32_t));
  std::memcpy(nodes, b.nodes.data(), (size_t)b.nodes_used * sizeof(BvhNode));
}

// rays: n x 6 floats (origin, direction); out prim (-1 = miss) and t,u,v
void oracle_closest_hit(void *sp, void *bp, const float *rays, size_t n, int brute, int threads, int32_t *prim, float *tuv) {
  Scene &s = *static_cast<Scene *>(sp);
  std::atomic<size_t> next{0};
  auto work = [&]() {
    for (;;) {
      size_t k0 = next.fetch_add(256);
      if (k0 >= n) break;
      for (size_t k = k0; k < k0 + 256 && k < n; k++) {
        vec3 o(rays[k * 6], rays[k * 6 + 1], rays[k * 6 + 2]), d(rays[k * 6 + 3], rays[k * 6 + 4], rays[k * 6 + 5]);
        float t = 0, u = 0, v = 0;
        int p = brute ? closest_hit_brute(s, o, d, 0.000001f, t, u, v) : closest_hit(s, *static_cast<Bvh *>(bp), o, d, 0.000001f, t, u, v);
        prim[k] = p;
        tuv[k * 3] = t;
        tuv[k * 3 + 1] = u;
        tuv[k * 3 + 2] = v;
      }
    }
  };
  std::vector<std::thread> pool;
  for (int i = 0; i < (threads > 1 ? threads : 1); i++) pool.emplace_back(work);
  for (auto &t : pool) t.join();
}

// Progressive CPU render: frames [frame_start, frame_start + frames) at 1 spp each,
// accumulated as Engine::writeFramebuffer does. accum/final: width*height*4 floats.
static RenderParams make_params(int width, int height, int frame_start, int frames, int max_depth, uint32_t max_samples_offset,
                                int vertical_invert, int rr_start, const float *inv_P, const float *inv_V, int threads, int sampling_roughness) {
  RenderParams p;
  p.width = width;
  p.height = height;
  p.frame_start = frame_start;
  p.frames = frames;
  p.max_depth = max_depth;
  p.max_samples_offset = max_samples_offset;
  p.vertical_invert = vertical_invert != 0;
  p.rr_start = rr_start;
  for (int i = 0; i < 16; i++) {
    p.inv_P.m[i] = inv_P[i];
    p.inv_V.m[i] = inv_V[i];
  }
  p.threads = threads;
  p.sampling_roughness = sampling_roughness;
  p.index_policy = g_index_policy;
  return p;
}

// Renders with the sampling roughness taken from FP32 texture sampling_roughness (-1: off).
void oracle_render_sampling(void *sp, void *bp, int width, int height, int frame_start, int frames, int max_depth, uint32_t max_samples_offset,
                            int vertical_invert, int rr_start, const float *inv_P, const float *inv_V, int threads, int sampling_roughness,
                            float *accum, float *final_out) {
  RenderParams p = make_params(width, height, frame_start, frames, max_depth, max_samples_offset, vertical_invert, rr_start, inv_P, inv_V, threads,
                               sampling_roughness);
  render(*static_cast<Scene *>(sp), *static_cast<Bvh *>(bp), p, accum, final_out);
}

void oracle_render_adjoint(void *sp, void *bp, int width, int height, int frame_start, int frames, int max_depth, uint32_t max_samples_offset,
                           int vertical_invert, int rr_start, const float *inv_P, const float *inv_V, int threads, int sampling_roughness,
                           const float *d_image, float scale, double *grad_pool) {
  RenderParams p = make_params(width, height, frame_start, frames, max_depth, max_samples_offset, vertical_invert, rr_start, inv_P, inv_V, threads,
                               sampling_roughness);
  render_adjoint(*static_cast<Scene *>(sp), *static_cast<Bvh *>(bp), p, d_image, scale, grad_pool);
}

// 1: the host libm's float cosf/sinf/asinf/powf as the reference's CPU build calls them;
// 0 (default): evaluated in double and rounded once (what the HIP path computes).
void oracle_set_float_libm(int on) { set_float_libm(on != 0); }

void oracle_render_pass(void *sp, void *bp, int width, int height, int frame_start, int frames, int vertical_invert, const float *inv_P,
                        const float *inv_V, int threads, int pass, float z_near, float z_far, float *accum, float *final_out) {
  RenderParams p = make_params(width, height, frame_start, frames, 0, 0, vertical_invert, -1, inv_P, inv_V, threads, -1);
  render_pass(*static_cast<Scene *>(sp), *static_cast<Bvh *>(bp), p, pass, z_near, z_far, accum, final_out);
}

// Index policy of the renders the entry points below build (0: CPU, 1: the GPU raygen's).
void oracle_set_index_policy(int policy) { g_index_policy = policy; }

// Sobol sample index of pixel (x, y) in `frame` for a width x height render.
uint32_t oracle_sample_index(int width, int height, int vertical_invert, int policy, uint32_t offset, int x, int y, int frame) {
  RenderParams p;
  p.width = width;
  p.height = height;
  p.vertical_invert = vertical_invert != 0;
  p.index_policy = policy;
  p.max_samples_offset = offset;
  return sample_index(p, x, y, frame);
}

uint64_t oracle_f32_pool_offset(void *sp, int texture) { return f32_pool_offset(*static_cast<Scene *>(sp), texture); }

void oracle_render(void *sp, void *bp, int width, int height, int frame_start, int frames, int max_depth, uint32_t max_samples_offset,
                   int vertical_invert, int rr_start, const float *inv_P, const float *inv_V, int threads, float *accum, float *final_out) {
  RenderParams p;
  p.width = width;
  p.height = height;
  p.frame_start = frame_start;
  p.frames = frames;
  p.max_depth = max_depth;
  p.max_samples_offset = max_samples_offset;
  p.vertical_invert = vertical_invert != 0;
  p.rr_start = rr_start;
  for (int i = 0; i < 16; i++) {
    p.inv_P.m[i] = inv_P[i];
    p.inv_V.m[i] = inv_V[i];
  }
  p.threads = threads;
  p.index_policy = g_index_policy;
  render(*static_cast<Scene *>(sp), *static_cast<Bvh *>(bp), p, accum, final_out);
}

// Rays entering scatter index `bounce` of every path of frame `frame` (width*height*6 floats).
void oracle_path_rays(void *sp, void *bp, int width, int height, int frame, int bounce, int max_depth, int rr_start, const float *inv_P,
                      const float *inv_V, int threads, float *rays) {
  RenderParams p;
  p.width = width;
  p.height = height;
  p.max_depth = max_depth;
  p.rr_start = rr_start;
  for (int i = 0; i < 16; i++) {
    p.inv_P.m[i] = inv_P[i];
    p.inv_V.m[i] = inv_V[i];
  }
  p.threads = threads;
  p.index_policy = g_index_policy;
  path_rays(*static_cast<Scene *>(sp), *static_cast<Bvh *>(bp), p, frame, bounce, rays);
}

// Rows [y0, y1) of frame `frame` into partial (width*height*4 floats), `threads` workers.
void oracle_render_rows(void *sp, void *bp, int width, int height, int frame, int y0, int y1, int max_depth, int rr_start, const float *inv_P,
                        const float *inv_V, int threads, float *partial) {
  RenderParams p;
  p.width = width;
  p.height = height;
  p.max_depth = max_depth;
  p.rr_start = rr_start;
  for (int i = 0; i < 16; i++) {
    p.inv_P.m[i] = inv_P[i];
    p.inv_V.m[i] = inv_V[i];
  }
  p.threads = threads;
  render_rows(*static_cast<Scene *>(sp), *static_cast<Bvh *>(bp), p, frame, y0, y1, partial);
}

void oracle_bake_irradiance(const float *env, int w, int h, int c, int ow, int oh, int samples, int threads, float *out) {
  bake_irradiance(env, w, h, c, ow, oh, samples, out, threads);
}

// sampler: 0 pgc3d (reference), 1 Owen-scrambled Sobol
void oracle_bake_irradiance_sampler(const float *env, int w, int h, int c, int ow, int oh, int samples, int threads, int sampler, float *out) {
  bake_irradiance(env, w, h, c, ow, oh, samples, out, threads, sampler);
}

void oracle_brdf_lut_value(float ndotv, float roughness, int samples, float *ab) { brdf_lut_value(ndotv, roughness, samples, ab[0], ab[1]); }
void oracle_bake_brdf_lut(int ow, int oh, int samples, int threads, float *out) { bake_brdf_lut(ow, oh, samples, out, threads); }
void oracle_bake_cubemap(const float *env, int w, int h, int c, int size, int threads, float *out) { bake_cubemap(env, w, h, c, size, out, threads); }
void oracle_bake_irradiance_riemann(const float *env, int w, int h, int c, int ow, int oh, int threads, float *out) {
  bake_irradiance_riemann(env, w, h, c, ow, oh, out, threads);
}
void oracle_cube_texel_direction(int face, int x, int y, int size, float *out) {
  const vec3 d = cube_texel_direction(face, x, y, size);
  out[0] = d.x;
  out[1] = d.y;
  out[2] = d.z;
}

void oracle_bake_prefilter(const float *env, int w, int h, int c, int ow, int oh, float roughness, int samples, int threads, float *out) {
  bake_prefilter(env, w, h, c, ow, oh, roughness, samples, out, threads);
}

}  // extern "C"

namespace oracle {
float unit_fresnel_real(float eta, float cos_i);
vec3 unit_correct_shading_normal(const vec3 &ng, const vec3 &wo, const vec3 &ns);
vec3 unit_vndf_sample(float aniso, float rough, const vec3 &wo, const float u[2]);
float unit_vndf_G1(float aniso, float rough, const vec3 &v);
float unit_vndf_D(float aniso, float rough, const vec3 &wm);
int unit_principled_sample(const vec3 &wo, const vec3 &albedo, float roughness, float metal, float eta, float uc, float u0, float u1, float uc2,
                           vec3 &f, float &pdf, vec3 &wi, float &costheta, int &valid);
int unit_principled_eval(const vec3 &wo, const vec3 &wi, const vec3 &albedo, float roughness, float metal, float eta, float uc, vec3 &f, float &pdf);
void unit_make_hit(const Scene &s, int prim, const vec3 &ro, const vec3 &rd, float t, float bu, float bv, float *out);
}  // namespace oracle

extern "C" {
float oracle_fresnel_real(float eta, float cos_i) { return unit_fresnel_real(eta, cos_i); }
void oracle_correct_shading_normal(const float *ng, const float *wo, const float *ns, float *out) {
  vec3 r = unit_correct_shading_normal(vec3(ng[0], ng[1], ng[2]), vec3(wo[0], wo[1], wo[2]), vec3(ns[0], ns[1], ns[2]));
  out[0] = r.x;
  out[1] = r.y;
  out[2] = r.z;
}
void oracle_vndf_sample(float aniso, float rough, const float *wo, const float *u, float *out) {
  vec3 r = unit_vndf_sample(aniso, rough, vec3(wo[0], wo[1], wo[2]), u);
  out[0] = r.x;
  out[1] = r.y;
  out[2] = r.z;
}
float oracle_vndf_G1(float aniso, float rough, const float *v) { return unit_vndf_G1(aniso, rough, vec3(v[0], v[1], v[2])); }
float oracle_vndf_D(float aniso, float rough, const float *v) { return unit_vndf_D(aniso, rough, vec3(v[0], v[1], v[2])); }
// out: f3, pdf, wi3, costheta, valid, lobe (10 floats)
void oracle_principled_sample(const float *wo, const float *albedo, float roughness, float metal, float eta, const float *u4, float *out) {
  vec3 f, wi;
  float pdf = 0, ct = 0;
  int valid = 0;
  int lobe = unit_principled_sample(vec3(wo[0], wo[1], wo[2]), vec3(albedo[0], albedo[1], albedo[2]), roughness, metal, eta, u4[0], u4[1], u4[2],
                                    u4[3], f, pdf, wi, ct, valid);
  float vals[10] = {f.x, f.y, f.z, pdf, wi.x, wi.y, wi.z, ct, (float)valid, (float)lobe};
  for (int i = 0; i < 10; i++) out[i] = vals[i];
}
void oracle_make_hit(void *sp, int prim, const float *ray, float t, float bu, float bv, float *out) {
  unit_make_hit(*static_cast<Scene *>(sp), prim, vec3(ray[0], ray[1], ray[2]), vec3(ray[3], ray[4], ray[5]), t, bu, bv, out);
}
// out: f3, pdf, lobe (5 floats)
void oracle_principled_eval(const float *wo, const float *wi, const float *albedo, float roughness, float metal, float eta, float uc, float *out) {
  vec3 f;
  float pdf = 0;
  int lobe = unit_principled_eval(vec3(wo[0], wo[1], wo[2]), vec3(wi[0], wi[1], wi[2]), vec3(albedo[0], albedo[1], albedo[2]), roughness, metal, eta, uc,
                                  f, pdf);
  out[0] = f.x;
  out[1] = f.y;
  out[2] = f.z;
  out[3] = pdf;
  out[4] = (float)lobe;
}
}  // extern "C"
