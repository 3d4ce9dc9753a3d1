// Oracle (test infrastructure only): CPU restatement of the reference's Nova
// hot path. Only tests/, the smoke check and bench.py's CPU-baseline leg load
// liboracle.so, and only as the checker of the HIP library or as the timed CPU
// baseline; nothing in the product links or calls it.
//
// Parity status per function (how each one is pinned):
//   sobol_sample / direction table ........ pinned (scipy unscrambled Sobol, net property, golden vectors)
//   acos_lut ............................... pinned (closed form math.acos)
//   atan2_approx ........................... pinned (error bound vs atan2, quadrant identities)
//   bvh_build (BvhtlBuilder) ............... pinned (brute-force partition, leaf invariants, SAH exhaustive small n)
//   closest hit (MT + native traversal) .... pinned (brute force over all triangles)
//   hit record (Embree path) ............... pinned by closed forms (unit-UV derivatives, the UV solve, hit
//                                            point, transformed normals, degenerate-UV basis); Embree supplies
//                                            only the barycentrics, with the same u->v1, v->v2 convention as MT
//   BxDFs / PrincipledBSDF ................. pinned by the reference BxDFTest invariants (sampled f/pdf =
//                                            separately written f/pdf, reciprocity, albedo <= 1, pdf closed
//                                            form, eta-1 transmission) + closed forms
//   PathIntegrator::Li ..................... pinned by analytic scenes (constant envmap, emission only)
//   irradiance bake ........................ pinned (constant map => constant, brute force of the texel rule)
//   prefilter bake ......................... pinned (no reference CPU implementation exists): linearity,
//                                            roughness 0 = the map, narrow lobes stay in a region, a cosine sky
//                                            against the GGX integral from the distribution
//   BRDF table (brdf_lut_value) ............. pinned (roughness 0 closed form; NdotV = 1 as a 1-D sum)
//   render passes (render_pass) ............. pinned (wall scene: plane distances, unit normals, texel / 255)
//   render_adjoint (extension) ............. pinned (central finite differences of the oracle's own renders)
//
// Deviations, kept identical in the HIP path: a 0/0 slab in BoundingBox::intersect spans the
// whole line (bvh.cpp), and the float cos/sin/asin/pow calls are evaluated in double and
// rounded once (fmath.hpp; set_float_libm(true) restores the host libm's float functions).
#pragma once
#include "vecmath.hpp"
#include <cstddef>
#include <cstdint>
#include <vector>

namespace oracle {

// ----------------------------------------------------------------- sampler
extern const uint32_t kSobolTable[256][32];
float sobol_sample(uint32_t index, unsigned dim, uint32_t seed = 0xDEADBEEFu);
uint32_t sobol_sample_bits(uint32_t index, unsigned dim, uint32_t seed = 0xDEADBEEFu);

struct SobolSampler {  // Sampler.h:16-55
  uint32_t index = 0;
  unsigned dim = 0;
  void reset(uint32_t idx) { index = idx; dim = 0; }
  float sample1D() { return sobol_sample(index, dim++); }
};

// ----------------------------------------------------------- approx trig
constexpr int kAcosLutPrecision = 500000;  // AXOMAE_LUT_MEDIUMP (math/CMakeLists.txt:26-35)
const std::vector<float> &acos_lut();
float acos_lt(float a);
float atan2_approx(float y, float x);

// ----------------------------------------------------------------- scene
struct Texture {
  int width = 0, height = 0, channels = 4;
  bool invert_u = false, invert_v = false;
  std::vector<uint32_t> u32;  // packed (a<<24)|(r<<16)|(g<<8)|b
  std::vector<float> f32;     // channels floats per texel
};

struct Material {
  int albedo = -1, metallic = -1, normal = -1, emissive = -1;  // u32 texture ids, -1 = absent
  int albedo_f32 = -1, roughness_f32 = -1;                     // FP32 texture extension, -1 = absent
  float eta[3] = {0, 0, 0}, k[3] = {0, 0, 0};
  float anisotropy = 0.f;
  float emission_scale = 1.f;  // FP32 emission scale extension (1 = reference behaviour)
};

struct Mesh {
  std::vector<float> vertices, normals, tangents, bitangents, uv;
  std::vector<uint32_t> indices;
  mat4 transform;  // model matrix
  int material = 0;
};

struct Scene {
  std::vector<Mesh> meshes;
  std::vector<Material> materials;
  std::vector<Texture> textures;
  int envmap = -1;  // texture id (f32)
  // flattened primitive table: global prim -> (mesh, first index offset)
  std::vector<uint32_t> prim_mesh, prim_tri;
  std::vector<mat4> inv, normal_mat;  // per mesh
  void finalize();
};

// ------------------------------------------------------------------- BVH
struct BvhNode {  // Bvhnl (BvhBuilder.h:10-15)
  float min[3], max[3];
  int32_t left, primitive_count;
};
struct Bvh {
  std::vector<BvhNode> nodes;
  std::vector<int32_t> prim_idx;
  int32_t nodes_used = 0;
};
void bvh_build(const Scene &scene, Bvh &out);
// native iter_traverse + Triangle::hit: returns prim id or -1, fills t/u/v (object-space barycentrics)
int closest_hit(const Scene &scene, const Bvh &bvh, const vec3 &o, const vec3 &d, float tnear, float &t, float &u, float &v);
int closest_hit_brute(const Scene &scene, const vec3 &o, const vec3 &d, float tnear, float &t, float &u, float &v);

// ------------------------------------------------------------- rendering
struct RenderParams {
  int width = 0, height = 0, frames = 1, max_depth = 0;
  int frame_start = 0;
  uint32_t max_samples_offset = 0;  // renderer_max_samples (0 on the Engine CPU path)
  bool vertical_invert = false;
  int rr_start = -1;  // Russian roulette extension: first scatter index that rolls; -1 = off
  mat4 inv_P, inv_V;
  int threads = 1;
  int tiles_w = 0, tiles_h = 0;  // 0 = one tile row per thread
  int sampling_roughness = -1;   // gradient-check harness: FP32 texture of the sampling roughness
  int index_policy = 0;          // 0: 4*p*frame + max_samples_offset (Integrator.h:80); 1: (x*H+y)*max_samples_offset + frame (integrator_optix.cu:43)
};
// Sobol index of pixel (x, y) (row 0 at the bottom) in `frame` under p.index_policy.
uint32_t sample_index(const RenderParams &p, int x, int y, int frame);
// Index policy the C entry points give the renders they build (test harness setting).
extern int g_index_policy;
vec4 path_Li(const Scene &scene, const Bvh &bvh, vec3 o, vec3 d, int depth, SobolSampler &sampler, const RenderParams &p, int scatter_index);
void render_frame(const Scene &scene, const Bvh &bvh, const RenderParams &p, int frame, float *partial);
void render(const Scene &scene, const Bvh &bvh, const RenderParams &p, float *accum, float *final_out);
void render_rows(const Scene &scene, const Bvh &bvh, const RenderParams &p, int frame, int y0, int y1, float *partial);
void path_rays(const Scene &scene, const Bvh &bvh, const RenderParams &p, int frame, int bounce, float *rays);
// Render passes (pass 1 normal, 2 depth, 3 emissive) accumulated over p's frames.
void render_pass(const Scene &scene, const Bvh &bvh, const RenderParams &p, int pass, float z_near, float z_far, float *accum, float *final_out);
// Reverse-mode gradient of a loss with d loss / d(frame radiance) = scale * d_image (RGBA per
// pixel, RGB used) for frames [frame_start, frame_start+frames), accumulated into grad_pool
// (one double per FP32 texel channel, textures concatenated in scene order).
void render_adjoint(const Scene &scene, const Bvh &bvh, const RenderParams &p, const float *d_image, float scale, double *grad_pool);
size_t f32_pool_offset(const Scene &scene, int texture);

// ---------------------------------------------------------------- bakes
void bake_irradiance(const float *env, int env_w, int env_h, int env_c, int out_w, int out_h, int samples, float *out, int threads, int sampler = 0);
void bake_prefilter(const float *env, int env_w, int env_h, int env_c, int out_w, int out_h, float roughness, int samples, float *out, int threads);
void brdf_lut_value(float ndotv, float roughness, int samples, float &A, float &B);
void bake_brdf_lut(int out_w, int out_h, int samples, float *out, int threads);
// The GL viewer's equirect -> cube map capture and Riemann-sum irradiance, on equirect maps.
vec3 cube_texel_direction(int face, int x, int y, int size);
vec3 equirect_texel_direction(int i, int j, int w, int h);
void bake_cubemap(const float *env, int env_w, int env_h, int env_c, int size, float *out, int threads);
vec3 riemann_irradiance(const float *env, int env_w, int env_h, int env_c, const vec3 &n);
void bake_irradiance_riemann(const float *env, int env_w, int env_h, int env_c, int out_w, int out_h, float *out, int threads);

}  // namespace oracle
