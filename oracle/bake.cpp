// Oracle (test infrastructure only): environment-map bakes.
//  * diffuse irradiance: TextureOperations<float>::computeDiffuseIrradiance and
//    helpers (TextureProcessing.cpp:9-64, TextureProcessing.h:173-301);
//  * GGX prefiltered specular: the GLSL baker envmap_prefilter.frag:1-116 restated
//    on an equirect source with level-0 sampling (no CPU version exists in the
//    reference, so this one is pinned only by closed forms).
#include "oracle.hpp"
#include "fmath.hpp"
#include <atomic>
#include <cmath>
#include <thread>

namespace oracle {

// math::importance_sampling::pgc3d (math_importance_sampling.h:7-21)
static vec3 pgc3d(unsigned x, unsigned y, unsigned z) {
  x = x * 1664525u + 1013904223u;
  y = y * 1664525u + 1013904223u;
  z = z * 1664525u + 1013904223u;
  x += y * z;
  y += z * x;
  z += x * y;
  x ^= x >> 16u;
  y ^= y >> 16u;
  z ^= z >> 16u;
  x += y * z;
  y += z * x;
  z += x * y;
  const float s = 1.f / float(0xFFFFFFFFu);
  return vec3((float)x * s, (float)y * s, (float)z * s);
}

static unsigned uv_to_pixel(float coord, unsigned dim) { return (unsigned)(int64_t)(coord * (float)dim); }

// sphericalToCartesian<float> (math_spherical.h:33-39)
static vec3 sph_to_cart(float phi, float theta) {
  const float z = cos_f(theta);
  const float x = sin_f(theta) * cos_f(phi);
  const float y = sin_f(theta) * sin_f(phi);
  return {x, y, z};
}

struct EnvView {
  const float *data;
  int w, h, c;
  // discreteSample + wrapAroundPixelCoords (TextureProcessing.h:173-180, 234-251)
  vec3 discrete(int x, int y) const {
    int xc, yc;
    if (x >= w) xc = x % w;
    else if (x < 0) xc = w + (x % w);
    else xc = x;
    if (y >= h) yc = y % h;
    else if (y < 0) yc = h + (y % h);
    else yc = y;
    int idx = (yc * w + xc) * c;
    return {data[idx], data[idx + 1], data[idx + 2]};
  }
  // uvSample (TextureProcessing.h:199-210): integer pixel coords give the top-left texel.
  vec3 uv_sample(float u, float v) const {
    float ui = std::floor(u), vi = std::floor(v);
    float up = (u > 1. || u < 0.) ? u - ui : u;
    float vp = (v > 1. || v < 0.) ? v - vi : v;
    float px = (float)uv_to_pixel(up, (unsigned)w), py = (float)uv_to_pixel(vp, (unsigned)h);
    float tlx = std::floor(px), tly = std::floor(py);
    float fu = (px - tlx) / ((tlx + 1) - tlx), fv = (py - tly) / ((tly + 1) - tly);
    vec3 tl = discrete((int)tlx, (int)tly), tr = discrete((int)(tlx + 1), (int)tly);
    vec3 bl = discrete((int)tlx, (int)(tly + 1)), br = discrete((int)(tlx + 1), (int)(tly + 1));
    float hd = 1 - fu;
    vec3 top = hd * tl + fu * tr;
    vec3 bot = hd * bl + fu * br;
    return (1 - fv) * top + fv * bot;
  }
  vec3 dir_sample(const vec3 &d) const {
    float phi = atan2_approx(d.y, d.x), theta = acos_lt(d.z);
    const double inv_2pi = 1.f / (2.f * M_PI), inv_pi = 1.f / M_PI;
    float u = (float)(phi * inv_2pi), v = (float)(theta * inv_pi);
    return uv_sample(u, v);
  }
};

template <class F>
static void parallel_rows(int n, int threads, F &&fn) {
  std::atomic<int> next{0};
  auto worker = [&]() {
    for (;;) {
      int k = next.fetch_add(1);
      if (k >= n) break;
      fn(k);
    }
  };
  if (threads <= 1) {
    worker();
    return;
  }
  std::vector<std::thread> pool;
  for (int i = 0; i < threads; i++) pool.emplace_back(worker);
  for (auto &t : pool) t.join();
}

// texel (i, j) of an out_w x out_h map: TextureProcessing.h:253-301
// sampler 0: pgc3d (the reference's); 1: Owen-scrambled Sobol dims 0-1 at index
// texel * samples + s (the Sobol-sampled bake extension).
static vec3 irradiance_texel(const EnvView &env, unsigned i, unsigned j, unsigned ow, unsigned oh, unsigned samples, int sampler) {
  float u = (float)((double)i / (double)ow), v = (float)((double)j / (double)oh);
  float phi = (float)(2 * M_PI * u), theta = (float)(M_PI * v);
  vec3 nrm = sph_to_cart(phi, theta);
  float aphi = atan2_approx(nrm.y, nrm.x), atheta = acos_lt(nrm.z);
  const double inv_2pi = 1.f / (2.f * M_PI), inv_pi = 1.f / M_PI;
  float uu = (float)(aphi * inv_2pi), vv = (float)(atheta * inv_pi);
  const unsigned px = uv_to_pixel(uu, ow), py = uv_to_pixel(vv, oh);
  vec3 right(1.f, 0.f, 0.f);
  float dd = dot(right, nrm);
  vec3 tangent(0.f, 1.f, 0.f);
  if (1.0 - std::fabs(dd) > 1e-6) tangent = normalize(cross(right, nrm));
  vec3 bitangent = cross(nrm, tangent);
  vec3 irr(0.f);
  for (unsigned s = 0; s < samples; s++) {
    vec3 r;
    if (sampler == 1) {
      const uint32_t idx = (j * ow + i) * samples + s;
      r = vec3(sobol_sample(idx, 0), sobol_sample(idx, 1), 0.f);
    } else {
      r = pgc3d(px, py, s);
    }
    float sphi = (float)(2 * M_PI * r.x);
    float stheta = asin_f(std::sqrt(r.y));
    vec3 c = sph_to_cart(sphi, stheta);
    c = c.x * tangent + c.y * bitangent + c.z * nrm;
    irr = irr + env.dir_sample(c);
  }
  return irr / (float)samples;
}

void bake_irradiance(const float *env, int env_w, int env_h, int env_c, int out_w, int out_h, int samples, float *out, int threads, int sampler) {
  EnvView ev{env, env_w, env_h, env_c};
  parallel_rows(out_w, threads, [&](int i) {
    for (int j = 0; j < out_h; j++) {
      vec3 irr = irradiance_texel(ev, (unsigned)i, (unsigned)j, (unsigned)out_w, (unsigned)out_h, (unsigned)samples, sampler);
      unsigned idx = ((unsigned)j * out_w + i) * env_c;
      out[idx] = irr.x;
      out[idx + 1] = irr.y;
      out[idx + 2] = irr.z;
      for (int k = 3; k < env_c; k++) out[idx + k] = 0.f;
    }
  });
}

// envmap_prefilter.frag:28-35 radicalInverse and Hammersley
static float radical_inverse(unsigned bits) {
  bits = (bits << 16u) | (bits >> 16u);
  bits = ((bits & 0x55555555u) << 1u) | ((bits & 0xAAAAAAAAu) >> 1u);
  bits = ((bits & 0x33333333u) << 2u) | ((bits & 0xCCCCCCCCu) >> 2u);
  bits = ((bits & 0x0F0F0F0Fu) << 4u) | ((bits & 0xF0F0F0F0u) >> 4u);
  bits = ((bits & 0x00FF00FFu) << 8u) | ((bits & 0xFF00FF00u) >> 8u);
  return (float)bits * 2.3283064365386963e-10f;
}

// ImportanceSampleGGX (envmap_prefilter.frag:52-70)
static vec3 importance_sample_ggx(float xi0, float xi1, const vec3 &N, float r) {
  const float PI = 3.14159265f;
  float a = r * r;
  float phi = 2.0f * PI * xi0;
  float ct = std::sqrt((1.0f - xi1) / (1.0f + (a * a - 1.0f) * xi1));
  float st = std::sqrt(1.0f - ct * ct);
  vec3 H(cos_f(phi) * st, sin_f(phi) * st, ct);
  vec3 up = std::fabs(N.z) < 0.999f ? vec3(0.f, 0.f, 1.f) : vec3(1.f, 0.f, 0.f);
  vec3 t = normalize(cross(up, N));
  vec3 b = cross(N, t);
  vec3 sv = t * H.x + b * H.y + N * H.z;
  return normalize(sv);
}

void bake_prefilter(const float *env, int env_w, int env_h, int env_c, int out_w, int out_h, float roughness, int samples, float *out, int threads) {
  EnvView ev{env, env_w, env_h, env_c};
  parallel_rows(out_h, threads, [&](int j) {
    for (int i = 0; i < out_w; i++) {
      float u = (float)((double)i / (double)out_w), v = (float)((double)j / (double)out_h);
      vec3 N = normalize(sph_to_cart((float)(2 * M_PI * u), (float)(M_PI * v)));
      vec3 col(0.f);
      if (roughness == 0.f) {
        col = ev.dir_sample(N);
      } else {
        float total = 0.f;
        for (int s = 0; s < samples; s++) {
          vec3 H = importance_sample_ggx((float)s / (float)samples, radical_inverse((unsigned)s), N, roughness);
          vec3 L = reflect(-N, H);
          float ndl = std::fmax(dot(N, L), 0.f);
          if (ndl > 0.f) {
            col = col + ev.dir_sample(L) * ndl;
            total += ndl;
          }
        }
        col = col / total;
      }
      unsigned idx = ((unsigned)j * out_w + i) * env_c;
      out[idx] = col.x;
      out[idx + 1] = col.y;
      out[idx + 2] = col.z;
      for (int k = 3; k < env_c; k++) out[idx + k] = 0.f;
    }
  });
}

}  // namespace oracle

namespace oracle {

// GeometrySchlickGGX with k = roughness^2 / 2 (brdf_lookup_table_baker.frag:9-17)
static float schlick_ggx_ibl(float ndotv, float roughness) {
  float a = roughness;
  float k = (a * a) / 2.0f;
  return ndotv / (ndotv * (1.0f - k) + k);
}

// computeLut (brdf_lookup_table_baker.frag:74-113): the split-sum scale A and bias B of
// F0 for one (NdotV, roughness), `samples` Hammersley GGX half vectors about N = +z.
void brdf_lut_value(float ndotv, float roughness, int samples, float &A, float &B) {
  const vec3 V(std::sqrt(1.0f - ndotv * ndotv), 0.0f, ndotv);
  const vec3 N(0.f, 0.f, 1.f);
  float a = 0.f, b = 0.f;
  for (int i = 0; i < samples; i++) {
    const vec3 H = importance_sample_ggx((float)i / (float)samples, radical_inverse((unsigned)i), N, roughness);
    const vec3 L = normalize(H * (2.0f * dot(V, H)) - V);
    const float ndl = std::fmax(L.z, 0.0f);
    const float ndh = std::fmax(H.z, 0.0f);
    const float vdh = std::fmax(dot(V, H), 0.0f);
    if (ndl > 0.0f) {
      const float G = schlick_ggx_ibl(ndl, roughness) * schlick_ggx_ibl(std::fmax(dot(N, V), 0.0f), roughness);
      const float g_vis = (G * vdh) / (ndh * ndotv);
      const float fc = pow5(1.0f - vdh);
      a += (1.0f - fc) * g_vis;
      b += fc * g_vis;
    }
  }
  A = a / (float)samples;
  B = b / (float)samples;
}

// The LUT texture: texel (i, j) at NdotV = (i + 0.5) / out_w, roughness = (j + 0.5) / out_h
// (the full-screen quad's texture coordinates at pixel centres); out is out_w x out_h x 2.
void bake_brdf_lut(int out_w, int out_h, int samples, float *out, int threads) {
  parallel_rows(out_h, threads, [&](int j) {
    for (int i = 0; i < out_w; i++) {
      float A, B;
      brdf_lut_value(((float)i + 0.5f) / (float)out_w, ((float)j + 0.5f) / (float)out_h, samples, A, B);
      out[((size_t)j * out_w + i) * 2] = A;
      out[((size_t)j * out_w + i) * 2 + 1] = B;
    }
  });
}

// ------------------------------------------------ GL viewer bakes on equirect maps
// The viewer's cube-map captures (RenderPipeline.cpp:14-26, renderToCubemap :186-206) and
// its two shaders, restated on the CPU:
//  * envmap_bake.frag:10-24: a face texel's direction -> sphericalToUv -> the equirect map;
//  * irradiance_baker.frag:15-39: the Riemann sum over azimuth and zenith steps of 0.0075.
// Lookup rule (GL sampler2D state of the equirect texture: RenderPipeline's glTexParameteri
// calls name the cube-map target, so the 2D texture keeps GL's default REPEAT wrap; level 0
// with GL_LINEAR filtering): texel centres at (i + 0.5) / w, bilinear weights, indices taken
// modulo the map size. atan, acos, cos and sin are evaluated in double and rounded once.

namespace {
struct GlEquirect {
  const float *data;
  int w, h, c;
  vec3 texel(int x, int y) const {
    x %= w;
    if (x < 0) x += w;
    y %= h;
    if (y < 0) y += h;
    const float *p = data + ((size_t)y * w + x) * c;
    return {p[0], p[1], p[2]};
  }
  // texture(environment_map, vec2(u, v)) with GL_LINEAR and GL_REPEAT
  vec3 bilinear(float u, float v) const {
    const float x = u * (float)w - 0.5f, y = v * (float)h - 0.5f;
    const float fx = std::floor(x), fy = std::floor(y);
    const float a = x - fx, b = y - fy;
    const int i0 = (int)fx, j0 = (int)fy;
    const vec3 t00 = texel(i0, j0), t10 = texel(i0 + 1, j0), t01 = texel(i0, j0 + 1), t11 = texel(i0 + 1, j0 + 1);
    return ((1.f - a) * (1.f - b)) * t00 + (a * (1.f - b)) * t10 + ((1.f - a) * b) * t01 + (a * b) * t11;
  }
  // sphericalToUv (envmap_bake.frag:10-19) of a unit direction, then the lookup
  vec3 lookup(const vec3 &d) const {
    const float PI = 3.14159265359f;
    const float theta = (float)std::atan2((double)-d.z, (double)d.x), phi = (float)std::acos((double)d.y);
    return bilinear(theta / (2.0f * PI), phi / PI);
  }
};

vec3 gl_normalize(const vec3 &v) {  // v * inversesqrt(dot(v, v))
  const float inv = 1.f / std::sqrt((v.x * v.x + v.y * v.y) + v.z * v.z);
  return v * inv;
}

// glm::lookAt basis with the eye at the origin: f = normalize(center), s = normalize(f x up),
// u = s x f; the face directions of RenderPipeline.cpp:14-26 in capture order
// (left, right, down, up, back, front).
void capture_basis(int face, vec3 &s, vec3 &u, vec3 &f) {
  static const float centers[6][3] = {{-1, 0, 0}, {1, 0, 0}, {0, -1, 0}, {0, 1, 0}, {0, 0, 1}, {0, 0, -1}};
  static const float ups[6][3] = {{0, 1, 0}, {0, 1, 0}, {0, 0, 1}, {0, 0, -1}, {0, 1, 0}, {0, 1, 0}};
  f = gl_normalize(vec3(centers[face][0], centers[face][1], centers[face][2]));
  const vec3 up(ups[face][0], ups[face][1], ups[face][2]);
  s = gl_normalize(cross(f, up));
  u = cross(s, f);
}
}  // namespace

// Face texel (x, y) of a size x size capture (y = 0 the bottom row, as GL's window
// coordinates): the 90-degree camera's ray through the pixel centre, ndc = 2 (x + 0.5) /
// size - 1, meets the unit cube at P = ndc_x s + ndc_y u + f, whose interpolated position the
// fragment shader normalizes.
vec3 cube_texel_direction(int face, int x, int y, int size) {
  vec3 s, u, f;
  capture_basis(face, s, u, f);
  const float nx = (float)(2 * x + 1) / (float)size - 1.f, ny = (float)(2 * y + 1) / (float)size - 1.f;
  return gl_normalize(nx * s + ny * u + f);
}

// Equirect -> cube map (bakeEnvmapToCubemap, RenderPipeline.cpp:28-62): out holds the six
// faces in capture order, face-major, rows bottom-up, `env_c` floats per texel (RGB, then
// zeros).
void bake_cubemap(const float *env, int env_w, int env_h, int env_c, int size, float *out, int threads) {
  const GlEquirect m{env, env_w, env_h, env_c};
  parallel_rows(6 * size, threads, [&](int row) {
    const int face = row / size, y = row % size;
    for (int x = 0; x < size; x++) {
      const vec3 col = m.lookup(cube_texel_direction(face, x, y, size));
      float *o = out + (((size_t)face * size + y) * size + x) * env_c;
      o[0] = col.x;
      o[1] = col.y;
      o[2] = col.z;
      for (int k = 3; k < env_c; k++) o[k] = 0.f;
    }
  });
}

// computeIrradiance (irradiance_baker.frag:15-34) at normal n, over the equirect map: the
// float loops phi = 0, 0.0075, ... <= 2 PI and theta = 0, ... <= PI / 2 exactly as GLSL
// steps them, each sample weighted by cos(theta) sin(theta), one running sum, then
// PI * sum / count.
vec3 riemann_irradiance(const float *env, int env_w, int env_h, int env_c, const vec3 &n) {
  const GlEquirect m{env, env_w, env_h, env_c};
  const float PI = 3.14159265358f, zenith_max = PI / 2, azimuth_max = 2 * PI, delta = 0.0075f;
  const vec3 UP(0.f, 1.f, 0.f);
  const vec3 right = gl_normalize(cross(UP, n));
  const vec3 up = gl_normalize(cross(n, right));
  vec3 irradiance(0.f);
  int count = 0;
  for (float phi = 0.f; phi <= azimuth_max; phi += delta) {
    const float cp = cos_f(phi), sp = sin_f(phi);
    for (float theta = 0.f; theta <= zenith_max; theta += delta) {
      const float ct = cos_f(theta), st = sin_f(theta);
      const float x = cp * st, y = st * sp, z = ct;
      const vec3 sample_vec = x * right + y * up + z * n;
      irradiance = irradiance + (m.lookup(gl_normalize(sample_vec)) * ct) * st;
      count++;
    }
  }
  return (PI * irradiance) * (1.f / (float)count);
}

// Equirect texel (i, j) -> direction, the inverse of sphericalToUv at the texel centre:
// azimuth 2 pi (i + 0.5) / w with x = cos, -z = sin; polar pi (j + 0.5) / h with y = cos.
vec3 equirect_texel_direction(int i, int j, int w, int h) {
  const double az = 2.0 * M_PI * ((double)i + 0.5) / (double)w, po = M_PI * ((double)j + 0.5) / (double)h;
  return vec3((float)(std::cos(az) * std::sin(po)), (float)std::cos(po), (float)(-std::sin(az) * std::sin(po)));
}

// The GL irradiance bake on equirect maps: out_w x out_h x env_c, texel (i, j) at
// equirect_texel_direction.
void bake_irradiance_riemann(const float *env, int env_w, int env_h, int env_c, int out_w, int out_h, float *out, int threads) {
  parallel_rows(out_h, threads, [&](int j) {
    for (int i = 0; i < out_w; i++) {
      const vec3 col = riemann_irradiance(env, env_w, env_h, env_c, equirect_texel_direction(i, j, out_w, out_h));
      float *o = out + ((size_t)j * out_w + i) * env_c;
      o[0] = col.x;
      o[1] = col.y;
      o[2] = col.z;
      for (int k = 3; k < env_c; k++) o[k] = 0.f;
    }
  });
}

}  // namespace oracle
