// Oracle (test infrastructure only): hit record (Embree-path semantics), texture
// sampling, Principled BSDF, the recursive PathIntegrator and the frame loop.
#include "oracle.hpp"
#include <limits>
#include "fmath.hpp"
#include <algorithm>
#include <atomic>
#include <cfloat>
#include <cmath>
#include <thread>

namespace oracle {

using Spec = vec3;  // RGB ColorSpectrum (Spectrum.h:18-59)

static inline float sqr(float v) { return v * v; }
static inline float msqrt(float v) { return std::sqrt(std::fmax(v, 0.f)); }  // math::sqrt<float> (math_utils.h:86-89)
static inline float smax(float a, float b) { return (a < b) ? b : a; }        // std::max
static inline float clampf(float v, float lo, float hi) { return (v < lo) ? lo : (hi < v) ? hi : v; }  // std::clamp
static inline bool spec_any(const Spec &s) { return s.x != 0.f || s.y != 0.f || s.z != 0.f; }        // operator bool
static inline bool spec_all_ne(const Spec &s, float v) { return s.x != v && s.y != v && s.z != v; }   // operator!=

// ---------------------------------------------------------------- textures

// NovaTextureInterface.h:14-25
static float wrap(float uv) {
  if (uv < 0.f) {
    float fp = std::ceil(uv) - uv;
    return 1.f - fp;
  }
  if (uv > 1.f) return uv - std::floor(uv);
  return uv;
}

// ImageTexture<uint32_t>::sample, CPU path: bilinear over 4 texels (TextureContext.h:129-167,
// NovaTextureInterface.h:56-145). Returns 0..255 channel values.
static vec4 sample_u32(const Texture &tx, float u, float v) {
  u = wrap(u);
  v = wrap(v);
  u = tx.invert_u ? 1 - u : u;
  v = tx.invert_v ? 1 - v : v;
  const int width = tx.width, height = tx.height;
  float pi = std::fabs(u * (float)(width - 1.f));
  float pj = std::fabs(v * (float)(height - 1.f));
  unsigned i = (unsigned)std::floor(pi), j = (unsigned)std::floor(pj);
  auto ensure_bound = [](unsigned idx, unsigned limit) {
    unsigned val = idx;
    if (val + 1 >= limit) val = limit - 2;
    if (val > limit) val = 0;
    return val;
  };
  i = ensure_bound(i, width);
  j = ensure_bound(j, height);
  float tx_ = (pi - (float)i) / (float)((i + 1) - i);
  float ty_ = (pj - (float)j) / (float)((j + 1) - j);
  uint32_t px[4] = {tx.u32[j * width + i], tx.u32[j * width + i + 1], tx.u32[(j + 1) * width + i], tx.u32[(j + 1) * width + i + 1]};
  vec4 c[4];
  for (int k = 0; k < 4; k++) {  // is_rgba = false: r = byte 2, g = byte 1, b = byte 0, a = byte 3
    uint32_t p = px[k];
    c[k] = vec4((float)((p >> 16) & 0xFF), (float)((p >> 8) & 0xFF), (float)(p & 0xFF), (float)(p >> 24));
    if (tx.channels == 3) c[k].w = 1.f;
  }
  float hd = 1 - tx_;
  vec4 top = hd * c[0] + tx_ * c[1];
  vec4 bot = hd * c[2] + tx_ * c[3];
  return (1 - ty_) * top + ty_ * bot;
}

// FP32 material texture extension: same bilinear rule on float texels (values used as-is).
static vec4 sample_f32_bilinear(const Texture &tx, float u, float v) {
  u = wrap(u);
  v = wrap(v);
  u = tx.invert_u ? 1 - u : u;
  v = tx.invert_v ? 1 - v : v;
  const int width = tx.width, height = tx.height, ch = tx.channels;
  float pi = std::fabs(u * (float)(width - 1.f));
  float pj = std::fabs(v * (float)(height - 1.f));
  unsigned i = (unsigned)std::floor(pi), j = (unsigned)std::floor(pj);
  if (i + 1 >= (unsigned)width) i = width - 2;
  if (j + 1 >= (unsigned)height) j = height - 2;
  float tx_ = pi - (float)i, ty_ = pj - (float)j;
  vec4 c[4];
  unsigned idx[4] = {j * width + i, j * width + i + 1, (j + 1) * width + i, (j + 1) * width + i + 1};
  for (int k = 0; k < 4; k++) {
    c[k] = vec4(0, 0, 0, 1);
    for (int q = 0; q < ch && q < 4; q++) c[k][q] = tx.f32[idx[k] * ch + q];
  }
  float hd = 1 - tx_;
  vec4 top = hd * c[0] + tx_ * c[1];
  vec4 bot = hd * c[2] + tx_ * c[3];
  return (1 - ty_) * top + ty_ * bot;
}

// uv2index (TextureContext.h:52-58) + math::texture::uvToPixel
static unsigned uv2index(float t, int dim) {
  float a = std::fabs(t);
  float rem = a - std::floor(a);
  rem = rem == 0 ? a : rem;
  return (unsigned)(int64_t)(rem * (float)(unsigned)dim);
}

// TextureCtx::f32pixel (TextureContext.h:196-223), nearest texel.
static vec4 f32pixel(const Texture &tx, float u, float v) {
  u = tx.invert_u ? 1 - u : u;
  v = tx.invert_v ? 1 - v : v;
  unsigned i = uv2index(u, tx.width - 1), j = uv2index(v, tx.height - 1);
  unsigned idx = (j * tx.width + i) * tx.channels;
  vec4 rgba(0.f);
  for (int p = 0; p < tx.channels && p < 4; p++) rgba[p] = tx.f32[idx + p];
  return rgba;
}

// EnvmapTexture::sample (NovaTextureInterface.h:161-171): z-up spherical mapping, approx trig.
static vec4 sample_envmap(const Scene &s, const vec3 &dir) {
  if (s.envmap < 0) return vec4(0.f);
  vec3 n = normalize(dir);
  float phi = atan2_approx(n.y, n.x);
  float theta = acos_lt(n.z);
  const double inv_2pi = 1.f / (2.f * M_PI), inv_pi = 1.f / M_PI;
  float u = (float)(phi * inv_2pi), v = (float)(theta * inv_pi);
  return f32pixel(s.textures[s.envmap], wrap(u), wrap(v));
}

// ------------------------------------------------------------- hit record

struct Hit {
  vec3 ng, dpdu, dpdv, position;
  float u = 0, v = 0, t = 0;
  mat3 frame;  // shading TBN (columns T, B, N)
};

// math::gram_shmidt 3-argument form (utils_3D.h:63-71)
static mat3 gram_shmidt3(const vec3 &normal, const vec3 &tangent, const vec3 &bitangent) {
  vec3 N = normalize(normal);
  vec3 T = normalize(tangent - dot(tangent, N) * N);
  vec3 B = cross(N, T);
  if (dot(B, bitangent) < 0.0f) B = -B;
  return mat3(T, B, N);
}

// math::make_onb (utils_3D.h:74-82)
static void make_onb(vec3 n, vec3 &b1, vec3 &b2) {
  n = normalize(n);
  float sign = std::copysign(1.f, n.z);
  float a = -1.f / (sign + n.z);
  float b = n.x * n.y * a;
  b1 = vec3(1.f + sign * n.x * n.x * a, sign * b, -sign * n.x);
  b2 = vec3(b, sign + n.y * n.y * a, -n.y);
}

template <class V>
static V bary(const V &p0, const V &p1, const V &p2, float w, float u, float v) {
  return p0 * w + p1 * u + p2 * v;
}
static float baryf(float p0, float p1, float p2, float w, float u, float v) { return p0 * w + p1 * u + p2 * v; }

static vec3 attr3(const std::vector<float> &a, uint32_t i) { return {a[i * 3], a[i * 3 + 1], a[i * 3 + 2]}; }

// computeDerivativesTriangle + fillShadingStructure + constructTbnFrame (embree.cpp:287-390)
static Hit make_hit(const Scene &s, int prim, const vec3 &ro, const vec3 &rd, float t, float bu, float bv) {
  const uint32_t mi = s.prim_mesh[prim];
  const Mesh &m = s.meshes[mi];
  const uint32_t tr = s.prim_tri[prim];
  const uint32_t i0 = m.indices[tr], i1 = m.indices[tr + 1], i2 = m.indices[tr + 2];
  const float w = 1 - (bu + bv);
  Hit h;
  float uv0[2] = {m.uv[i0 * 2], m.uv[i0 * 2 + 1]}, uv1[2] = {m.uv[i1 * 2], m.uv[i1 * 2 + 1]}, uv2[2] = {m.uv[i2 * 2], m.uv[i2 * 2 + 1]};
  h.u = baryf(uv0[0], uv1[0], uv2[0], w, bu, bv);
  h.v = baryf(uv0[1], uv1[1], uv2[1], w, bu, bv);

  const mat3 l = upper3(m.transform);
  const vec3 v0 = attr3(m.vertices, i0), v1 = attr3(m.vertices, i1), v2 = attr3(m.vertices, i2);
  float duv02x = uv0[0] - uv2[0], duv02y = uv0[1] - uv2[1];
  float duv12x = uv1[0] - uv2[0], duv12y = uv1[1] - uv2[1];
  vec3 dp02 = l * vec3(v0.x - v2.x, v0.y - v2.y, v0.z - v2.z);
  vec3 dp12 = l * vec3(v1.x - v2.x, v1.y - v2.y, v1.z - v2.z);
  float det = duv02x * duv12y - duv02y * duv12x;
  bool degen = std::fabs(det) < 1e-9f;
  float invdet = 1.f / det;
  vec3 dpdu = (duv12y * dp02 - duv02y * dp12) * invdet;
  vec3 dpdv = (duv02x * dp12 - duv12x * dp02) * invdet;
  vec3 n;
  if (degen || length(cross(dpdu, dpdv)) == 0) {
    n = normalize(cross(dp12, dp02));
    make_onb(n, dpdu, dpdv);
  } else {
    n = cross(dp02, dp12);
  }
  h.ng = normalize(n);
  h.position = ro + t * rd;
  h.dpdu = dpdu;
  h.dpdv = dpdv;
  h.t = t;

  vec3 nn = bary(attr3(m.normals, i0), attr3(m.normals, i1), attr3(m.normals, i2), w, bu, bv);
  vec3 tt = bary(attr3(m.tangents, i0), attr3(m.tangents, i1), attr3(m.tangents, i2), w, bu, bv);
  vec3 bb = bary(attr3(m.bitangents, i0), attr3(m.bitangents, i1), attr3(m.bitangents, i2), w, bu, bv);
  const mat4 &nm4 = s.normal_mat[mi];
  mat3 nm = upper3(nm4);
  vec3 tn = normalize(nm * nn);
  vec3 tt2 = normalize(l * tt);
  vec3 tb = normalize(l * bb);
  h.frame = gram_shmidt3(tn, tt2, tb);
  return h;
}

// --------------------------------------------------------------- BxDFs

struct BsdfParams {  // bsdf_params_s (BxDF.h:22-39)
  vec3 ns, ng, dpdu, dpdv;
  float wo_dot_ng = 0, wo_dot_ns = 0, roughness = 0, metal = 0, transmission = 0, anisotropy = 0;
  float subsurface = 1.f, sheen = 0.f, clearcoat = 0.f, specular = 1.f;
  bool thin = false;
  Spec eta, k, albedo;
};

struct Sample {  // BSDFSample (BxDF_flags.h:26-36)
  Spec f{0.f};
  vec3 wi;
  float pdf = 0, costheta = 0;
  // for the adjoint: the local-frame directions and microfacet normal of the sample
  vec3 lwo, lwi, wm;
  bool microfacet = false;
};

enum : unsigned { REFLECTION = 1, TRANSMISSION = 2, DIFFUSE = 4, GLOSSY = 8, SPECULAR = 16 };

static float schlick(float c, float f0) { return f0 + (1.f - f0) * pow5(1.f - c); }
static Spec schlick(float c, const Spec &f0) {
  float p = pow5(1.f - c);
  return {f0.x + (1.f - f0.x) * p, f0.y + (1.f - f0.y) * p, f0.z + (1.f - f0.z) * p};
}

// bxdf:: helpers (BxDF_math.h:57-124)
static inline float costheta(const vec3 &v) { return v.z; }
static inline float abscos(const vec3 &v) { return std::fabs(v.z); }
static inline float absdot(const vec3 &a, const vec3 &b) { return std::fabs(dot(a, b)); }
static inline bool same_hemi(const vec3 &a, const vec3 &b) { return a.z * b.z > 0; }
static inline float cos2theta(const vec3 &v) { return sqr(v.z); }
static inline float sin2theta(const vec3 &v) { return smax(0.f, 1.f - cos2theta(v)); }
static inline float sin2theta_c(float c) { return smax(0.f, 1.f - sqr(c)); }
static inline float sintheta(const vec3 &v) { return msqrt(sin2theta(v)); }
static inline float tan2theta(const vec3 &v) { return sin2theta(v) / cos2theta(v); }
static inline float cosphi(const vec3 &v) {
  if (sintheta(v) == 0.f) return 1.f;
  return clampf(v.x / sintheta(v), -1.f, 1.f);
}
static inline float sinphi(const vec3 &v) {
  if (sintheta(v) == 0.f) return 0.f;
  return clampf(v.y / sintheta(v), -1.f, 1.f);
}
static inline float cos2phi(const vec3 &v) { return sqr(cosphi(v)); }
static inline float sin2phi(const vec3 &v) { return sqr(sinphi(v)); }

static float refract(const vec3 &wo, vec3 n, float eta, bool &valid, vec3 &wt) {
  valid = true;
  float ci = dot(wo, n);
  if (ci < 0) {
    n = -n;
    eta = 1.f / eta;
    ci = -ci;
  }
  float s2i = sin2theta_c(ci);
  float s2t = s2i / sqr(eta);
  if (s2t >= 1.f) {
    valid = false;
    return 0.f;
  }
  float ct = msqrt(1.f - s2t);
  wt = -wo / eta + (ci / eta - ct) * n;
  return eta;
}

static vec2 shirley_chiu(const float u[2]) {
  float sx = 2 * u[0] - 1, sy = 2 * u[1] - 1;
  if (sx == 0 && sy == 0) return vec2(0.f, 0.f);
  float r, theta;
  if (std::fabs(sx) > std::fabs(sy)) {
    r = sx;
    theta = (float)M_PI_4 * (sy / sx);
  } else {
    r = sy;
    theta = (float)M_PI_2 - (float)M_PI_4 * (sx / sy);
  }
  return vec2(r * cos_f(theta), r * sin_f(theta));
}
static vec3 cosine_hemisphere(const float u[2]) {
  vec2 s = shirley_chiu(u);
  float z = std::sqrt(smax(0.f, 1 - s.x * s.x - s.y * s.y));
  return {s.x, s.y, z};
}

// Fresnel::real (BxDF_math.h:193-208)
static float fresnel_real(float etar, float ci) {
  float eta = etar;
  ci = gmin(gmax(ci, -1.f), 1.f);
  if (ci < 0) {
    eta = 1.f / eta;
    ci = -ci;
  }
  float s2i = sin2theta_c(ci);
  float s2t = s2i / sqr(eta);
  if (s2t >= 1) return 1.f;
  float ct = msqrt(1.f - s2t);
  float rpar = (eta * ci - ct) / (eta * ci + ct);
  float rperp = (ci - eta * ct) / (ci + eta * ct);
  return (sqr(rpar) + sqr(rperp)) / 2.f;
}

struct VNDF {  // BxDF_math.h:231-342
  float ax = 0, ax_inv = 0, ay = 0, ay_inv = 0;
  static float r2a(float r) { return r != 0 ? r * r : 1e-6f; }
  VNDF(float aniso, float rough) {
    aniso = clampf(aniso, -1.f, 1.f);
    float aspect = msqrt(1.f - 0.9f * aniso);
    ax = r2a(rough) / aspect;
    ay = r2a(rough) * aspect;
    ax_inv = 1.f / ax;
    ay_inv = 1.f / ay;
  }
  float D(const vec3 &wm) const {
    if (costheta(wm) < 1e-6f) return 0.f;
    float c2 = cos2phi(wm) * sqr(ax_inv);
    float s2 = sin2phi(wm) * sqr(ay_inv);
    float denom = (float)(M_PI * ax * ay * sqr(cos2theta(wm)) * sqr(1.f + tan2theta(wm) * (c2 + s2)));
    return 1.f / denom;
  }
  float lambda(const vec3 &v) const {
    float t2 = tan2theta(v);
    if (std::isinf(t2) || std::isnan(t2)) return 0;
    float a2 = sqr(ax) * cos2phi(v) + sqr(ay) * sin2phi(v);
    return (msqrt(1 + a2 * t2) - 1) * 0.5f;
  }
  float G1(vec3 v) const {
    v.z = smax(v.z, 1e-6f);
    return 1.f / (1.f + lambda(v));
  }
  float G(const vec3 &wo, const vec3 &wi) const { return 1.f / (1.f + lambda(wo) + lambda(wi)); }
  float pdf(const vec3 &wo, const vec3 &wm) const { return D(wm) * absdot(wo, wm) * G1(wo) / abscos(wo); }
  vec3 sample(const vec3 &wo, const float uc[2]) const {
    vec3 woh = normalize(vec3(wo.x * ax, wo.y * ay, wo.z));
    if (woh.z < 0) woh = -woh;
    vec3 t, b;
    if (woh.z < 1.f)
      t = normalize(cross(vec3(0, 0, 1), woh));
    else
      t = vec3(1, 0, 0);
    b = cross(woh, t);
    float r = msqrt(uc[0]);
    float ang = (float)(2.f * M_PI * uc[1]);  // float t = 2.f * M_PI * u[1] (BxDF_math.h:145-149)
    vec2 p(r * cos_f(ang), r * sin_f(ang));
    float s = 0.5f * (1.f + woh.z);
    float a = s;
    p.y = msqrt(1.f - sqr(p.x)) * (1.f - a) + a * p.y;
    vec3 nh = p.x * t + p.y * b + msqrt(std::fmax(0.f, 1.f - sqr(p.x) - sqr(p.y))) * woh;
    return normalize(vec3(ax * nh.x, ay * nh.y, smax(1e-6f, nh.z)));
  }
  bool full_specular() const { return smax(ax, ay) < 1e-5f; }
};

enum Lobe { LOBE_NONE, LOBE_DIFFUSE, LOBE_SPECULAR, LOBE_DIELECTRIC };

// PrincipledBSDF (BSDF.h:22-130) with its selected lobe.
struct Principled {
  BsdfParams params;
  mat3 frame;
  float p_diffuse = 0, p_specular = 0, p_transmission = 0;
  float w_diffuse = 0, w_specular = 0, w_transmission = 0;
  Lobe lobe = LOBE_NONE;
  float lobe_weight = 1, lobe_prob = 1;
  float vndf_aniso = 0, vndf_rough = 0;
  float vndf_rough_eval = 0;  // roughness of the BSDF value (vndf_rough samples; they differ only in gradient checks)

  void configure(const BsdfParams &p) {
    params = p;
    frame = gram_shmidt3(params.ns, params.dpdu, params.dpdv);
    if (params.wo_dot_ng < 0) {
      frame.c[0] = -frame.c[0];
      frame.c[1] = -frame.c[1];
      frame.c[2] = -frame.c[2];
      if (!params.thin && spec_all_ne(params.eta, 0.f))
        params.eta = vec3(1.f / params.eta.x, 1.f / params.eta.y, 1.f / params.eta.z);
    }
    params.ng = normalize(mul_transpose(frame, params.ng));
    // generateWeightsFromParam (BSDF.h:43-66)
    float ac = std::fabs(params.wo_dot_ng);
    w_diffuse = (1 - params.transmission) * (1 - params.metal);
    float w_sheen = (1 - params.metal) * params.sheen * (1.f - params.transmission);
    float w_cc = 0.25f * params.clearcoat;
    w_specular = params.specular;
    w_transmission = (1.f - params.metal) * params.transmission;
    float F = schlick(ac, params.specular * 0.04f);
    float pd = w_diffuse * (1.f - F);
    float ps = w_specular * F;
    float W = pd + ps + w_sheen + w_cc + w_transmission;
    if (W == 0) W = 1.f;
    p_diffuse = pd / W;
    p_specular = ps / W;
    p_transmission = w_transmission / W;
  }

  void construct(float uc) {
    float idx = 0.f;
    if (uc < (idx += p_diffuse)) {
      lobe = LOBE_DIFFUSE;
      lobe_weight = w_diffuse;
      lobe_prob = p_diffuse;
    } else if (uc < (idx += p_specular)) {
      lobe = LOBE_SPECULAR;
      lobe_weight = w_specular;
      lobe_prob = p_specular;
    } else if (uc < (idx += p_transmission)) {
      lobe = LOBE_DIELECTRIC;
      lobe_weight = w_transmission;
      lobe_prob = p_transmission;
    } else {
      lobe = LOBE_NONE;
    }
    vndf_aniso = params.anisotropy;
    vndf_rough = params.roughness;
    vndf_rough_eval = params.roughness;
  }

  unsigned flags() const {
    VNDF g(vndf_aniso, vndf_rough);
    switch (lobe) {
      case LOBE_DIFFUSE: return DIFFUSE | REFLECTION;
      case LOBE_SPECULAR: return g.full_specular() ? (SPECULAR | REFLECTION) : (GLOSSY | REFLECTION);
      case LOBE_DIELECTRIC: {
        unsigned f = params.eta.x == 1.f ? TRANSMISSION : (REFLECTION | TRANSMISSION);
        return f | (g.full_specular() ? SPECULAR : GLOSSY);
      }
      default: return 0;
    }
  }

  bool sample_diffuse(const vec3 &wo, const float u[2], Sample &s) const {
    auto fres = [](const vec3 &w, float value) { return 1.f + (value - 1) * pow5(1.f - abscos(w)); };
    vec3 wi = normalize(cosine_hemisphere(u));
    vec3 h = normalize(wi + wo);
    float ci = costheta(wi), co = costheta(wo);
    if (!same_hemi(wo, wi) || ci <= 0.f || co <= 0.f) return false;
    float rough = sqr(params.roughness);
    float fd90 = 0.5f + 2 * rough * sqr(absdot(wi, h));
    Spec fd = params.albedo * (float)M_1_PI * fres(wi, fd90) * fres(wo, fd90);
    float fss90 = rough * sqr(absdot(h, wi));
    float fi = fres(wi, fss90), fo = fres(wo, fss90);
    float tail = 0.5f + fi * fo * (-0.5f + 1 / (co + ci));
    Spec fss = 1.25f * params.albedo * (float)M_1_PI * tail;
    Spec f = lobe_weight * ((1 - params.subsurface) * fd + params.subsurface * fss);
    s.costheta = ci;
    s.f = f;
    s.pdf = lobe_prob * (ci * (float)M_1_PI);
    s.wi = wi;
    return true;
  }

  bool sample_specular(const vec3 &wo, const float u[2], Sample &s) const {
    VNDF g(vndf_aniso, vndf_rough), ge(vndf_aniso, vndf_rough_eval);
    vec3 wm = g.sample(wo, u);
    vec3 wi = reflect(-wo, wm);
    float co = costheta(wo), ci = costheta(wi);
    if (co <= 0 || ci <= 0) return false;
    float wm_dot_wo = absdot(wm, wo);
    Spec F0(0.04f);
    Spec mm(params.metal);
    F0 = F0 * (Spec(1.f) - mm) + mm * params.albedo;
    F0 = schlick(wm_dot_wo, F0);
    Spec f = F0 * ge.D(wm) * ge.G(wo, wi) / (4 * co * ci);
    float pdf = g.pdf(wo, wm) / (4 * wm_dot_wo);
    s.f = lobe_weight * f;
    s.pdf = lobe_prob * pdf;
    s.wi = wi;
    s.costheta = ci;
    s.wm = wm;
    s.microfacet = true;
    return true;
  }

  bool sample_dielectric(const vec3 &wo, float uc, const float u[2], Sample &s) const {
    VNDF g(vndf_aniso, vndf_rough), ge(vndf_aniso, vndf_rough_eval);
    const float ior = params.eta.x;
    if (g.full_specular() || params.eta.x == 1.f) {
      float R = fresnel_real(ior, costheta(wo)), T = 1 - R;
      float pr = R, pt = T;
      if (pr == 0.f && pt == 0.f) return false;
      if (uc < (pr / (pr + pt))) {  // specularReflection
        vec3 wi = reflect(-wo, vec3(0, 0, 1.f));
        float ci = abscos(wi);
        s.costheta = ci;
        s.f = lobe_weight * params.albedo * R / ci;
        s.pdf = lobe_prob * pr / (pr + pt);
        s.wi = wi;
        return true;
      }
      bool ok = true;  // specularTransmission
      vec3 wi;
      refract(wo, vec3(0.f, 0.f, 1.f), ior, ok, wi);
      if (!ok) return false;
      float ci = abscos(wi);
      Spec f = T * params.albedo / ci;
      f = f / sqr(ior);
      s.f = lobe_weight * f;
      s.costheta = ci;
      s.pdf = lobe_prob * pt / (pr + pt);
      s.wi = wi;
      return true;
    }
    vec3 wm = g.sample(wo, u);
    float wo_dot_wm = dot(wo, wm);
    float R = fresnel_real(ior, wo_dot_wm), T = 1 - R;
    float pr = R, pt = T;
    if (pr == 0.f && pt == 0.f) return false;
    if (uc < (pr / (pr + pt))) {  // glossyReflection
      vec3 wi = reflect(-wo, wm);
      if (!same_hemi(wo, wi)) return false;
      Spec f = params.albedo * ge.D(wm) * ge.G(wo, wi) * R / (4 * costheta(wi) * costheta(wo));
      float pdf = g.pdf(wo, wm) / (4 * absdot(wo, wm)) * pr / (pr + pt);
      s.f = lobe_weight * f;
      s.pdf = lobe_prob * pdf;
      s.wi = wi;
      s.costheta = abscos(wi);
      s.wm = wm;
      s.microfacet = true;
      return true;
    }
    // glossyTransmission
    vec3 wi(0.f);
    bool no_tir = true;
    float eta = refract(wo, wm, ior, no_tir, wi);
    if (!no_tir) return false;
    wi = normalize(wi);
    if (same_hemi(wi, wo) || wi.z == 0.f) return false;
    float ci = costheta(wi), co = costheta(wo);
    float denom = sqr(dot(wi, wm) + dot(wo, wm) / eta);
    float jac = absdot(wi, wm) / denom;
    float pdf = g.pdf(wo, wm) * jac * pt / (pt + pr);
    Spec f = params.albedo * T * ge.D(wm) * ge.G(wo, wi) * std::fabs(dot(wi, wm) * dot(wo, wm) / (ci * co * denom));
    f = f / sqr(eta);
    s.f = lobe_weight * f;
    s.costheta = abscos(wi);
    s.pdf = lobe_prob * pdf;
    s.wi = wi;
    s.wm = wm;
    s.microfacet = true;
    return true;
  }

  // Evaluation of the selected lobe for a given pair of local directions, the lobe's f and
  // pdf weighted like weightLobe / weightPdf. Not on the integrator's path (Li only samples);
  // written apart from the samplers above so the unit tests can check sample_f against it.
  // DisneyDiffuseBxDF::f / pdf (BxDF.h:257-286)
  void eval_diffuse(const vec3 &wo, const vec3 &wi, Spec &f, float &pdf) const {
    auto fres = [](const vec3 &w, float value) { return 1.f + (value - 1) * pow5(1.f - abscos(w)); };
    f = Spec(0.f);
    pdf = same_hemi(wo, wi) ? lobe_prob * (abscos(wi) * (float)M_1_PI) : 0.f;
    vec3 h = normalize(wi + wo);
    if (!(same_hemi(wo, wi) && same_hemi(wo, h))) return;
    float rough = sqr(params.roughness);
    float fd90 = 0.5f + 2 * rough * sqr(absdot(wi, h));
    Spec fd = params.albedo * (float)M_1_PI * fres(wi, fd90) * fres(wo, fd90);
    float fss90 = rough * sqr(absdot(h, wi));
    Spec fss = 1.25f * params.albedo * (float)M_1_PI * (0.5f + fres(wi, fss90) * fres(wo, fss90) * (-0.5f + 1 / (abscos(wo) + abscos(wi))));
    f = lobe_weight * ((1 - params.subsurface) * fd + params.subsurface * fss);
  }
  // SpecularBxDF::f / pdf (BxDF.h:343-364)
  void eval_specular(const vec3 &wo, const vec3 &wi, Spec &f, float &pdf) const {
    f = Spec(0.f);
    pdf = 0.f;
    if (!same_hemi(wi, wo)) return;
    VNDF g(vndf_aniso, vndf_rough);
    vec3 wm = normalize(wi + wo);
    float wm_dot_wo = absdot(wm, wo);
    Spec mm(params.metal);
    Spec F0 = Spec(0.04f) * (Spec(1.f) - mm) + mm * params.albedo;
    F0 = schlick(wm_dot_wo, F0);
    f = lobe_weight * (F0 * g.D(wm) * g.G(wo, wi) / (4 * abscos(wo) * abscos(wi)));
    pdf = lobe_prob * (g.pdf(wo, wm) / (4 * wm_dot_wo));
  }
  // DielectricBxDF::f / pdf (BxDF.h:497-551), radiance transport
  void eval_dielectric(const vec3 &wo, const vec3 &wi, Spec &f, float &pdf) const {
    VNDF g(vndf_aniso, vndf_rough);
    f = Spec(0.f);
    pdf = 0.f;
    const float eta = params.eta.x;
    if (g.full_specular() || eta == 1.f) return;
    if (same_hemi(wi, wo)) {
      vec3 wm = normalize(wi + wo);
      float wo_dot_wm = absdot(wo, wm);
      float R = fresnel_real(eta, wo_dot_wm), T = 1.f - R;
      f = lobe_weight * (params.albedo * g.D(wm) * g.G(wo, wi) * R / (4 * costheta(wi) * costheta(wo)));
      pdf = lobe_prob * (g.pdf(wo, wm) / (4 * wo_dot_wm) * R / (R + T));
      return;
    }
    vec3 wm = normalize(wi * eta + wo);
    if (dot(wo, wm) < 0) wm = -wm;
    float wo_dot_wm = absdot(wo, wm);
    float R = fresnel_real(eta, wo_dot_wm), T = 1.f - R;
    float denom = sqr(dot(wi, wm) + dot(wo, wm) / eta);
    Spec v = params.albedo * T * g.D(wm) * g.G(wo, wi) * std::fabs(dot(wi, wm) * dot(wo, wm) / (costheta(wi) * costheta(wo) * denom));
    f = lobe_weight * (v / sqr(eta));
    pdf = lobe_prob * (g.pdf(wo, wm) * (absdot(wi, wm) / denom) * T / (T + R));
  }
  void eval(const vec3 &lwo, const vec3 &lwi, Spec &f, float &pdf) const {
    f = Spec(0.f);
    pdf = 0.f;
    switch (lobe) {
      case LOBE_DIFFUSE: eval_diffuse(lwo, lwi, f, pdf); break;
      case LOBE_SPECULAR: eval_specular(lwo, lwi, f, pdf); break;
      case LOBE_DIELECTRIC: eval_dielectric(lwo, lwi, f, pdf); break;
      default: break;
    }
  }

  // PrincipledBSDF::sample_f (BSDF.h:107-130)
  bool sample_f(const vec3 &wo, float uc, const float u[2], Sample &s) const {
    if (lobe == LOBE_NONE) return false;
    vec3 lwo = normalize(mul_transpose(frame, wo));
    if (!(flags() & (REFLECTION | TRANSMISSION)) || lwo.z == 0.f) return false;
    bool ok = false;
    switch (lobe) {
      case LOBE_DIFFUSE: ok = sample_diffuse(lwo, u, s); break;
      case LOBE_SPECULAR: ok = sample_specular(lwo, u, s); break;
      case LOBE_DIELECTRIC: ok = sample_dielectric(lwo, uc, u, s); break;
      default: break;
    }
    if (!ok) return false;
    if (!spec_any(s.f) || s.pdf == 0.f || s.wi.z == 0.f) return false;
    s.lwo = lwo;
    s.lwi = s.wi;
    s.wi = normalize(frame * s.wi);
    return true;
  }
};

// bsdf::correct_specular_shading_normal (BxDF_math.h:26-53)
static vec3 correct_specular_shading_normal(const vec3 &ng, const vec3 &wo, vec3 ns) {
  vec3 R = reflect(-wo, ns);
  const float woz = dot(wo, ng);
  const float threshold = std::fmin(0.9f * woz, 0.01f);
  if (dot(ng, R) >= threshold) return ns;
  vec3 X = normalize(ns - dot(ns, ng) * ng);
  if (length2(X) == 0.f) X = ns;
  const float wox = dot(wo, X);
  const float a = sqr(wox) + sqr(woz);
  const float b = 2.0f * (a + woz * threshold);
  const float c = sqr(threshold + woz);
  const float Nz2 = (wox < 0) ? 0.25f * (b + msqrt(sqr(b) - 4.0f * a * c)) / a : 0.25f * (b - msqrt(sqr(b) - 4.0f * a * c)) / a;
  if (Nz2 > 1.f || Nz2 < 0.f) return ng;
  float Nx = msqrt(1.0f - Nz2), Nz = msqrt(Nz2);
  return Nx * X + Nz * ng;
}

struct ScatterResult {
  Spec attenuation, emissive;
  vec3 normal;  // mat_rec.normal: the shading normal from the normal map
  Sample lobe;
  Principled bsdf;  // the configured BSDF (adjoint)
  float u = 0, v = 0;
};

// PrincipledMaterial::scatter + sampleBSDF (NovaMaterials.h:45-137), with the
// FP32 texture and emission-scale extensions.
static bool scatter(const Scene &s, const Material &mat, const vec3 &in_dir, const Hit &h, SobolSampler &sampler, ScatterResult &out,
                    int sampling_roughness = -1) {
  const float u = h.u, v = h.v;
  const float ONE_255 = 1.f / 255.f;
  vec3 nrm(0.f, 0.f, 1.f);
  if (mat.normal >= 0) {
    vec4 t = sample_u32(s.textures[mat.normal], u, v);
    nrm = vec3(t.x * ONE_255, t.y * ONE_255, t.z * ONE_255);
  }
  vec3 local_n = nrm * 2.f - 1.f;
  vec3 ns = normalize(h.frame * local_n);
  out.normal = ns;
  BsdfParams bp;
  bp.dpdu = h.dpdu;
  bp.dpdv = h.dpdv;
  bp.ns = ns;
  bp.ng = h.ng;
  vec3 wo = normalize(-in_dir);
  bp.wo_dot_ng = dot(wo, h.ng);
  bp.wo_dot_ns = dot(wo, ns);
  if (mat.albedo_f32 >= 0) {
    vec4 a = sample_f32_bilinear(s.textures[mat.albedo_f32], u, v);
    bp.albedo = vec3(a.x, a.y, a.z);
  } else if (mat.albedo >= 0) {
    vec4 a = sample_u32(s.textures[mat.albedo], u, v);
    bp.albedo = vec3(a.x * ONE_255, a.y * ONE_255, a.z * ONE_255);
  } else {
    bp.albedo = vec3(1.f);
  }
  vec3 met(0.f);
  if (mat.metallic >= 0) {
    vec4 mt = sample_u32(s.textures[mat.metallic], u, v);
    met = vec3(mt.x * ONE_255, mt.y * ONE_255, mt.z * ONE_255);
  }
  bp.roughness = met.y;
  if (mat.roughness_f32 >= 0) bp.roughness = sample_f32_bilinear(s.textures[mat.roughness_f32], u, v).x;
  bp.metal = met.z;
  Spec eta(mat.eta[0], mat.eta[1], mat.eta[2]), k(mat.k[0], mat.k[1], mat.k[2]);
  bp.transmission = (spec_any(eta) && !spec_any(k)) ? 1.f : 0.f;
  bp.anisotropy = mat.anisotropy;
  bp.eta = eta;
  bp.k = k;

  float uc = sampler.sample1D();
  float u0[2];
  u0[0] = sampler.sample1D();
  u0[1] = sampler.sample1D();
  Principled bsdf;
  bsdf.configure(bp);
  bsdf.construct(uc);
  // gradient-check harness: another FP32 texture drives the sampling roughness
  if (sampling_roughness >= 0 && mat.roughness_f32 >= 0) bsdf.vndf_rough = sample_f32_bilinear(s.textures[sampling_roughness], u, v).x;
  if (bsdf.flags() & (GLOSSY | REFLECTION)) {
    if (bp.wo_dot_ng >= 0) {
      vec3 nsc = normalize(correct_specular_shading_normal(bp.ng, wo, bp.ns));
      bp.ns = nsc;
      bp.wo_dot_ns = dot(wo, nsc);
      // configureBSDFParams re-runs with the corrected normal; the lobe and its weights stay.
      Lobe keep = bsdf.lobe;
      float lw = bsdf.lobe_weight, lp = bsdf.lobe_prob, va = bsdf.vndf_aniso, vr = bsdf.vndf_rough, ve = bsdf.vndf_rough_eval;
      bsdf.configure(bp);
      bsdf.lobe = keep;
      bsdf.lobe_weight = lw;
      bsdf.lobe_prob = lp;
      bsdf.vndf_aniso = va;
      bsdf.vndf_rough = vr;
      bsdf.vndf_rough_eval = ve;
    }
  }
  uc = sampler.sample1D();
  Sample smp;
  bool valid = bsdf.sample_f(wo, uc, u0, smp);
  out.attenuation = smp.f;
  vec3 em(0.f);
  if (mat.emissive >= 0) {
    vec4 e = sample_u32(s.textures[mat.emissive], u, v);
    em = vec3(e.x * ONE_255, e.y * ONE_255, e.z * ONE_255);
  }
  out.emissive = em * mat.emission_scale;
  out.lobe = smp;
  out.bsdf = bsdf;
  out.u = u;
  out.v = v;
  return valid;
}

// -------------------------------------------------------------- integrator

static vec4 denan(const vec4 &v) { return vec_nonfinite(v) ? vec4(0.f) : v; }

// PathIntegrator::Li (Integrator.cpp:59-98), recursive; with the Russian-roulette
// extension (rr_start >= 0) rolling one extra sampler dimension after each scatter
// whose index is >= rr_start, with survival q = min(1, max component of throughput).
// Ray recorder for the closest-hit tests: when set, the ray entering scatter index
// g_ray_bounce is written to g_ray_out (o, d: six floats).
static thread_local float *g_ray_out = nullptr;
static thread_local int g_ray_bounce = -1;

static vec4 Li_rec(const Scene &s, const Bvh &bvh, vec3 o, vec3 d, int depth, SobolSampler &sampler, const RenderParams &p, int scatter_index,
                   Spec throughput) {
  if (g_ray_out && scatter_index == g_ray_bounce) {
    const float r[6] = {o.x, o.y, o.z, d.x, d.y, d.z};
    for (int k = 0; k < 6; k++) g_ray_out[k] = r[k];
  }
  if (depth >= 0) {
    float t = 0, bu = 0, bv = 0;
    int prim = closest_hit(s, bvh, o, d, 0.000001f, t, bu, bv);
    if (prim >= 0) {
      Hit h = make_hit(s, prim, o, d, t, bu, bv);
      const Material &mat = s.materials[s.meshes[s.prim_mesh[prim]].material];
      ScatterResult sr;
      if (!scatter(s, mat, d, h, sampler, sr, p.sampling_roughness)) return vec4(0.f);
      // Ray::spawn (Ray.h:22-27)
      vec3 wi = sr.lobe.wi;
      vec3 nd = normalize(wi);
      const float eps = 1e-3f;
      vec3 no = h.position + h.ng * (eps + eps * gsign(dot(h.ng, wi)));
      Spec color = sr.attenuation * sr.lobe.costheta / sr.lobe.pdf;
      vec4 next;
      bool terminated = false;
      if (p.rr_start >= 0 && scatter_index >= p.rr_start) {
        Spec tp = throughput * color;
        float q = std::fmin(1.f, std::fmax(tp.x, std::fmax(tp.y, tp.z)));
        float ur = sampler.sample1D();
        if (!(ur < q)) {
          terminated = true;
        } else {
          color = color / q;
          throughput = throughput * color;
        }
      } else {
        throughput = throughput * color;
      }
      if (terminated)
        next = vec4(0.f);
      else
        next = Li_rec(s, bvh, no, nd, depth - 1, sampler, p, scatter_index + 1, throughput);
      vec4 e(sr.emissive, 1.f);
      vec4 c(color, 1.f);
      vec4 fin = e + c * next;
      return denan(fin);
    }
  }
  return sample_envmap(s, d);
}

vec4 path_Li(const Scene &s, const Bvh &bvh, vec3 o, vec3 d, int depth, SobolSampler &sampler, const RenderParams &p, int scatter_index) {
  return Li_rec(s, bvh, o, d, depth, sampler, p, scatter_index, Spec(1.f));
}

// math::camera::ray_inv_mat (math_camera.h:46-58)
static void camera_ray(const RenderParams &p, float nx, float ny, vec3 &o, vec3 &d) {
  vec4 oo = p.inv_P * vec4(nx, ny, -1.f, 1.f);
  oo = oo / oo.w;
  oo = p.inv_V * vec4(oo.x, oo.y, oo.z, 1.f);
  vec4 dd = p.inv_P * vec4(nx, ny, -1.f, 0.f);
  vec4 dw = p.inv_V * vec4(dd.x, dd.y, -1.f, 0.f);
  float xx = dw.x * dw.x, yy = dw.y * dw.y, zz = dw.z * dw.z, ww = dw.w * dw.w;
  float len2 = (xx + yy) + (zz + ww);
  dw = dw * (1.f / std::sqrt(len2));
  o = oo.xyz();
  d = dw.xyz();
}

struct Tile {
  int x0, y0, x1, y1;
};

// divideByTiles (nova_utils.h:33-52)
static std::vector<Tile> divide_by_tiles(int W, int H, int nw, int nh) {
  std::vector<Tile> tiles;
  int w = W / nw, h = H / nh;
  int wr = w + (W % w), hr = h + (H % h);
  for (int j = 0; j < nh; ++j)
    for (int i = 0; i < nw; ++i) {
      Tile t;
      t.x0 = i * w;
      t.x1 = t.x0 + ((i == nw - 1) ? wr : w);
      t.y0 = j * h;
      t.y1 = t.y0 + ((j == nh - 1) ? hr : h);
      tiles.push_back(t);
    }
  return tiles;
}

int g_index_policy = 0;

// Integrator.h:38-42,80 (CPU: idx = 4 * pixel, row flipped with vertical_invert) or the
// GPU raygen's (x * H + y) * sample_max + frame (integrator_optix.cu:43); u32 wraparound.
uint32_t sample_index(const RenderParams &p, int x, int y, int frame) {
  if (p.index_policy == 1) return ((unsigned)x * (unsigned)p.height + (unsigned)y) * p.max_samples_offset + (unsigned)frame;
  const unsigned idx = p.vertical_invert ? (unsigned)(((p.height - 1 - y) * p.width + x) * 4) : (unsigned)((y * p.width + x) * 4);
  return idx * (unsigned)frame + p.max_samples_offset;
}

// AbstractIntegrator::render for one tile (Integrator.h:53-95), spp = 1 per frame.
static void render_tile(const Scene &s, const Bvh &bvh, const RenderParams &p, int frame, const Tile &t, float *partial) {
  const float RAND_DX = 0.0005f, RAND_DY = 0.0005f;
  SobolSampler sampler;
  for (int y = t.y1 - 1; y >= t.y0; y--)
    for (int x = t.x0; x < t.x1; x++) {
      const unsigned idx = p.vertical_invert ? (unsigned)(((p.height - 1 - y) * p.width + x) * 4) : (unsigned)((y * p.width + x) * 4);
      sampler.reset(sample_index(p, x, y, frame));
      float s0 = sampler.sample1D(), s1 = sampler.sample1D();
      const float dx = s0 * RAND_DX, dy = s1 * RAND_DY;
      int sy = p.height - y;
      float ndc_x = (float)(2 * x - p.width) / (float)p.width;
      float ndc_y = (float)(p.height - 2 * sy) / (float)p.height;
      vec3 o, d;
      camera_ray(p, ndc_x + dx, ndc_y + dy, o, d);
      vec4 rgb = path_Li(s, bvh, o, d, p.max_depth, sampler, p, 0);
      partial[idx] = rgb.x;
      partial[idx + 1] = rgb.y;
      partial[idx + 2] = rgb.z;
      partial[idx + 3] = rgb.w;
    }
}

// The rays that enter scatter index `bounce` (0: camera rays) of every path of one frame,
// pixel (x, y) at rays[(y * width + x) * 6]; NaN where the path ended before that bounce.
// The paths are the renderer's own (same sample indices, jitter, scatter and roulette);
// they are cut after that bounce, which changes none of the rays before it.
void path_rays(const Scene &s, const Bvh &bvh, const RenderParams &p, int frame, int bounce, float *rays) {
  const float RAND_DX = 0.0005f, RAND_DY = 0.0005f;
  std::atomic<int> next{0};
  auto worker = [&]() {
    SobolSampler sampler;
    g_ray_bounce = bounce;
    for (;;) {
      const int y = next.fetch_add(1);
      if (y >= p.height) break;
      for (int x = 0; x < p.width; x++) {
        float *out = rays + ((size_t)y * p.width + x) * 6;
        for (int k = 0; k < 6; k++) out[k] = std::numeric_limits<float>::quiet_NaN();
        g_ray_out = out;
        sampler.reset(sample_index(p, x, y, frame));
        float s0 = sampler.sample1D(), s1 = sampler.sample1D();
        const float dx = s0 * RAND_DX, dy = s1 * RAND_DY;
        int sy = p.height - y;
        float ndc_x = (float)(2 * x - p.width) / (float)p.width;
        float ndc_y = (float)(p.height - 2 * sy) / (float)p.height;
        vec3 o, d;
        camera_ray(p, ndc_x + dx, ndc_y + dy, o, d);
        path_Li(s, bvh, o, d, std::min(p.max_depth, bounce), sampler, p, 0);
      }
    }
    g_ray_out = nullptr;
    g_ray_bounce = -1;
  };
  const int nthreads = std::max(1, p.threads);
  std::vector<std::thread> pool;
  for (int i = 0; i < nthreads; i++) pool.emplace_back(worker);
  for (auto &th : pool) th.join();
}

void render_frame(const Scene &s, const Bvh &bvh, const RenderParams &p, int frame, float *partial) {
  int nthreads = std::max(1, p.threads);
  // divideByTiles only covers the image when the tile count per axis is at most
  // sqrt(extent); the defaults stay inside that bound.
  int tw = p.tiles_w > 0 ? p.tiles_w : std::max(1, std::min((int)std::sqrt((double)p.width), nthreads * 4));
  int th = p.tiles_h > 0 ? p.tiles_h : std::max(1, std::min((int)std::sqrt((double)p.height), nthreads * 4));
  std::vector<Tile> tiles = divide_by_tiles(p.width, p.height, tw, th);
  std::atomic<size_t> next{0};
  auto worker = [&]() {
    for (;;) {
      size_t k = next.fetch_add(1);
      if (k >= tiles.size()) break;
      render_tile(s, bvh, p, frame, tiles[k], partial);
    }
  };
  if (nthreads == 1) {
    worker();
    return;
  }
  std::vector<std::thread> pool;
  for (int i = 0; i < nthreads; i++) pool.emplace_back(worker);
  for (auto &th2 : pool) th2.join();
}

// Rows [y0, y1) of one frame, one row per task (the CPU-baseline sample).
void render_rows(const Scene &s, const Bvh &bvh, const RenderParams &p, int frame, int y0, int y1, float *partial) {
  std::atomic<int> next{y0};
  auto worker = [&]() {
    for (;;) {
      int y = next.fetch_add(1);
      if (y >= y1) break;
      Tile t{0, y, p.width, y + 1};
      render_tile(s, bvh, p, frame, t, partial);
    }
  };
  int nthreads = std::max(1, p.threads);
  std::vector<std::thread> pool;
  for (int i = 0; i < nthreads; i++) pool.emplace_back(worker);
  for (auto &th : pool) th.join();
}

// Engine::renderCpu frame loop + writeFramebuffer (engine.cpp:351-363, 498-525)
void render(const Scene &s, const Bvh &bvh, const RenderParams &p, float *accum, float *final_out) {
  const size_t n = (size_t)p.width * p.height * 4;
  std::vector<float> partial(n, 0.f);
  for (int f = p.frame_start; f < p.frame_start + p.frames; f++) {
    render_frame(s, bvh, p, f, partial.data());
    for (size_t i = 0; i < n; i++) {
      accum[i] += partial[i];
      final_out[i] = accum[i] / float(f + 1);
    }
  }
}


// Render passes: NormalIntegrator / DepthIntegrator / EmissiveIntegrator (NormalIntegrator.cpp,
// DepthIntegrator.cpp, EmissiveIntegrator.cpp). One ray per pixel through screen2ndc without
// jitter, the sampler reset to the pixel offset (no frame factor); the frame is the same each
// time and is accumulated as the path integrator's frames are. pass: 1 normal, 2 depth,
// 3 emissive.
void render_pass(const Scene &s, const Bvh &bvh, const RenderParams &p, int pass, float z_near, float z_far, float *accum, float *final_out) {
  const size_t n = (size_t)p.width * p.height * 4;
  std::vector<float> partial(n, 0.f);
  std::atomic<int> next{0};
  auto worker = [&]() {
    SobolSampler sampler;
    for (;;) {
      const int y = next.fetch_add(1);
      if (y >= p.height) break;
      for (int x = 0; x < p.width; x++) {
        const unsigned idx = p.vertical_invert ? (unsigned)(((p.height - 1 - y) * p.width + x) * 4) : (unsigned)((y * p.width + x) * 4);
        sampler.reset(idx);
        const int sy = p.height - y;
        const float ndc_x = (float)(2 * x - p.width) / (float)p.width;
        const float ndc_y = (float)(p.height - 2 * sy) / (float)p.height;
        vec3 o, d;
        camera_ray(p, ndc_x, ndc_y, o, d);
        float t = 0, bu = 0, bv = 0;
        const int prim = closest_hit(s, bvh, o, d, 0.000001f, t, bu, bv);
        vec4 rgb(0.f);
        if (pass == 2) {  // DepthIntegrator::render; Li gives MAXFLOAT on a miss
          const float dist = prim >= 0 ? t : 3.402823466e+38f;
          float depth = 1 - (dist - z_near) / (z_far - z_near);
          depth = (2.0f * z_near * z_far) / (z_far + z_near - depth * (z_far - z_near)) * 2.f - 1.f;
          rgb = vec4(depth / z_far);
          if (dist == 3.402823466e+38f) rgb = vec4(z_far);
        } else if (prim >= 0) {
          Hit h = make_hit(s, prim, o, d, t, bu, bv);
          const Material &mat = s.materials[s.meshes[s.prim_mesh[prim]].material];
          ScatterResult sr;
          const bool valid = scatter(s, mat, d, h, sampler, sr);
          if (pass == 1)
            rgb = vec4(sr.normal.x, sr.normal.y, sr.normal.z, 1.f);
          else if (valid)
            rgb = vec4(sr.emissive.x, sr.emissive.y, sr.emissive.z, 1.f);
        }
        partial[idx] = rgb.x;
        partial[idx + 1] = rgb.y;
        partial[idx + 2] = rgb.z;
        partial[idx + 3] = rgb.w;
      }
    }
  };
  std::vector<std::thread> pool;
  for (int i = 0; i < std::max(1, p.threads); i++) pool.emplace_back(worker);
  for (auto &th : pool) th.join();
  for (int f = p.frame_start; f < p.frame_start + p.frames; f++)
    for (size_t i = 0; i < n; i++) {
      accum[i] += partial[i];
      final_out[i] = accum[i] / float(f + 1);
    }
}

// --------------------------------------------------------------- adjoint
// CPU gradient of an image loss with respect to FP32 albedo / roughness texels, written as
// plain reverse mode over the recursion V_k = Le_k + c_k V_{k+1}: the adjoint of V_k is
// passed down (dV_{k+1} = dV_k * c_k) and on the way back up the weight c_k receives
// dV_k * V_{k+1}. The derivative of c_k with the sample held fixed is taken by forward-mode
// dual numbers in double over a restatement of the lobe's BSDF value (the GPU uses
// hand-derived closed forms), so the two are independent.

namespace {
struct Dual {
  double v, d;
};
Dual operator+(Dual a, Dual b) { return {a.v + b.v, a.d + b.d}; }
Dual operator-(Dual a, Dual b) { return {a.v - b.v, a.d - b.d}; }
Dual operator*(Dual a, Dual b) { return {a.v * b.v, a.d * b.v + a.v * b.d}; }
Dual operator/(Dual a, Dual b) { return {a.v / b.v, (a.d * b.v - a.v * b.d) / (b.v * b.v)}; }
Dual cst(double x) { return {x, 0.0}; }
Dual dsqrt(Dual a) {
  const double r = std::sqrt(std::max(a.v, 0.0));
  return {r, r > 0 ? a.d / (2 * r) : 0.0};
}

// GGX with alpha_x = a / aspect, alpha_y = a * aspect (BxDF_math.h:231-342), in duals.
struct DualGGX {
  Dual ax, ay;
  DualGGX(float aniso, Dual r) {
    aniso = clampf(aniso, -1.f, 1.f);
    const double aspect = msqrt(1.f - 0.9f * aniso);
    const Dual a = r.v != 0 ? r * r : cst(1e-6);
    ax = a / cst(aspect);
    ay = a * cst(aspect);
  }
  Dual D(const vec3 &wm) const {
    const double c2t = (double)wm.z * wm.z, s2t = std::max(0.0, 1.0 - c2t), t2 = s2t / c2t;
    const double c2p = cos2phi(wm), s2p = sin2phi(wm);
    const Dual e = cst(t2) * (cst(c2p) / (ax * ax) + cst(s2p) / (ay * ay));
    const Dual one_e = cst(1.0) + e;
    return cst(1.0) / (cst(M_PI * c2t * c2t) * ax * ay * one_e * one_e);
  }
  Dual lambda(const vec3 &v) const {
    const double c2t = (double)v.z * v.z, t2 = std::max(0.0, 1.0 - c2t) / c2t;
    if (!std::isfinite(t2)) return cst(0.0);
    const Dual a2 = ax * ax * cst(cos2phi(v)) + ay * ay * cst(sin2phi(v));
    return (dsqrt(cst(1.0) + a2 * cst(t2)) - cst(1.0)) * cst(0.5);
  }
  Dual G(const vec3 &wo, const vec3 &wi) const { return cst(1.0) / (cst(1.0) + lambda(wo) + lambda(wi)); }
};

// The sampled lobe's BSDF value for one channel, as a function of that channel's albedo and
// of the roughness (both duals); direction-only factors are constants.
Dual lobe_value(const ScatterResult &sr, int ch, Dual alb, Dual r) {
  const Principled &b = sr.bsdf;
  const Sample &sm = sr.lobe;
  const vec3 wo = sm.lwo, wi = sm.lwi;
  const double co = wo.z, ci = wi.z;
  if (b.lobe == LOBE_DIFFUSE) {
    const vec3 h = normalize(wi + wo);
    const double hw = std::fabs(dot(h, wi));
    const Dual fss90 = r * r * cst(hw * hw);
    const Dual fi = cst(1.0) + (fss90 - cst(1.0)) * cst(std::pow(1.0 - std::fabs(ci), 5));
    const Dual fo = cst(1.0) + (fss90 - cst(1.0)) * cst(std::pow(1.0 - std::fabs(co), 5));
    const Dual tail = cst(0.5) + fi * fo * cst(-0.5 + 1.0 / (co + ci));
    return cst(b.lobe_weight * 1.25 * M_1_PI) * alb * tail;
  }
  if (!sm.microfacet) return cst(b.lobe_weight) * alb;  // delta lobes: albedo times direction terms
  const DualGGX g(b.vndf_aniso, r);
  const Dual dg = g.D(sm.wm) * g.G(wo, wi);
  if (b.lobe == LOBE_SPECULAR) {
    const double m = b.params.metal;
    const double p = std::pow(1.0 - std::fabs((double)dot(sm.wm, wo)), 5);
    const Dual F0 = cst(0.04 * (1 - m)) + cst(m) * alb;
    const Dual F = F0 + (cst(1.0) - F0) * cst(p);
    return cst(b.lobe_weight / (4 * co * ci)) * F * dg;
  }
  return cst(b.lobe_weight) * alb * dg;  // dielectric glossy reflection / transmission
}

void lobe_log_derivatives(const ScatterResult &sr, double dlna[3], double &dlnr) {
  const Spec &a = sr.bsdf.params.albedo;
  const double av[3] = {a.x, a.y, a.z};
  const double r = sr.bsdf.vndf_rough_eval;
  for (int ch = 0; ch < 3; ch++) {
    const Dual f = lobe_value(sr, ch, Dual{av[ch], 1.0}, cst(r));
    dlna[ch] = f.v != 0 ? f.d / f.v : 0.0;
  }
  const Dual f = lobe_value(sr, 0, cst(av[0]), Dual{r, 1.0});
  dlnr = f.v != 0 ? f.d / f.v : 0.0;
}

// Adds g[c] * (bilinear weight) to channel c of the four texels sample_f32_bilinear blends.
void splat(const Texture &tx, double *grad, float u, float v, const double *g, int nch) {
  u = wrap(u);
  v = wrap(v);
  u = tx.invert_u ? 1 - u : u;
  v = tx.invert_v ? 1 - v : v;
  const int width = tx.width, height = tx.height;
  float pi = std::fabs(u * (float)(width - 1.f));
  float pj = std::fabs(v * (float)(height - 1.f));
  unsigned i = (unsigned)std::floor(pi), j = (unsigned)std::floor(pj);
  if (i + 1 >= (unsigned)width) i = width - 2;
  if (j + 1 >= (unsigned)height) j = height - 2;
  const float tx_ = pi - (float)i, ty_ = pj - (float)j;
  const unsigned idx[4] = {j * width + i, j * width + i + 1, (j + 1) * width + i, (j + 1) * width + i + 1};
  const float hd = 1 - tx_, vy = 1 - ty_;
  const float w[4] = {vy * hd, vy * tx_, ty_ * hd, ty_ * tx_};
  for (int k = 0; k < 4; k++)
    for (int c = 0; c < nch && c < tx.channels; c++) grad[(size_t)idx[k] * tx.channels + c] += (double)w[k] * g[c];
}

struct AdjCtx {
  const Scene &s;
  const Bvh &bvh;
  const RenderParams &p;
  double *grad;  // f32 pool layout
};

vec4 Li_adj(const AdjCtx &A, vec3 o, vec3 d, int depth, SobolSampler &sampler, int scatter_index, Spec throughput, const double dV[3]) {
  const Scene &s = A.s;
  const RenderParams &p = A.p;
  if (depth >= 0) {
    float t = 0, bu = 0, bv = 0;
    int prim = closest_hit(s, A.bvh, o, d, 0.000001f, t, bu, bv);
    if (prim >= 0) {
      Hit h = make_hit(s, prim, o, d, t, bu, bv);
      const Material &mat = s.materials[s.meshes[s.prim_mesh[prim]].material];
      ScatterResult sr;
      if (!scatter(s, mat, d, h, sampler, sr, p.sampling_roughness)) return vec4(0.f);
      vec3 wi = sr.lobe.wi;
      vec3 nd = normalize(wi);
      const float eps = 1e-3f;
      vec3 no = h.position + h.ng * (eps + eps * gsign(dot(h.ng, wi)));
      Spec color = sr.attenuation * sr.lobe.costheta / sr.lobe.pdf;
      bool terminated = false;
      int rr_channel = -1;
      if (p.rr_start >= 0 && scatter_index >= p.rr_start) {
        Spec tp = throughput * color;
        float q = std::fmin(1.f, std::fmax(tp.x, std::fmax(tp.y, tp.z)));
        float ur = sampler.sample1D();
        if (!(ur < q)) {
          terminated = true;
        } else {
          color = color / q;
          throughput = throughput * color;
          if (q < 1.f) rr_channel = (tp.x == q) ? 0 : (tp.y == q) ? 1 : 2;
        }
      } else {
        throughput = throughput * color;
      }
      const bool finite = std::isfinite(color.x) && std::isfinite(color.y) && std::isfinite(color.z);
      const double cv[3] = {color.x, color.y, color.z};
      const double zero[3] = {0, 0, 0};
      double dnext[3] = {dV[0] * cv[0], dV[1] * cv[1], dV[2] * cv[2]};
      vec4 next = terminated ? vec4(0.f) : Li_adj(A, no, nd, depth - 1, sampler, scatter_index + 1, throughput, finite ? dnext : zero);
      vec4 e(sr.emissive, 1.f);
      vec4 c(color, 1.f);
      vec4 fin = e + c * next;
      if (vec_nonfinite(fin)) return vec4(0.f);
      if (!terminated) {
        // d loss / d c_k per channel, times c_k: the weight of d ln c_k / d theta
        const double nv[3] = {next.x, next.y, next.z};
        double wc[3], wsum = 0;
        for (int k = 0; k < 3; k++) {
          wc[k] = dV[k] * nv[k] * cv[k];
          wsum += wc[k];
        }
        double dlna[3], dlnr;
        lobe_log_derivatives(sr, dlna, dlnr);
        double ga[3] = {wc[0] * dlna[0], wc[1] * dlna[1], wc[2] * dlna[2]};
        double gr = wsum * dlnr;
        if (rr_channel >= 0) {  // color = c / q, q = T_m c_m
          ga[rr_channel] -= wsum * dlna[rr_channel];
          gr = 0;
        }
        if (mat.albedo_f32 >= 0) splat(s.textures[mat.albedo_f32], A.grad + f32_pool_offset(s, mat.albedo_f32), sr.u, sr.v, ga, 3);
        if (mat.roughness_f32 >= 0) splat(s.textures[mat.roughness_f32], A.grad + f32_pool_offset(s, mat.roughness_f32), sr.u, sr.v, &gr, 1);
      }
      return fin;
    }
  }
  return sample_envmap(s, d);
}
}  // namespace

size_t f32_pool_offset(const Scene &s, int texture) {
  size_t off = 0;
  for (int t = 0; t < texture; t++)
    if (!s.textures[t].f32.empty()) off += (size_t)s.textures[t].width * s.textures[t].height * s.textures[t].channels;
  return off;
}

void render_adjoint(const Scene &s, const Bvh &bvh, const RenderParams &p, const float *d_image, float scale, double *grad_pool) {
  size_t pool = 0;
  for (const Texture &t : s.textures)
    if (!t.f32.empty()) pool += (size_t)t.width * t.height * t.channels;
  const int nthreads = std::max(1, p.threads);
  std::vector<std::vector<double>> part(nthreads, std::vector<double>(pool, 0.0));
  std::atomic<int> next_row{0};
  const int rows = p.height * p.frames;
  auto worker = [&](int tid) {
    AdjCtx A{s, bvh, p, part[tid].data()};
    SobolSampler sampler;
    for (;;) {
      const int r = next_row.fetch_add(1);
      if (r >= rows) break;
      const int frame = p.frame_start + r / p.height, y = r % p.height;
      for (int x = 0; x < p.width; x++) {  // AbstractIntegrator::render for one pixel (Integrator.h:53-95)
        const unsigned idx = p.vertical_invert ? (unsigned)(((p.height - 1 - y) * p.width + x) * 4) : (unsigned)((y * p.width + x) * 4);
      sampler.reset(sample_index(p, x, y, frame));
        float s0 = sampler.sample1D(), s1 = sampler.sample1D();
        const float dx = s0 * 0.0005f, dy = s1 * 0.0005f;
        int sy = p.height - y;
        float ndc_x = (float)(2 * x - p.width) / (float)p.width;
        float ndc_y = (float)(p.height - 2 * sy) / (float)p.height;
        vec3 o, d;
        camera_ray(p, ndc_x + dx, ndc_y + dy, o, d);
        const double dV[3] = {(double)scale * d_image[idx], (double)scale * d_image[idx + 1], (double)scale * d_image[idx + 2]};
        Li_adj(A, o, d, p.max_depth, sampler, 0, Spec(1.f), dV);
      }
    }
  };
  std::vector<std::thread> pool_threads;
  for (int i = 0; i < nthreads; i++) pool_threads.emplace_back(worker, i);
  for (auto &th : pool_threads) th.join();
  for (int i = 0; i < nthreads; i++)
    for (size_t k = 0; k < pool; k++) grad_pool[k] += part[i][k];
}

}  // namespace oracle

// ------------------------------------------------------------ unit access
// Entry points for the BxDF property tests (the reference's BxDFTest invariants).
namespace oracle {
float unit_fresnel_real(float eta, float cos_i) { return fresnel_real(eta, cos_i); }
vec3 unit_correct_shading_normal(const vec3 &ng, const vec3 &wo, const vec3 &ns) { return correct_specular_shading_normal(ng, wo, ns); }
vec3 unit_vndf_sample(float aniso, float rough, const vec3 &wo, const float u[2]) { return VNDF(aniso, rough).sample(wo, u); }
float unit_vndf_G1(float aniso, float rough, const vec3 &v) { return VNDF(aniso, rough).G1(v); }
float unit_vndf_D(float aniso, float rough, const vec3 &wm) { return VNDF(aniso, rough).D(wm); }
// One Principled sample in a frame where ns = ng = +z, dpdu = +x, dpdv = +y (world = local).
// Returns the lobe (0 none, 1 diffuse, 2 specular, 3 dielectric) and fills f, pdf, wi, costheta
// when the sample is valid (valid = 0 otherwise).
static BsdfParams unit_params(const vec3 &wo, const vec3 &albedo, float roughness, float metal, float eta) {
  BsdfParams bp;
  bp.ns = vec3(0, 0, 1);
  bp.ng = vec3(0, 0, 1);
  bp.dpdu = vec3(1, 0, 0);
  bp.dpdv = vec3(0, 1, 0);
  bp.wo_dot_ng = dot(wo, bp.ng);
  bp.wo_dot_ns = bp.wo_dot_ng;
  bp.albedo = albedo;
  bp.roughness = roughness;
  bp.metal = metal;
  bp.eta = vec3(eta);
  bp.transmission = eta != 0.f ? 1.f : 0.f;
  return bp;
}
int unit_principled_sample(const vec3 &wo, const vec3 &albedo, float roughness, float metal, float eta, float uc, float u0, float u1, float uc2,
                           vec3 &f, float &pdf, vec3 &wi, float &costheta, int &valid) {
  Principled b;
  b.configure(unit_params(wo, albedo, roughness, metal, eta));
  b.construct(uc);
  const float u[2] = {u0, u1};
  Sample s;
  valid = b.sample_f(wo, uc2, u, s) ? 1 : 0;
  f = s.f;
  pdf = s.pdf;
  wi = s.wi;
  costheta = s.costheta;
  return (int)b.lobe;
}
// The hit record of primitive prim at (t, bu, bv) on ray (ro, rd): ng, dpdu, dpdv, position,
// u, v, then the shading frame's T, B, N columns (23 floats).
void unit_make_hit(const Scene &s, int prim, const vec3 &ro, const vec3 &rd, float t, float bu, float bv, float *out) {
  const Hit h = make_hit(s, prim, ro, rd, t, bu, bv);
  const vec3 vs[4] = {h.ng, h.dpdu, h.dpdv, h.position};
  for (int k = 0; k < 4; k++) {
    out[3 * k] = vs[k].x;
    out[3 * k + 1] = vs[k].y;
    out[3 * k + 2] = vs[k].z;
  }
  out[12] = h.u;
  out[13] = h.v;
  for (int c = 0; c < 3; c++) {
    out[14 + 3 * c] = h.frame.c[c].x;
    out[15 + 3 * c] = h.frame.c[c].y;
    out[16 + 3 * c] = h.frame.c[c].z;
  }
}
// The lobe that uc selects, evaluated at (wo, wi) in the same frame: f and pdf of the lobe,
// weighted by its lobe weight and selection probability as sample_f's are.
int unit_principled_eval(const vec3 &wo, const vec3 &wi, const vec3 &albedo, float roughness, float metal, float eta, float uc, vec3 &f, float &pdf) {
  Principled b;
  b.configure(unit_params(wo, albedo, roughness, metal, eta));
  b.construct(uc);
  b.eval(wo, wi, f, pdf);
  return (int)b.lobe;
}
}  // namespace oracle
