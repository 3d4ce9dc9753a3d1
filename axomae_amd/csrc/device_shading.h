// Device shading: texture sampling, the hit record and the Principled BSDF.
// Every formula follows the reference's CPU path term by term (file:line per function)
// so that, compiled without contraction, the GPU rounds like the reference.
#pragma once
#include "device_math.h"
#include "nova_internal.h"

namespace nova_hip {

struct ShadeCtx {
  const TexRec *tex;
  const uint32_t *tex_u32;
  const float *tex_f32;
  const MatRec *mats;
  const MeshXf *mesh_xf;
  const ShadeTri *shade;
  const float *acos_lut;
  int32_t envmap;
  int32_t sampling_roughness;  // -1, or the FP32 texture that drives sampling (gradient checks)
};

// ImageTexture<uint32_t>::sample, CPU branch: bilinear over 4 texels with ensure_bound
// (NovaTextureInterface.h:56-145, TextureContext.h:129-167). Returns 0..255 channel values.
NOVA_HD float4 sample_u32(const ShadeCtx &c, int id, float u, float v) {
  const TexRec t = c.tex[id];
  u = wrap_uv(u);
  v = wrap_uv(v);
  u = t.invert_u ? 1 - u : u;
  v = t.invert_v ? 1 - v : v;
  const float pi = fabsf(u * (float)(t.width - 1.f));
  const float pj = fabsf(v * (float)(t.height - 1.f));
  unsigned i = (unsigned)floorf(pi), j = (unsigned)floorf(pj);
  if (i + 1 >= (unsigned)t.width) i = t.width - 2;
  if (i > (unsigned)t.width) i = 0;
  if (j + 1 >= (unsigned)t.height) j = t.height - 2;
  if (j > (unsigned)t.height) j = 0;
  const float tx = (pi - (float)i) / 1.f;
  const float ty = (pj - (float)j) / 1.f;
  const uint32_t *px = c.tex_u32 + t.offset;
  uint32_t q[4] = {px[j * t.width + i], px[j * t.width + i + 1], px[(j + 1) * t.width + i], px[(j + 1) * t.width + i + 1]};
  float4 cc[4];
  for (int k = 0; k < 4; k++) {
    cc[k] = make_float4((float)((q[k] >> 16) & 0xFF), (float)((q[k] >> 8) & 0xFF), (float)(q[k] & 0xFF), (float)(q[k] >> 24));
    if (t.channels == 3) cc[k].w = 1.f;
  }
  const float hd = 1 - tx;
  float4 top, bot, r;
  top.x = hd * cc[0].x + tx * cc[1].x;
  top.y = hd * cc[0].y + tx * cc[1].y;
  top.z = hd * cc[0].z + tx * cc[1].z;
  top.w = hd * cc[0].w + tx * cc[1].w;
  bot.x = hd * cc[2].x + tx * cc[3].x;
  bot.y = hd * cc[2].y + tx * cc[3].y;
  bot.z = hd * cc[2].z + tx * cc[3].z;
  bot.w = hd * cc[2].w + tx * cc[3].w;
  const float vy = 1 - ty;
  r.x = vy * top.x + ty * bot.x;
  r.y = vy * top.y + ty * bot.y;
  r.z = vy * top.z + ty * bot.z;
  r.w = vy * top.w + ty * bot.w;
  return r;
}

// FP32 material-texture extension: the same bilinear rule on float texels.
NOVA_HD float4 sample_f32_bilinear(const ShadeCtx &c, int id, float u, float v) {
  const TexRec t = c.tex[id];
  u = wrap_uv(u);
  v = wrap_uv(v);
  u = t.invert_u ? 1 - u : u;
  v = t.invert_v ? 1 - v : v;
  const float pi = fabsf(u * (float)(t.width - 1.f));
  const float pj = fabsf(v * (float)(t.height - 1.f));
  unsigned i = (unsigned)floorf(pi), j = (unsigned)floorf(pj);
  if (i + 1 >= (unsigned)t.width) i = t.width - 2;
  if (j + 1 >= (unsigned)t.height) j = t.height - 2;
  const float tx = pi - (float)i, ty = pj - (float)j;
  const float *px = c.tex_f32 + t.offset;
  unsigned idx[4] = {j * t.width + i, j * t.width + i + 1, (j + 1) * t.width + i, (j + 1) * t.width + i + 1};
  float cc[4][4];
  for (int k = 0; k < 4; k++)
    for (int q = 0; q < 4; q++) cc[k][q] = (q < t.channels) ? px[(size_t)idx[k] * t.channels + q] : (q == 3 ? 1.f : 0.f);
  float out[4];
  const float hd = 1 - tx, vy = 1 - ty;
  for (int q = 0; q < 4; q++) {
    float top = hd * cc[0][q] + tx * cc[1][q];
    float bot = hd * cc[2][q] + tx * cc[3][q];
    out[q] = vy * top + ty * bot;
  }
  return make_float4(out[0], out[1], out[2], out[3]);
}

// The four texel indices and weights sample_f32_bilinear blends (for scattering gradients).
NOVA_HD void bilinear_taps(const ShadeCtx &c, int id, float u, float v, unsigned idx[4], float w[4]) {
  const TexRec t = c.tex[id];
  u = wrap_uv(u);
  v = wrap_uv(v);
  u = t.invert_u ? 1 - u : u;
  v = t.invert_v ? 1 - v : v;
  const float pi = fabsf(u * (float)(t.width - 1.f));
  const float pj = fabsf(v * (float)(t.height - 1.f));
  unsigned i = (unsigned)floorf(pi), j = (unsigned)floorf(pj);
  if (i + 1 >= (unsigned)t.width) i = t.width - 2;
  if (j + 1 >= (unsigned)t.height) j = t.height - 2;
  const float tx = pi - (float)i, ty = pj - (float)j;
  const float hd = 1 - tx, vy = 1 - ty;
  idx[0] = j * t.width + i;
  idx[1] = j * t.width + i + 1;
  idx[2] = (j + 1) * t.width + i;
  idx[3] = (j + 1) * t.width + i + 1;
  w[0] = vy * hd;
  w[1] = vy * tx;
  w[2] = ty * hd;
  w[3] = ty * tx;
}

// uv2index + uvToPixel (TextureContext.h:52-58)
NOVA_HD unsigned uv2index(float t, int dim) {
  float a = fabsf(t);
  float rem = a - floorf(a);
  rem = rem == 0 ? a : rem;
  return (unsigned)(long long)(rem * (float)(unsigned)dim);
}

// EnvmapTexture::sample + TextureCtx::f32pixel (NovaTextureInterface.h:161-171,
// TextureContext.h:196-223): z-up spherical coordinates with atan2_approx / acos LUT,
// nearest texel.
NOVA_HD float4 sample_envmap(const ShadeCtx &c, f3 dir) {
  if (c.envmap < 0) return make_float4(0.f, 0.f, 0.f, 0.f);
  f3 n = normalize(dir);
  const float phi = atan2_approx(n.y, n.x);
  const float theta = acos_lt(c.acos_lut, n.z);
  const double inv_2pi = 1.f / (2.f * 3.14159265358979323846), inv_pi = 1.f / 3.14159265358979323846;
  float u = wrap_uv((float)(phi * inv_2pi)), v = wrap_uv((float)(theta * inv_pi));
  const TexRec t = c.tex[c.envmap];
  u = t.invert_u ? 1 - u : u;
  v = t.invert_v ? 1 - v : v;
  unsigned i = uv2index(u, t.width - 1), j = uv2index(v, t.height - 1);
  i = i < (unsigned)t.width ? i : 0u;  // only a NaN direction can leave the texture
  j = j < (unsigned)t.height ? j : 0u;
  const float *px = c.tex_f32 + t.offset + (size_t)(j * t.width + i) * t.channels;
  float4 r = make_float4(0.f, 0.f, 0.f, 0.f);
  r.x = px[0];
  if (t.channels > 1) r.y = px[1];
  if (t.channels > 2) r.z = px[2];
  if (t.channels > 3) r.w = px[3];
  return r;
}

// ------------------------------------------------------------- hit record

struct HitRec {
  f3 ng, dpdu, dpdv, pos;
  float u, v;
  m3 frame;
};

// math::gram_shmidt 3-argument form (utils_3D.h:63-71)
NOVA_HD m3 gram_shmidt3(f3 normal, f3 tangent, f3 bitangent) {
  f3 N = normalize(normal);
  f3 T = normalize(tangent - dot(tangent, N) * N);
  f3 B = cross(N, T);
  if (dot(B, bitangent) < 0.0f) B = -B;
  return m3{T, B, N};
}

// math::make_onb (utils_3D.h:74-82)
NOVA_HD void make_onb(f3 n, f3 &b1, f3 &b2) {
  n = normalize(n);
  float sign = copysignf(1.f, n.z);
  float a = -1.f / (sign + n.z);
  float b = n.x * n.y * a;
  b1 = mk3(1.f + sign * n.x * n.x * a, sign * b, -sign * n.x);
  b2 = mk3(b, sign + n.y * n.y * a, -n.y);
}

NOVA_HD f3 ld3(const float *p) { return mk3(p[0], p[1], p[2]); }
NOVA_HD f3 bary3(f3 p0, f3 p1, f3 p2, float w, float u, float v) { return p0 * w + p1 * u + p2 * v; }
NOVA_HD f3 mul9(const float *m, f3 v) {  // column-major 3x3
  return mk3((m[0] * v.x + m[3] * v.y) + m[6] * v.z, (m[1] * v.x + m[4] * v.y) + m[7] * v.z, (m[2] * v.x + m[5] * v.y) + m[8] * v.z);
}

// computeDerivativesTriangle + fillShadingStructure + constructTbnFrame (embree.cpp:287-390)
// with the closest hit's (t, u, v) standing in for Embree's.
NOVA_HD HitRec make_hit(const ShadeCtx &c, const ShadeTri &tr, f3 ro, f3 rd, float t, float bu, float bv) {
  HitRec h;
  const float w = 1 - (bu + bv);
  h.u = tr.uv0[0] * w + tr.uv1[0] * bu + tr.uv2[0] * bv;
  h.v = tr.uv0[1] * w + tr.uv1[1] * bu + tr.uv2[1] * bv;
  const MeshXf &xf = c.mesh_xf[tr.mesh];
  const float duv02x = tr.uv0[0] - tr.uv2[0], duv02y = tr.uv0[1] - tr.uv2[1];
  const float duv12x = tr.uv1[0] - tr.uv2[0], duv12y = tr.uv1[1] - tr.uv2[1];
  const f3 dp02 = mul9(xf.m3, mk3(tr.p0[0] - tr.p2[0], tr.p0[1] - tr.p2[1], tr.p0[2] - tr.p2[2]));
  const f3 dp12 = mul9(xf.m3, mk3(tr.p1[0] - tr.p2[0], tr.p1[1] - tr.p2[1], tr.p1[2] - tr.p2[2]));
  const float det = duv02x * duv12y - duv02y * duv12x;
  const bool degen = fabsf(det) < 1e-9f;
  const float invdet = 1.f / det;
  f3 dpdu = (duv12y * dp02 - duv02y * dp12) * invdet;
  f3 dpdv = (duv02x * dp12 - duv12x * dp02) * invdet;
  f3 n;
  if (degen || length(cross(dpdu, dpdv)) == 0) {
    n = normalize(cross(dp12, dp02));
    make_onb(n, dpdu, dpdv);
  } else {
    n = cross(dp02, dp12);
  }
  h.ng = normalize(n);
  h.pos = ro + t * rd;
  h.dpdu = dpdu;
  h.dpdv = dpdv;
  const f3 nn = bary3(ld3(tr.n0), ld3(tr.n1), ld3(tr.n2), w, bu, bv);
  const f3 tt = bary3(ld3(tr.t0), ld3(tr.t1), ld3(tr.t2), w, bu, bv);
  const f3 bb = bary3(ld3(tr.b0), ld3(tr.b1), ld3(tr.b2), w, bu, bv);
  const f3 tn = normalize(mul9(xf.n3, nn));
  const f3 tt2 = normalize(mul9(xf.m3, tt));
  const f3 tb = normalize(mul9(xf.m3, bb));
  h.frame = gram_shmidt3(tn, tt2, tb);
  return h;
}

// ------------------------------------------------------------------ BxDFs
// bxdf:: helpers (BxDF_math.h:57-124)
NOVA_HD float abscos(f3 v) { return fabsf(v.z); }
NOVA_HD float absdot(f3 a, f3 b) { return fabsf(dot(a, b)); }
NOVA_HD bool same_hemi(f3 a, f3 b) { return a.z * b.z > 0; }
NOVA_HD float cos2theta(f3 v) { return sqr(v.z); }
NOVA_HD float sin2theta(f3 v) { return smax(0.f, 1.f - cos2theta(v)); }
NOVA_HD float sin2theta_c(float c) { return smax(0.f, 1.f - sqr(c)); }
NOVA_HD float sintheta(f3 v) { return msqrt(sin2theta(v)); }
NOVA_HD float tan2theta(f3 v) { return sin2theta(v) / cos2theta(v); }
NOVA_HD float cosphi(f3 v) {
  float st = sintheta(v);
  if (st == 0.f) return 1.f;
  return clampf(v.x / st, -1.f, 1.f);
}
NOVA_HD float sinphi(f3 v) {
  float st = sintheta(v);
  if (st == 0.f) return 0.f;
  return clampf(v.y / st, -1.f, 1.f);
}
NOVA_HD float cos2phi(f3 v) { return sqr(cosphi(v)); }
NOVA_HD float sin2phi(f3 v) { return sqr(sinphi(v)); }
NOVA_HD float schlick1(float c, float f0) { return f0 + (1.f - f0) * pow5(1.f - c); }

NOVA_HD float refract(f3 wo, f3 n, float eta, bool &valid, f3 &wt) {
  valid = true;
  float ci = dot(wo, n);
  if (ci < 0) {
    n = -n;
    eta = 1.f / eta;
    ci = -ci;
  }
  float s2t = sin2theta_c(ci) / sqr(eta);
  if (s2t >= 1.f) {
    valid = false;
    return 0.f;
  }
  float ct = msqrt(1.f - s2t);
  wt = -wo / eta + (ci / eta - ct) * n;
  return eta;
}

// Fresnel::real (BxDF_math.h:193-208)
NOVA_HD float fresnel_real(float etar, float ci) {
  float eta = etar;
  ci = gmin(gmax(ci, -1.f), 1.f);
  if (ci < 0) {
    eta = 1.f / eta;
    ci = -ci;
  }
  float s2t = sin2theta_c(ci) / sqr(eta);
  if (s2t >= 1) return 1.f;
  float ct = msqrt(1.f - s2t);
  float rpar = (eta * ci - ct) / (eta * ci + ct);
  float rperp = (ci - eta * ct) / (ci + eta * ct);
  return (sqr(rpar) + sqr(rperp)) / 2.f;
}

// shirley_chiu + hemisphere_cosine_sample (BxDF_math.h:129-157)
NOVA_HD f3 cosine_hemisphere(float u0, float u1) {
  float sx = 2 * u0 - 1, sy = 2 * u1 - 1;
  float x = 0.f, y = 0.f;
  if (!(sx == 0 && sy == 0)) {
    float r, th;
    if (fabsf(sx) > fabsf(sy)) {
      r = sx;
      th = 0.785398163397448309616f * (sy / sx);
    } else {
      r = sy;
      th = 1.57079632679489661923f - 0.785398163397448309616f * (sx / sy);
    }
    x = r * cos_f(th);
    y = r * sin_f(th);
  }
  float z = sqrtf(smax(0.f, 1 - x * x - y * y));
  return mk3(x, y, z);
}

// VNDF (BxDF_math.h:231-342)
struct VNDF {
  float ax, ay, ax_inv, ay_inv;
  float a, da_dr;  // alpha = roughness^2 (1e-6 at roughness 0) and its derivative
  NOVA_HD VNDF(float aniso, float rough) {
    aniso = clampf(aniso, -1.f, 1.f);
    float aspect = msqrt(1.f - 0.9f * aniso);
    a = rough != 0 ? rough * rough : 1e-6f;
    da_dr = rough != 0 ? 2.f * rough : 0.f;
    ax = a / aspect;
    ay = a * aspect;
    ax_inv = 1.f / ax;
    ay_inv = 1.f / ay;
  }
  NOVA_HD float D(f3 wm) const {
    if (wm.z < 1e-6f) return 0.f;
    float c2 = cos2phi(wm) * sqr(ax_inv);
    float s2 = sin2phi(wm) * sqr(ay_inv);
    float denom = (float)(3.14159265358979323846 * (double)ax * (double)ay * (double)sqr(cos2theta(wm)) *
                          (double)sqr(1.f + tan2theta(wm) * (c2 + s2)));
    return 1.f / denom;
  }
  NOVA_HD float lambda(f3 v) const {
    float t2 = tan2theta(v);
    if (isinf(t2) || isnan(t2)) return 0;
    float a2 = sqr(ax) * cos2phi(v) + sqr(ay) * sin2phi(v);
    return (msqrt(1 + a2 * t2) - 1) * 0.5f;
  }
  NOVA_HD float G1(f3 v) const {
    v.z = smax(v.z, 1e-6f);
    return 1.f / (1.f + lambda(v));
  }
  NOVA_HD float G(f3 wo, f3 wi) const { return 1.f / (1.f + lambda(wo) + lambda(wi)); }
  NOVA_HD float pdf(f3 wo, f3 wm) const { return D(wm) * absdot(wo, wm) * G1(wo) / abscos(wo); }
  NOVA_HD f3 sample(f3 wo, float u0, float u1) const {
    f3 woh = normalize(mk3(wo.x * ax, wo.y * ay, wo.z));
    if (woh.z < 0) woh = -woh;
    f3 t = woh.z < 1.f ? normalize(cross(mk3(0, 0, 1), woh)) : mk3(1, 0, 0);
    f3 b = cross(woh, t);
    float r = msqrt(u0);
    float ang = (float)(2.0 * 3.14159265358979323846 * (double)u1);
    float px = r * cos_f(ang), py = r * sin_f(ang);
    float s = 0.5f * (1.f + woh.z);
    py = msqrt(1.f - sqr(px)) * (1.f - s) + s * py;
    f3 nh = px * t + py * b + msqrt(fmaxf(0.f, 1.f - sqr(px) - sqr(py))) * woh;
    return normalize(mk3(ax * nh.x, ay * nh.y, smax(1e-6f, nh.z)));
  }
  NOVA_HD bool full_specular() const { return smax(ax, ay) < 1e-5f; }
  // d ln D / d alpha: with ax, ay both proportional to alpha, D = 1 / (pi a^2 cos^4 (1+e)^2)
  // where e = tan^2 (cos^2phi/ax^2 + sin^2phi/ay^2) scales as 1/a^2.
  NOVA_HD float dlnD_da(f3 wm) const {
    const float e = tan2theta(wm) * (cos2phi(wm) * sqr(ax_inv) + sin2phi(wm) * sqr(ay_inv));
    return 2.f * (e - 1.f) / (a * (1.f + e));
  }
  // d lambda / d alpha: lambda = (sqrt(1 + a2 t2) - 1) / 2 with a2 scaling as alpha^2.
  NOVA_HD float dlambda_da(f3 v) const {
    const float t2 = tan2theta(v);
    if (isinf(t2) || isnan(t2)) return 0.f;
    const float a2 = sqr(ax) * cos2phi(v) + sqr(ay) * sin2phi(v);
    return a2 * t2 / (2.f * a * msqrt(1 + a2 * t2));
  }
  NOVA_HD float dlnG_da(f3 wo, f3 wi) const { return -(dlambda_da(wo) + dlambda_da(wi)) / (1.f + lambda(wo) + lambda(wi)); }
};

enum : unsigned { BX_REFLECTION = 1, BX_TRANSMISSION = 2, BX_DIFFUSE = 4, BX_GLOSSY = 8, BX_SPECULAR = 16 };
enum : int { LOBE_NONE = 0, LOBE_DIFFUSE = 1, LOBE_SPECULAR = 2, LOBE_DIELECTRIC = 3 };

struct BsdfIn {
  f3 ns, ng, dpdu, dpdv;
  float wo_dot_ng, roughness, metal, transmission, anisotropy;
  f3 eta, albedo;
};

struct Principled {
  m3 frame;
  f3 eta, albedo;
  float roughness, metal, anisotropy;
  float p_diffuse, p_specular, p_transmission, w_diffuse, w_specular, w_transmission;
  int lobe;
  float lobe_weight, lobe_prob;

  // configureBSDFParams + generateWeightsFromParam (BSDF.h:43-97); subsurface 1, sheen 0,
  // clearcoat 0, specular 1, thin_surface false (BxDF.h:22-39 defaults).
  NOVA_HD void configure(const BsdfIn &p) {
    frame = gram_shmidt3(p.ns, p.dpdu, p.dpdv);
    eta = p.eta;
    albedo = p.albedo;
    roughness = p.roughness;
    metal = p.metal;
    anisotropy = p.anisotropy;
    if (p.wo_dot_ng < 0) {
      frame.c0 = -frame.c0;
      frame.c1 = -frame.c1;
      frame.c2 = -frame.c2;
      if (eta.x != 0.f && eta.y != 0.f && eta.z != 0.f) eta = mk3(1.f / eta.x, 1.f / eta.y, 1.f / eta.z);
    }
    const float ac = fabsf(p.wo_dot_ng);
    w_diffuse = (1 - p.transmission) * (1 - p.metal);
    const float w_sheen = (1 - p.metal) * 0.f * (1.f - p.transmission);
    const float w_cc = 0.25f * 0.f;
    w_specular = 1.f;
    w_transmission = (1.f - p.metal) * p.transmission;
    const float F = schlick1(ac, 1.f * 0.04f);
    const float pd = w_diffuse * (1.f - F);
    const float ps = w_specular * F;
    float W = pd + ps + w_sheen + w_cc + w_transmission;
    if (W == 0) W = 1.f;
    p_diffuse = pd / W;
    p_specular = ps / W;
    p_transmission = w_transmission / W;
  }
  // constructBXDF (BSDF.h:68-79)
  NOVA_HD void construct(float uc) {
    float idx = 0.f;
    lobe = LOBE_NONE;
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
    }
  }
  NOVA_HD unsigned flags(const VNDF &g) const {
    if (lobe == LOBE_DIFFUSE) return BX_DIFFUSE | BX_REFLECTION;
    if (lobe == LOBE_SPECULAR) return g.full_specular() ? (BX_SPECULAR | BX_REFLECTION) : (BX_GLOSSY | BX_REFLECTION);
    if (lobe == LOBE_DIELECTRIC) {
      unsigned f = eta.x == 1.f ? BX_TRANSMISSION : (BX_REFLECTION | BX_TRANSMISSION);
      return f | (g.full_specular() ? BX_SPECULAR : BX_GLOSSY);
    }
    return 0;
  }
};

enum : int { SAMPLE_DIFFUSE = 0, SAMPLE_MICROFACET = 1, SAMPLE_DELTA = 2 };

struct BsdfSample {
  f3 f, wi;
  float pdf, costheta;
  f3 wm;     // microfacet normal of a microfacet sample (local frame)
  int kind;  // SAMPLE_*
};

// DisneyDiffuseBxDF::sample_f (BxDF.h:295-325)
NOVA_HD bool sample_disney(const Principled &b, f3 wo, float u0, float u1, BsdfSample &s) {
  f3 wi = normalize(cosine_hemisphere(u0, u1));
  f3 h = normalize(wi + wo);
  const float ci = wi.z, co = wo.z;
  if (!same_hemi(wo, wi) || ci <= 0.f || co <= 0.f) return false;
  const float rough = sqr(b.roughness);
  const float fd90 = 0.5f + 2 * rough * sqr(absdot(wi, h));
  const float fi_d = 1.f + (fd90 - 1) * pow5(1.f - abscos(wi)), fo_d = 1.f + (fd90 - 1) * pow5(1.f - abscos(wo));
  const f3 fd = b.albedo * 0.318309886183790671538f * fi_d * fo_d;
  const float fss90 = rough * sqr(absdot(h, wi));
  const float fi = 1.f + (fss90 - 1) * pow5(1.f - abscos(wi)), fo = 1.f + (fss90 - 1) * pow5(1.f - abscos(wo));
  const float tail = 0.5f + fi * fo * (-0.5f + 1 / (co + ci));
  const f3 fss = 1.25f * b.albedo * 0.318309886183790671538f * tail;
  s.f = b.lobe_weight * ((1 - 1.f) * fd + 1.f * fss);
  s.pdf = b.lobe_prob * (ci * 0.318309886183790671538f);
  s.costheta = ci;
  s.wi = wi;
  s.kind = SAMPLE_DIFFUSE;
  return true;
}

// SpecularBxDF::sample_f (BxDF.h:367-397)
// g samples and weights by its pdf; ge gives the BSDF value (the same distribution except
// in the detached-sampling gradient checks).
NOVA_HD bool sample_specular(const Principled &b, const VNDF &g, const VNDF &ge, f3 wo, float u0, float u1, BsdfSample &s) {
  f3 wm = g.sample(wo, u0, u1);
  f3 wi = reflect(-wo, wm);
  const float co = wo.z, ci = wi.z;
  if (co <= 0 || ci <= 0) return false;
  const float wmo = absdot(wm, wo);
  const float m = b.metal;
  f3 F0 = mk3(0.04f * (1.f - m) + m * b.albedo.x, 0.04f * (1.f - m) + m * b.albedo.y, 0.04f * (1.f - m) + m * b.albedo.z);
  const float p = pow5(1.f - wmo);
  F0 = mk3(F0.x + (1.f - F0.x) * p, F0.y + (1.f - F0.y) * p, F0.z + (1.f - F0.z) * p);
  const f3 f = F0 * ge.D(wm) * ge.G(wo, wi) / (4 * co * ci);
  const float pdf = g.pdf(wo, wm) / (4 * wmo);
  s.f = b.lobe_weight * f;
  s.pdf = b.lobe_prob * pdf;
  s.wi = wi;
  s.costheta = ci;
  s.wm = wm;
  s.kind = SAMPLE_MICROFACET;
  return true;
}

// DielectricBxDF::sample_f and its four branches (BxDF.h:403-585). The branches fill
// locals and the record is written once: per-branch field stores made the compiler keep
// the record in scratch memory.
NOVA_HD bool sample_dielectric(const Principled &b, const VNDF &g, const VNDF &ge, f3 wo, float uc, float u0, float u1, BsdfSample &s) {
  const float ior = b.eta.x;
  f3 f, wi, wm = mk3(0.f, 0.f, 0.f);
  float pdf, ct;
  int kind;
  if (g.full_specular() || ior == 1.f) {
    const float R = fresnel_real(ior, wo.z), T = 1 - R;
    const float pr = R, pt = T;
    if (pr == 0.f && pt == 0.f) return false;
    if (uc < (pr / (pr + pt))) {
      wi = reflect(-wo, mk3(0, 0, 1.f));
      ct = abscos(wi);
      f = b.lobe_weight * b.albedo * R / ct;
      pdf = b.lobe_prob * pr / (pr + pt);
    } else {
      bool ok = true;
      wi = mk3(0, 0, 0);
      refract(wo, mk3(0.f, 0.f, 1.f), ior, ok, wi);
      if (!ok) return false;
      ct = abscos(wi);
      f3 ft = T * b.albedo / ct;
      ft = ft / sqr(ior);
      f = b.lobe_weight * ft;
      pdf = b.lobe_prob * pt / (pr + pt);
    }
    kind = SAMPLE_DELTA;
  } else {
    wm = g.sample(wo, u0, u1);
    const float wo_dot_wm = dot(wo, wm);
    const float R = fresnel_real(ior, wo_dot_wm), T = 1 - R;
    const float pr = R, pt = T;
    if (pr == 0.f && pt == 0.f) return false;
    if (uc < (pr / (pr + pt))) {
      wi = reflect(-wo, wm);
      if (!same_hemi(wo, wi)) return false;
      const f3 fr = b.albedo * ge.D(wm) * ge.G(wo, wi) * R / (4 * wi.z * wo.z);
      const float pr_ = g.pdf(wo, wm) / (4 * absdot(wo, wm)) * pr / (pr + pt);
      f = b.lobe_weight * fr;
      pdf = b.lobe_prob * pr_;
      ct = abscos(wi);
    } else {
      wi = mk3(0, 0, 0);
      bool no_tir = true;
      const float eta = refract(wo, wm, ior, no_tir, wi);
      if (!no_tir) return false;
      wi = normalize(wi);
      if (same_hemi(wi, wo) || wi.z == 0.f) return false;
      const float ci = wi.z, co = wo.z;
      const float denom = sqr(dot(wi, wm) + dot(wo, wm) / eta);
      const float jac = absdot(wi, wm) / denom;
      const float pt_ = g.pdf(wo, wm) * jac * pt / (pt + pr);
      f3 ft = b.albedo * T * ge.D(wm) * ge.G(wo, wi) * fabsf(dot(wi, wm) * dot(wo, wm) / (ci * co * denom));
      ft = ft / sqr(eta);
      f = b.lobe_weight * ft;
      ct = abscos(wi);
      pdf = b.lobe_prob * pt_;
    }
    kind = SAMPLE_MICROFACET;
  }
  s.f = f;
  s.wi = wi;
  s.pdf = pdf;
  s.costheta = ct;
  s.wm = wm;
  s.kind = kind;
  return true;
}

// Derivatives of ln c, c = f cos / pdf, for a sample taken with everything but the BSDF
// value held fixed: dlna[ch] with respect to the albedo channel ch, dlnr with respect to
// the roughness. Diffuse: f = w 1.25 albedo/pi tail(r), tail = 0.5 + fi fo (1/(co+ci) - 0.5),
// fi = 1 + (r^2 (h.wi)^2 - 1) (1-ci)^5. Microfacet lobes: f proportional to D G in alpha = r^2;
// the specular lobe's Fresnel F = F0 + (1-F0) p with F0 = 0.04 (1-m) + m albedo. Delta lobes
// do not depend on the roughness. wo and s.wi are in the local frame.
NOVA_HD void bsdf_log_derivatives(const Principled &b, const VNDF &ge, f3 wo, const BsdfSample &s, f3 &dlna, float &dlnr) {
  const f3 inv_alb = mk3(b.albedo.x != 0.f ? 1.f / b.albedo.x : 0.f, b.albedo.y != 0.f ? 1.f / b.albedo.y : 0.f,
                         b.albedo.z != 0.f ? 1.f / b.albedo.z : 0.f);
  dlna = inv_alb;
  dlnr = 0.f;
  if (b.lobe == LOBE_DIFFUSE) {
    const f3 wi = s.wi;
    const f3 h = normalize(wi + wo);
    const float ci = wi.z, co = wo.z;
    const float hw2 = sqr(absdot(h, wi));
    const float fss90 = sqr(b.roughness) * hw2;
    const float pi5 = pow5(1.f - abscos(wi)), po5 = pow5(1.f - abscos(wo));
    const float fi = 1.f + (fss90 - 1) * pi5, fo = 1.f + (fss90 - 1) * po5;
    const float k = -0.5f + 1 / (co + ci);
    const float tail = 0.5f + fi * fo * k;
    const float dtail = (pi5 * fo + fi * po5) * k * 2.f * b.roughness * hw2;
    dlnr = tail != 0.f ? dtail / tail : 0.f;
    return;
  }
  if (s.kind == SAMPLE_MICROFACET) dlnr = (ge.dlnD_da(s.wm) + ge.dlnG_da(wo, s.wi)) * ge.da_dr;
  if (b.lobe == LOBE_SPECULAR) {
    const float m = b.metal;
    const float p = pow5(1.f - absdot(s.wm, wo));
    const f3 F0 = mk3(0.04f * (1.f - m) + m * b.albedo.x, 0.04f * (1.f - m) + m * b.albedo.y, 0.04f * (1.f - m) + m * b.albedo.z);
    const f3 F = mk3(F0.x + (1.f - F0.x) * p, F0.y + (1.f - F0.y) * p, F0.z + (1.f - F0.z) * p);
    const float dF = m * (1.f - p);
    dlna = mk3(F.x != 0.f ? dF / F.x : 0.f, F.y != 0.f ? dF / F.y : 0.f, F.z != 0.f ? dF / F.z : 0.f);
  }
}

// bsdf::correct_specular_shading_normal (BxDF_math.h:26-53)
NOVA_HD f3 correct_specular_shading_normal(f3 ng, f3 wo, f3 ns) {
  f3 R = reflect(-wo, ns);
  const float woz = dot(wo, ng);
  const float threshold = fminf(0.9f * woz, 0.01f);
  if (dot(ng, R) >= threshold) return ns;
  f3 X = normalize(ns - dot(ns, ng) * ng);
  if (length2(X) == 0.f) X = ns;
  const float wox = dot(wo, X);
  const float a = sqr(wox) + sqr(woz);
  const float b = 2.0f * (a + woz * threshold);
  const float c = sqr(threshold + woz);
  const float Nz2 = (wox < 0) ? 0.25f * (b + msqrt(sqr(b) - 4.0f * a * c)) / a : 0.25f * (b - msqrt(sqr(b) - 4.0f * a * c)) / a;
  if (Nz2 > 1.f || Nz2 < 0.f) return ng;
  const float Nx = msqrt(1.0f - Nz2), Nz = msqrt(Nz2);
  return Nx * X + Nz * ng;
}

}  // namespace nova_hip
