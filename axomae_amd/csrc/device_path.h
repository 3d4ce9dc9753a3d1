// Per-path stages of the integrator as host+device functions: the kernels in
// integrator.hip run them per lane, and a host build of the same functions can be
// checked against the oracle on a CPU.
#pragma once
#include "device_math.h"
#include "device_shading.h"

namespace nova_hip {

struct FrameParams {
  int32_t width, height, frame0, frames;
  uint32_t offset;
  int32_t vinvert;
  int32_t policy;  // NOVA_HIP_INDEX_CPU or NOVA_HIP_INDEX_OPTIX
  int32_t frame_minor;  // integrator path numbering: 1 pixel-major / frame-minor, 0 frame-major (= slot)
  int32_t tiled;        // with frame_minor: pixels taken in 8x8 tiles (pixel_of_rank), 0 row by row
  float inv_P[16], inv_V[16];
};

struct ShadeParams {
  FrameParams fp;
  int32_t bounce, last, rr;  // rr: this scatter rolls Russian roulette
  uint32_t dim0;             // first sampler dimension of this scatter
};

// Output offset of pixel p (floats): AbstractIntegrator::generateImageOffset (Integrator.h:38-42)
NOVA_HD uint32_t pixel_offset(const FrameParams &fp, uint32_t p) {
  const uint32_t x = p % (uint32_t)fp.width, y = p / (uint32_t)fp.width;
  return fp.vinvert ? (((uint32_t)fp.height - 1 - y) * (uint32_t)fp.width + x) * 4u : (y * (uint32_t)fp.width + x) * 4u;
}

// Pixel of rank r in the order the paths take them. Row by row, or with fp.tiled: each band
// of 8 full rows in 8x8 tiles from left to right (row by row inside a tile), then the band's
// last width % 8 columns row by row, and the rows below the last full band row by row. With
// few frames per pixel a wave's camera rays then come from a compact block of pixels
// instead of a strip of one row. A bijection on [0, width * height).
NOVA_HD uint32_t pixel_of_rank(const FrameParams &fp, uint32_t r) {
  if (!fp.tiled) return r;
  const uint32_t W = (uint32_t)fp.width, H8 = (uint32_t)fp.height & ~7u, W8 = W & ~7u;
  if (r >= H8 * W) return r;
  const uint32_t b = r / (8u * W), j = r - b * (8u * W);
  if (j < 8u * W8) return (8u * b + ((j & 63u) >> 3)) * W + 8u * (j >> 6) + (j & 7u);
  const uint32_t rem = W - W8, k = j - 8u * W8;
  return (8u * b + k / rem) * W + W8 + k % rem;
}

// Renders number paths pixel-major, frame-minor: path id = rank * frames + frame (rank: the
// pixel's place in pixel_of_rank's order), which is also the path's entry in the first
// queue. So a wave traces one pixel's camera rays of consecutive frames (they differ only by
// the jitter and walk the same nodes), and the path state and the per-path radiance
// (partial) are indexed by path id. The sampler's slot keeps the frame-major numbering,
// slot = frame * npix + pixel. The adjoint's replay numbers its paths the same way; its
// workgroups sum the texel gradients of a pixel's frames in LDS before the device atomics.
// (frame_minor = 0, path id = slot, is the frame-major numbering.)
NOVA_HD uint32_t path_slot(const FrameParams &fp, uint32_t pid) {
  if (!fp.frame_minor) return pid;
  const uint32_t npix = (uint32_t)fp.width * (uint32_t)fp.height;
  return (pid % (uint32_t)fp.frames) * npix + pixel_of_rank(fp, pid / (uint32_t)fp.frames);
}

// Sampler index of a path slot (frame-major slots), u32 wraparound. CPU policy:
// idx * frame + renderer_max_samples with idx = 4 * pixel (Integrator.h:80). OptiX policy:
// (x * H + y) * sample_max + frame (integrator_optix.cu:43).
NOVA_HD uint32_t sample_index(const FrameParams &fp, uint32_t slot) {
  const uint32_t npix = (uint32_t)fp.width * (uint32_t)fp.height;
  const uint32_t f = (uint32_t)fp.frame0 + slot / npix, p = slot % npix;
  if (fp.policy == NOVA_HIP_INDEX_OPTIX) return ((p % (uint32_t)fp.width) * (uint32_t)fp.height + p / (uint32_t)fp.width) * fp.offset + f;
  return pixel_offset(fp, p) * f + fp.offset;
}

NOVA_HD float sobol_dim(const uint32_t *table, uint32_t dim, uint32_t gray) { return u32_to_unit(owen_scramble(sobol_xor(table + (size_t)dim * 32, gray))); }

// The same samples from per-nibble tables of a few consecutive dimensions: entry
// [(dim - dim0) * 128 + k * 16 + v] is the XOR of the direction numbers of dimension dim
// selected by the bits of nibble value v at bits 4k..4k+3, so a sample is 8 table reads
// instead of 32 conditional XORs (XOR is order-free: the bits are identical).
struct SobolNibbles {
  const uint32_t *tab;
  uint32_t dim0;
};
NOVA_HD uint32_t sobol_nibble_xor(const uint32_t *t, uint32_t gray) {
  uint32_t acc = 0;
#pragma unroll
  for (int k = 0; k < 8; k++) acc ^= t[k * 16 + ((gray >> (4 * k)) & 15u)];
  return acc;
}
NOVA_HD float sobol_dim(const SobolNibbles &s, uint32_t dim, uint32_t gray) {
  return u32_to_unit(owen_scramble(sobol_nibble_xor(s.tab + (size_t)(dim - s.dim0) * 128, gray)));
}
// Fills the nibble tables of dims [dim0, dim0 + ndims) (entries of dims >= 256 are 0);
// entries e = first, first + step, ... (one thread's share).
NOVA_HD void sobol_nibble_fill(const uint32_t *table, uint32_t dim0, int ndims, uint32_t *out, int first, int step) {
  for (int e = first; e < ndims * 128; e += step) {
    const uint32_t dim = dim0 + (uint32_t)(e / 128);
    const int k = (e % 128) / 16, v = e % 16;
    uint32_t x = 0;
    if (dim < 256)
      for (int i = 0; i < 4; i++)
        if ((v >> i) & 1) x ^= table[(size_t)dim * 32 + 4 * k + i];
    out[e] = x;
  }
}

// Camera ray of a path slot: jitter from dims 0-1, screen2ndc(x, H - y) and ray_inv_mat
// (Integrator.h:80-88, math_camera.h:28-58) in glm's operation order.
// ray_inv_mat at NDC (nx, ny).
NOVA_HD void camera_ray_ndc(const FrameParams &fp, float nx, float ny, f3 &o, f3 &d) {
  const float *P = fp.inv_P, *V = fp.inv_V;
  float ox = mat4_row(P, 0, nx, ny, -1.f, 1.f), oy = mat4_row(P, 1, nx, ny, -1.f, 1.f), oz = mat4_row(P, 2, nx, ny, -1.f, 1.f),
        ow = mat4_row(P, 3, nx, ny, -1.f, 1.f);
  ox = ox / ow;
  oy = oy / ow;
  oz = oz / ow;
  o = mk3(mat4_row(V, 0, ox, oy, oz, 1.f), mat4_row(V, 1, ox, oy, oz, 1.f), mat4_row(V, 2, ox, oy, oz, 1.f));
  const float ddx = mat4_row(P, 0, nx, ny, -1.f, 0.f), ddy = mat4_row(P, 1, nx, ny, -1.f, 0.f);
  const float vx = mat4_row(V, 0, ddx, ddy, -1.f, 0.f), vy = mat4_row(V, 1, ddx, ddy, -1.f, 0.f), vz = mat4_row(V, 2, ddx, ddy, -1.f, 0.f),
              vw = mat4_row(V, 3, ddx, ddy, -1.f, 0.f);
  const float xx = vx * vx, yy = vy * vy, zz = vz * vz, ww = vw * vw;
  const float inv_len = 1.f / sqrtf((xx + yy) + (zz + ww));  // glm::normalize(vec4)
  d = mk3(vx * inv_len, vy * inv_len, vz * inv_len);
}

// Sob: the direction table (const uint32_t *) or nibble tables holding dims 0-1 (SobolNibbles).
template <class Sob>
NOVA_HD void camera_ray(const FrameParams &fp, const Sob &sobol, uint32_t slot, f3 &o, f3 &d) {
  const uint32_t npix = (uint32_t)fp.width * (uint32_t)fp.height;
  const uint32_t p = slot % npix;
  const int x = (int)(p % (uint32_t)fp.width), y = (int)(p / (uint32_t)fp.width);
  const uint32_t idx = sample_index(fp, slot);
  const uint32_t g = idx ^ (idx >> 1);
  const float dx = sobol_dim(sobol, 0, g) * 0.0005f, dy = sobol_dim(sobol, 1, g) * 0.0005f;
  const int sy = fp.height - y;
  const float nx = (float)(2 * x - fp.width) / (float)fp.width + dx;
  const float ny = (float)(fp.height - 2 * sy) / (float)fp.height + dy;
  camera_ray_ndc(fp, nx, ny, o, d);
}

// Camera ray of pixel p for the render passes: the pixel corner, no jitter
// (NormalIntegrator.cpp / DepthIntegrator.cpp / EmissiveIntegrator.cpp render loops).
NOVA_HD void camera_ray_pass(const FrameParams &fp, uint32_t p, f3 &o, f3 &d) {
  const int x = (int)(p % (uint32_t)fp.width), y = (int)(p / (uint32_t)fp.width);
  const int sy = fp.height - y;
  const float nx = (float)(2 * x - fp.width) / (float)fp.width;
  const float ny = (float)(fp.height - 2 * sy) / (float)fp.height;
  camera_ray_ndc(fp, nx, ny, o, d);
}

// Moller-Trumbore exactly as Triangle::hit orders it (Triangle.h:158-175); accepts
// tnear <= t <= tbest, so a later triangle at an equal t replaces the earlier one.
NOVA_HD bool mt_test(const TriMT &tr, f3 o, f3 d, float tnear, float tbest, float &t, float &u, float &v) {
  const f3 v0 = mk3(tr.a.x, tr.a.y, tr.a.z), e1 = mk3(tr.b.x, tr.b.y, tr.b.z), e2 = mk3(tr.c.x, tr.c.y, tr.c.z);
  const f3 P = cross(d, e2);
  const float det = dot(P, e1);
  if (det == 0.f) return false;
  const float inv_det = 1.f / det;
  const f3 T = o - v0;
  const float uu = dot(P, T) * inv_det;
  if (uu < 0 || uu > 1.f) return false;
  const f3 Q = cross(T, e1);
  const float vv = dot(Q, d) * inv_det;
  if (vv < 0.f || (uu + vv) > 1.f) return false;
  const float tt = dot(Q, e2) * inv_det;
  if (tt < tnear || tt > tbest) return false;
  t = tt;
  u = uu;
  v = vv;
  return true;
}

// Triangle::hit for a mesh with a non-identity transform (Triangle.h:150-175): the ray is
// taken to object space by the inverse model matrix (glm mat4 * vec4 order,
// (m0 x + m1 y) + (m2 z + m3 w)), then tested against the object-space record. t is the
// same ray parameter in both spaces. Inlined: a call in the traversal loop would give the
// kernels a stack frame (96 B of scratch per lane, extend 1.8x slower).
NOVA_HD bool mt_test_object(const TriMT &tr, const float4 *xf_inv, f3 o, f3 d, float tnear, float tbest, float &t, float &u, float &v) {
  const uint32_t k = (uint32_t)__builtin_bit_cast(int32_t, tr.c.w) - 1u;
  const float4 r0 = xf_inv[3 * k], r1 = xf_inv[3 * k + 1], r2 = xf_inv[3 * k + 2];
  const f3 oo = mk3((r0.x * o.x + r0.y * o.y) + (r0.z * o.z + r0.w * 1.f), (r1.x * o.x + r1.y * o.y) + (r1.z * o.z + r1.w * 1.f),
                    (r2.x * o.x + r2.y * o.y) + (r2.z * o.z + r2.w * 1.f));
  const f3 dd = mk3((r0.x * d.x + r0.y * d.y) + (r0.z * d.z + r0.w * 0.f), (r1.x * d.x + r1.y * d.y) + (r1.z * d.z + r1.w * 0.f),
                    (r2.x * d.x + r2.y * d.y) + (r2.z * d.z + r2.w * 0.f));
  return mt_test(tr, oo, dd, tnear, tbest, t, u, v);
}

// The triangle test of a traversal record: world-space MT for meshes with the identity
// transform, object-space MT otherwise. tie() returns the TieTables record holding the
// inverse matrices (read only in the second case).
template <class TieGet>
NOVA_HD bool tri_hit(const TriMT &tr, TieGet tie, f3 o, f3 d, float tnear, float tbest, float &t, float &u, float &v) {
  if (__builtin_bit_cast(int32_t, tr.c.w) != 0) return mt_test_object(tr, tie()->xf_inv, o, d, tnear, tbest, t, u, v);
  return mt_test(tr, o, d, tnear, tbest, t, u, v);
}

// Path-replay adjoint inputs: the primal radiance of every path slot (the batch's partial
// buffers), the loss gradient per pixel (scaled) and the gradient pool (FP32 texel layout).
struct AdjointCtx {
  const float *ltot;  // RGBA primal radiance of the path being replayed
  const float *dimg;
  float scale;
  float *grad;
  // Hits of the primal pass by (bounce, pixel-major path id): (prim bits, t, u, v), for
  // bounces below cache_bounces. The primal pass writes them, the replay reads them instead
  // of tracing its rays again (they are the same rays, so the hits are the same); deeper
  // bounces are traced again. Null: the replay traces every bounce.
  float4 *hit_cache;
  uint32_t cache_paths;    // paths per bounce in the cache
  uint32_t cache_bounces;  // bounces 0 .. cache_bounces-1 are cached
  // The replay kernel's LDS table (kGradSlots keys, 3 values each) that sums its workgroup's
  // texel gradients before they go to device memory; null: straight to device memory.
  uint32_t *table_keys = nullptr;
  float *table_vals = nullptr;
  uint16_t *table_list = nullptr;   // slots taken since the last flush, in claim order
  uint32_t *table_count = nullptr;  // entries of table_list
};

// Texel-gradient table of a replay workgroup: neighbouring pixels' paths splat mostly onto
// the same texels, so each distinct texel gets one device atomic per channel and workgroup
// iteration instead of one per lane. Open addressing on the texel's index in the gradient
// pool; kGradEmpty marks a free slot.
constexpr int kGradSlots = 2048;
constexpr uint32_t kGradEmpty = 0xFFFFFFFFu;
#ifdef __HIP_DEVICE_COMPILE__
// Adds v[0..n) at `key` to the table, listing a newly taken slot; false when no slot was
// found within the probe limit (the caller then adds to device memory).
__device__ __forceinline__ bool grad_table_add(const AdjointCtx *adj, uint32_t key, const float *v, int n) {
  uint32_t *keys = adj->table_keys;
  float *vals = adj->table_vals;
  uint32_t h = (key * 2654435761u) >> 21;  // 11 bits: kGradSlots
  for (int probe = 0; probe < 32; probe++) {
    const uint32_t old = atomicCAS(&keys[h], kGradEmpty, key);
    if (old == kGradEmpty) adj->table_list[atomicAdd(adj->table_count, 1u)] = (uint16_t)h;
    if (old == kGradEmpty || old == key) {
      for (int c = 0; c < n; c++)
        if (v[c] != 0.f) atomicAdd(&vals[h * 3 + c], v[c]);
      return true;
    }
    h = (h + 1) & (kGradSlots - 1);
  }
  return false;
}
#endif

NOVA_HD void add_grad(float *p, float v) {
#ifdef __HIP_DEVICE_COMPILE__
  unsafeAtomicAdd(p, v);
#else
  *p += v;
#endif
}

// Adds g[ch] * (bilinear weight) to the gradient of channel ch of the four texels an FP32
// texture lookup blended (through the workgroup's table when the replay gives one).
NOVA_HD void splat_texture_grad(const ShadeCtx &sc, const AdjointCtx *adj, int tex, float u, float v, const float *g, int nch) {
  unsigned idx[4];
  float w[4];
  bilinear_taps(sc, tex, u, v, idx, w);
  const TexRec t = sc.tex[tex];
  const int nc = nch < t.channels ? nch : t.channels;  // at most 3
  for (int k = 0; k < 4; k++) {
    const uint64_t key = t.offset + (uint64_t)idx[k] * (uint32_t)t.channels;  // the texel's channel 0 in the pool
    float val[3] = {0.f, 0.f, 0.f};
    for (int ch = 0; ch < nc; ch++) val[ch] = w[k] * g[ch];
#ifdef __HIP_DEVICE_COMPILE__
    if (adj->table_keys && key < kGradEmpty && grad_table_add(adj, (uint32_t)key, val, nc)) continue;
#endif
    for (int ch = 0; ch < nc; ch++)
      if (g[ch] != 0.f) add_grad(adj->grad + key + ch, val[ch]);
  }
}

// The surface half of PrincipledMaterial::scatter (NovaMaterials.h:89-137) at a hit: the
// hit record, the material's textures, the shading normal and the Principled BSDF sample
// (sampleBSDF, NovaMaterials.h:63-79: dims dim0 (lobe), dim0+1..2 (u0), dim0+3 (uc)).
struct SurfaceSample {
  HitRec h;
  int32_t material;
  f3 ns;           // shading normal from the normal map, before the specular correction (mat_rec.normal)
  BsdfSample smp;  // wi in world space
  bool ok;         // PrincipledBSDF::sample_f succeeded
  f3 dlna;         // with kAdj: d ln c / d albedo channel
  float dlnr;      // with kAdj: d ln c / d roughness
};

template <bool kAdj = false, class Sobol = const uint32_t *>
NOVA_HD void sample_surface(const ShadeParams &sp, const ShadeCtx &sc, Sobol sobol, uint32_t slot, f3 o, f3 d, int prim, float ht, float hu, float hv,
                            SurfaceSample &out) {
  const ShadeTri &tr = sc.shade[prim];
  out.h = make_hit(sc, tr, o, d, ht, hu, hv);
  const HitRec &h = out.h;
  out.material = tr.material;
  const MatRec &mat = sc.mats[tr.material];
  const float ONE_255 = 1.f / 255.f;
  f3 nrm = mk3(0.f, 0.f, 1.f);
  if (mat.normal >= 0) {
    float4 t4 = sample_u32(sc, mat.normal, h.u, h.v);
    nrm = mk3(t4.x * ONE_255, t4.y * ONE_255, t4.z * ONE_255);
  }
  const f3 ns = normalize(mul(h.frame, nrm * 2.f - 1.f));
  out.ns = ns;
  const f3 wo = normalize(-d);
  BsdfIn bi;
  bi.ns = ns;
  bi.ng = h.ng;
  bi.dpdu = h.dpdu;
  bi.dpdv = h.dpdv;
  bi.wo_dot_ng = dot(wo, h.ng);
  if (mat.albedo_f32 >= 0) {
    float4 a4 = sample_f32_bilinear(sc, mat.albedo_f32, h.u, h.v);
    bi.albedo = mk3(a4.x, a4.y, a4.z);
  } else if (mat.albedo >= 0) {
    float4 a4 = sample_u32(sc, mat.albedo, h.u, h.v);
    bi.albedo = mk3(a4.x * ONE_255, a4.y * ONE_255, a4.z * ONE_255);
  } else {
    bi.albedo = mk3(1.f, 1.f, 1.f);
  }
  f3 met = mk3(0.f, 0.f, 0.f);
  if (mat.metallic >= 0) {
    float4 m4 = sample_u32(sc, mat.metallic, h.u, h.v);
    met = mk3(m4.x * ONE_255, m4.y * ONE_255, m4.z * ONE_255);
  }
  bi.roughness = met.y;
  if (mat.roughness_f32 >= 0) bi.roughness = sample_f32_bilinear(sc, mat.roughness_f32, h.u, h.v).x;
  bi.metal = met.z;
  const bool any_eta = mat.eta[0] != 0.f || mat.eta[1] != 0.f || mat.eta[2] != 0.f;
  const bool any_k = mat.k[0] != 0.f || mat.k[1] != 0.f || mat.k[2] != 0.f;
  bi.transmission = (any_eta && !any_k) ? 1.f : 0.f;
  bi.anisotropy = mat.anisotropy;
  bi.eta = mk3(mat.eta[0], mat.eta[1], mat.eta[2]);

  const uint32_t idx = sample_index(sp.fp, slot);
  const uint32_t g = idx ^ (idx >> 1);
  const float uc = sobol_dim(sobol, sp.dim0, g);
  const float u0 = sobol_dim(sobol, sp.dim0 + 1, g);
  const float u1 = sobol_dim(sobol, sp.dim0 + 2, g);
  Principled b;
  b.configure(bi);
  b.construct(uc);
  float r_sampling = b.roughness;
  if (sc.sampling_roughness >= 0 && mat.roughness_f32 >= 0) r_sampling = sample_f32_bilinear(sc, sc.sampling_roughness, h.u, h.v).x;
  const VNDF ggx(b.anisotropy, r_sampling);  // samples and weights by its pdf
  const VNDF ggx_e(b.anisotropy, b.roughness);  // BSDF value
  if ((b.flags(ggx) & (BX_GLOSSY | BX_REFLECTION)) && bi.wo_dot_ng >= 0) {
    // the corrected shading normal reconfigures the frame; the chosen lobe and its weights stay
    bi.ns = normalize(correct_specular_shading_normal(bi.ng, wo, bi.ns));
    const int lobe = b.lobe;
    const float lw = b.lobe_weight, lp = b.lobe_prob;
    b.configure(bi);
    b.lobe = lobe;
    b.lobe_weight = lw;
    b.lobe_prob = lp;
  }
  const float uc2 = sobol_dim(sobol, sp.dim0 + 3, g);
  // PrincipledBSDF::sample_f (BSDF.h:107-130)
  BsdfSample smp{};
  bool ok = false;
  out.dlna = mk3(0.f, 0.f, 0.f);
  out.dlnr = 0.f;
  if (b.lobe != LOBE_NONE) {
    const f3 lwo = normalize(mul_t(b.frame, wo));
    if ((b.flags(ggx) & (BX_REFLECTION | BX_TRANSMISSION)) && lwo.z != 0.f) {
      if (b.lobe == LOBE_DIFFUSE)
        ok = sample_disney(b, lwo, u0, u1, smp);
      else if (b.lobe == LOBE_SPECULAR)
        ok = sample_specular(b, ggx, ggx_e, lwo, u0, u1, smp);
      else
        ok = sample_dielectric(b, ggx, ggx_e, lwo, uc2, u0, u1, smp);
      if (ok && ((smp.f.x == 0.f && smp.f.y == 0.f && smp.f.z == 0.f) || smp.pdf == 0.f || smp.wi.z == 0.f)) ok = false;
      if (kAdj && ok) bsdf_log_derivatives(b, ggx_e, lwo, smp, out.dlna, out.dlnr);
      if (ok) smp.wi = normalize(mul(b.frame, smp.wi));
    }
  }
  out.smp = smp;
  out.ok = ok;
}

// Emitted radiance of a material at (u, v): the emissive texture / 255 (TexturePackSampler
// emissive, NovaMaterials.h:133) times the emission-scale extension (1 by default).
NOVA_HD f3 emitted(const ShadeCtx &sc, const MatRec &mat, float u, float v) {
  if (mat.emissive < 0) return mk3(0.f, 0.f, 0.f);
  const float ONE_255 = 1.f / 255.f;
  const float4 e4 = sample_u32(sc, mat.emissive, u, v);
  return mk3(e4.x * ONE_255, e4.y * ONE_255, e4.z * ONE_255) * mat.emission_scale;
}

// One path segment after its closest hit: the miss/envmap rule, the hit record, the
// Principled scatter, the Russian-roulette extension and Ray::spawn. Updates the path
// state (T, L, A) front to back (see integrator.hip) and returns true when the path
// continues with the spawned ray (no, nd); on false the path is finished and L, A hold
// its radiance and alpha.
// With kAdj the segment is replayed: after a scatter the radiance still to come,
// R = L_path - L (everything after this vertex, T c V_next), gives the loss gradient of
// the vertex's weight c; it is pushed to the albedo and roughness texels through
// d ln c / d(param) and the bilinear weights.
template <bool kAdj = false, class Sobol = const uint32_t *>
NOVA_HD bool shade_path(const ShadeParams &sp, const ShadeCtx &sc, Sobol sobol, uint32_t slot, f3 o, f3 d, int prim, float ht, float hu,
                        float hv, f3 &T, f3 &L, float &A, f3 &no, f3 &nd, bool &escaped, const AdjointCtx *adj = nullptr) {
  escaped = false;
  if (prim < 0) {  // Li on a miss: envmap(ray.direction) (Integrator.cpp:93-97)
    escaped = true;
    const float4 env = sample_envmap(sc, d);
    L = L + T * mk3(env.x, env.y, env.z);
    A = A + env.w;
    return false;
  }
  SurfaceSample ss;
  sample_surface<kAdj>(sp, sc, sobol, slot, o, d, prim, ht, hu, hv, ss);
  if (!ss.ok) return false;  // scatter failure: this level contributes 0 (Integrator.cpp:72-78)
  const HitRec &h = ss.h;
  const BsdfSample &smp = ss.smp;
  const MatRec &mat = sc.mats[ss.material];
  f3 c = smp.f * smp.costheta / smp.pdf;
  if (!(isfinite(c.x) && isfinite(c.y) && isfinite(c.z))) return false;  // DENAN zeroes this level
  const f3 Le = emitted(sc, mat, h.u, h.v);
  const uint32_t idx = sample_index(sp.fp, slot);
  const uint32_t g = idx ^ (idx >> 1);
  bool survive = true;
  int rr_channel = -1;  // channel whose T c set q < 1 (the rescale c / q depends on it)
  if (sp.rr) {  // Russian-roulette extension: one more dimension after the scatter's four
    const f3 tp = T * c;
    const float qv = fminf(1.f, fmaxf(tp.x, fmaxf(tp.y, tp.z)));
    const float ur = sobol_dim(sobol, sp.dim0 + 4, g);
    if (!(ur < qv)) {
      survive = false;
    } else {
      c = c / qv;
      if (qv < 1.f) rr_channel = (tp.x == qv) ? 0 : (tp.y == qv) ? 1 : 2;
    }
  }
  L = L + T * Le;
  A = A + 1.f;
  if (!survive) return false;
  if (kAdj) {
    const f3 dlna = ss.dlna;
    const float dlnr = ss.dlnr;
    const uint32_t npix = (uint32_t)sp.fp.width * (uint32_t)sp.fp.height;
    const uint32_t px = pixel_offset(sp.fp, slot % npix);
    const float *lt = adj->ltot;  // the path's primal radiance: the caller points ltot at it
    const float *di = adj->dimg + px;
    // loss gradient of this vertex's weight, per channel: dimg * (L_path - L)
    const float G[3] = {adj->scale * di[0] * (lt[0] - L.x), adj->scale * di[1] * (lt[1] - L.y), adj->scale * di[2] * (lt[2] - L.z)};
    float ga[3] = {G[0] * dlna.x, G[1] * dlna.y, G[2] * dlna.z};
    float gr[1] = {(G[0] + G[1] + G[2]) * dlnr};
    if (rr_channel >= 0) {  // c_ch / q with q = T_m c_m: ln q moves with channel m only
      const float dm[3] = {dlna.x, dlna.y, dlna.z};
      ga[rr_channel] -= (G[0] + G[1] + G[2]) * dm[rr_channel];
      gr[0] = 0.f;
    }
    if (mat.albedo_f32 >= 0) splat_texture_grad(sc, adj, mat.albedo_f32, h.u, h.v, ga, 3);
    if (mat.roughness_f32 >= 0) splat_texture_grad(sc, adj, mat.roughness_f32, h.u, h.v, gr, 1);
  }
  T = T * c;
  // Ray::spawn (Ray.h:22-27)
  const f3 wi = smp.wi;
  nd = normalize(wi);
  no = h.pos + h.ng * (1e-3f + 1e-3f * gsign(dot(h.ng, wi)));
  if (sp.last) {  // depth budget spent: V = envmap(direction) without traversal
    const float4 env = sample_envmap(sc, nd);
    L = L + T * mk3(env.x, env.y, env.z);
    A = A + env.w;
    return false;
  }
  return true;
}

// Render passes, one value per pixel (the same for every frame):
//  NORMAL   (NormalIntegrator::Li): the material's shading normal at the hit, (ns, 1); 0 on a miss.
//  DEPTH    (DepthIntegrator::render): z = 1 - (t - near)/(far - near),
//           2 near far / (far + near - z (far - near)) * 2 - 1, divided by far, in all four
//           channels; far in all four on a miss.
//  EMISSIVE (EmissiveIntegrator::Li): the emissive texel / 255 and alpha 1 when the scatter
//           succeeds, else 0.
// The passes reset the sampler to the pixel offset (4 * pixel, no frame factor) and scatter
// with dims 0-3; `fp` is the render's, pixel p in [0, W*H).
enum : int { PASS_NORMAL = 1, PASS_DEPTH = 2, PASS_EMISSIVE = 3 };
NOVA_HD float4 pass_value(int pass, const FrameParams &fp, const ShadeCtx &sc, const uint32_t *sobol, uint32_t p, f3 o, f3 d, int prim, float t,
                          float u, float v, float z_near, float z_far) {
  if (pass == PASS_DEPTH) {
    if (prim < 0) return make_float4(z_far, z_far, z_far, z_far);
    float z = 1 - (t - z_near) / (z_far - z_near);
    z = (2.0f * z_near * z_far) / (z_far + z_near - z * (z_far - z_near)) * 2.f - 1.f;
    const float r = z / z_far;
    return make_float4(r, r, r, r);
  }
  if (prim < 0) return make_float4(0.f, 0.f, 0.f, 0.f);
  ShadeParams sp;
  sp.fp = fp;
  sp.fp.frame0 = 1;  // sample index = pixel offset * 1 + 0
  sp.fp.offset = 0;
  sp.fp.policy = NOVA_HIP_INDEX_CPU;
  sp.bounce = 0;
  sp.last = 0;
  sp.rr = 0;
  sp.dim0 = 0;
  SurfaceSample ss;
  sample_surface<false>(sp, sc, sobol, p, o, d, prim, t, u, v, ss);
  if (pass == PASS_NORMAL) return make_float4(ss.ns.x, ss.ns.y, ss.ns.z, 1.f);
  if (!ss.ok) return make_float4(0.f, 0.f, 0.f, 0.f);
  const f3 e = emitted(sc, sc.mats[ss.material], ss.h.u, ss.h.v);
  return make_float4(e.x, e.y, e.z, 1.f);
}

}  // namespace nova_hip
