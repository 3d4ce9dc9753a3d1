// Batched environment-map bakes: one launch covers every output texel of every map.
//  * diffuse irradiance: TextureOperations<float>::computeDiffuseIrradiance
//    (TextureProcessing.cpp:9-64, TextureProcessing.h:253-301) - pgc3d-hashed cosine
//    samples per texel, looked up with the nearest-texel rule uvSample degenerates to;
//  * GGX prefiltered specular: envmap_prefilter.frag:1-116 on equirect, level 0;
//  * the split-sum BRDF table: brdf_lookup_table_baker.frag:1-117.
// Each output texel is one lane; samples are summed in the reference's order.
#include <algorithm>
#include <vector>

#include "device_math.h"
#include "nova_internal.h"

namespace nova_hip {

struct BakeMap {
  const float *env;
  float *out;
  int32_t w, h, c, ow, oh;
  int32_t kind;  // NOVA_HIP_BAKE_*
  float roughness;
  uint64_t first;  // first global texel index of this map
};

__device__ __forceinline__ unsigned uv_to_pixel(float coord, unsigned dim) { return (unsigned)(long long)(coord * (float)dim); }

// sphericalToCartesian<float> (math_spherical.h:33-39)
__device__ __forceinline__ f3 sph_to_cart(float phi, float theta) {
  const float z = cos_f(theta);
  const float x = sin_f(theta) * cos_f(phi);
  const float y = sin_f(theta) * sin_f(phi);
  return mk3(x, y, z);
}

// discreteSample + uvSample (TextureProcessing.h:173-210)
__device__ __forceinline__ f3 discrete(const BakeMap &m, int x, int y) {
  int xc, yc;
  x = x < 0 ? 0 : x;  // coordinates come from unsigned texel indices; a NaN direction must stay in bounds
  y = y < 0 ? 0 : y;
  if (x >= m.w) xc = x % m.w;
  else if (x < 0) xc = m.w + (x % m.w);
  else xc = x;
  if (y >= m.h) yc = y % m.h;
  else if (y < 0) yc = m.h + (y % m.h);
  else yc = y;
  const float *p = m.env + ((size_t)yc * m.w + xc) * m.c;
  if (m.c == 4) {  // RGBA texels of a 16-byte aligned map: one vector load
    const float4 t = *reinterpret_cast<const float4 *>(p);
    return mk3(t.x, t.y, t.z);
  }
  return mk3(p[0], p[1], p[2]);
}

// uvSample: the pixel coordinates px, py are whole numbers (an unsigned texel index turned
// back into a float), so its bilinear weights fu = (px - floor(px)) / 1 and fv are exactly
// zero and the blend (1-fu)(1-fv) tl + fu tr + ... is the top-left texel (for finite maps;
// a neighbour holding Inf or NaN would turn the reference's sum into NaN). Only that texel
// is read: one gather instead of four.
__device__ __forceinline__ f3 uv_sample(const BakeMap &m, float u, float v) {
  const float ui = floorf(u), vi = floorf(v);
  const float up = (u > 1. || u < 0.) ? u - ui : u;
  const float vp = (v > 1. || v < 0.) ? v - vi : v;
  const unsigned px = min(uv_to_pixel(up, (unsigned)m.w), (unsigned)m.w), py = min(uv_to_pixel(vp, (unsigned)m.h), (unsigned)m.h);
  return discrete(m, (int)px, (int)py);
}

__device__ __forceinline__ f3 dir_sample(const BakeMap &m, const float *lut, f3 d) {
  const float phi = atan2_approx(d.y, d.x), theta = acos_lt(lut, d.z);
  const double inv_2pi = 1.f / (2.f * 3.14159265358979323846), inv_pi = 1.f / 3.14159265358979323846;
  return uv_sample(m, (float)(phi * inv_2pi), (float)(theta * inv_pi));
}

// sin and cos of 2*pi*x for x in [0, 1): one float sincospi, no range reduction. The
// reference rounds 2*pi*x to float and calls the float libm (TextureProcessing.h:270-276,
// envmap_prefilter.frag); the two differ by a few ulp, far inside the bakes' 1e-3 gate,
// and the double-precision cos/sin of the render path would make these sample loops
// FP64-bound.
__device__ __forceinline__ void sincos_2pi(float x, float &s, float &c) { sincospif(2.f * x, &s, &c); }

// math::importance_sampling::pgc3d (math_importance_sampling.h:7-21)
__device__ __forceinline__ f3 pgc3d(unsigned x, unsigned y, unsigned z) {
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
  const float s = 1.f / 4294967296.f;
  return mk3((float)x * s, (float)y * s, (float)z * s);
}

__device__ __forceinline__ int find_map(const BakeMap *maps, int count, uint64_t g) {
  int lo = 0, hi = count - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (maps[mid].first <= g)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

// One irradiance texel. The reference draws its cosine samples from pgc3d of the texel's
// recomputed pixel and the sample number; the Sobol variant takes Owen-scrambled Sobol
// points (dims 0 and 1) at indices texel * samples + s, so each texel gets its own
// stratified block of the sequence.
__device__ __forceinline__ f3 irradiance_texel(const BakeMap &m, const float *__restrict__ lut, const uint32_t *__restrict__ sobol, unsigned i,
                                               unsigned j, int samples) {
  {
    const float u = (float)((double)i / (double)m.ow), v = (float)((double)j / (double)m.oh);
    const f3 nrm = sph_to_cart((float)(2 * 3.14159265358979323846 * u), (float)(3.14159265358979323846 * v));
    const float aphi = atan2_approx(nrm.y, nrm.x), atheta = acos_lt(lut, nrm.z);
    const double inv_2pi = 1.f / (2.f * 3.14159265358979323846), inv_pi = 1.f / 3.14159265358979323846;
    const unsigned px = uv_to_pixel((float)(aphi * inv_2pi), (unsigned)m.ow), py = uv_to_pixel((float)(atheta * inv_pi), (unsigned)m.oh);
    const f3 right = mk3(1.f, 0.f, 0.f);
    const float dd = dot(right, nrm);
    f3 tangent = mk3(0.f, 1.f, 0.f);
    if (1.0 - fabs((double)dd) > 1e-6) tangent = normalize(cross(right, nrm));
    const f3 bitangent = cross(nrm, tangent);
    f3 irr = mk3(0.f, 0.f, 0.f);
    const uint32_t base = (j * (uint32_t)m.ow + i) * (uint32_t)samples;
    // Sobol points of consecutive indices: the gray code of idx differs from idx-1's in bit
    // ctz(idx), so each sample XORs one direction number into the running values (equal to
    // the direct gray-code sum bit for bit).
    uint32_t ax = 0, ay = 0;
    if (m.kind == NOVA_HIP_BAKE_IRRADIANCE_SOBOL) {
      const uint32_t g0 = base ^ (base >> 1);
      ax = sobol_xor(sobol, g0);
      ay = sobol_xor(sobol + 32, g0);
    }
    for (int s = 0; s < samples; s++) {
      float rx, ry;
      if (m.kind == NOVA_HIP_BAKE_IRRADIANCE_SOBOL) {
        if (s > 0) {
          const uint32_t bit = (uint32_t)__builtin_ctz(base + (uint32_t)s);
          ax ^= sobol[bit];
          ay ^= sobol[32 + bit];
        }
        rx = u32_to_unit(owen_scramble(ax));
        ry = u32_to_unit(owen_scramble(ay));
      } else {
        const f3 r = pgc3d(px, py, (unsigned)s);
        rx = r.x;
        ry = r.y;
      }
      // cosine sample at phi = 2 pi rx, theta = asin(sqrt(ry)) (TextureProcessing.h:270-276):
      // sin(theta) = sqrt(ry), cos(theta) = sqrt(1 - ry)
      float sp, cp;
      sincos_2pi(rx, sp, cp);
      const float st = sqrtf(ry), ct = sqrtf(1.f - ry);
      f3 c = mk3(st * cp, st * sp, ct);
      c = c.x * tangent + c.y * bitangent + c.z * nrm;
      irr = irr + dir_sample(m, lut, c);
    }
    return irr / (float)samples;
  }
}

__device__ __forceinline__ float radical_inverse(unsigned bits) {
  bits = (bits << 16u) | (bits >> 16u);
  bits = ((bits & 0x55555555u) << 1u) | ((bits & 0xAAAAAAAAu) >> 1u);
  bits = ((bits & 0x33333333u) << 2u) | ((bits & 0xCCCCCCCCu) >> 2u);
  bits = ((bits & 0x0F0F0F0Fu) << 4u) | ((bits & 0xF0F0F0F0u) >> 4u);
  bits = ((bits & 0x00FF00FFu) << 8u) | ((bits & 0xFF00FF00u) >> 8u);
  return (float)bits * 2.3283064365386963e-10f;
}

// One prefiltered-specular texel: N = V = R, Hammersley GGX half vectors, NdotL weights.
__device__ __forceinline__ f3 prefilter_texel(const BakeMap &m, const float *__restrict__ lut, unsigned i, unsigned j, int samples) {
  {
    const float u = (float)((double)i / (double)m.ow), v = (float)((double)j / (double)m.oh);
    const f3 N = normalize(sph_to_cart((float)(2 * 3.14159265358979323846 * u), (float)(3.14159265358979323846 * v)));
    f3 col = mk3(0.f, 0.f, 0.f);
    if (m.roughness == 0.f) {
      col = dir_sample(m, lut, N);
    } else {
      float total_w = 0.f;
      const float a = m.roughness * m.roughness;
      const f3 up = fabsf(N.z) < 0.999f ? mk3(0.f, 0.f, 1.f) : mk3(1.f, 0.f, 0.f);
      const f3 t = normalize(cross(up, N));
      const f3 b = cross(N, t);
      for (int s = 0; s < samples; s++) {
        const float xi0 = (float)s / (float)samples, xi1 = radical_inverse((unsigned)s);
        const float ct = sqrtf((1.0f - xi1) / (1.0f + (a * a - 1.0f) * xi1));
        const float st = sqrtf(1.0f - ct * ct);
        float sp, cp;
        sincos_2pi(xi0, sp, cp);
        const f3 H = normalize(t * (cp * st) + b * (sp * st) + N * ct);
        const f3 L = reflect(-N, H);
        const float ndl = fmaxf(dot(N, L), 0.f);
        if (ndl > 0.f) {
          col = col + dir_sample(m, lut, L) * ndl;
          total_w += ndl;
        }
      }
      col = col / total_w;
    }
    return col;
  }
}

// computeLut (brdf_lookup_table_baker.frag:74-113) at texel (i, j): NdotV = (i + 0.5) / ow,
// roughness = (j + 0.5) / oh; the split-sum scale A and bias B over `samples` Hammersley GGX
// half vectors about N = +z, with GeometrySmith's k = roughness^2 / 2.
__device__ __forceinline__ float schlick_ggx_ibl(float ndotv, float roughness) {
  const float a = roughness;
  const float k = (a * a) / 2.0f;
  return ndotv / (ndotv * (1.0f - k) + k);
}

__device__ __forceinline__ float2 brdf_lut_texel(const BakeMap &m, unsigned i, unsigned j, int samples) {
  const float ndotv = ((float)i + 0.5f) / (float)m.ow, roughness = ((float)j + 0.5f) / (float)m.oh;
  const f3 V = mk3(sqrtf(1.0f - ndotv * ndotv), 0.0f, ndotv);
  const f3 N = mk3(0.f, 0.f, 1.f);
  const float a = roughness * roughness;
  const f3 up = fabsf(N.z) < 0.999f ? mk3(0.f, 0.f, 1.f) : mk3(1.f, 0.f, 0.f);
  const f3 t = normalize(cross(up, N));
  const f3 b = cross(N, t);
  const float gv = schlick_ggx_ibl(fmaxf(dot(N, V), 0.0f), roughness);
  float A = 0.f, B = 0.f;
  for (int s = 0; s < samples; s++) {
    const float xi0 = (float)s / (float)samples, xi1 = radical_inverse((unsigned)s);
    const float ct = sqrtf((1.0f - xi1) / (1.0f + (a * a - 1.0f) * xi1));
    const float st = sqrtf(1.0f - ct * ct);
    float sp, cp;
    sincos_2pi(xi0, sp, cp);
    const f3 H = normalize(t * (cp * st) + b * (sp * st) + N * ct);
    const f3 L = normalize(H * (2.0f * dot(V, H)) - V);
    const float ndl = fmaxf(L.z, 0.0f), ndh = fmaxf(H.z, 0.0f), vdh = fmaxf(dot(V, H), 0.0f);
    if (ndl > 0.0f) {
      const float G = schlick_ggx_ibl(ndl, roughness) * gv;
      const float g_vis = (G * vdh) / (ndh * ndotv);
      const float fc = pow5(1.0f - vdh);
      A += (1.0f - fc) * g_vis;
      B += fc * g_vis;
    }
  }
  return make_float2(A / (float)samples, B / (float)samples);
}

// Every texel of every job in one grid-stride launch; jobs are laid out back to back, so a
// wave almost always works on a single map and kind.
constexpr unsigned kBakeBlock = 256;
__global__ __launch_bounds__(kBakeBlock) __attribute__((amdgpu_waves_per_eu(4, 4))) void bake_kernel(const BakeMap *__restrict__ maps, int count, uint64_t total, const float *__restrict__ lut,
                            const uint32_t *__restrict__ sobol, int samples) {
  for (uint64_t g = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; g < total; g += (uint64_t)gridDim.x * blockDim.x) {
    const BakeMap m = maps[find_map(maps, count, g)];
    const uint64_t local = g - m.first;
    const unsigned i = (unsigned)(local % (unsigned)m.ow), j = (unsigned)(local / (unsigned)m.ow);
    if (m.kind == NOVA_HIP_BAKE_BRDF_LUT) {
      const float2 ab = brdf_lut_texel(m, i, j, samples);
      float *o = m.out + ((size_t)j * m.ow + i) * 2;
      o[0] = ab.x;
      o[1] = ab.y;
      continue;
    }
    const f3 col = m.kind == NOVA_HIP_BAKE_PREFILTER ? prefilter_texel(m, lut, i, j, samples) : irradiance_texel(m, lut, sobol, i, j, samples);
    float *o = m.out + ((size_t)j * m.ow + i) * m.c;
    o[0] = col.x;
    o[1] = col.y;
    o[2] = col.z;
    for (int k = 3; k < m.c; k++) o[k] = 0.f;
  }
}

static int launch_batch(const nova_hip_bake_job *jobs, uint32_t count, int32_t samples, hipStream_t s) {
  if (count == 0) return NOVA_HIP_SUCCESS;
  if (!jobs || samples <= 0) return NOVA_HIP_INVALID_ARGUMENT;
  std::vector<BakeMap> maps(count);
  uint64_t total = 0;
  for (uint32_t i = 0; i < count; i++) {
    const nova_hip_bake_job &jb = jobs[i];
    if (jb.kind == NOVA_HIP_BAKE_BRDF_LUT) {  // no environment map; two output channels
      if (!jb.out || jb.out_width < 1 || jb.out_height < 1 || jb.channels != 2) return NOVA_HIP_INVALID_ARGUMENT;
    } else {
      if (!jb.env || !jb.out || jb.width < 1 || jb.height < 1 || jb.channels < 3 || jb.channels > 4 || jb.out_width < 1 || jb.out_height < 1)
        return NOVA_HIP_INVALID_ARGUMENT;
      if (jb.channels == 4 && (reinterpret_cast<uintptr_t>(jb.env) & 15u)) return NOVA_HIP_INVALID_ARGUMENT;  // RGBA maps are read as float4
      if (jb.kind != NOVA_HIP_BAKE_IRRADIANCE && jb.kind != NOVA_HIP_BAKE_IRRADIANCE_SOBOL && jb.kind != NOVA_HIP_BAKE_PREFILTER)
        return NOVA_HIP_INVALID_ARGUMENT;
    }
    // Sobol indices texel * samples + s must stay below 2^32
    if (jb.kind == NOVA_HIP_BAKE_IRRADIANCE_SOBOL && (uint64_t)jb.out_width * jb.out_height * (uint64_t)samples > 0xFFFFFFFFull)
      return NOVA_HIP_INVALID_ARGUMENT;
    maps[i] = BakeMap{jb.env, jb.out, jb.width, jb.height, jb.channels, jb.out_width, jb.out_height, jb.kind, jb.roughness, total};
    total += (uint64_t)jb.out_width * jb.out_height;
  }
  const uint32_t *sobol = device_sobol_table();
  if (!sobol) return NOVA_HIP_OUT_OF_MEMORY;
  // acos table: one device copy per process and device
  static float *d_lut = nullptr;
  static int lut_dev = -1;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return NOVA_HIP_DEVICE_ERROR;
  if (!d_lut || lut_dev != dev) {
    std::vector<float> lut(kAcosLut);
    for (int i = 0; i < kAcosLut; i++) lut[i] = (float)std::acos(((double)i / (double)kAcosLut) * 2.0 - 1.0);
    if (hipMalloc(&d_lut, kAcosLut * sizeof(float)) != hipSuccess) return NOVA_HIP_OUT_OF_MEMORY;
    if (hipMemcpy(d_lut, lut.data(), kAcosLut * sizeof(float), hipMemcpyHostToDevice) != hipSuccess) return NOVA_HIP_DEVICE_ERROR;
    lut_dev = dev;
  }
  BakeMap *d_maps = nullptr;
  if (hipMalloc(&d_maps, count * sizeof(BakeMap)) != hipSuccess) return NOVA_HIP_OUT_OF_MEMORY;
  if (hipMemcpyAsync(d_maps, maps.data(), count * sizeof(BakeMap), hipMemcpyHostToDevice, s) != hipSuccess) {
    hipFree(d_maps);
    return NOVA_HIP_DEVICE_ERROR;
  }
  const unsigned block = kBakeBlock;
  const unsigned grid = (unsigned)std::min<uint64_t>((total + block - 1) / block, 1u << 20);
  hipLaunchKernelGGL(bake_kernel, dim3(grid), dim3(block), 0, s, d_maps, (int)count, total, d_lut, sobol, samples);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  hipFree(d_maps);
  return e == hipSuccess ? NOVA_HIP_SUCCESS : NOVA_HIP_DEVICE_ERROR;
}

static int launch_legacy(int kind, const float *const *env, const int32_t *w, const int32_t *h, const int32_t *c, uint32_t count, float *const *out,
                         const int32_t *ow, const int32_t *oh, const float *rough, int32_t samples, hipStream_t s) {
  if (count == 0) return NOVA_HIP_SUCCESS;
  if (!env || !w || !h || !c || !out || !ow || !oh) return NOVA_HIP_INVALID_ARGUMENT;
  std::vector<nova_hip_bake_job> jobs(count);
  for (uint32_t i = 0; i < count; i++) jobs[i] = nova_hip_bake_job{env[i], w[i], h[i], c[i], out[i], ow[i], oh[i], kind, rough ? rough[i] : 0.f};
  return launch_batch(jobs.data(), count, samples, s);
}

}  // namespace nova_hip

extern "C" int nova_hip_bake_irradiance(const float *const *env_dev, const int32_t *w, const int32_t *h, const int32_t *c, uint32_t count,
                                        float *const *out_dev, const int32_t *out_w, const int32_t *out_h, int32_t samples, void *stream) {
  return nova_hip::launch_legacy(NOVA_HIP_BAKE_IRRADIANCE, env_dev, w, h, c, count, out_dev, out_w, out_h, nullptr, samples, (hipStream_t)stream);
}

extern "C" int nova_hip_bake_batch(const nova_hip_bake_job *jobs, uint32_t count, int32_t samples, void *stream) {
  return nova_hip::launch_batch(jobs, count, samples, (hipStream_t)stream);
}

extern "C" int nova_hip_bake_prefilter(const float *const *env_dev, const int32_t *w, const int32_t *h, const int32_t *c, uint32_t count,
                                       float *const *out_dev, const int32_t *out_w, const int32_t *out_h, const float *roughness, int32_t samples,
                                       void *stream) {
  return nova_hip::launch_legacy(NOVA_HIP_BAKE_PREFILTER, env_dev, w, h, c, count, out_dev, out_w, out_h, roughness, samples, (hipStream_t)stream);
}
