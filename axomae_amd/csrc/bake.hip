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
// ----------------------------------------------- GL viewer bakes on equirect maps
// envmap_bake.frag:10-24 with the capture views of RenderPipeline.cpp:14-26, and
// irradiance_baker.frag:15-39. Lookup rule: the equirect texture's GL sampler state (the
// viewer sets its parameters on the cube-map target, so the 2D texture keeps GL's default
// REPEAT wrap), level 0 with GL_LINEAR filtering: texel centres at (i + 0.5) / w, bilinear
// weights, indices modulo the map size. atan and acos in double, rounded once.
__device__ __forceinline__ f3 gl_normalize(f3 v) {
  const float inv = 1.f / sqrtf((v.x * v.x + v.y * v.y) + v.z * v.z);
  return mk3(v.x * inv, v.y * inv, v.z * inv);
}

__device__ __forceinline__ f3 gl_texel(const BakeMap &m, int x, int y) {
  x %= m.w;
  if (x < 0) x += m.w;
  y %= m.h;
  if (y < 0) y += m.h;
  const float *p = m.env + ((size_t)y * m.w + x) * m.c;
  if (m.c == 4) {
    const float4 t = *reinterpret_cast<const float4 *>(p);
    return mk3(t.x, t.y, t.z);
  }
  return mk3(p[0], p[1], p[2]);
}

// texture(environment_map, sphericalToUv(d)) for a unit direction d
__device__ __forceinline__ f3 gl_lookup(const BakeMap &m, f3 d) {
  const float PI = 3.14159265359f;
  const float theta = (float)atan2(-(double)d.z, (double)d.x), phi = (float)acos((double)d.y);
  const float u = theta / (2.0f * PI), v = phi / PI;
  const float x = u * (float)m.w - 0.5f, y = v * (float)m.h - 0.5f;
  const float fx = floorf(x), fy = floorf(y);
  const float a = x - fx, b = y - fy;
  const int i0 = (int)fx, j0 = (int)fy;
  const f3 t00 = gl_texel(m, i0, j0), t10 = gl_texel(m, i0 + 1, j0), t01 = gl_texel(m, i0, j0 + 1), t11 = gl_texel(m, i0 + 1, j0 + 1);
  const float w00 = (1.f - a) * (1.f - b), w10 = a * (1.f - b), w01 = (1.f - a) * b, w11 = a * b;
  return ((w00 * t00 + w10 * t10) + w01 * t01) + w11 * t11;
}

// Capture basis of face `face` (glm::lookAt from the origin: f = normalize(center),
// s = normalize(f x up), u = s x f), faces in capture order: left, right, down, up, back, front.
__device__ __forceinline__ void capture_basis(int face, f3 &s, f3 &u, f3 &f) {
  const float cx[6] = {-1, 1, 0, 0, 0, 0}, cy[6] = {0, 0, -1, 1, 0, 0}, cz[6] = {0, 0, 0, 0, 1, -1};
  const float ux[6] = {0, 0, 0, 0, 0, 0}, uy[6] = {1, 1, 0, 0, 1, 1}, uz[6] = {0, 0, 1, -1, 0, 0};
  f = gl_normalize(mk3(cx[face], cy[face], cz[face]));
  s = gl_normalize(cross(f, mk3(ux[face], uy[face], uz[face])));
  u = cross(s, f);
}

// Face texel (x, y) of a size x size capture, y = 0 the bottom row: the 90-degree camera's
// ray through the pixel centre meets the unit cube at ndc_x s + ndc_y u + f.
__device__ __forceinline__ f3 cube_texel_direction(int face, int x, int y, int size) {
  f3 s, u, f;
  capture_basis(face, s, u, f);
  const float nx = (float)(2 * x + 1) / (float)size - 1.f, ny = (float)(2 * y + 1) / (float)size - 1.f;
  return gl_normalize((nx * s + ny * u) + f);
}

// Equirect texel (i, j) -> its centre direction (the inverse of sphericalToUv), in double.
__device__ __forceinline__ f3 equirect_texel_direction(int i, int j, int w, int h) {
  const double az = 2.0 * 3.14159265358979323846 * ((double)i + 0.5) / (double)w, po = 3.14159265358979323846 * ((double)j + 0.5) / (double)h;
  return mk3((float)(cos(az) * sin(po)), (float)cos(po), (float)(-sin(az) * sin(po)));
}

constexpr unsigned kBakeBlock = 256;
// Workgroups go to the MI355X's 8 XCDs (each with its own L2) round robin. Workgroup b works
// on texel block (b % 8) * (grid / 8) + b / 8 instead of b, so each XCD's L2 sees a
// contiguous run of texels -- one or two maps at a time -- rather than every eighth block
// of all of them. The grid is a multiple of 8, so this is a permutation of the blocks.
constexpr unsigned kXcds = 8;
__global__ __launch_bounds__(kBakeBlock) __attribute__((amdgpu_waves_per_eu(4, 4))) void bake_kernel(const BakeMap *__restrict__ maps, int count, uint64_t total, const float *__restrict__ lut,
                            const uint32_t *__restrict__ sobol, int samples) {
  const uint64_t block = (uint64_t)(blockIdx.x % kXcds) * (gridDim.x / kXcds) + blockIdx.x / kXcds;
  for (uint64_t g = block * blockDim.x + threadIdx.x; g < total; g += (uint64_t)gridDim.x * blockDim.x) {
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

// The Riemann irradiance of irradiance_baker.frag, one workgroup per output texel: lane t
// sums the zenith loop of azimuth steps t, t + 256, ... (the steps' cos and sin, stepped in
// float as GLSL steps them, come from host tables), then the lanes' sums are added in a
// fixed tree. The shader keeps one running sum; the sum here is regrouped by azimuth,
// which moves the result by float rounding only.
constexpr unsigned kRiemannBlock = 256;
__global__ __launch_bounds__(kRiemannBlock) void riemann_kernel(const BakeMap *__restrict__ maps, int count, const float2 *__restrict__ phi_tab,
                                                               int nphi, const float2 *__restrict__ theta_tab, int ntheta) {
  __shared__ float red[3][kRiemannBlock];
  const uint64_t g = blockIdx.x;
  const BakeMap m = maps[find_map(maps, count, g)];
  const uint64_t local = g - m.first;
  const int i = (int)(local % (unsigned)m.ow), j = (int)(local / (unsigned)m.ow);
  const f3 n = equirect_texel_direction(i, j, m.ow, m.oh);
  const f3 right = gl_normalize(cross(mk3(0.f, 1.f, 0.f), n));
  const f3 up = gl_normalize(cross(n, right));
  f3 acc = mk3(0.f, 0.f, 0.f);
  for (int k = threadIdx.x; k < nphi; k += kRiemannBlock) {
    const float2 ph = phi_tab[k];
    for (int t = 0; t < ntheta; t++) {
      const float2 th = theta_tab[t];  // (cos, sin)
      const float x = ph.x * th.y, y = th.y * ph.y, z = th.x;
      const f3 sv = (x * right + y * up) + z * n;
      acc = acc + (gl_lookup(m, gl_normalize(sv)) * th.x) * th.y;
    }
  }
  red[0][threadIdx.x] = acc.x;
  red[1][threadIdx.x] = acc.y;
  red[2][threadIdx.x] = acc.z;
  __syncthreads();
  for (unsigned w = kRiemannBlock / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w)
      for (int c = 0; c < 3; c++) red[c][threadIdx.x] += red[c][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const float PI = 3.14159265358f;
    const float inv = 1.f / (float)(nphi * ntheta);
    float *o = m.out + ((size_t)j * m.ow + i) * m.c;
    for (int c = 0; c < 3; c++) o[c] = (PI * red[c][0]) * inv;
    for (int c = 3; c < m.c; c++) o[c] = 0.f;
  }
}

// The zenith and azimuth steps of irradiance_baker.frag's loops, stepped in float from 0 by
// 0.0075 while <= the bound, with their cos and sin (double, rounded once).
static void riemann_steps(float bound, std::vector<float2> &tab) {
  tab.clear();
  for (float a = 0.f; a <= bound; a += 0.0075f) tab.push_back(make_float2((float)std::cos((double)a), (float)std::sin((double)a)));
}

// Cube-map captures, one texel per lane: six faces of ow x ow, face-major, rows bottom-up.
// A kernel of its own: its double-precision atan and acos would raise the register budget of
// the sampled bakes.
__global__ __launch_bounds__(kBakeBlock) void cubemap_kernel(const BakeMap *__restrict__ maps, int count, uint64_t total) {
  for (uint64_t g = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; g < total; g += (uint64_t)gridDim.x * blockDim.x) {
    const BakeMap m = maps[find_map(maps, count, g)];
    const uint64_t local = g - m.first;
    const unsigned S = (unsigned)m.ow, face = (unsigned)(local / ((uint64_t)S * S)), r = (unsigned)(local % ((uint64_t)S * S));
    const f3 col = gl_lookup(m, cube_texel_direction((int)face, (int)(r % S), (int)(r / S), (int)S));
    float *o = m.out + local * m.c;
    o[0] = col.x;
    o[1] = col.y;
    o[2] = col.z;
    for (int k = 3; k < m.c; k++) o[k] = 0.f;
  }
}

static int launch_batch(const nova_hip_bake_job *jobs, uint32_t count, int32_t samples, hipStream_t s) {
  if (count == 0) return NOVA_HIP_SUCCESS;
  if (!jobs || samples <= 0) return NOVA_HIP_INVALID_ARGUMENT;
  std::vector<BakeMap> maps, cube, riemann;  // sampled bakes; cube captures; Riemann jobs (a workgroup per texel)
  uint64_t total = 0, cube_total = 0, riemann_total = 0;
  for (uint32_t i = 0; i < count; i++) {
    const nova_hip_bake_job &jb = jobs[i];
    if (jb.kind == NOVA_HIP_BAKE_BRDF_LUT) {  // no environment map; two output channels
      if (!jb.out || jb.out_width < 1 || jb.out_height < 1 || jb.channels != 2) return NOVA_HIP_INVALID_ARGUMENT;
    } else {
      if (!jb.env || !jb.out || jb.width < 1 || jb.height < 1 || jb.channels < 3 || jb.channels > 4 || jb.out_width < 1 || jb.out_height < 1)
        return NOVA_HIP_INVALID_ARGUMENT;
      if (jb.channels == 4 && (reinterpret_cast<uintptr_t>(jb.env) & 15u)) return NOVA_HIP_INVALID_ARGUMENT;  // RGBA maps are read as float4
      if (jb.kind != NOVA_HIP_BAKE_IRRADIANCE && jb.kind != NOVA_HIP_BAKE_IRRADIANCE_SOBOL && jb.kind != NOVA_HIP_BAKE_PREFILTER &&
          jb.kind != NOVA_HIP_BAKE_CUBEMAP && jb.kind != NOVA_HIP_BAKE_IRRADIANCE_RIEMANN)
        return NOVA_HIP_INVALID_ARGUMENT;
      if (jb.kind == NOVA_HIP_BAKE_CUBEMAP && jb.out_width != jb.out_height) return NOVA_HIP_INVALID_ARGUMENT;  // square faces
    }
    // Sobol indices texel * samples + s must stay below 2^32
    if (jb.kind == NOVA_HIP_BAKE_IRRADIANCE_SOBOL && (uint64_t)jb.out_width * jb.out_height * (uint64_t)samples > 0xFFFFFFFFull)
      return NOVA_HIP_INVALID_ARGUMENT;
    const uint64_t texels = (uint64_t)jb.out_width * jb.out_height * (jb.kind == NOVA_HIP_BAKE_CUBEMAP ? 6u : 1u);
    const bool is_cube = jb.kind == NOVA_HIP_BAKE_CUBEMAP, is_riemann = jb.kind == NOVA_HIP_BAKE_IRRADIANCE_RIEMANN;
    uint64_t &first = is_riemann ? riemann_total : is_cube ? cube_total : total;
    (is_riemann ? riemann : is_cube ? cube : maps)
        .push_back(BakeMap{jb.env, jb.out, jb.width, jb.height, jb.channels, jb.out_width, jb.out_height, jb.kind, jb.roughness, first});
    first += texels;
  }
  if (riemann_total > 0x7FFFFFFFull) return NOVA_HIP_INVALID_ARGUMENT;  // one workgroup per texel
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
  // job tables and the Riemann step tables in one device buffer
  std::vector<float2> phi_tab, theta_tab;
  if (!riemann.empty()) {
    const float PI = 3.14159265358f, zenith_max = PI / 2, azimuth_max = 2 * PI;
    riemann_steps(azimuth_max, phi_tab);
    riemann_steps(zenith_max, theta_tab);
  }
  const size_t maps_bytes = maps.size() * sizeof(BakeMap), cube_bytes = cube.size() * sizeof(BakeMap), riemann_bytes = riemann.size() * sizeof(BakeMap);
  const size_t tab_bytes = (phi_tab.size() + theta_tab.size()) * sizeof(float2);
  char *buf = nullptr;
  if (hipMalloc(&buf, maps_bytes + cube_bytes + riemann_bytes + tab_bytes + 16) != hipSuccess) return NOVA_HIP_OUT_OF_MEMORY;
  BakeMap *d_maps = reinterpret_cast<BakeMap *>(buf), *d_cube = reinterpret_cast<BakeMap *>(buf + maps_bytes);
  BakeMap *d_riemann = reinterpret_cast<BakeMap *>(buf + maps_bytes + cube_bytes);
  float2 *d_phi = reinterpret_cast<float2 *>(buf + maps_bytes + cube_bytes + riemann_bytes), *d_theta = d_phi + phi_tab.size();
  hipError_t e = hipSuccess;
  if (maps_bytes) e = hipMemcpyAsync(d_maps, maps.data(), maps_bytes, hipMemcpyHostToDevice, s);
  if (e == hipSuccess && cube_bytes) e = hipMemcpyAsync(d_cube, cube.data(), cube_bytes, hipMemcpyHostToDevice, s);
  if (e == hipSuccess && riemann_bytes) e = hipMemcpyAsync(d_riemann, riemann.data(), riemann_bytes, hipMemcpyHostToDevice, s);
  if (e == hipSuccess && !phi_tab.empty()) e = hipMemcpyAsync(d_phi, phi_tab.data(), phi_tab.size() * sizeof(float2), hipMemcpyHostToDevice, s);
  if (e == hipSuccess && !theta_tab.empty()) e = hipMemcpyAsync(d_theta, theta_tab.data(), theta_tab.size() * sizeof(float2), hipMemcpyHostToDevice, s);
  if (e == hipSuccess && total > 0) {
    const unsigned block = kBakeBlock;
    const unsigned grid = (unsigned)std::min<uint64_t>(((total + block - 1) / block + kXcds - 1) / kXcds * kXcds, 1u << 20);  // a multiple of kXcds
    hipLaunchKernelGGL(bake_kernel, dim3(grid), dim3(block), 0, s, d_maps, (int)maps.size(), total, d_lut, sobol, samples);
    e = hipGetLastError();
  }
  if (e == hipSuccess && cube_total > 0) {
    const unsigned grid = (unsigned)std::min<uint64_t>((cube_total + kBakeBlock - 1) / kBakeBlock, 1u << 20);
    hipLaunchKernelGGL(cubemap_kernel, dim3(grid), dim3(kBakeBlock), 0, s, d_cube, (int)cube.size(), cube_total);
    e = hipGetLastError();
  }
  if (e == hipSuccess && riemann_total > 0) {
    hipLaunchKernelGGL(riemann_kernel, dim3((unsigned)riemann_total), dim3(kRiemannBlock), 0, s, d_riemann, (int)riemann.size(), d_phi,
                       (int)phi_tab.size(), d_theta, (int)theta_tab.size());
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  hipFree(buf);
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
