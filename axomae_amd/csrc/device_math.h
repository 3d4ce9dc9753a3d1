// Device-side float math in the reference's (glm) evaluation order.
// Built with -ffp-contract=off so that no a*b+c is fused: every product and sum
// rounds exactly as the reference's x86-64 build does.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

// Helpers are host+device so that a host build of the same code can be checked on a CPU.
#define NOVA_HD __host__ __device__ __forceinline__

namespace nova_hip {

struct f3 {
  float x, y, z;
};
struct f2 {
  float x, y;
};

NOVA_HD f3 mk3(float x, float y, float z) { return f3{x, y, z}; }
NOVA_HD f3 operator+(f3 a, f3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
NOVA_HD f3 operator-(f3 a, f3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
NOVA_HD f3 operator*(f3 a, f3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
NOVA_HD f3 operator*(f3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
NOVA_HD f3 operator*(float s, f3 a) { return {s * a.x, s * a.y, s * a.z}; }
NOVA_HD f3 operator/(f3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
NOVA_HD f3 operator-(f3 a) { return {-a.x, -a.y, -a.z}; }
NOVA_HD f3 operator-(f3 a, float s) { return {a.x - s, a.y - s, a.z - s}; }

// glm::dot for vec3: (x + y) + z
NOVA_HD float dot(f3 a, f3 b) {
  float xx = a.x * b.x, yy = a.y * b.y, zz = a.z * b.z;
  return (xx + yy) + zz;
}
NOVA_HD f3 cross(f3 a, f3 b) { return {a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y}; }
NOVA_HD float length2(f3 v) { return dot(v, v); }
NOVA_HD float length(f3 v) { return sqrtf(dot(v, v)); }
NOVA_HD f3 normalize(f3 v) { return v * (1.f / sqrtf(dot(v, v))); }
NOVA_HD f3 reflect(f3 i, f3 n) { return i - n * dot(n, i) * 2.f; }
NOVA_HD float gmin(float a, float b) { return b < a ? b : a; }
NOVA_HD float gmax(float a, float b) { return a < b ? b : a; }
NOVA_HD f3 gmin(f3 a, f3 b) { return {gmin(a.x, b.x), gmin(a.y, b.y), gmin(a.z, b.z)}; }
NOVA_HD f3 gmax(f3 a, f3 b) { return {gmax(a.x, b.x), gmax(a.y, b.y), gmax(a.z, b.z)}; }
NOVA_HD float gsign(float x) { return (float)(0.f < x) - (float)(x < 0.f); }
NOVA_HD float sqr(float v) { return v * v; }
NOVA_HD float msqrt(float v) { return sqrtf(fmaxf(v, 0.f)); }  // math::sqrt<float>: NaN -> 0
NOVA_HD float smax(float a, float b) { return (a < b) ? b : a; }  // std::max
NOVA_HD float clampf(float v, float lo, float hi) { return (v < lo) ? lo : (hi < v) ? hi : v; }

// 3x3 frame stored as columns (T, B, N) like glm::mat3.
struct m3 {
  f3 c0, c1, c2;
};
NOVA_HD f3 mul(const m3 &m, f3 v) {
  return {(m.c0.x * v.x + m.c1.x * v.y) + m.c2.x * v.z, (m.c0.y * v.x + m.c1.y * v.y) + m.c2.y * v.z,
          (m.c0.z * v.x + m.c1.z * v.y) + m.c2.z * v.z};
}
NOVA_HD f3 mul_t(const m3 &m, f3 v) {  // transpose(m) * v
  return {(m.c0.x * v.x + m.c0.y * v.y) + m.c0.z * v.z, (m.c1.x * v.x + m.c1.y * v.y) + m.c1.z * v.z,
          (m.c2.x * v.x + m.c2.y * v.y) + m.c2.z * v.z};
}

// glm mat4 * vec4: (m0*x + m1*y) + (m2*z + m3*w); M column-major (M[col*4+row]).
NOVA_HD float mat4_row(const float *M, int r, float x, float y, float z, float w) {
  float a = M[0 + r] * x, b = M[4 + r] * y, c = M[8 + r] * z, d = M[12 + r] * w;
  return (a + b) + (c + d);
}

// ------------------------------------------------------------ Sobol / Owen
// SobolGenerator (math_random.h:29-97): gray-code XOR of 32 direction numbers,
// reverse bits, Laine-Karras hash with seed 0xDEADBEEF, reverse bits, times 2^-32.
NOVA_HD uint32_t lk_hash(uint32_t x, uint32_t seed) {
  x ^= x * 0x3d20adeau;
  x += seed;
  x *= (seed >> 16) | 0x1u;
  x ^= x * 0x05526c56u;
  x ^= x * 0x53a22864u;
  return x;
}
NOVA_HD uint32_t owen_scramble(uint32_t x) {
  return __builtin_bitreverse32(lk_hash(__builtin_bitreverse32(x), 0xDEADBEEFu));
}
NOVA_HD float u32_to_unit(uint32_t v) { return (float)v * (1.f / 4294967296.f); }

// Float cos/sin/pow/asin evaluated in double and rounded once: a correctly rounded result
// in all but vanishingly rare cases, which is what the reference's x86-64 libm float
// functions return, so the device agrees with the reference bit for bit almost always.
NOVA_HD float cos_f(float x) { return (float)cos((double)x); }
NOVA_HD float sin_f(float x) { return (float)sin((double)x); }
NOVA_HD float asin_f(float x) { return (float)asin((double)x); }
NOVA_HD float pow5(float x) {
  double d = (double)x, d2 = d * d;
  return (float)(d2 * d2 * d);
}

// Direction row `dv` is wave-uniform in the integrator (one dimension per launch stage),
// so the 32 loads become scalar loads and the loop is 32 conditional XORs.
NOVA_HD uint32_t sobol_xor(const uint32_t *__restrict__ dv, uint32_t gray) {
  uint32_t acc = 0;
#pragma unroll
  for (int i = 0; i < 32; i++) acc ^= ((gray >> i) & 1u) ? dv[i] : 0u;
  return acc;
}

// ------------------------------------------------------- approximate trig
NOVA_HD float horner_approx(float x) {
  const float a1 = 0.99997726f, a3 = -0.33262347f, a5 = 0.19354346f;
  const float a7 = -0.11643287f, a9 = 0.05265332f, a11 = -0.01172120f;
  float xs = x * x;
  return x * (a1 + xs * (a3 + xs * (a5 + xs * (a7 + xs * (a9 + xs * a11)))));
}
// math::atan2_approx (math_utils_approx.h:24-40)
NOVA_HD float atan2_approx(float y, float x) {
  const float pi = 3.14159265358979323846f, pi_2 = 1.57079632679489661923f;
  bool swap = fabsf(x) < fabsf(y);
  float in = (swap ? x : y) / (swap ? y : x);
  float res = horner_approx(in);
  res = swap ? (in >= 0.0f ? pi_2 : -pi_2) - res : res;
  if (x >= 0.0f && y >= 0.0f) {
  } else if (x < 0.0f && y >= 0.0f) {
    res = pi + res;
  } else if (x < 0.0f && y < 0.0f) {
    res = -pi + res;
  }
  return res;
}
constexpr int kAcosLut = 500000;
// math::spherical::acos_lt (math_acos_lt.cpp:36-39)
NOVA_HD float acos_lt(const float *__restrict__ lut, float a) {
  const float f = (a + 1.f) * (float)kAcosLut * 0.5f;
  int s = (f == f) ? (int)fminf(fmaxf(f, -1.f), 2147483520.f) : 0;  // NaN -> 0  // NaN -> 0 as x86's clamp of INT_MIN
  s = s < 0 ? 0 : (s > kAcosLut - 1 ? kAcosLut - 1 : s);
  return lut[s];
}

NOVA_HD float wrap_uv(float uv) {  // NovaTextureInterface.h:14-25
  if (uv < 0.f) return 1.f - (ceilf(uv) - uv);
  if (uv > 1.f) return uv - floorf(uv);
  return uv;
}

}  // namespace nova_hip
