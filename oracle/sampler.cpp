// Oracle (test infrastructure only): Owen-scrambled Sobol generator and the
// approximate trigonometry used by the envmap lookup.
#include "oracle.hpp"
#include "fmath.hpp"
#include <algorithm>
#include <cmath>

namespace oracle {
bool g_float_libm = false;

const uint32_t kSobolTable[256][32] = {
#include "../axomae_amd/csrc/sobol_table.inc"
};

// math_random.h:48-54
static uint32_t reverse_bits(uint32_t x) {
  x = ((x & 0xaaaaaaaau) >> 1) | ((x & 0x55555555u) << 1);
  x = ((x & 0xccccccccu) >> 2) | ((x & 0x33333333u) << 2);
  x = ((x & 0xf0f0f0f0u) >> 4) | ((x & 0x0f0f0f0fu) << 4);
  x = ((x & 0xff00ff00u) >> 8) | ((x & 0x00ff00ffu) << 8);
  return (x >> 16) | (x << 16);
}

// Laine-Karras style hash, math_random.h:56-63 (32-bit wraparound)
static uint32_t lk_hash(uint32_t x, uint32_t seed) {
  x ^= x * 0x3d20adeau;
  x += seed;
  x *= (seed >> 16) | 0x1u;
  x ^= x * 0x05526c56u;
  x ^= x * 0x53a22864u;
  return x;
}

// math_random.h:77-89: gray-code XOR of the direction numbers, then Owen scramble.
uint32_t sobol_sample_bits(uint32_t index, unsigned dim, uint32_t seed) {
  uint32_t accum = 0;
  uint32_t gc = index ^ (index >> 1);
  for (unsigned i = 0; i < 32; i++)
    if ((gc >> i) & 1u) accum ^= kSobolTable[dim][i];
  return reverse_bits(lk_hash(reverse_bits(accum), seed));
}

// ONE_UINT_MAX = 1.f / float(0xFFFFFFFF) == 2^-32 exactly (float(0xFFFFFFFF) rounds to 2^32).
float sobol_sample(uint32_t index, unsigned dim, uint32_t seed) {
  const float one_uint_max = 1.f / float(0xFFFFFFFFu);
  return float(sobol_sample_bits(index, dim, seed)) * one_uint_max;
}

// scripts/generate_luts_trigonometrics.py:11-20: (float)acos((i/P)*2 - 1) in double.
const std::vector<float> &acos_lut() {
  static const std::vector<float> lut = [] {
    std::vector<float> t(kAcosLutPrecision);
    for (int i = 0; i < kAcosLutPrecision; i++) {
      double scaled = ((double)i / (double)kAcosLutPrecision) * 2.0 - 1.0;
      t[i] = (float)std::acos(scaled);
    }
    return t;
  }();
  return lut;
}

// math_acos_lt.cpp:36-39
float acos_lt(float a) {
  int scaled = std::clamp((int)((a + 1.f) * (float)kAcosLutPrecision * 0.5f), 0, kAcosLutPrecision - 1);
  return acos_lut()[scaled];
}

// math_utils_approx.h:14-40
static float horner_approx(float x) {
  const float a1 = 0.99997726f, a3 = -0.33262347f, a5 = 0.19354346f;
  const float a7 = -0.11643287f, a9 = 0.05265332f, a11 = -0.01172120f;
  float x_sq = x * x;
  return x * (a1 + x_sq * (a3 + x_sq * (a5 + x_sq * (a7 + x_sq * (a9 + x_sq * a11)))));
}

float atan2_approx(float y, float x) {
  const float pi = (float)M_PI, pi_2 = (float)M_PI_2;
  bool swap = std::fabs(x) < std::fabs(y);
  float atan_input = (swap ? x : y) / (swap ? y : x);
  float res = horner_approx(atan_input);
  res = swap ? (atan_input >= 0.0f ? pi_2 : -pi_2) - res : res;
  if (x >= 0.0f && y >= 0.0f) {
  } else if (x < 0.0f && y >= 0.0f) {
    res = pi + res;
  } else if (x < 0.0f && y < 0.0f) {
    res = -pi + res;
  }
  return res;
}

}  // namespace oracle
