// Float cos / sin / asin / x^5 of the oracle (test infrastructure).
//
// The reference's CPU build calls the float overloads of libm (cosf, sinf, asinf, powf) on
// these paths (BxDF_math.h:129-152 and 145-149, Fresnel::schlick, DisneyDiffuse's
// (1-cos)^5, TextureProcessing.h:253-301). Their last-bit results depend on the libm the
// reference is linked with: glibc's cosf, for one, is not correctly rounded on every
// input. By default the oracle therefore evaluates them in double and rounds once, which
// is the correctly rounded value except in vanishing cases and is what the HIP path
// computes. set_float_libm(true) switches to the host libm's float functions, to measure
// how much the images depend on that choice.
#pragma once
#include <cmath>

namespace oracle {
extern bool g_float_libm;
inline void set_float_libm(bool on) { g_float_libm = on; }
inline float cos_f(float x) { return g_float_libm ? std::cos(x) : (float)std::cos((double)x); }
inline float sin_f(float x) { return g_float_libm ? std::sin(x) : (float)std::sin((double)x); }
inline float asin_f(float x) { return g_float_libm ? std::asin(x) : (float)std::asin((double)x); }
inline float pow5(float x) {
  if (g_float_libm) return powf(x, 5);
  const double d = (double)x, d2 = d * d;
  return (float)(d2 * d2 * d);
}
}  // namespace oracle
