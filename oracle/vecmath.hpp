// Oracle (test infrastructure only): glm-order float vector math.
//
// The reference evaluates all hot-path arithmetic through glm. This header fixes
// the evaluation order of every glm operation the path uses so that the oracle,
// compiled with -O2 -ffp-contract=off, rounds exactly like a glm build on x86-64:
//   mat4*vec4 = (m0*x + m1*y) + (m2*z + m3*w)       (glm func_matrix / type_mat4x4.inl)
//   mat3*vec3 = (m0*x + m1*y) + m2*z
//   dot3      = (x*x' + y*y') + z*z'
//   cross     = (a.y*b.z - b.y*a.z, a.z*b.x - b.z*a.x, a.x*b.y - b.x*a.y)
//   normalize = v * (1 / sqrt(dot(v, v)))
//   reflect   = I - N * dot(N, I) * 2
//   min(a, b) = b < a ? b : a ;  max(a, b) = a < b ? b : a
// Only tests/, the smoke check and bench.py's CPU-baseline leg use the oracle.
#pragma once
#include <cmath>
#include <cstdint>

namespace oracle {

struct vec2 {
  float x = 0, y = 0;
  vec2() = default;
  vec2(float a, float b) : x(a), y(b) {}
};

struct vec3 {
  float x = 0, y = 0, z = 0;
  vec3() = default;
  explicit vec3(float s) : x(s), y(s), z(s) {}
  vec3(float a, float b, float c) : x(a), y(b), z(c) {}
  float &operator[](int i) { return (&x)[i]; }
  float operator[](int i) const { return (&x)[i]; }
};

struct vec4 {
  float x = 0, y = 0, z = 0, w = 0;
  vec4() = default;
  explicit vec4(float s) : x(s), y(s), z(s), w(s) {}
  vec4(float a, float b, float c, float d) : x(a), y(b), z(c), w(d) {}
  vec4(const vec3 &v, float d) : x(v.x), y(v.y), z(v.z), w(d) {}
  vec3 xyz() const { return {x, y, z}; }
  float &operator[](int i) { return (&x)[i]; }
  float operator[](int i) const { return (&x)[i]; }
};

inline vec3 operator+(const vec3 &a, const vec3 &b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline vec3 operator-(const vec3 &a, const vec3 &b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline vec3 operator*(const vec3 &a, const vec3 &b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
inline vec3 operator*(const vec3 &a, float s) { return {a.x * s, a.y * s, a.z * s}; }
inline vec3 operator*(float s, const vec3 &a) { return {s * a.x, s * a.y, s * a.z}; }
inline vec3 operator/(const vec3 &a, float s) { return {a.x / s, a.y / s, a.z / s}; }
inline vec3 operator-(const vec3 &a) { return {-a.x, -a.y, -a.z}; }
inline vec3 operator+(const vec3 &a, float s) { return {a.x + s, a.y + s, a.z + s}; }
inline vec3 operator-(const vec3 &a, float s) { return {a.x - s, a.y - s, a.z - s}; }

inline vec4 operator+(const vec4 &a, const vec4 &b) { return {a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w}; }
inline vec4 operator*(const vec4 &a, const vec4 &b) { return {a.x * b.x, a.y * b.y, a.z * b.z, a.w * b.w}; }
inline vec4 operator*(float s, const vec4 &a) { return {s * a.x, s * a.y, s * a.z, s * a.w}; }
inline vec4 operator*(const vec4 &a, float s) { return {a.x * s, a.y * s, a.z * s, a.w * s}; }
inline vec4 operator/(const vec4 &a, float s) { return {a.x / s, a.y / s, a.z / s, a.w / s}; }

inline float dot(const vec3 &a, const vec3 &b) {
  float xx = a.x * b.x, yy = a.y * b.y, zz = a.z * b.z;
  return (xx + yy) + zz;
}
inline vec3 cross(const vec3 &a, const vec3 &b) {
  return {a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y};
}
inline float length2(const vec3 &v) { return dot(v, v); }
inline float length(const vec3 &v) { return std::sqrt(dot(v, v)); }
inline vec3 normalize(const vec3 &v) { return v * (1.f / std::sqrt(dot(v, v))); }
inline vec3 reflect(const vec3 &i, const vec3 &n) { return i - n * dot(n, i) * 2.f; }
inline float gmin(float a, float b) { return b < a ? b : a; }
inline float gmax(float a, float b) { return a < b ? b : a; }
inline vec3 gmin(const vec3 &a, const vec3 &b) { return {gmin(a.x, b.x), gmin(a.y, b.y), gmin(a.z, b.z)}; }
inline vec3 gmax(const vec3 &a, const vec3 &b) { return {gmax(a.x, b.x), gmax(a.y, b.y), gmax(a.z, b.z)}; }
inline float gsign(float x) { return (float)(0.f < x) - (float)(x < 0.f); }
inline vec3 vabs(const vec3 &v) { return {std::fabs(v.x), std::fabs(v.y), std::fabs(v.z)}; }

// Column-major 3x3 (c[i] is column i), as glm::mat3.
struct mat3 {
  vec3 c[3];
  mat3() { c[0] = {1, 0, 0}; c[1] = {0, 1, 0}; c[2] = {0, 0, 1}; }
  mat3(const vec3 &a, const vec3 &b, const vec3 &d) { c[0] = a; c[1] = b; c[2] = d; }
};
inline vec3 operator*(const mat3 &m, const vec3 &v) {
  vec3 r;
  for (int k = 0; k < 3; k++) r[k] = (m.c[0][k] * v.x + m.c[1][k] * v.y) + m.c[2][k] * v.z;
  return r;
}
// transpose(m) * v : row k of the result is dot(column k, v) in x,y,z order.
inline vec3 mul_transpose(const mat3 &m, const vec3 &v) {
  vec3 r;
  for (int k = 0; k < 3; k++) r[k] = (m.c[k].x * v.x + m.c[k].y * v.y) + m.c[k].z * v.z;
  return r;
}

// Column-major 4x4 (m[col*4 + row]), as glm::mat4 / value_ptr.
struct mat4 {
  float m[16];
  mat4() {
    for (int i = 0; i < 16; i++) m[i] = (i % 5 == 0) ? 1.f : 0.f;
  }
  float at(int col, int row) const { return m[col * 4 + row]; }
};
inline vec4 operator*(const mat4 &M, const vec4 &v) {
  vec4 r;
  for (int k = 0; k < 4; k++) {
    float mul0 = M.at(0, k) * v.x, mul1 = M.at(1, k) * v.y;
    float mul2 = M.at(2, k) * v.z, mul3 = M.at(3, k) * v.w;
    float add0 = mul0 + mul1, add1 = mul2 + mul3;
    r[k] = add0 + add1;
  }
  return r;
}
// glm::mat3(mat4): upper-left 3x3.
inline mat3 upper3(const mat4 &M) {
  return mat3({M.at(0, 0), M.at(0, 1), M.at(0, 2)}, {M.at(1, 0), M.at(1, 1), M.at(1, 2)}, {M.at(2, 0), M.at(2, 1), M.at(2, 2)});
}

inline bool vec_nonfinite(const vec4 &v) {
  return !std::isfinite(v.x) || !std::isfinite(v.y) || !std::isfinite(v.z) || !std::isfinite(v.w);
}

}  // namespace oracle
