// Per-triangle flattening as a host+device function: world-space AABB and centroid as
// the reference builder's cache computes them (Triangle::computeAABB / centroid,
// Triangle.h:108-125; Object3D.h:194-199), the world-space MT record and the shading record.
#pragma once
#include "device_math.h"
#include "nova_internal.h"

namespace nova_hip {

struct TriRef {
  uint32_t i0, i1, i2;  // global vertex indices
  uint32_t mesh;
};

NOVA_HD void flatten_triangle(uint32_t p, const TriRef &r, const float *pos, const float *nrm, const float *tan, const float *bit, const float *uv,
                              const float *xforms, const int32_t *mesh_mat, TriMT *tri, ShadeTri *shade, float *aabb, float *centroid) {
  const float *M = xforms + (size_t)r.mesh * 16;
  uint32_t id[3] = {r.i0, r.i1, r.i2};
  f3 v[3], w[3];
  for (int k = 0; k < 3; k++) {
    v[k] = mk3(pos[id[k] * 3], pos[id[k] * 3 + 1], pos[id[k] * 3 + 2]);
    w[k] = mk3(mat4_row(M, 0, v[k].x, v[k].y, v[k].z, 1.f), mat4_row(M, 1, v[k].x, v[k].y, v[k].z, 1.f),
               mat4_row(M, 2, v[k].x, v[k].y, v[k].z, 1.f));
  }
  const f3 mn = gmin(w[0], gmin(w[1], w[2]));
  const f3 mx = gmax(w[0], gmax(w[1], w[2]));
  aabb[(size_t)p * 6 + 0] = mn.x;
  aabb[(size_t)p * 6 + 1] = mn.y;
  aabb[(size_t)p * 6 + 2] = mn.z;
  aabb[(size_t)p * 6 + 3] = mx.x;
  aabb[(size_t)p * 6 + 4] = mx.y;
  aabb[(size_t)p * 6 + 5] = mx.z;
  const f3 c = mk3((v[0].x + v[1].x + v[2].x) * 0.333333f, (v[0].y + v[1].y + v[2].y) * 0.333333f, (v[0].z + v[1].z + v[2].z) * 0.333333f);
  centroid[(size_t)p * 3 + 0] = mat4_row(M, 0, c.x, c.y, c.z, 1.f);
  centroid[(size_t)p * 3 + 1] = mat4_row(M, 1, c.x, c.y, c.z, 1.f);
  centroid[(size_t)p * 3 + 2] = mat4_row(M, 2, c.x, c.y, c.z, 1.f);

  const bool degenerate = r.i0 == r.i1 || r.i2 == r.i1 || r.i2 == r.i0;
  // identity transform: the world record; otherwise the object-space record the reference's
  // Triangle::hit tests after taking the ray to object space
  bool identity = true;
  for (int k = 0; k < 16; k++) identity = identity && M[k] == ((k % 5 == 0) ? 1.f : 0.f);
  const f3 *c0 = identity ? w : v;
  f3 e1 = c0[1] - c0[0], e2 = c0[2] - c0[0];
  if (degenerate) e1 = e2 = mk3(0.f, 0.f, 0.f);
  TriMT t;
  t.a = make_float4(c0[0].x, c0[0].y, c0[0].z, __builtin_bit_cast(float, (int32_t)p));
  t.b = make_float4(e1.x, e1.y, e1.z, 0.f);
  t.c = make_float4(e2.x, e2.y, e2.z, __builtin_bit_cast(float, identity ? 0 : (int32_t)r.mesh + 1));
  tri[p] = t;

  ShadeTri s;
  float *pp[3] = {s.p0, s.p1, s.p2}, *pn[3] = {s.n0, s.n1, s.n2}, *pt[3] = {s.t0, s.t1, s.t2}, *pb[3] = {s.b0, s.b1, s.b2};
  float *pu[3] = {s.uv0, s.uv1, s.uv2};
  for (int k = 0; k < 3; k++) {
    for (int q = 0; q < 3; q++) {
      pp[k][q] = pos[id[k] * 3 + q];
      pn[k][q] = nrm[id[k] * 3 + q];
      pt[k][q] = tan[id[k] * 3 + q];
      pb[k][q] = bit[id[k] * 3 + q];
    }
    pu[k][0] = uv[id[k] * 2];
    pu[k][1] = uv[id[k] * 2 + 1];
  }
  s.material = mesh_mat[r.mesh];
  s.mesh = (int32_t)r.mesh;
  s.pad[0] = degenerate ? 1 : 0;
  s.pad[1] = 0;
  shade[p] = s;
}

// Host side of the upload: rows 0-2 of the inverse of a model matrix (column-major in),
// as three float4 (row r: the r-th components of the inverse's four columns). The inverse
// is the adjugate over the determinant in double, rounded once to float.
inline void mesh_inverse_rows(const float *M, float4 rows[3]) {
  double a[4][4], inv[4][4];
  for (int c = 0; c < 4; c++)
    for (int r = 0; r < 4; r++) a[r][c] = M[c * 4 + r];
  auto minor3 = [&](int skip_r, int skip_c) {
    double m[3][3];
    for (int r = 0, rr = 0; r < 4; r++) {
      if (r == skip_r) continue;
      for (int c = 0, cc = 0; c < 4; c++) {
        if (c == skip_c) continue;
        m[rr][cc++] = a[r][c];
      }
      rr++;
    }
    return m[0][0] * (m[1][1] * m[2][2] - m[1][2] * m[2][1]) - m[0][1] * (m[1][0] * m[2][2] - m[1][2] * m[2][0]) +
           m[0][2] * (m[1][0] * m[2][1] - m[1][1] * m[2][0]);
  };
  double det = 0.0;
  for (int c = 0; c < 4; c++) det += ((c & 1) ? -1.0 : 1.0) * a[0][c] * minor3(0, c);
  for (int r = 0; r < 4; r++)
    for (int c = 0; c < 4; c++) inv[r][c] = (((r + c) & 1) ? -1.0 : 1.0) * minor3(c, r) / det;
  for (int r = 0; r < 3; r++) rows[r] = make_float4((float)inv[r][0], (float)inv[r][1], (float)inv[r][2], (float)inv[r][3]);
}

// Host side of the upload: mat3(transform.m) and the normal matrix
// transpose(inverse(mat3(m))) (utils_3D.h:28 compute_normal_mat), column-major. The inverse
// is taken in double and rounded once.
inline void mesh_xf_blocks(const float *M, MeshXf &out) {
  double a[3][3];
  for (int c = 0; c < 3; c++)
    for (int r = 0; r < 3; r++) {
      a[r][c] = M[c * 4 + r];
      out.m3[c * 3 + r] = M[c * 4 + r];
    }
  double det = a[0][0] * (a[1][1] * a[2][2] - a[1][2] * a[2][1]) - a[0][1] * (a[1][0] * a[2][2] - a[1][2] * a[2][0]) +
               a[0][2] * (a[1][0] * a[2][1] - a[1][1] * a[2][0]);
  double inv[3][3];
  inv[0][0] = (a[1][1] * a[2][2] - a[1][2] * a[2][1]) / det;
  inv[0][1] = (a[0][2] * a[2][1] - a[0][1] * a[2][2]) / det;
  inv[0][2] = (a[0][1] * a[1][2] - a[0][2] * a[1][1]) / det;
  inv[1][0] = (a[1][2] * a[2][0] - a[1][0] * a[2][2]) / det;
  inv[1][1] = (a[0][0] * a[2][2] - a[0][2] * a[2][0]) / det;
  inv[1][2] = (a[0][2] * a[1][0] - a[0][0] * a[1][2]) / det;
  inv[2][0] = (a[1][0] * a[2][1] - a[1][1] * a[2][0]) / det;
  inv[2][1] = (a[0][1] * a[2][0] - a[0][0] * a[2][1]) / det;
  inv[2][2] = (a[0][0] * a[1][1] - a[0][1] * a[1][0]) / det;
  // n = transpose(inverse(m3)): element (row r, col c) = inv[c][r]; stored column-major.
  for (int c = 0; c < 3; c++)
    for (int r = 0; r < 3; r++) out.n3[c * 3 + r] = (float)inv[c][r];
}

}  // namespace nova_hip
