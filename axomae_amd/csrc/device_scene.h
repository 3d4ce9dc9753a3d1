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
  f3 e1 = w[1] - w[0], e2 = w[2] - w[0];
  if (degenerate) e1 = e2 = mk3(0.f, 0.f, 0.f);
  union {
    int32_t i;
    float f;
  } pid;
  pid.i = (int32_t)p;
  TriMT t;
  t.a = make_float4(w[0].x, w[0].y, w[0].z, pid.f);
  t.b = make_float4(e1.x, e1.y, e1.z, 0.f);
  t.c = make_float4(e2.x, e2.y, e2.z, 0.f);
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

}  // namespace nova_hip
