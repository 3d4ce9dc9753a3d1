// Oracle (test infrastructure only): scene flattening, the reference's binary
// SAH builder, its stack traversal and Moller-Trumbore triangle test.
#include "oracle.hpp"
#include <algorithm>
#include <cfloat>
#include <climits>
#include <stdexcept>
#include <utility>

namespace oracle {

static mat4 invert4(const mat4 &M) {
  // Gauss-Jordan in double. Only the identity is used on parity paths, where this is exact.
  double a[4][8];
  for (int r = 0; r < 4; r++)
    for (int c = 0; c < 4; c++) {
      a[r][c] = M.at(c, r);
      a[r][c + 4] = (r == c) ? 1.0 : 0.0;
    }
  for (int c = 0; c < 4; c++) {
    int piv = c;
    for (int r = c + 1; r < 4; r++)
      if (std::fabs(a[r][c]) > std::fabs(a[piv][c])) piv = r;
    for (int k = 0; k < 8; k++) std::swap(a[c][k], a[piv][k]);
    double d = a[c][c];
    for (int k = 0; k < 8; k++) a[c][k] /= d;
    for (int r = 0; r < 4; r++)
      if (r != c) {
        double f = a[r][c];
        for (int k = 0; k < 8; k++) a[r][k] -= f * a[c][k];
      }
  }
  mat4 out;
  for (int r = 0; r < 4; r++)
    for (int c = 0; c < 4; c++) out.m[c * 4 + r] = (float)a[r][c + 4];
  return out;
}

// Global primitive order: mesh-major, triangle-minor (engine.cpp:159-175).
void Scene::finalize() {
  prim_mesh.clear();
  prim_tri.clear();
  inv.clear();
  normal_mat.clear();
  for (size_t m = 0; m < meshes.size(); m++) {
    for (size_t t = 0; t < meshes[m].indices.size(); t += 3) {
      prim_mesh.push_back((uint32_t)m);
      prim_tri.push_back((uint32_t)t);
    }
    mat4 iv = invert4(meshes[m].transform);
    inv.push_back(iv);
    // n = transpose(inverse(mat3(m))) stored as a mat4 holding the 3x3 block.
    mat4 nm;
    for (int c = 0; c < 3; c++)
      for (int r = 0; r < 3; r++) nm.m[c * 4 + r] = iv.at(r, c);
    normal_mat.push_back(nm);
  }
}

static inline vec3 vtx(const Mesh &m, uint32_t vi) { return {m.vertices[vi * 3], m.vertices[vi * 3 + 1], m.vertices[vi * 3 + 2]}; }

struct Box {  // geometry::BoundingBox (BoundingBox.h:24-25 defaults)
  vec3 mn{1e32f}, mx{-1e32f};
  Box() = default;
  Box(const vec3 &a, const vec3 &b) : mn(a), mx(b) {}
  Box operator+(const Box &o) const { return {gmin(mn, o.mn), gmax(mx, o.mx)}; }
  float half_area() const {  // BoundingBox.cpp:162-165
    vec3 l = mx - mn;
    return (l.x * l.y + l.y * l.z) + l.z * l.x;
  }
};

// Triangle::computeAABB / centroid (Triangle.h:108-125, Object3D.h:194-199)
static void prim_bounds(const Scene &s, int p, Box &box, vec3 &centroid) {
  const Mesh &m = s.meshes[s.prim_mesh[p]];
  uint32_t t = s.prim_tri[p];
  vec3 v0 = vtx(m, m.indices[t]), v1 = vtx(m, m.indices[t + 1]), v2 = vtx(m, m.indices[t + 2]);
  vec3 w0 = (m.transform * vec4(v0, 1.f)).xyz(), w1 = (m.transform * vec4(v1, 1.f)).xyz(), w2 = (m.transform * vec4(v2, 1.f)).xyz();
  box = Box(gmin(w0, gmin(w1, w2)), gmax(w0, gmax(w1, w2)));
  vec3 c((v0.x + v1.x + v2.x) * 0.333333f, (v0.y + v1.y + v2.y) * 0.333333f, (v0.z + v1.z + v2.z) * 0.333333f);
  centroid = (m.transform * vec4(c, 1.f)).xyz();
}

namespace {
struct Builder {
  const std::vector<Box> &aabbs;
  const std::vector<vec3> &centroids;
  Bvh &bvh;

  // update_aabb (BvhBuilder.cpp:77-101): seeded with float(INT_MAX) / float(INT_MIN).
  void update_aabb(int32_t id) {
    BvhNode &node = bvh.nodes[id];
    Box nb(vec3((float)INT_MAX), vec3((float)INT_MIN));
    for (int32_t i = 0; i < node.primitive_count; i++) {
      nb = nb + aabbs[bvh.prim_idx[node.left + i]];
      for (int k = 0; k < 3; k++) {
        node.min[k] = nb.mn[k];
        node.max[k] = nb.mx[k];
      }
    }
  }

  // eval_sah (BvhBuilder.cpp:117-140)
  float eval_sah(const BvhNode &node, int axis, float pos) const {
    Box l, r;
    int nl = 0, nr = 0;
    for (int32_t i = 0; i < node.primitive_count; i++) {
      int32_t p = bvh.prim_idx[node.left + i];
      if (centroids[p][axis] < pos) {
        nl++;
        l = l + aabbs[p];
      } else {
        nr++;
        r = r + aabbs[p];
      }
    }
    const float cost = l.half_area() * (float)nl + r.half_area() * (float)nr;
    return cost > 0.f ? cost : (float)INT_MAX;
  }

  // axis_subdiv_sah with PERFORMANCE = 4 segments (BvhBuilder.cpp:154-184)
  int axis_subdiv_sah(int32_t id, float &best_cost_r, float &best_position_r) const {
    const unsigned SEG = 4;
    int best_axis = 0;
    float best_position = 0.f;
    float best_cost = (float)INT_MAX;
    const BvhNode &node = bvh.nodes[id];
    for (int axis = 0; axis < 3; axis++) {
      const float bmin = node.min[axis], bmax = node.max[axis];
      float dist = bmax - bmin;
      if (dist == 0) continue;
      float seg = dist / (float)SEG;
      for (int i = 0; i < (int)SEG; i++) {
        float cand = seg * (float)i + bmin;
        const float cost = eval_sah(node, axis, cand);
        if (cost < best_cost) {
          best_cost = cost;
          best_axis = axis;
          best_position = cand;
        }
      }
    }
    best_cost_r = best_cost;
    best_position_r = best_position;
    return best_axis;
  }

  // create_nodes (BvhBuilder.cpp:186-203): in-place swap partition.
  int32_t partition(const BvhNode &node, int axis, float pos) {
    int32_t i = node.left, j = i + node.primitive_count - 1;
    while (i <= j) {
      if (centroids[bvh.prim_idx[i]][axis] < pos)
        i++;
      else
        std::swap(bvh.prim_idx[i], bvh.prim_idx[j--]);
    }
    return i;
  }

  // subdivide (BvhBuilder.cpp:205-253), iterative with the same DFS allocation order.
  void run() {
    std::vector<int32_t> stack{0};
    while (!stack.empty()) {
      int32_t id = stack.back();
      stack.pop_back();
      float best_cost = 0, best_pos = 0;
      int axis = axis_subdiv_sah(id, best_cost, best_pos);
      BvhNode &node = bvh.nodes[id];
      Box nb(vec3(node.min[0], node.min[1], node.min[2]), vec3(node.max[0], node.max[1], node.max[2]));
      const float cost = nb.half_area() * (float)node.primitive_count;
      if (best_cost >= cost) continue;
      int32_t i = partition(node, axis, best_pos);
      int32_t left_count = i - node.left;
      if (left_count == 0 || left_count == node.primitive_count) continue;
      int32_t li = bvh.nodes_used++, ri = bvh.nodes_used++;
      BvhNode &l = bvh.nodes[li];
      BvhNode &r = bvh.nodes[ri];
      BvhNode &nd = bvh.nodes[id];
      l.left = nd.left;
      l.primitive_count = left_count;
      r.left = i;
      r.primitive_count = nd.primitive_count - left_count;
      nd.primitive_count = 0;
      nd.left = li;
      update_aabb(li);
      update_aabb(ri);
      stack.push_back(ri);
      stack.push_back(li);
    }
  }
};
}  // namespace

void bvh_build(const Scene &s, Bvh &bvh) {
  const size_t n = s.prim_mesh.size();
  if (n == 0) throw std::runtime_error("Primitives list is empty.");
  bvh.nodes.assign(2 * n - 1, BvhNode{{0, 0, 0}, {0, 0, 0}, 0, 0});
  bvh.prim_idx.resize(n);
  for (size_t i = 0; i < n; i++) bvh.prim_idx[i] = (int32_t)i;
  std::vector<Box> aabbs(n);
  std::vector<vec3> centroids(n);
  for (size_t p = 0; p < n; p++) prim_bounds(s, (int)p, aabbs[p], centroids[p]);
  bvh.nodes[0].left = 0;
  bvh.nodes[0].primitive_count = (int32_t)n;
  Builder b{aabbs, centroids, bvh};
  b.update_aabb(0);
  bvh.nodes_used = 1;
  b.run();
}

// BoundingBox::intersect(dir, origin, dist_min) (BoundingBox.cpp:150-160) with one
// deviation: a slab whose bounds give 0/0 (the ray lies in the box face's plane with a zero
// direction component) is ignored instead of poisoning the test with NaN. The reference's
// std::min/std::max chain drops such boxes and misses hits on the face; its default CPU
// path (Embree) finds them, and the closest hit is defined here as the nearest triangle
// over the whole scene, so the oracle keeps those boxes.
static void slab(float lo_plane, float hi_plane, float o, float d, float &t1, float &t2) {
  t1 = (lo_plane - o) / d;
  t2 = (hi_plane - o) / d;
  if (std::isnan(t1) || std::isnan(t2)) {  // 0/0: the ray runs inside a face plane of this slab
    t1 = -INFINITY;
    t2 = INFINITY;
  }
}

static float box_intersect(const BvhNode &nd, const vec3 &d, const vec3 &o, float dist_min) {
  float tx1, tx2, ty1, ty2, tz1, tz2;
  slab(nd.min[0], nd.max[0], o.x, d.x, tx1, tx2);
  slab(nd.min[1], nd.max[1], o.y, d.y, ty1, ty2);
  slab(nd.min[2], nd.max[2], o.z, d.z, tz1, tz2);
  float tmin = std::min(tx1, tx2), tmax = std::max(tx1, tx2);
  tmin = std::max(tmin, std::min(ty1, ty2)), tmax = std::min(tmax, std::max(ty1, ty2));
  tmin = std::max(tmin, std::min(tz1, tz2)), tmax = std::min(tmax, std::max(tz1, tz2));
  if (tmax >= tmin && tmax > 0 && tmin <= dist_min) return tmin;
  return FLT_MAX;
}

// Triangle::hit (Triangle.h:146-175): ray taken to object space by transform.inv.
static bool tri_hit(const Scene &s, int p, const vec3 &wo, const vec3 &wd, float tmin, float tmax, float &t_out, float &u_out, float &v_out) {
  const uint32_t mi = s.prim_mesh[p];
  const Mesh &m = s.meshes[mi];
  const uint32_t t = s.prim_tri[p];
  const uint32_t i0 = m.indices[t], i1 = m.indices[t + 1], i2 = m.indices[t + 2];
  if (i0 == i1 || i2 == i1 || i2 == i0) return false;
  const mat4 &inv = s.inv[mi];
  const vec3 o = (inv * vec4(wo, 1.f)).xyz();
  const vec3 d = (inv * vec4(wd, 0.f)).xyz();
  const vec3 v0 = vtx(m, i0), v1 = vtx(m, i1), v2 = vtx(m, i2);
  const vec3 e1 = v1 - v0, e2 = v2 - v0;
  const vec3 P = cross(d, e2);
  const float det = dot(P, e1);
  const float inv_det = 1.f / det;
  if (det == 0.f) return false;
  const vec3 T = o - v0;
  const float u = dot(P, T) * inv_det;
  if (u < 0 || u > 1.f) return false;
  const vec3 Q = cross(T, e1);
  const float v = dot(Q, d) * inv_det;
  if (v < 0.f || (u + v) > 1.f) return false;
  float tt = dot(Q, e2) * inv_det;
  if (tt < tmin || tt > tmax) return false;
  t_out = tt;
  u_out = u;
  v_out = v;
  return true;
}

// iter_traverse (native.cpp:56-114)
int closest_hit(const Scene &s, const Bvh &bvh, const vec3 &o, const vec3 &d, float tnear, float &t, float &u, float &v) {
  constexpr int MAX_STACK = 2048;
  float tmax = FLT_MAX;
  const BvhNode *stack[MAX_STACK];
  int idx = 0;
  stack[0] = &bvh.nodes[0];
  int best = -1;
  while (idx != -1) {
    const BvhNode *node = stack[idx--];
    if (node->primitive_count == 0) {
      const BvhNode *l = &bvh.nodes[node->left];
      const BvhNode *r = &bvh.nodes[node->left + 1];
      float li = box_intersect(*l, d, o, tmax), ri = box_intersect(*r, d, o, tmax);
      if (li > ri) {
        std::swap(li, ri);
        std::swap(l, r);
      }
      if (li == FLT_MAX) continue;
      if (ri == FLT_MAX) {
        if (idx + 1 < MAX_STACK) stack[++idx] = l;
        continue;
      }
      if (idx + 1 < MAX_STACK) stack[++idx] = r;
      if (idx + 1 < MAX_STACK) stack[++idx] = l;
      continue;
    }
    for (int32_t i = 0; i < node->primitive_count; i++) {
      int32_t p = bvh.prim_idx[node->left + i];
      float tt, uu, vv;
      if (!tri_hit(s, p, o, d, tnear, tmax, tt, uu, vv)) continue;
      if (tt <= tmax) {
        best = p;
        tmax = tt;
        t = tt;
        u = uu;
        v = vv;
      }
    }
  }
  return best;
}

int closest_hit_brute(const Scene &s, const vec3 &o, const vec3 &d, float tnear, float &t, float &u, float &v) {
  float tmax = FLT_MAX;
  int best = -1;
  for (int p = 0; p < (int)s.prim_mesh.size(); p++) {
    float tt, uu, vv;
    if (tri_hit(s, p, o, d, tnear, tmax, tt, uu, vv) && tt <= tmax) {
      best = p;
      tmax = tt;
      t = tt;
      u = uu;
      v = vv;
    }
  }
  return best;
}

}  // namespace oracle
