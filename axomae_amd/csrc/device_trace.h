// Closest-hit traversal of the 4-wide BVH and the binary-to-4-wide collapse step, written
// once for the device kernels and for a host build of the same code.
#pragma once
#include "device_math.h"
#include "device_path.h"
#include "nova_internal.h"

namespace nova_hip {

// Node and triangle records live in device global memory: loading them through the
// global address space gives global_load instructions instead of flat ones.
#ifdef __HIP_DEVICE_COMPILE__
#define NOVA_GLOBAL __attribute__((address_space(1)))
#else
#define NOVA_GLOBAL
#endif

NOVA_HD TriMT load_tri(const TriMT *p) {
  const NOVA_GLOBAL float4 *q = (const NOVA_GLOBAL float4 *)p;
  TriMT t;
  t.a = q[0];
  t.b = q[1];
  t.c = q[2];
  return t;
}

struct TraceCtx {
  const Node4 *nodes;
  const TriMT *tris;
  uint32_t *gstack;  // overflow entries: [level - lds levels][global thread]
  uint32_t gstride;  // number of threads in the grid
  const TieTables *tie;  // binary tree, only read to settle exact equal-t ties
  const Node4F *nodesf = nullptr;  // decoded-plane copy of the nodes (the wave-coherent walk reads these)
};

// BoundingBox::halfArea (BoundingBox.cpp:162-165)
NOVA_HD float half_area(float mnx, float mny, float mnz, float mxx, float mxy, float mxz) {
  float lx = mxx - mnx, ly = mxy - mny, lz = mxz - mnz;
  return (lx * ly + ly * lz) + lz * lx;
}

// Children of one 4-wide node: start from the two children of binary inner node b and
// open the inner child with the largest half area until there are four (or only leaves
// remain). A leaf b becomes a single leaf child. Returns the child count.
NOVA_HD int collapse_children(const BNode *bn, int32_t b, int32_t ch[4]) {
  const BNode root = bn[b];
  if (root.count > 0) {
    ch[0] = b;
    return 1;
  }
  int m = 0;
  ch[m++] = root.left;
  ch[m++] = root.left + 1;
  while (m < 4) {
    int pick = -1;
    float best = -1.f;
    for (int k = 0; k < m; k++) {
      const BNode &c = bn[ch[k]];
      if (c.count > 0) continue;
      const float ha = half_area(c.mn[0], c.mn[1], c.mn[2], c.mx[0], c.mx[1], c.mx[2]);
      if (ha > best) {
        best = ha;
        pick = k;
      }
    }
    if (pick < 0) break;
    const int32_t open = ch[pick];
    ch[pick] = bn[open].left;
    ch[m++] = bn[open].left + 1;
  }
  return m;
}

// Plane value of grid byte q on an axis with origin org and step s (a power of two).
NOVA_HD float grid_plane(uint32_t q, float s, float org) { return fmaf((float)q, s, org); }

// Grid step of one axis of a node whose children span [lo, hi]: the smallest power of two
// with 250 steps covering the span, and at least twice the float spacing at the span's
// largest magnitude (so each step moves the decoded value).
NOVA_HD float grid_step(float lo, float hi) {
  int e_span = -160, e_mag = 0;
  const float span = hi - lo;
  if (span > 0.f) (void)frexpf(span / 250.f, &e_span);  // span / 250 <= 2^e_span
  (void)frexpf(fmaxf(fabsf(lo), fabsf(hi)), &e_mag);     // magnitude < 2^e_mag, spacing 2^(e_mag - 24)
  const int e = e_span > e_mag - 23 ? e_span : e_mag - 23;
  return ldexpf(1.f, e < -149 ? -149 : e);
}

// Largest q in [0, 255] whose plane is <= v (q = 0 is the origin itself, which is <= v).
NOVA_HD uint32_t grid_floor(float v, float s, float org) {
  const float f = floorf((v - org) / s);
  int q = f >= 0.f ? (f <= 255.f ? (int)f : 255) : 0;
  while (q > 0 && grid_plane((uint32_t)q, s, org) > v) q--;
  while (q < 255 && grid_plane((uint32_t)q + 1, s, org) <= v) q++;
  return (uint32_t)q;
}

// Smallest q in [0, 255] whose plane is >= v; the step is chosen so that plane 255 is.
NOVA_HD uint32_t grid_ceil(float v, float s, float org) {
  const float f = ceilf((v - org) / s);
  int q = f >= 0.f ? (f <= 255.f ? (int)f : 255) : 0;
  while (q < 255 && grid_plane((uint32_t)q, s, org) < v) q++;
  while (q > 0 && grid_plane((uint32_t)q - 1, s, org) >= v) q--;
  return (uint32_t)q;
}

// Boxes of a 4-wide node from the binary nodes ch[0 .. m-1]: grid origin at the low corner of
// their union, one step per axis, each child's planes rounded outward onto the grid. Empty
// slots get planes 0. The caller sets the child references.
NOVA_HD void set_node_boxes(Node4 &nd, const BNode *bn, const int32_t ch[4], int m) {
  float lo[3], hi[3], org[3], step[3];
  uint32_t qmn[3] = {0, 0, 0}, qmx[3] = {0, 0, 0};
  for (int a = 0; a < 3; a++) {
    lo[a] = bn[ch[0]].mn[a];
    hi[a] = bn[ch[0]].mx[a];
    for (int k = 1; k < m; k++) {
      lo[a] = fminf(lo[a], bn[ch[k]].mn[a]);
      hi[a] = fmaxf(hi[a], bn[ch[k]].mx[a]);
    }
    org[a] = lo[a];
    step[a] = grid_step(lo[a], hi[a]);
    while (grid_plane(255u, step[a], org[a]) < hi[a]) step[a] *= 2.f;  // not taken for finite boxes; a guard
  }
  for (int k = 0; k < m; k++)
    for (int a = 0; a < 3; a++) {
      qmn[a] |= grid_floor(bn[ch[k]].mn[a], step[a], org[a]) << (8 * k);
      qmx[a] |= grid_ceil(bn[ch[k]].mx[a], step[a], org[a]) << (8 * k);
    }
  nd.orgx = org[0];
  nd.orgy = org[1];
  nd.orgz = org[2];
  nd.stepx = step[0];
  nd.stepy = step[1];
  nd.stepz = step[2];
  nd.qminx = qmn[0];
  nd.qmaxx = qmx[0];
  nd.qminy = qmn[1];
  nd.qmaxy = qmx[1];
  nd.qminz = qmn[2];
  nd.qmaxz = qmx[2];
}

// BoundingBox::intersect (BoundingBox.cpp:150-160) in its division form, std::min/std::max
// operand order kept; a 0/0 slab (ray inside a face plane) spans the whole line.
NOVA_HD float ref_box_entry(const BNode &b, f3 o, f3 d) {
  float t[6];
  const float lo[3] = {b.mn[0], b.mn[1], b.mn[2]}, hi[3] = {b.mx[0], b.mx[1], b.mx[2]};
  const float oo[3] = {o.x, o.y, o.z}, dd[3] = {d.x, d.y, d.z};
  for (int a = 0; a < 3; a++) {
    float t1 = (lo[a] - oo[a]) / dd[a], t2 = (hi[a] - oo[a]) / dd[a];
    if (t1 != t1 || t2 != t2) {
      t1 = -__builtin_inff();
      t2 = __builtin_inff();
    }
    t[2 * a] = t1;
    t[2 * a + 1] = t2;
  }
  float t0 = gmin(t[0], t[1]), t1 = smax(t[0], t[1]);
  t0 = smax(t0, gmin(t[2], t[3]));
  t1 = gmin(t1, smax(t[2], t[3]));
  t0 = smax(t0, gmin(t[4], t[5]));
  t1 = gmin(t1, smax(t[4], t[5]));
  return (t1 >= t0 && t1 > 0.f) ? t0 : 3.402823466e+38f;
}

// Settles a tie between two triangles with bit-identical t the way the reference's
// depth-first walk does (native.cpp:56-114, ties kept by `t <= tmax`): the one it reaches
// last wins. Within a binary leaf that is the higher position. Otherwise the walk enters
// the nearer child of the two leaves' lowest common ancestor first (the left one when the
// entry distances are equal) and the triangle under the other child wins -- provided the
// walk still reaches it: below that child every box is tested against tmax = t, and a box
// whose rounded entry distance exceeds t is culled there. Returns true when pos_new wins.
#ifdef __HIP_DEVICE_COMPILE__
__device__ __noinline__
#else
__host__ __device__ inline
#endif
    bool
    reference_tie_winner_is_new(const TieTables *tie, int32_t pos_new, int32_t pos_old, f3 o, f3 d, float t) {
  // one pointer to a record in device memory: the common path keeps no tree pointers live,
  // and no caller has to copy a context into scratch memory for this out-of-line call
  const TieTables tc = *tie;
  const int32_t leaf_new = tc.leaf_of_pos[pos_new], leaf_old = tc.leaf_of_pos[pos_old];
  int32_t a = leaf_new, b = leaf_old;
  if (a == b) return pos_new > pos_old;
  int32_t da = tc.bdepth[a], db = tc.bdepth[b];
  for (; da > db && a > 0; da--) a = tc.bparent[a];
  for (; db > da && b > 0; db--) b = tc.bparent[b];
  if (a == b) return pos_new > pos_old;
  while (a > 0 && b > 0 && tc.bparent[a] != tc.bparent[b]) {
    a = tc.bparent[a];
    b = tc.bparent[b];
  }
  if (a <= 0 || b <= 0) return pos_new > pos_old;  // malformed tree: keep the leaf order
  const int32_t left = tc.bnodes[tc.bparent[a]].left;
  const float li = ref_box_entry(tc.bnodes[left], o, d), ri = ref_box_entry(tc.bnodes[left + 1], o, d);
  const int32_t first = li > ri ? left + 1 : left;
  const bool new_later = a != first;
  // is the later triangle reached once the earlier one has set tmax = t?
  const int32_t top = new_later ? a : b;
  bool reached = (new_later ? (a == left ? li : ri) : (b == left ? li : ri)) < 3.402823466e+38f;
  for (int32_t v = new_later ? leaf_new : leaf_old; reached && v != top && v > 0; v = tc.bparent[v]) reached = ref_box_entry(tc.bnodes[v], o, d) <= t;
  return new_later ? reached : !reached;
}

// Per-ray traversal state, so that a lane can be handed a new ray as soon as its ray ends.
struct TraceState {
  f3 o, d, inv;
  bool px, py, pz, parallel;
  float tbest, bu, bv;
  int best;
  int32_t bpos, node;
  int sp;
  uint32_t tri_rounds;  // instrumented walks: triangle-loop rounds this lane led for its wave
};

NOVA_HD void trace_init(TraceState &st, f3 o, f3 d) {
  const float kInf = __builtin_inff();
  st.o = o;
  st.d = d;
  st.inv = mk3(1.f / d.x, 1.f / d.y, 1.f / d.z);
  // An axis whose reciprocal is infinite (zero or denormal direction component) is a
  // containment test instead: inside the slab (boundary included) the whole line, else
  // nothing -- what the reference's 0/0 slabs amount to in the oracle's reading.
  st.px = !(fabsf(st.inv.x) < kInf);
  st.py = !(fabsf(st.inv.y) < kInf);
  st.pz = !(fabsf(st.inv.z) < kInf);
  st.parallel = st.px || st.py || st.pz;
  st.tbest = 3.402823466e+38f;
  st.best = -1;
  st.bpos = -1;
  st.bu = st.bv = 0.f;
  st.node = 0;
  st.sp = 0;
  st.tri_rounds = 0;
}

// Where a walk finds its tie tables: the context's pointer by default.
struct TieFromCtx {
  NOVA_HD const TieTables *operator()(const TraceCtx &tc) const { return tc.tie; }
};

// One node of the walk: test the node's four children with a slightly widened slab (no box
// holding a hit the reference's division-form test accepts is culled), order them near to
// far, push the inner ones, intersect the leaves (exact Moller-Trumbore; exact equal-t ties
// settled in the reference's visiting order). Returns true when the walk is over; the hit is
// then st.best / st.tbest / st.bu / st.bv. Stack provides push(level, node) and at(level);
// tie_of(tc) gives the tie tables.
template <bool kCount, class Stack, class TieOf = TieFromCtx>
NOVA_HD bool trace_step(const TraceCtx &tc, Stack &stack, TraceState &st, float tnear, uint32_t &nodes_visited, uint32_t &tri_tests,
                        TieOf tie_of = TieOf()) {
  const float kUp = 1.000001f;
  const float kInf = __builtin_inff();
    if (kCount) nodes_visited++;
    const NOVA_GLOBAL float4 *np = (const NOVA_GLOBAL float4 *)(tc.nodes + st.node);
    const float4 g0 = np[0], g1 = np[1];
    const uint4 g2 = *(const NOVA_GLOBAL uint4 *)(np + 2);
    const int4 ch = *(const NOVA_GLOBAL int4 *)(np + 3);
    const float orgx = g0.x, orgy = g0.y, orgz = g0.z, sx = g0.w, sy = g1.x, sz = g1.y;
    const uint32_t qmnx = __builtin_bit_cast(uint32_t, g1.z), qmxx = __builtin_bit_cast(uint32_t, g1.w);
    const uint32_t qmny = g2.x, qmxy = g2.y, qmnz = g2.z, qmxz = g2.w;
    const int chv[4] = {ch.x, ch.y, ch.z, ch.w};
    // per child: entry distance (+inf when culled) and child reference (-1 when culled)
    float key[4];
    int32_t cref[4];
    // Two copies of the slab test: rays with an axis-parallel direction take the
    // containment form; the common copy keeps every min/max operand a product of the same
    // block, which spares the NaN-quieting copies the compiler adds to values merged from
    // branches.
    if (st.parallel) {
#pragma unroll
      for (int k = 0; k < 4; k++) {
        const float bmnx = grid_plane((qmnx >> (8 * k)) & 255u, sx, orgx), bmxx = grid_plane((qmxx >> (8 * k)) & 255u, sx, orgx);
        const float bmny = grid_plane((qmny >> (8 * k)) & 255u, sy, orgy), bmxy = grid_plane((qmxy >> (8 * k)) & 255u, sy, orgy);
        const float bmnz = grid_plane((qmnz >> (8 * k)) & 255u, sz, orgz), bmxz = grid_plane((qmxz >> (8 * k)) & 255u, sz, orgz);
        float tx1 = (bmnx - st.o.x) * st.inv.x, tx2 = (bmxx - st.o.x) * st.inv.x;
        float ty1 = (bmny - st.o.y) * st.inv.y, ty2 = (bmxy - st.o.y) * st.inv.y;
        float tz1 = (bmnz - st.o.z) * st.inv.z, tz2 = (bmxz - st.o.z) * st.inv.z;
        if (st.px) {
          const bool in = bmnx <= st.o.x && st.o.x <= bmxx;
          tx1 = -kInf;
          tx2 = in ? kInf : -kInf;
        }
        if (st.py) {
          const bool in = bmny <= st.o.y && st.o.y <= bmxy;
          ty1 = -kInf;
          ty2 = in ? kInf : -kInf;
        }
        if (st.pz) {
          const bool in = bmnz <= st.o.z && st.o.z <= bmxz;
          tz1 = -kInf;
          tz2 = in ? kInf : -kInf;
        }
        const float t0 = fmaxf(fmaxf(fminf(tx1, tx2), fminf(ty1, ty2)), fminf(tz1, tz2));
        const float t1 = fminf(fminf(fmaxf(tx1, tx2), fmaxf(ty1, ty2)), fmaxf(tz1, tz2)) * kUp;
        const bool hit = chv[k] != -1 && t0 <= t1 && t1 >= 0.f && t0 <= st.tbest * kUp;
        key[k] = hit ? t0 : kInf;
        cref[k] = hit ? chv[k] : -1;
      }
    } else {
      // With finite reciprocals the nearer plane of an axis is the min plane when the
      // direction is positive and the max plane when it is negative (a decoded min plane is
      // <= its max plane and the products keep that order), so the per-axis min/max of the
      // two slab distances becomes a choice of plane bytes, made once per node.
      const uint32_t nqx = st.inv.x < 0.f ? qmxx : qmnx, fqx = st.inv.x < 0.f ? qmnx : qmxx;
      const uint32_t nqy = st.inv.y < 0.f ? qmxy : qmny, fqy = st.inv.y < 0.f ? qmny : qmxy;
      const uint32_t nqz = st.inv.z < 0.f ? qmxz : qmnz, fqz = st.inv.z < 0.f ? qmnz : qmxz;
#pragma unroll
      for (int k = 0; k < 4; k++) {
        const float tnx = (grid_plane((nqx >> (8 * k)) & 255u, sx, orgx) - st.o.x) * st.inv.x;
        const float tfx = (grid_plane((fqx >> (8 * k)) & 255u, sx, orgx) - st.o.x) * st.inv.x;
        const float tny = (grid_plane((nqy >> (8 * k)) & 255u, sy, orgy) - st.o.y) * st.inv.y;
        const float tfy = (grid_plane((fqy >> (8 * k)) & 255u, sy, orgy) - st.o.y) * st.inv.y;
        const float tnz = (grid_plane((nqz >> (8 * k)) & 255u, sz, orgz) - st.o.z) * st.inv.z;
        const float tfz = (grid_plane((fqz >> (8 * k)) & 255u, sz, orgz) - st.o.z) * st.inv.z;
        const float t0 = fmaxf(fmaxf(tnx, tny), tnz);
        const float t1 = fminf(fminf(tfx, tfy), tfz) * kUp;
        const bool hit = chv[k] != -1 && t0 <= t1 && t1 >= 0.f && t0 <= st.tbest * kUp;
        key[k] = hit ? t0 : kInf;
        cref[k] = hit ? chv[k] : -1;
      }
    }
    // near-to-far order of the four children: a fixed 5-exchange sorting network, so the
    // arrays stay in registers (culled children carry +inf and end last)
    auto exch = [&](int i, int j) {
      const bool sw = key[j] < key[i];
      const float kf = sw ? key[j] : key[i], kl = sw ? key[i] : key[j];
      const int32_t rf = sw ? cref[j] : cref[i], rl = sw ? cref[i] : cref[j];
      key[i] = kf;
      key[j] = kl;
      cref[i] = rf;
      cref[j] = rl;
    };
    exch(0, 1);
    exch(2, 3);
    exch(0, 2);
    exch(1, 3);
    exch(1, 2);
    int32_t next = -1;
#pragma unroll
    for (int a = 3; a >= 0; a--) {  // far to near: push inner nodes, keep the nearest
      if (cref[a] < 0) continue;
      if (next >= 0) stack.push(st.sp++, (uint32_t)next);
      next = cref[a];
    }
#pragma unroll
    for (int a = 0; a < 4; a++) {  // leaves now, near to far
      if (cref[a] > -2) continue;
      if (key[a] > st.tbest * kUp) continue;
      const int32_t start = leaf_first(cref[a]);
      int32_t cnt = 1;  // the leaf size comes with its first record
      for (int32_t i = start; i < start + cnt; i++) {
        const TriMT tr = load_tri(tc.tris + i);
        if (i == start) cnt = __builtin_bit_cast(int32_t, tr.b.w);
        if (kCount) {
          tri_tests++;
#ifdef __HIP_DEVICE_COMPILE__
          // the lowest active lane counts the round for its wave
          if (__lane_id() == (unsigned)(__builtin_ffsll((long long)__ballot(1)) - 1)) st.tri_rounds++;
#endif
        }
        float tt, uu, vv;
        if (tri_hit(tr, [&]() { return tie_of(tc); }, st.o, st.d, tnear, st.tbest, tt, uu, vv) &&
            (tt < st.tbest || st.best < 0 || reference_tie_winner_is_new(tie_of(tc), i, st.bpos, st.o, st.d, tt))) {
          st.tbest = tt;
          st.bpos = i;
          st.best = __builtin_bit_cast(int32_t, tr.a.w);
          st.bu = uu;
          st.bv = vv;
        }
      }
    }
    if (next >= 0) {
      st.node = next;
      return false;
    }
    if (st.sp == 0) return true;
    st.node = (int32_t)stack.at(--st.sp);
    return false;
}

// Closest hit of one ray: the steps until the walk ends (each node is visited once; the cap
// only guards against a malformed tree).
template <bool kCount, class Stack>
NOVA_HD int trace_closest(const TraceCtx &tc, Stack &stack, f3 o, f3 d, float tnear, float &t_out, float &u_out, float &v_out,
                          uint32_t &nodes_visited, uint32_t &tri_tests) {
  TraceState st;
  trace_init(st, o, d);
  for (uint32_t visits = 0; visits < (1u << 24); visits++)
    if (trace_step<kCount>(tc, stack, st, tnear, nodes_visited, tri_tests)) break;
  t_out = st.tbest;
  u_out = st.bu;
  v_out = st.bv;
  return st.best;
}

// ----------------------------------------------------------- wave-coherent walk
//
// For waves whose 64 rays nearly coincide (one pixel's camera rays of 64 frames: the jitter
// is under half a pixel), the wave walks the tree once for all of them. The node to visit is
// wave-uniform and read through the scalar unit (a 64-byte node is two scalar loads, no
// vector gathers); each lane runs the slab test of its own ray, and the wave visits a child
// when any lane enters it, carrying the mask of those lanes. A lane thus tests exactly the
// boxes and triangles its own walk would reach -- only the order is the wave's (near to far
// for the first lane of the mask). The closest hit with the reference's tie rule does not
// depend on the order, with one exception: when Moller-Trumbore's t for a triangle lies
// before its box's entry (a ray grazing the edge two triangles share, where t carries the
// rounding of a small determinant), a walk that first finds the neighbouring triangle culls
// that box and one that reaches the box first keeps the closer hit. The reference's own
// walk is order-dependent there too; the per-ray walk follows each ray's near-to-far order. The stack is the wave's: (node, lane mask) entries, entry e held
// in lane e of three registers, so a push is a compare-and-select and a pop a lane read.
#ifdef __HIP_DEVICE_COMPILE__
typedef const __attribute__((address_space(4))) Node4F ConstNode4F;  // scalar-loadable: uniform address, never written by the walk
typedef const __attribute__((address_space(4))) float4 ConstFloat4;
typedef const __attribute__((address_space(4))) TriMT ConstTriMT;
#else
typedef const Node4F ConstNode4F;
typedef const float4 ConstFloat4;
typedef const TriMT ConstTriMT;
#endif

constexpr int kWaveStack = 64;  // (node, mask) entries per wave, one per lane; usable when 3 * wide depth + 4 < kWaveStack

template <bool kCount>
__device__ void coherent_walk(const TraceCtx &tc, TraceState &st, uint64_t lanes, float tnear, uint32_t &nodes_visited, uint32_t &tri_tests) {
  const float kUp = 1.000001f;
  const float kInf = __builtin_inff();
  ConstNode4F *nodes = (ConstNode4F *)tc.nodesf;
  ConstTriMT *tris = (ConstTriMT *)tc.tris;
  const uint32_t lane = __lane_id();
  int32_t node = 0;
  uint64_t mask = lanes;  // lanes that entered this node's box
  int sp = 0;
  // the wave's stack: entry e is lane e of three registers (node, mask low, mask high)
  int32_t stk_node = 0, stk_lo = 0, stk_hi = 0;
  // When the wave's rays agree on the sign of each direction component (one pixel's rays
  // nearly always do) and none is axis-parallel, the near and far plane rows of every axis
  // are the same for all lanes: wave-uniform row indices, no per-lane select.
  const uint64_t negx = __ballot(st.inv.x < 0.f) & lanes, negy = __ballot(st.inv.y < 0.f) & lanes, negz = __ballot(st.inv.z < 0.f) & lanes;
  const bool uniform = __ballot(st.parallel) == 0 && (negx == 0 || negx == lanes) && (negy == 0 || negy == lanes) && (negz == 0 || negz == lanes);
  // float4 rows of a node: 2 * axis + side (side 0 min, 1 max), then the children at 6
  const uint32_t rnx = negx ? 1u : 0u, rny = negy ? 3u : 2u, rnz = negz ? 5u : 4u;  // near rows
  const uint32_t rfx = rnx ^ 1u, rfy = rny ^ 1u, rfz = rnz ^ 1u;                     // far rows
  for (uint32_t guard = 0; guard < (1u << 24) && mask; guard++) {
    node = __builtin_amdgcn_readfirstlane(node);
    ConstNode4F &nd = nodes[node];
    const bool in = (mask >> lane) & 1u;
    float key[4] = {kInf, kInf, kInf, kInf};
    bool hit[4] = {false, false, false, false};
    if (in) {
      if (kCount) nodes_visited++;
      if (uniform) {
        ConstFloat4 *rows = (ConstFloat4 *)&nodes[node];
        const float4 qnx = rows[rnx], qfx = rows[rfx], qny = rows[rny], qfy = rows[rfy], qnz = rows[rnz], qfz = rows[rfz];
        const float nxs[4] = {qnx.x, qnx.y, qnx.z, qnx.w}, fxs[4] = {qfx.x, qfx.y, qfx.z, qfx.w};
        const float nys[4] = {qny.x, qny.y, qny.z, qny.w}, fys[4] = {qfy.x, qfy.y, qfy.z, qfy.w};
        const float nzs[4] = {qnz.x, qnz.y, qnz.z, qnz.w}, fzs[4] = {qfz.x, qfz.y, qfz.z, qfz.w};
#pragma unroll
        for (int k = 0; k < 4; k++) {
          const float tnx = (nxs[k] - st.o.x) * st.inv.x, tfx = (fxs[k] - st.o.x) * st.inv.x;
          const float tny = (nys[k] - st.o.y) * st.inv.y, tfy = (fys[k] - st.o.y) * st.inv.y;
          const float tnz = (nzs[k] - st.o.z) * st.inv.z, tfz = (fzs[k] - st.o.z) * st.inv.z;
          const float t0 = fmaxf(fmaxf(tnx, tny), tnz);
          const float t1 = fminf(fminf(tfx, tfy), tfz) * kUp;
          hit[k] = nd.child[k] != -1 && t0 <= t1 && t1 >= 0.f && t0 <= st.tbest * kUp;
          key[k] = hit[k] ? t0 : kInf;
        }
      } else {  // mixed signs or axis-parallel rays: both planes of each axis, then min/max (the same values)
#pragma unroll
        for (int k = 0; k < 4; k++) {
          const float bmnx = nd.plane[0][0][k], bmxx = nd.plane[0][1][k];
          const float bmny = nd.plane[1][0][k], bmxy = nd.plane[1][1][k];
          const float bmnz = nd.plane[2][0][k], bmxz = nd.plane[2][1][k];
          float tx1 = (bmnx - st.o.x) * st.inv.x, tx2 = (bmxx - st.o.x) * st.inv.x;
          float ty1 = (bmny - st.o.y) * st.inv.y, ty2 = (bmxy - st.o.y) * st.inv.y;
          float tz1 = (bmnz - st.o.z) * st.inv.z, tz2 = (bmxz - st.o.z) * st.inv.z;
          if (st.px) {
            const bool inside = bmnx <= st.o.x && st.o.x <= bmxx;
            tx1 = -kInf;
            tx2 = inside ? kInf : -kInf;
          }
          if (st.py) {
            const bool inside = bmny <= st.o.y && st.o.y <= bmxy;
            ty1 = -kInf;
            ty2 = inside ? kInf : -kInf;
          }
          if (st.pz) {
            const bool inside = bmnz <= st.o.z && st.o.z <= bmxz;
            tz1 = -kInf;
            tz2 = inside ? kInf : -kInf;
          }
          const float t0 = fmaxf(fmaxf(fminf(tx1, tx2), fminf(ty1, ty2)), fminf(tz1, tz2));
          const float t1 = fminf(fminf(fmaxf(tx1, tx2), fmaxf(ty1, ty2)), fmaxf(tz1, tz2)) * kUp;
          hit[k] = nd.child[k] != -1 && t0 <= t1 && t1 >= 0.f && t0 <= st.tbest * kUp;
          key[k] = hit[k] ? t0 : kInf;
        }
      }
    }
    // The wave's view of each child: the lanes that enter it, and an ordering key -- the
    // first lane's entry distance as an orderable integer with the child's index in its low
    // two bits (so the four keys differ; order among near-equal keys does not matter).
    const int leader = __builtin_ffsll((long long)mask) - 1;
    uint64_t m[4];
    int32_t pk[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
      m[k] = __ballot(hit[k]);
      const int32_t b = __builtin_amdgcn_readlane(__builtin_bit_cast(int32_t, key[k]), leader);
      pk[k] = ((m[k] ? (b ^ ((b >> 31) & 0x7fffffff)) : 0x7fffffff) & ~3) | k;
    }
    // leaves, in child order, by the lanes that enter each leaf's box
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const int32_t c = nd.child[k];
      if (c > -2 || !m[k]) continue;
      const bool mine = ((m[k] >> lane) & 1u) && key[k] <= st.tbest * kUp;
      const int32_t start = leaf_first(c);
      const int32_t cnt = __builtin_bit_cast(int32_t, tris[start].b.w);
      for (int32_t i = start; i < start + cnt; i++) {
        const TriMT tr = tris[i];
        if (!mine) continue;
        if (kCount) tri_tests++;
        float tt, uu, vv;
        if (tri_hit(tr, [&]() { return tc.tie; }, st.o, st.d, tnear, st.tbest, tt, uu, vv) &&
            (tt < st.tbest || st.best < 0 || reference_tie_winner_is_new(tc.tie, i, st.bpos, st.o, st.d, tt))) {
          st.tbest = tt;
          st.bpos = i;
          st.best = __builtin_bit_cast(int32_t, tr.a.w);
          st.bu = uu;
          st.bv = vv;
        }
      }
    }
    // Inner children entered by some lane: the nearest is visited next, the others go onto
    // the stack far to near -- child k to slot sp + (entered inner children farther than k).
    bool inner[4];
    int ninner = 0;
#pragma unroll
    for (int k = 0; k < 4; k++) {
      inner[k] = nd.child[k] >= 0 && m[k] != 0;
      ninner += inner[k] ? 1 : 0;
    }
    int32_t next = -1;
    uint64_t next_mask = 0;
#pragma unroll
    for (int k = 0; k < 4; k++) {
      int farther = 0;
#pragma unroll
      for (int j = 0; j < 4; j++) farther += (j != k && inner[j] && pk[j] > pk[k]) ? 1 : 0;
      const bool nearest = inner[k] && farther == ninner - 1;
      next = nearest ? nd.child[k] : next;
      next_mask = nearest ? m[k] : next_mask;
      // pushed entries land at sp .. sp + ninner - 2; the others write the spare last slot
      const int slot = (inner[k] && !nearest) ? sp + farther : kWaveStack - 1;
      const bool here = (int)lane == slot;  // a lane write: compare and select
      stk_node = here ? nd.child[k] : stk_node;
      stk_lo = here ? (int32_t)(uint32_t)m[k] : stk_lo;
      stk_hi = here ? (int32_t)(uint32_t)(m[k] >> 32) : stk_hi;
    }
    if (next >= 0) {
      sp += ninner - 1;
      node = next;
      mask = next_mask;
      continue;
    }
    if (sp == 0) break;
    sp--;
    node = __builtin_amdgcn_readlane(stk_node, sp);
    mask = ((uint64_t)(uint32_t)__builtin_amdgcn_readlane(stk_hi, sp) << 32) | (uint32_t)__builtin_amdgcn_readlane(stk_lo, sp);
  }
}

#ifndef NOVA_LDS_STACK
#define NOVA_LDS_STACK 16
#endif
constexpr int kTraceBlock = 256;
constexpr int kLdsStack = NOVA_LDS_STACK;  // traversal stack levels held in LDS per lane

// Traversal stack of one lane of a kTraceBlock-thread block: the first kLdsStack levels in
// LDS (lane-interleaved), the rest in a global overflow area strided by the grid size.
struct LdsStack {
  uint32_t *lds;
  uint32_t *gstack;
  uint32_t gstride, gtid, lane;
  __device__ __forceinline__ void push(int sp, uint32_t v) {
    if (sp < kLdsStack)
      lds[sp * kTraceBlock + lane] = v;
    else
      gstack[(size_t)(sp - kLdsStack) * gstride + gtid] = v;
  }
  __device__ __forceinline__ uint32_t at(int sp) const {
    return sp < kLdsStack ? lds[sp * kTraceBlock + lane] : gstack[(size_t)(sp - kLdsStack) * gstride + gtid];
  }
};

}  // namespace nova_hip
