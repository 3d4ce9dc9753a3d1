// GPU SAH build with the reference builder's exact decisions, then a collapse to a
// 4-wide BVH.
//
// Reference: BvhtlBuilder::buildTriangleBasedScene (BvhBuilder.cpp:37-253), PERFORMANCE
// build type (4 candidate planes per axis), SAH segmentation.
//
// The build is level-parallel: one workgroup per node of the current level.
//  * Binning. A candidate plane k on axis a sits at seg*k + bmin (seg = extent/4). The
//    candidates are non-decreasing in k, so "centroid < plane_k" holds exactly for the
//    primitives whose bin b = #{k : plane_k <= centroid} is <= k. Per axis the node keeps
//    5 bins (count, min, max); the left side of candidate k is the union of bins 0..k and the
//    right side the union of bins k+1..4. Min/max unions are exact, so the reference's
//    costs hA(L)*nL + hA(R)*nR are reproduced bit for bit, including its NaN -> INT_MAX
//    sentinel for an empty side and its strict "first minimum below float(INT_MAX)" rule.
//  * Partition. The reference partitions with an in-place swap loop (BvhBuilder.cpp:186-203)
//    whose output order the children inherit. With nL the left count, H = [0,nL) and
//    T = [nL,n) (positions relative to the node), h_r the r-th right-class position in H
//    (ascending), t_r the r-th left-class position in T (descending, t_-1 = n) and
//    q = t_{k-1} (n if H holds no right-class element), the loop's result is:
//      left-class x in H stays; right-class h_r moves to t_{r-1}-1; left-class t_r moves
//      to h_r; right-class x_p in T moves to p-1, except x_nL which moves to q-1.
//    Ranks come from block-wide ballot/mbcnt prefix sums; the partition is applied even
//    when the split is then rejected (nL = 0 or n), as in the reference.
//  * Children are allocated in pairs (left, left+1), so node.left / node.left+1 address
//    them as the reference's traversal expects. Node numbering is breadth-first per level;
//    the primitive order prim_idx does not depend on it.
#include <algorithm>
#include <vector>

#include "device_math.h"
#include "device_trace.h"
#include "nova_internal.h"

namespace nova_hip {

constexpr int kBuildBlock = 256;
constexpr float kIntMaxF = 2147483648.f;  // float(INT_MAX)

__device__ __forceinline__ int ford(float f) {  // order-preserving float -> int
  int i = __float_as_int(f);
  return i >= 0 ? i : i ^ 0x7FFFFFFF;
}
__device__ __forceinline__ float unford(int i) { return __int_as_float(i >= 0 ? i : i ^ 0x7FFFFFFF); }

struct BlockScan {
  int wave_tot[kBuildBlock / 64];
  int total;
};

// Exclusive block prefix of `bit`; returns the rank and sets `chunk_total`.
__device__ __forceinline__ int block_rank(int bit, BlockScan &sc, int &chunk_total) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  uint64_t m = __ballot(bit);
  int in_wave = __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0));
  if (lane == 0) sc.wave_tot[wave] = __popcll(m);
  __syncthreads();
  int before = 0, tot = 0;
  for (int w = 0; w < kBuildBlock / 64; w++) {
    int t = sc.wave_tot[w];
    before += (w < wave) ? t : 0;
    tot += t;
  }
  __syncthreads();
  chunk_total = tot;
  return before + in_wave;
}

struct LevelShared {
  int cnt[3][5];
  int bmn[3][5][3];
  int bmx[3][5][3];
  int cmn[2][3], cmx[2][3];  // child bounds
  int nl;
  int split, axis;
  float pos;
  BlockScan sc;
};

// Candidate planes of a node: seg*k + bmin per axis (axis_subdiv_sah, BvhBuilder.cpp:117-150).
struct NodeCands {
  float ext[3], cand[3][4];
};
__device__ __forceinline__ NodeCands node_cands(const BNode &nd) {
  NodeCands c;
  for (int a = 0; a < 3; a++) {
    c.ext[a] = nd.mx[a] - nd.mn[a];
    const float seg = c.ext[a] / 4.f;
    for (int k = 0; k < 4; k++) c.cand[a][k] = seg * (float)k + nd.mn[a];
  }
  return c;
}

// Adds primitive p to the 3 x 5 bins (count, min, max as ordered ints; layout [axis*5+bin]).
__device__ __forceinline__ void bin_primitive(const NodeCands &c, const float *__restrict__ aabb, const float *__restrict__ centroid, int p, int *cnt,
                                              int *bmn, int *bmx) {
  const float cc[3] = {centroid[(size_t)p * 3], centroid[(size_t)p * 3 + 1], centroid[(size_t)p * 3 + 2]};
  int lo[3], hi[3];
  for (int q = 0; q < 3; q++) {
    lo[q] = ford(aabb[(size_t)p * 6 + q]);
    hi[q] = ford(aabb[(size_t)p * 6 + 3 + q]);
  }
  for (int a = 0; a < 3; a++) {
    if (c.ext[a] == 0) continue;
    int b = 0;
    for (int k = 0; k < 4; k++) b += (cc[a] < c.cand[a][k]) ? 0 : 1;
    atomicAdd(&cnt[a * 5 + b], 1);
    for (int q = 0; q < 3; q++) {
      atomicMin(&bmn[(a * 5 + b) * 3 + q], lo[q]);
      atomicMax(&bmx[(a * 5 + b) * 3 + q], hi[q]);
    }
  }
}

struct SplitChoice {
  int split, axis;
  float pos;
};

// eval_sah over the 4 candidates of each axis from the bins, the first strict minimum
// below float(INT_MAX) (empty side -> NaN -> INT_MAX), then the leaf test against the
// node's own cost (BvhBuilder.cpp:117-184, 206-253).
__device__ __forceinline__ SplitChoice decide_split(const BNode &nd, const NodeCands &c, const int *cnt, const int *bmn, const int *bmx, int n) {
  float best_cost = kIntMaxF, best_pos = 0.f;
  int best_axis = 0;
  for (int a = 0; a < 3; a++) {
    if (c.ext[a] == 0) continue;
    for (int k = 0; k < 4; k++) {
      float lmn[3] = {1e32f, 1e32f, 1e32f}, lmx[3] = {-1e32f, -1e32f, -1e32f};
      float rmn[3] = {1e32f, 1e32f, 1e32f}, rmx[3] = {-1e32f, -1e32f, -1e32f};
      int nl = 0, nr = 0;
      for (int b = 0; b < 5; b++) {
        const bool left = b <= k;
        if (left)
          nl += cnt[a * 5 + b];
        else
          nr += cnt[a * 5 + b];
        for (int q = 0; q < 3; q++) {
          const float bn = unford(bmn[(a * 5 + b) * 3 + q]), bx = unford(bmx[(a * 5 + b) * 3 + q]);
          if (left) {
            lmn[q] = gmin(lmn[q], bn);
            lmx[q] = gmax(lmx[q], bx);
          } else {
            rmn[q] = gmin(rmn[q], bn);
            rmx[q] = gmax(rmx[q], bx);
          }
        }
      }
      float cost = half_area(lmn[0], lmn[1], lmn[2], lmx[0], lmx[1], lmx[2]) * (float)nl +
                   half_area(rmn[0], rmn[1], rmn[2], rmx[0], rmx[1], rmx[2]) * (float)nr;
      cost = cost > 0.f ? cost : kIntMaxF;
      if (cost < best_cost) {
        best_cost = cost;
        best_axis = a;
        best_pos = c.cand[a][k];
      }
    }
  }
  const float node_cost = half_area(nd.mn[0], nd.mn[1], nd.mn[2], nd.mx[0], nd.mx[1], nd.mx[2]) * (float)n;
  return SplitChoice{best_cost >= node_cost ? 0 : 1, best_axis, best_pos};
}

// Final position (relative to the node) of the element at rel under the swap loop's closed form.
__device__ __forceinline__ int partition_dest(int rel, int bit, int rankL, int nL, int n, int q, const int32_t *__restrict__ hr_pos_s,
                                              const int32_t *__restrict__ tl_pos_s) {
  if (rel < nL) {
    if (bit) return rel;
    const int r = rel - rankL;
    return (r == 0 ? n : tl_pos_s[r - 1]) - 1;
  }
  if (bit) return hr_pos_s[nL - 1 - rankL];
  return (rel == nL && nL < q) ? q - 1 : rel - 1;
}

__global__ __launch_bounds__(kBuildBlock) void bvh_level_kernel(const int32_t *__restrict__ active, BNode *__restrict__ nodes,
                                                                int32_t *__restrict__ prim_idx, int32_t *__restrict__ tmp_idx,
                                                                const float *__restrict__ aabb, const float *__restrict__ centroid,
                                                                uint8_t *__restrict__ cls, int32_t *__restrict__ rk, int32_t *__restrict__ hr_pos,
                                                                int32_t *__restrict__ tl_pos, int32_t *__restrict__ node_counter,
                                                                int32_t *__restrict__ next_active, int32_t *__restrict__ next_count,
                                                                int32_t *__restrict__ parent, int32_t *__restrict__ depth) {
  __shared__ LevelShared sh;
  const int tid = threadIdx.x;
  const int32_t id = active[blockIdx.x];
  const BNode nd = nodes[id];
  const int s = nd.left, n = nd.count;
  const NodeCands nc = node_cands(nd);
  for (int i = tid; i < 15; i += kBuildBlock) {
    int a = i / 5, b = i % 5;
    sh.cnt[a][b] = 0;
    for (int q = 0; q < 3; q++) {
      sh.bmn[a][b][q] = ford(1e32f);
      sh.bmx[a][b][q] = ford(-1e32f);
    }
  }
  __syncthreads();

  // ---- binning (axis_subdiv_sah + eval_sah, BvhBuilder.cpp:117-184)
  for (int i = s + tid; i < s + n; i += kBuildBlock) bin_primitive(nc, aabb, centroid, prim_idx[i], &sh.cnt[0][0], &sh.bmn[0][0][0], &sh.bmx[0][0][0]);
  __syncthreads();

  if (tid == 0) {
    const SplitChoice sc = decide_split(nd, nc, &sh.cnt[0][0], &sh.bmn[0][0][0], &sh.bmx[0][0][0], n);
    sh.split = sc.split;
    sh.axis = sc.axis;
    sh.pos = sc.pos;
    sh.nl = 0;
    for (int q = 0; q < 3; q++)
      for (int c2 = 0; c2 < 2; c2++) {
        sh.cmn[c2][q] = ford(1e32f);
        sh.cmx[c2][q] = ford(-1e32f);
      }
  }
  __syncthreads();
  if (!sh.split) return;  // leaf: the reference returns before create_nodes
  const int axis = sh.axis;
  const float pos = sh.pos;

  // ---- classify and count (create_nodes predicate, BvhBuilder.cpp:194-199)
  int my_l = 0;
  for (int i = s + tid; i < s + n; i += kBuildBlock) {
    const int p = prim_idx[i];
    const int b = centroid[(size_t)p * 3 + axis] < pos ? 1 : 0;
    cls[i] = (uint8_t)b;
    my_l += b;
  }
  atomicAdd(&sh.nl, my_l);
  __syncthreads();
  const int nL = sh.nl;

  // ---- ranks and the h_r / t_r maps
  int running = 0;
  for (int base = s; base < s + n; base += kBuildBlock) {
    const int i = base + tid;
    const int bit = (i < s + n) ? cls[i] : 0;
    int tot;
    const int r_in = block_rank(bit, sh.sc, tot);
    if (i < s + n) {
      const int rankL = running + r_in;
      rk[i] = rankL;
      const int rel = i - s;
      if (rel < nL && !bit) hr_pos[s + (rel - rankL)] = rel;
      if (rel >= nL && bit) tl_pos[s + (nL - 1 - rankL)] = rel;
    }
    running += tot;
  }
  __syncthreads();
  const int LH = (nL < n) ? rk[s + nL] : nL;  // left-class elements inside H
  const int k = nL - LH;                         // right-class elements inside H
  const int q = (k == 0) ? n : tl_pos[s + k - 1];

  // ---- scatter to the swap loop's final order
  for (int i = s + tid; i < s + n; i += kBuildBlock) {
    const int rel = i - s;
    const int bit = cls[i];
    const int rankL = rk[i];
    tmp_idx[s + partition_dest(rel, bit, rankL, nL, n, q, hr_pos + s, tl_pos + s)] = prim_idx[i];
  }
  __syncthreads();

  // ---- copy back and reduce the children's bounds (update_aabb, BvhBuilder.cpp:77-101)
  int lmn[2][3], lmx[2][3];
  for (int c2 = 0; c2 < 2; c2++)
    for (int qq = 0; qq < 3; qq++) {
      lmn[c2][qq] = ford(1e32f);
      lmx[c2][qq] = ford(-1e32f);
    }
  for (int i = s + tid; i < s + n; i += kBuildBlock) {
    const int p = tmp_idx[i];
    prim_idx[i] = p;
    const int side = (i - s) < nL ? 0 : 1;
    for (int qq = 0; qq < 3; qq++) {
      lmn[side][qq] = min(lmn[side][qq], ford(aabb[(size_t)p * 6 + qq]));
      lmx[side][qq] = max(lmx[side][qq], ford(aabb[(size_t)p * 6 + 3 + qq]));
    }
  }
  for (int c2 = 0; c2 < 2; c2++)
    for (int qq = 0; qq < 3; qq++) {
      atomicMin(&sh.cmn[c2][qq], lmn[c2][qq]);
      atomicMax(&sh.cmx[c2][qq], lmx[c2][qq]);
    }
  __syncthreads();
  if (tid != 0 || nL == 0 || nL == n) return;
  const int32_t li = atomicAdd(node_counter, 2);
  BNode L, R;
  for (int qq = 0; qq < 3; qq++) {
    L.mn[qq] = unford(sh.cmn[0][qq]);
    L.mx[qq] = unford(sh.cmx[0][qq]);
    R.mn[qq] = unford(sh.cmn[1][qq]);
    R.mx[qq] = unford(sh.cmx[1][qq]);
  }
  L.left = s;
  L.count = nL;
  R.left = s + nL;
  R.count = n - nL;
  nodes[li] = L;
  nodes[li + 1] = R;
  parent[li] = parent[li + 1] = id;
  depth[li] = depth[li + 1] = depth[id] + 1;
  BNode P = nd;
  P.left = li;
  P.count = 0;
  nodes[id] = P;
  const int32_t slot = atomicAdd(next_count, 2);
  next_active[slot] = li;
  next_active[slot + 1] = li + 1;
}

// ------------------------------------------------- nodes too large for one workgroup
// A node with more than kBigNode primitives is processed by many workgroups, kBigChunk
// primitives each, in eight steps that compute exactly what bvh_level_kernel computes:
// bin (per-chunk LDS bins merged with atomics: min/max/count unions are order-free),
// decide, classify + count per chunk, scan of the chunk counts in chunk order, ranks
// (chunk base + in-chunk rank = the node-wide rank), scatter by the same closed form,
// copy back with child bounds, allocate the children.
constexpr int kBigNode = 16384;
constexpr int kBigChunk = 4096;

struct BigNode {
  int32_t id, s, n;
  int32_t chunk_first, chunk_count;
  int32_t split, axis, nL;
  float pos;
  int32_t cnt[15], bmn[45], bmx[45];  // bins
  int32_t cmn[2][3], cmx[2][3];       // child bounds
};

struct BigChunk {
  int32_t big, start, end;  // prim_idx positions [start, end)
};

// Splits the level's nodes: big ones into `big` (id, s, n), the rest into `small`.
__global__ void split_active_kernel(const int32_t *__restrict__ active, int32_t count, const BNode *__restrict__ nodes, BigNode *big,
                                    int32_t *big_count, int32_t *small, int32_t *small_count) {
  for (int32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += gridDim.x * blockDim.x) {
    const int32_t id = active[i];
    const BNode nd = nodes[id];
    if (nd.count > kBigNode) {
      const int32_t k = atomicAdd(big_count, 1);
      big[k].id = id;
      big[k].s = nd.left;
      big[k].n = nd.count;
    } else {
      small[atomicAdd(small_count, 1)] = id;
    }
  }
}

__global__ void big_init_kernel(BigNode *big, int32_t count) {
  for (int32_t b = blockIdx.x; b < count; b += gridDim.x) {
    BigNode &B = big[b];
    for (int i = threadIdx.x; i < 15; i += blockDim.x) B.cnt[i] = 0;
    for (int i = threadIdx.x; i < 45; i += blockDim.x) {
      B.bmn[i] = ford(1e32f);
      B.bmx[i] = ford(-1e32f);
    }
    if (threadIdx.x < 6) {
      B.cmn[threadIdx.x / 3][threadIdx.x % 3] = ford(1e32f);
      B.cmx[threadIdx.x / 3][threadIdx.x % 3] = ford(-1e32f);
    }
  }
}

__global__ __launch_bounds__(kBuildBlock) void big_bin_kernel(BigNode *big, const BigChunk *__restrict__ chunks, const BNode *__restrict__ nodes,
                                                              const int32_t *__restrict__ prim_idx, const float *__restrict__ aabb,
                                                              const float *__restrict__ centroid) {
  __shared__ int cnt[15], bmn[45], bmx[45];
  const BigChunk ch = chunks[blockIdx.x];
  BigNode &B = big[ch.big];
  const NodeCands nc = node_cands(nodes[B.id]);
  for (int i = threadIdx.x; i < 45; i += blockDim.x) {
    if (i < 15) cnt[i] = 0;
    bmn[i] = ford(1e32f);
    bmx[i] = ford(-1e32f);
  }
  __syncthreads();
  for (int i = ch.start + threadIdx.x; i < ch.end; i += kBuildBlock) bin_primitive(nc, aabb, centroid, prim_idx[i], cnt, bmn, bmx);
  __syncthreads();
  for (int i = threadIdx.x; i < 45; i += blockDim.x) {
    if (i < 15 && cnt[i]) atomicAdd(&B.cnt[i], cnt[i]);
    atomicMin(&B.bmn[i], bmn[i]);
    atomicMax(&B.bmx[i], bmx[i]);
  }
}

__global__ void big_decide_kernel(BigNode *big, int32_t count, const BNode *__restrict__ nodes) {
  const int32_t b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= count) return;
  BigNode &B = big[b];
  const BNode nd = nodes[B.id];
  const SplitChoice sc = decide_split(nd, node_cands(nd), B.cnt, B.bmn, B.bmx, B.n);
  B.split = sc.split;
  B.axis = sc.axis;
  B.pos = sc.pos;
}

__global__ __launch_bounds__(kBuildBlock) void big_classify_kernel(const BigNode *__restrict__ big, const BigChunk *__restrict__ chunks,
                                                                   const int32_t *__restrict__ prim_idx, const float *__restrict__ centroid,
                                                                   uint8_t *__restrict__ cls, int32_t *__restrict__ chunk_nl) {
  __shared__ int nl;
  const BigChunk ch = chunks[blockIdx.x];
  const BigNode &B = big[ch.big];
  if (!B.split) return;
  if (threadIdx.x == 0) nl = 0;
  __syncthreads();
  int my = 0;
  for (int i = ch.start + threadIdx.x; i < ch.end; i += kBuildBlock) {
    const int bit = centroid[(size_t)prim_idx[i] * 3 + B.axis] < B.pos ? 1 : 0;
    cls[i] = (uint8_t)bit;
    my += bit;
  }
  atomicAdd(&nl, my);
  __syncthreads();
  if (threadIdx.x == 0) chunk_nl[blockIdx.x] = nl;
}

// Exclusive prefix of the chunk counts in chunk order, and the node's left count.
__global__ void big_scan_kernel(BigNode *big, int32_t count, const int32_t *__restrict__ chunk_nl, int32_t *__restrict__ chunk_base) {
  const int32_t b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= count || !big[b].split) return;
  int running = 0;
  for (int32_t c = big[b].chunk_first; c < big[b].chunk_first + big[b].chunk_count; c++) {
    chunk_base[c] = running;
    running += chunk_nl[c];
  }
  big[b].nL = running;
}

__global__ __launch_bounds__(kBuildBlock) void big_rank_kernel(const BigNode *__restrict__ big, const BigChunk *__restrict__ chunks,
                                                               const uint8_t *__restrict__ cls, const int32_t *__restrict__ chunk_base,
                                                               int32_t *__restrict__ rk, int32_t *__restrict__ hr_pos, int32_t *__restrict__ tl_pos) {
  __shared__ BlockScan sc;
  const BigChunk ch = chunks[blockIdx.x];
  const BigNode &B = big[ch.big];
  if (!B.split) return;
  const int s = B.s, nL = B.nL;
  int running = chunk_base[blockIdx.x];
  for (int base = ch.start; base < ch.end; base += kBuildBlock) {
    const int i = base + threadIdx.x;
    const int bit = (i < ch.end) ? cls[i] : 0;
    int tot;
    const int r_in = block_rank(bit, sc, tot);
    if (i < ch.end) {
      const int rankL = running + r_in;
      rk[i] = rankL;
      const int rel = i - s;
      if (rel < nL && !bit) hr_pos[s + (rel - rankL)] = rel;
      if (rel >= nL && bit) tl_pos[s + (nL - 1 - rankL)] = rel;
    }
    running += tot;
  }
}

__global__ __launch_bounds__(kBuildBlock) void big_scatter_kernel(const BigNode *__restrict__ big, const BigChunk *__restrict__ chunks,
                                                                  const uint8_t *__restrict__ cls, const int32_t *__restrict__ rk,
                                                                  const int32_t *__restrict__ hr_pos, const int32_t *__restrict__ tl_pos,
                                                                  const int32_t *__restrict__ prim_idx, int32_t *__restrict__ tmp_idx) {
  const BigChunk ch = chunks[blockIdx.x];
  const BigNode &B = big[ch.big];
  if (!B.split) return;
  const int s = B.s, n = B.n, nL = B.nL;
  const int LH = (nL < n) ? rk[s + nL] : nL;  // left-class elements inside H
  const int k = nL - LH;                         // right-class elements inside H
  const int q = (k == 0) ? n : tl_pos[s + k - 1];
  for (int i = ch.start + threadIdx.x; i < ch.end; i += kBuildBlock)
    tmp_idx[s + partition_dest(i - s, cls[i], rk[i], nL, n, q, hr_pos + s, tl_pos + s)] = prim_idx[i];
}

__global__ __launch_bounds__(kBuildBlock) void big_copyback_kernel(BigNode *big, const BigChunk *__restrict__ chunks,
                                                                   const int32_t *__restrict__ tmp_idx, int32_t *__restrict__ prim_idx,
                                                                   const float *__restrict__ aabb) {
  __shared__ int cmn[2][3], cmx[2][3];
  const BigChunk ch = chunks[blockIdx.x];
  BigNode &B = big[ch.big];
  if (!B.split) return;
  if (threadIdx.x < 6) {
    cmn[threadIdx.x / 3][threadIdx.x % 3] = ford(1e32f);
    cmx[threadIdx.x / 3][threadIdx.x % 3] = ford(-1e32f);
  }
  __syncthreads();
  int lmn[2][3], lmx[2][3];
  for (int c2 = 0; c2 < 2; c2++)
    for (int qq = 0; qq < 3; qq++) {
      lmn[c2][qq] = ford(1e32f);
      lmx[c2][qq] = ford(-1e32f);
    }
  for (int i = ch.start + threadIdx.x; i < ch.end; i += kBuildBlock) {
    const int p = tmp_idx[i];
    prim_idx[i] = p;
    const int side = (i - B.s) < B.nL ? 0 : 1;
    for (int qq = 0; qq < 3; qq++) {
      lmn[side][qq] = min(lmn[side][qq], ford(aabb[(size_t)p * 6 + qq]));
      lmx[side][qq] = max(lmx[side][qq], ford(aabb[(size_t)p * 6 + 3 + qq]));
    }
  }
  for (int c2 = 0; c2 < 2; c2++)
    for (int qq = 0; qq < 3; qq++) {
      atomicMin(&cmn[c2][qq], lmn[c2][qq]);
      atomicMax(&cmx[c2][qq], lmx[c2][qq]);
    }
  __syncthreads();
  if (threadIdx.x < 6) {
    atomicMin(&B.cmn[threadIdx.x / 3][threadIdx.x % 3], cmn[threadIdx.x / 3][threadIdx.x % 3]);
    atomicMax(&B.cmx[threadIdx.x / 3][threadIdx.x % 3], cmx[threadIdx.x / 3][threadIdx.x % 3]);
  }
}

__global__ void big_alloc_kernel(const BigNode *__restrict__ big, int32_t count, BNode *__restrict__ nodes, int32_t *__restrict__ node_counter,
                                 int32_t *__restrict__ next_active, int32_t *__restrict__ next_count, int32_t *__restrict__ parent,
                                 int32_t *__restrict__ depth) {
  const int32_t b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= count) return;
  const BigNode &B = big[b];
  if (!B.split || B.nL == 0 || B.nL == B.n) return;
  const int32_t li = atomicAdd(node_counter, 2);
  BNode L, R;
  for (int qq = 0; qq < 3; qq++) {
    L.mn[qq] = unford(B.cmn[0][qq]);
    L.mx[qq] = unford(B.cmx[0][qq]);
    R.mn[qq] = unford(B.cmn[1][qq]);
    R.mx[qq] = unford(B.cmx[1][qq]);
  }
  L.left = B.s;
  L.count = B.nL;
  R.left = B.s + B.nL;
  R.count = B.n - B.nL;
  nodes[li] = L;
  nodes[li + 1] = R;
  parent[li] = parent[li + 1] = B.id;
  depth[li] = depth[li + 1] = depth[B.id] + 1;
  BNode P = nodes[B.id];
  P.left = li;
  P.count = 0;
  nodes[B.id] = P;
  const int32_t slot = atomicAdd(next_count, 2);
  next_active[slot] = li;
  next_active[slot + 1] = li + 1;
}

// Root bounds: min/max over every primitive AABB (update_aabb of node 0).
__global__ void root_bounds_kernel(const float *__restrict__ aabb, uint32_t n, int *__restrict__ out) {
  __shared__ int smn[3], smx[3];
  if (threadIdx.x < 3) {
    smn[threadIdx.x] = ford(1e32f);
    smx[threadIdx.x] = ford(-1e32f);
  }
  __syncthreads();
  int mn[3] = {ford(1e32f), ford(1e32f), ford(1e32f)}, mx[3] = {ford(-1e32f), ford(-1e32f), ford(-1e32f)};
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    for (int q = 0; q < 3; q++) {
      mn[q] = min(mn[q], ford(aabb[(size_t)i * 6 + q]));
      mx[q] = max(mx[q], ford(aabb[(size_t)i * 6 + 3 + q]));
    }
  for (int q = 0; q < 3; q++) {
    atomicMin(&smn[q], mn[q]);
    atomicMax(&smx[q], mx[q]);
  }
  __syncthreads();
  if (threadIdx.x < 3) {
    atomicMin(&out[threadIdx.x], smn[threadIdx.x]);
    atomicMax(&out[3 + threadIdx.x], smx[threadIdx.x]);
  }
}

// Binary leaf holding each prim_idx position (the tie rule of the traversal needs it).
// Per binary leaf: the leaf of each of its positions, and its size in the first traversal
// record's b.w.
__global__ void leaf_of_position_kernel(const BNode *__restrict__ nodes, uint32_t count, int32_t *__restrict__ leaf_of_pos,
                                        TriMT *__restrict__ tri_leaf) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += gridDim.x * blockDim.x) {
    const BNode nd = nodes[i];
    for (int32_t k = 0; k < nd.count; k++) leaf_of_pos[nd.left + k] = (int32_t)i;
    if (nd.count > 0) tri_leaf[nd.left].b.w = __builtin_bit_cast(float, nd.count);
  }
}

__global__ void init_root_kernel(BNode *nodes, const int *bounds, uint32_t n, int32_t *prim_idx, int32_t *counters, int32_t *parent, int32_t *depth) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) prim_idx[i] = (int32_t)i;
  if (i == 0) {
    BNode r;
    for (int q = 0; q < 3; q++) {
      r.mn[q] = unford(bounds[q]);
      r.mx[q] = unford(bounds[3 + q]);
    }
    r.left = 0;
    r.count = (int32_t)n;
    nodes[0] = r;
    parent[0] = -1;
    depth[0] = 0;
    counters[0] = 1;  // node counter
    counters[1] = 0;  // next level size
  }
}

__global__ void reset_level_counters_kernel(int32_t *counters) {
  counters[1] = 0;
  counters[4] = 0;
  counters[5] = 0;
}

__global__ void init_bounds_kernel(int *b) {
  if (threadIdx.x < 3) {
    b[threadIdx.x] = ford(1e32f);
    b[3 + threadIdx.x] = ford(-1e32f);
  }
}

// ---------------------------------------------------------------- collapse

// One thread per frontier entry: binary inner node -> 4-wide node by repeatedly opening
// the child with the largest half area.
__global__ void collapse_kernel(const int32_t *__restrict__ frontier_bin, const int32_t *__restrict__ frontier_slot, int32_t count,
                                const BNode *__restrict__ bn, Node4 *__restrict__ out, int32_t *__restrict__ wide_counter,
                                int32_t *__restrict__ next_bin, int32_t *__restrict__ next_slot, int32_t *__restrict__ next_count) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= count) return;
  const int32_t b = frontier_bin[t];
  const int32_t slot = frontier_slot[t];
  int32_t ch[4];
  const int m = collapse_children(bn, b, ch);
  Node4 nd;
  set_node_boxes(nd, bn, ch, m);
  for (int k = 0; k < 4; k++) {
    if (k >= m) {
      nd.child[k] = -1;  // empty slot, never entered
      continue;
    }
    const BNode &c = bn[ch[k]];
    if (c.count > 0) {
      nd.child[k] = leaf_ref(c.left);
    } else {
      const int32_t w = atomicAdd(wide_counter, 1);
      nd.child[k] = w;
      const int32_t j = atomicAdd(next_count, 1);
      next_bin[j] = ch[k];
      next_slot[j] = w;
    }
  }
  out[slot] = nd;
}

// Node4 -> Node4F: every plane decoded with grid_plane, the value the walks compute.
__global__ void decode_nodes_kernel(const Node4 *__restrict__ in, Node4F *__restrict__ out, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const Node4 nd = in[i];
  const float org[3] = {nd.orgx, nd.orgy, nd.orgz}, step[3] = {nd.stepx, nd.stepy, nd.stepz};
  const uint32_t qmn[3] = {nd.qminx, nd.qminy, nd.qminz}, qmx[3] = {nd.qmaxx, nd.qmaxy, nd.qmaxz};
  Node4F f;
  for (int a = 0; a < 3; a++)
    for (int k = 0; k < 4; k++) {
      f.plane[a][0][k] = grid_plane((qmn[a] >> (8 * k)) & 255u, step[a], org[a]);
      f.plane[a][1][k] = grid_plane((qmx[a] >> (8 * k)) & 255u, step[a], org[a]);
    }
  for (int k = 0; k < 4; k++) {
    f.child[k] = nd.child[k];
    f.pad[k] = 0;
  }
  out[i] = f;
}

__global__ void gather_leaf_tris(const int32_t *__restrict__ prim_idx, const TriMT *__restrict__ tri, TriMT *__restrict__ out, uint32_t n) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = tri[prim_idx[i]];
}

uint32_t traversal_stack_bound(const nova_hip_ctx *ctx) { return 3u * ctx->bvh_info.wide_depth + 4u; }

}  // namespace nova_hip

using namespace nova_hip;

extern "C" int nova_hip_bvh_build(nova_hip_ctx *ctx, int build_candidates, nova_hip_bvh_info *info, void *stream) {
  if (!ctx) return NOVA_HIP_INVALID_ARGUMENT;
  if (!ctx->scene_ready) return set_error(ctx, NOVA_HIP_SCENE_NOT_PROCESSED, "scene not uploaded");
  if (build_candidates != 4)
    return set_error(ctx, NOVA_HIP_INVALID_ARGUMENT, "only the PERFORMANCE build type (4 candidates per axis) is supported");
  hipStream_t s = (hipStream_t)stream;
  NOVA_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  const uint32_t n = ctx->prim_count;
  const size_t maxnodes = 2 * (size_t)n - 1;
  (void)hipDeviceSynchronize();  // render work queued without a wait may still read the old tree
  for (void **p : {(void **)&ctx->prim_idx, (void **)&ctx->bnodes, (void **)&ctx->nodes4, (void **)&ctx->nodes4f, (void **)&ctx->tri_leaf,
                   (void **)&ctx->bparent, (void **)&ctx->bdepth, (void **)&ctx->leaf_of_pos, (void **)&ctx->tie})
    if (*p) {
      hipFree(*p);
      *p = nullptr;
    }
  ctx->bvh_ready = false;
  NOVA_HIP_CHECK(ctx, hipMalloc(&ctx->prim_idx, n * sizeof(int32_t)));
  NOVA_HIP_CHECK(ctx, hipMalloc(&ctx->bnodes, maxnodes * sizeof(BNode)));
  NOVA_HIP_CHECK(ctx, hipMalloc(&ctx->bparent, maxnodes * sizeof(int32_t)));
  NOVA_HIP_CHECK(ctx, hipMalloc(&ctx->bdepth, maxnodes * sizeof(int32_t)));
  NOVA_HIP_CHECK(ctx, hipMalloc(&ctx->leaf_of_pos, (size_t)n * sizeof(int32_t)));
  NOVA_HIP_CHECK(ctx, hipMalloc(&ctx->tie, sizeof(TieTables)));
  const TieTables tie{ctx->bnodes, ctx->bparent, ctx->bdepth, ctx->leaf_of_pos, ctx->xf_inv};
  NOVA_HIP_CHECK(ctx, hipMemcpy(ctx->tie, &tie, sizeof(TieTables), hipMemcpyHostToDevice));

  // temporaries
  int32_t *tmp_idx, *rk, *hr, *tl, *act0, *act1, *counters;
  uint8_t *cls;
  int *bounds;
  NOVA_HIP_CHECK(ctx, hipMalloc(&tmp_idx, n * sizeof(int32_t)));
  NOVA_HIP_CHECK(ctx, hipMalloc(&rk, n * sizeof(int32_t)));
  NOVA_HIP_CHECK(ctx, hipMalloc(&hr, n * sizeof(int32_t)));
  NOVA_HIP_CHECK(ctx, hipMalloc(&tl, n * sizeof(int32_t)));
  NOVA_HIP_CHECK(ctx, hipMalloc(&act0, (n + 2) * sizeof(int32_t)));
  NOVA_HIP_CHECK(ctx, hipMalloc(&act1, (n + 2) * sizeof(int32_t)));
  NOVA_HIP_CHECK(ctx, hipMalloc(&counters, 8 * sizeof(int32_t)));
  NOVA_HIP_CHECK(ctx, hipMalloc(&cls, n));
  NOVA_HIP_CHECK(ctx, hipMalloc(&bounds, 6 * sizeof(int)));

  hipEvent_t ev0, ev1;
  NOVA_HIP_CHECK(ctx, hipEventCreate(&ev0));
  NOVA_HIP_CHECK(ctx, hipEventCreate(&ev1));
  NOVA_HIP_CHECK(ctx, hipEventRecord(ev0, s));

  hipLaunchKernelGGL(init_bounds_kernel, dim3(1), dim3(64), 0, s, bounds);
  hipLaunchKernelGGL(root_bounds_kernel, dim3(std::min<uint32_t>(1024, (n + 255) / 256)), dim3(256), 0, s, ctx->prim_aabb, n, bounds);
  hipLaunchKernelGGL(init_root_kernel, dim3((n + 255) / 256), dim3(256), 0, s, ctx->bnodes, bounds, n, ctx->prim_idx, counters, ctx->bparent,
                     ctx->bdepth);
  const int32_t zero = 0;
  NOVA_HIP_CHECK(ctx, hipMemcpyAsync(act0, &zero, sizeof(int32_t), hipMemcpyHostToDevice, s));
  int32_t active = 1;
  uint32_t depth = 0;
  int32_t *cur = act0, *nxt = act1;
  // per-level work lists of the big-node path
  const int32_t max_big = (int32_t)(n / kBigNode + 1), max_chunks = (int32_t)(n / kBigChunk + max_big + 1);
  BigNode *big = nullptr;
  BigChunk *chunks = nullptr;
  int32_t *small = nullptr, *chunk_nl = nullptr, *chunk_base = nullptr;
  NOVA_HIP_CHECK(ctx, hipMalloc(&big, max_big * sizeof(BigNode)));
  NOVA_HIP_CHECK(ctx, hipMalloc(&chunks, max_chunks * sizeof(BigChunk)));
  NOVA_HIP_CHECK(ctx, hipMalloc(&small, (n + 2) * sizeof(int32_t)));
  NOVA_HIP_CHECK(ctx, hipMalloc(&chunk_nl, max_chunks * sizeof(int32_t)));
  NOVA_HIP_CHECK(ctx, hipMalloc(&chunk_base, max_chunks * sizeof(int32_t)));
  std::vector<BigNode> hbig;
  std::vector<BigChunk> hchunks;
  while (active > 0) {
    if (depth > n) return set_error(ctx, NOVA_HIP_DEVICE_ERROR, "BVH build did not converge");
    // counters: [1] next level size, [4] big nodes, [5] small nodes
    hipLaunchKernelGGL(reset_level_counters_kernel, dim3(1), dim3(1), 0, s, counters);
    hipLaunchKernelGGL(split_active_kernel, dim3(std::min<int32_t>((active + 255) / 256, 4096)), dim3(256), 0, s, cur, active, ctx->bnodes, big,
                       counters + 4, small, counters + 5);
    NOVA_HIP_CHECK(ctx, hipGetLastError());
    int32_t split_counts[2] = {0, 0};
    NOVA_HIP_CHECK(ctx, hipMemcpyAsync(split_counts, counters + 4, 2 * sizeof(int32_t), hipMemcpyDeviceToHost, s));
    NOVA_HIP_CHECK(ctx, hipStreamSynchronize(s));
    const int32_t nbig = split_counts[0], nsmall = split_counts[1];
    if (nbig > max_big || nsmall > active) return set_error(ctx, NOVA_HIP_DEVICE_ERROR, "BVH build: inconsistent level split");
    if (nsmall > 0)
      hipLaunchKernelGGL(bvh_level_kernel, dim3(nsmall), dim3(kBuildBlock), 0, s, small, ctx->bnodes, ctx->prim_idx, tmp_idx, ctx->prim_aabb,
                         ctx->prim_centroid, cls, rk, hr, tl, counters, nxt, counters + 1, ctx->bparent, ctx->bdepth);
    if (nbig > 0) {
      hbig.resize(nbig);
      NOVA_HIP_CHECK(ctx, hipMemcpy(hbig.data(), big, nbig * sizeof(BigNode), hipMemcpyDeviceToHost));
      hchunks.clear();
      for (int32_t b = 0; b < nbig; b++) {
        hbig[b].chunk_first = (int32_t)hchunks.size();
        for (int32_t c0 = hbig[b].s; c0 < hbig[b].s + hbig[b].n; c0 += kBigChunk)
          hchunks.push_back(BigChunk{b, c0, std::min(c0 + kBigChunk, hbig[b].s + hbig[b].n)});
        hbig[b].chunk_count = (int32_t)hchunks.size() - hbig[b].chunk_first;
      }
      const int32_t nch = (int32_t)hchunks.size();
      if (nch > max_chunks) return set_error(ctx, NOVA_HIP_DEVICE_ERROR, "BVH build: chunk table overflow");
      NOVA_HIP_CHECK(ctx, hipMemcpyAsync(big, hbig.data(), nbig * sizeof(BigNode), hipMemcpyHostToDevice, s));
      NOVA_HIP_CHECK(ctx, hipMemcpyAsync(chunks, hchunks.data(), nch * sizeof(BigChunk), hipMemcpyHostToDevice, s));
      const unsigned bb = (unsigned)((nbig + 63) / 64);
      hipLaunchKernelGGL(big_init_kernel, dim3(nbig), dim3(64), 0, s, big, nbig);
      hipLaunchKernelGGL(big_bin_kernel, dim3(nch), dim3(kBuildBlock), 0, s, big, chunks, ctx->bnodes, ctx->prim_idx, ctx->prim_aabb,
                         ctx->prim_centroid);
      hipLaunchKernelGGL(big_decide_kernel, dim3(bb), dim3(64), 0, s, big, nbig, ctx->bnodes);
      hipLaunchKernelGGL(big_classify_kernel, dim3(nch), dim3(kBuildBlock), 0, s, big, chunks, ctx->prim_idx, ctx->prim_centroid, cls, chunk_nl);
      hipLaunchKernelGGL(big_scan_kernel, dim3(bb), dim3(64), 0, s, big, nbig, chunk_nl, chunk_base);
      hipLaunchKernelGGL(big_rank_kernel, dim3(nch), dim3(kBuildBlock), 0, s, big, chunks, cls, chunk_base, rk, hr, tl);
      hipLaunchKernelGGL(big_scatter_kernel, dim3(nch), dim3(kBuildBlock), 0, s, big, chunks, cls, rk, hr, tl, ctx->prim_idx, tmp_idx);
      hipLaunchKernelGGL(big_copyback_kernel, dim3(nch), dim3(kBuildBlock), 0, s, big, chunks, tmp_idx, ctx->prim_idx, ctx->prim_aabb);
      hipLaunchKernelGGL(big_alloc_kernel, dim3(bb), dim3(64), 0, s, big, nbig, ctx->bnodes, counters, nxt, counters + 1, ctx->bparent,
                         ctx->bdepth);
    }
    NOVA_HIP_CHECK(ctx, hipGetLastError());
    NOVA_HIP_CHECK(ctx, hipMemcpyAsync(&active, counters + 1, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    NOVA_HIP_CHECK(ctx, hipStreamSynchronize(s));
    std::swap(cur, nxt);
    depth++;
  }
  for (void *p : {(void *)big, (void *)chunks, (void *)small, (void *)chunk_nl, (void *)chunk_base}) hipFree(p);
  int32_t used = 0;
  NOVA_HIP_CHECK(ctx, hipMemcpy(&used, counters, sizeof(int32_t), hipMemcpyDeviceToHost));
  ctx->bnode_count = (uint32_t)used;

  // ---- collapse to 4-wide
  NOVA_HIP_CHECK(ctx, hipMalloc(&ctx->nodes4, std::max<size_t>(1, used) * sizeof(Node4)));
  int32_t *fb0 = act0, *fb1 = act1, *fs0 = rk, *fs1 = hr;  // reuse temporaries
  int32_t h_counters[2] = {1, 0};                        // wide counter, next count
  NOVA_HIP_CHECK(ctx, hipMemcpyAsync(counters + 2, h_counters, 2 * sizeof(int32_t), hipMemcpyHostToDevice, s));
  NOVA_HIP_CHECK(ctx, hipMemcpyAsync(fb0, &zero, sizeof(int32_t), hipMemcpyHostToDevice, s));
  NOVA_HIP_CHECK(ctx, hipMemcpyAsync(fs0, &zero, sizeof(int32_t), hipMemcpyHostToDevice, s));
  int32_t frontier = 1;
  uint32_t wdepth = 0;
  while (frontier > 0) {
    if (wdepth > depth + 1) return set_error(ctx, NOVA_HIP_DEVICE_ERROR, "BVH collapse did not converge");
    NOVA_HIP_CHECK(ctx, hipMemsetAsync(counters + 3, 0, sizeof(int32_t), s));
    hipLaunchKernelGGL(collapse_kernel, dim3((frontier + 127) / 128), dim3(128), 0, s, fb0, fs0, frontier, ctx->bnodes, ctx->nodes4, counters + 2,
                       fb1, fs1, counters + 3);
    NOVA_HIP_CHECK(ctx, hipGetLastError());
    NOVA_HIP_CHECK(ctx, hipMemcpyAsync(&frontier, counters + 3, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    NOVA_HIP_CHECK(ctx, hipStreamSynchronize(s));
    std::swap(fb0, fb1);
    std::swap(fs0, fs1);
    wdepth++;
  }
  int32_t wide = 0;
  NOVA_HIP_CHECK(ctx, hipMemcpy(&wide, counters + 2, sizeof(int32_t), hipMemcpyDeviceToHost));
  NOVA_HIP_CHECK(ctx, hipMalloc(&ctx->nodes4f, std::max<size_t>(1, (size_t)wide) * sizeof(Node4F)));
  hipLaunchKernelGGL(decode_nodes_kernel, dim3((wide + 255) / 256), dim3(256), 0, s, ctx->nodes4, ctx->nodes4f, (uint32_t)wide);
  NOVA_HIP_CHECK(ctx, hipMalloc(&ctx->tri_leaf, (size_t)n * sizeof(TriMT)));
  hipLaunchKernelGGL(gather_leaf_tris, dim3((n + 255) / 256), dim3(256), 0, s, ctx->prim_idx, ctx->tri_world, ctx->tri_leaf, n);
  hipLaunchKernelGGL(leaf_of_position_kernel, dim3(std::min<uint32_t>(4096, (used + 255) / 256)), dim3(256), 0, s, ctx->bnodes, (uint32_t)used,
                     ctx->leaf_of_pos, ctx->tri_leaf);
  NOVA_HIP_CHECK(ctx, hipGetLastError());
  NOVA_HIP_CHECK(ctx, hipEventRecord(ev1, s));
  NOVA_HIP_CHECK(ctx, hipEventSynchronize(ev1));
  float ms = 0.f;
  hipEventElapsedTime(&ms, ev0, ev1);
  hipEventDestroy(ev0);
  hipEventDestroy(ev1);

  // leaves = binary nodes with count > 0 among the used ones
  std::vector<BNode> host(used);
  NOVA_HIP_CHECK(ctx, hipMemcpy(host.data(), ctx->bnodes, used * sizeof(BNode), hipMemcpyDeviceToHost));
  uint32_t leaves = 0;
  for (const BNode &b : host) leaves += b.count > 0 ? 1 : 0;

  for (void *p : {(void *)tmp_idx, (void *)rk, (void *)hr, (void *)tl, (void *)act0, (void *)act1, (void *)counters, (void *)cls, (void *)bounds})
    hipFree(p);

  ctx->bvh_info.primitive_count = n;
  ctx->bvh_info.binary_nodes = (uint32_t)used;
  ctx->bvh_info.binary_depth = depth;
  ctx->bvh_info.wide_nodes = (uint32_t)wide;
  ctx->bvh_info.wide_depth = wdepth;
  ctx->bvh_info.leaves = leaves;
  ctx->bvh_info.build_ms = ms;
  if (info) *info = ctx->bvh_info;
  ctx->tail_hint = 0;
  ctx->bvh_ready = true;
  ctx->scene_gen++;
  return NOVA_HIP_SUCCESS;
}

extern "C" int nova_hip_bvh_read(nova_hip_ctx *ctx, int32_t *prim_idx, void *binary_nodes) {
  if (!ctx) return NOVA_HIP_INVALID_ARGUMENT;
  if (!ctx->bvh_ready) return set_error(ctx, NOVA_HIP_SCENE_NOT_PROCESSED, "BVH not built");
  NOVA_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  if (prim_idx) NOVA_HIP_CHECK(ctx, hipMemcpy(prim_idx, ctx->prim_idx, ctx->prim_count * sizeof(int32_t), hipMemcpyDeviceToHost));
  if (binary_nodes) NOVA_HIP_CHECK(ctx, hipMemcpy(binary_nodes, ctx->bnodes, ctx->bnode_count * sizeof(BNode), hipMemcpyDeviceToHost));
  return NOVA_HIP_SUCCESS;
}
