// Wavefront path integrator and 4-wide BVH closest-hit traversal.
//
// Per batch of frames (1 spp each, as the reference's Engine renders them):
//   generate  : camera rays for every (frame, pixel) path slot, Sobol dims 0-1
//   per bounce: extend (closest hit of the active queue)
//               shade  (hit record, Principled BSDF sample, Russian roulette,
//                       spawn; survivors are compacted into the next queue with
//                       ballot/mbcnt ranks and one atomic per workgroup and iteration)
//   accumulate: accum += partial(frame); final = accum / (frame + 1), in frame order.
//
// The reference evaluates Li recursively (Integrator.cpp:59-98):
//   V_k = DENAN(Le_k + c_k * V_{k+1}),  c_k = f * |cos| / pdf,
// with V = envmap(dir) on a miss or once the depth budget is spent, and V_k = 0 when
// the BSDF sample fails. The shade stage evaluates the same sum front to back,
// L += T * Le_k; T *= c_k, where DENAN at level k is equivalent to ending the path with
// L unchanged as soon as c_k has a non-finite component (V_{k+1} is always finite).
// Alpha follows the same recursion (1 per scattering level plus the terminal alpha).
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <vector>

#include "device_math.h"
#include "device_path.h"
#include "device_queue.h"
#include "device_shading.h"
#include "device_trace.h"
#include "nova_internal.h"

namespace nova_hip {

// ------------------------------------------------------------------ kernels

constexpr int kWorkCounter = 4;  // count[] slot of the extend work counter, zeroed before each extend
constexpr int kTailSlot = 5;     // count[] slot: 0, or 1 + the bounce from which the batch's paths finish in tail_kernel

struct Queues {
  float *ray[2];      // 6 planes of cap floats each: ox oy oz dx dy dz
  int32_t *slot[2];
  uint32_t *count;    // [0], [1]: queue sizes; [2]: segments traced
  unsigned long long *stats64;  // [0] node visits, [1] triangle tests, [2] escaped segments (camera-ray launches count into [4], [5]);
                                // per-ray walk: [8] wave loop turns, [9] active lanes summed over turns, [10] wave rounds of the triangle loop;
                                // tail_kernel: [16] segments, [17] node visits, [18] triangle tests
  float4 *hit;        // closest hit per queue entry: (prim id bits, t, u, v), one 16-byte record
  float *T, *L, *A;   // path state by slot: T,L 3 planes of cap each; A alpha
  float *partial;     // frames x W x H x 4
  uint32_t cap;
};


// AbstractIntegrator::render per pixel (Integrator.h:53-95): one camera ray per path slot.
// The first shade takes the path id from the queue position and starts every path at T = 1,
// L = 0, A = 0, so neither path ids nor path state are written here.
__global__ __launch_bounds__(256) void generate_kernel(FrameParams fp, const uint32_t *__restrict__ sobol, Queues q, uint32_t active) {
  // the jitter dimensions 0-1 from per-nibble tables: 8 LDS reads per sample instead of a
  // 32-step conditional XOR over the index bits
  __shared__ uint32_t sob_tab[2 * 128];
  sobol_nibble_fill(sobol, 0, 2, sob_tab, threadIdx.x, blockDim.x);
  __syncthreads();
  const SobolNibbles sobn{sob_tab, 0};
  const uint32_t n = q.cap;  // plane stride
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < active; i += gridDim.x * blockDim.x) {
    const uint32_t slot = path_slot(fp, i);
    f3 o, d;
    camera_ray(fp, sobn, slot, o, d);
    float *r = q.ray[0];
    r[0 * (size_t)n + i] = o.x;
    r[1 * (size_t)n + i] = o.y;
    r[2 * (size_t)n + i] = o.z;
    r[3 * (size_t)n + i] = d.x;
    r[4 * (size_t)n + i] = d.y;
    r[5 * (size_t)n + i] = d.z;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    q.count[0] = active;
    q.count[1] = 0;
    q.count[kWorkCounter] = 0;
    q.count[kTailSlot] = 0;
  }
}

#ifndef NOVA_TRACE_REFILL
#define NOVA_TRACE_REFILL 56  // idle lanes that trigger a refill in the extend kernel (0: grid-stride, one ray per lane at a time)
#endif
#ifndef NOVA_TRACE_WAVES
#define NOVA_TRACE_WAVES 6  // waves per SIMD the extend kernel is compiled for (0: compiler's choice)
#endif
// Extend workgroups per CU: as many as are resident at once (a 256-thread workgroup is one
// wave on each of the CU's 4 SIMDs). A lane's first ray is fixed by its thread index, so a
// workgroup that waits for a free slot would start its rays only after the queue has drained
// and add a whole walk's latency at the end of the launch.
#ifndef NOVA_TRACE_BLOCKS_PER_CU
#define NOVA_TRACE_BLOCKS_PER_CU (NOVA_TRACE_WAVES > 0 ? NOVA_TRACE_WAVES : 6)
#endif
static_assert(kTraceBlock == 256, "the workgroups-per-CU count assumes one wave per SIMD per workgroup");
#if NOVA_TRACE_WAVES > 0
#define NOVA_TRACE_OCCUPANCY __attribute__((amdgpu_waves_per_eu(NOVA_TRACE_WAVES, NOVA_TRACE_WAVES)))
#else
#define NOVA_TRACE_OCCUPANCY
#endif

template <bool kCount>
__global__ __launch_bounds__(kTraceBlock) NOVA_TRACE_OCCUPANCY void extend_kernel(TraceCtx tc, Queues q, int which, int bounce, uint32_t tail_rays,
                                                                                  int tail_force) {
  __shared__ uint32_t lds[kLdsStack * kTraceBlock];
  uint32_t nv = 0, nt = 0;
  const uint32_t n = q.count[which];
  const float *r = q.ray[which];
  const size_t cap = q.cap;
  const uint32_t gtid = blockIdx.x * blockDim.x + threadIdx.x;
  if (tail_rays) {  // this bounce may hand the batch's remaining paths to tail_kernel
    if (q.count[kTailSlot] != 0) return;  // an earlier bounce did
    if (n <= tail_rays || tail_force) {   // forced: the host launches no later bounce (bit 31: more than tail_rays rays)
      if (gtid == 0) q.count[kTailSlot] = ((uint32_t)bounce + 1u) | (n > tail_rays ? 0x80000000u : 0u);
      return;
    }
  }
  if (gtid == 0) q.count[2] += n;
  LdsStack stk{lds, tc.gstack, tc.gstride, gtid, threadIdx.x};
#if NOVA_TRACE_REFILL > 0
  // One traversal step per loop turn; a lane whose ray has ended takes the next unclaimed
  // ray as soon as NOVA_TRACE_REFILL lanes of its wave are idle (one atomic per refill for
  // the wave), so lanes do not wait for the wave's longest ray.
  // The first rays go one per lane by thread index, without an atomic: a launch over a
  // nearly empty queue (the last bounces) then makes no claims at all, where one claim per
  // wave on the shared counter cost about 90 us per launch. Refills claim rays past those.
  const uint32_t lane = threadIdx.x & 63;
  const uint32_t nthreads = gridDim.x * blockDim.x;
  TraceState ts;
  bool have = false, exhausted = nthreads >= n;
  uint32_t ray = 0;
  uint32_t wave_turns = 0, lane_turns = 0, tri_rounds = 0;  // instrumented launches only
  for (uint32_t turns = 0; turns < (1u << 30); turns++) {
    const uint64_t idle = __ballot(!have);
    uint32_t base = gtid - lane;  // first turn: every lane idle, lane l takes ray gtid
    bool fill = turns == 0;
    if (!fill && !exhausted && __popcll(idle) >= NOVA_TRACE_REFILL) {
      const int leader = __builtin_ffsll((long long)idle) - 1;
      if ((int)lane == leader) base = nthreads + atomicAdd(&q.count[kWorkCounter], (uint32_t)__popcll(idle));
      base = __shfl(base, leader, 64);
      if (base >= n) exhausted = true;
      fill = true;
    }
    if (fill && !have) {
      const uint32_t i = base + __builtin_amdgcn_mbcnt_hi((uint32_t)(idle >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)idle, 0));
      if (i < n) {
        trace_init(ts, mk3(r[i], r[cap + i], r[2 * cap + i]), mk3(r[3 * cap + i], r[4 * cap + i], r[5 * cap + i]));
        ray = i;
        have = true;
      }
    }
    const uint64_t busy = __ballot(have);
    if (busy == 0) break;
    if (kCount && lane == 0) {
      wave_turns++;
      lane_turns += (uint32_t)__popcll(busy);
    }
    if (have && trace_step<kCount>(tc, stk, ts, 0.000001f, nv, nt)) {
      q.hit[ray] = make_float4(__builtin_bit_cast(float, (int32_t)(ts.best)), ts.tbest, ts.bu, ts.bv);
      have = false;
      if (kCount) tri_rounds += ts.tri_rounds;
    }
  }
  if (kCount) {
    atomicAdd(&q.stats64[8], (unsigned long long)wave_turns);
    atomicAdd(&q.stats64[9], (unsigned long long)lane_turns);
    atomicAdd(&q.stats64[10], (unsigned long long)tri_rounds);
  }
#else
  for (uint32_t i = gtid; i < n; i += gridDim.x * blockDim.x) {
    const f3 o = mk3(r[i], r[cap + i], r[2 * cap + i]);
    const f3 d = mk3(r[3 * cap + i], r[4 * cap + i], r[5 * cap + i]);
    float t, u, v;
    const int prim = trace_closest<kCount>(tc, stk, o, d, 0.000001f, t, u, v, nv, nt);
    q.hit[i] = make_float4(__builtin_bit_cast(float, (int32_t)(prim)), t, u, v);
  }
#endif
  if (kCount) {
    atomicAdd(&q.stats64[0], (unsigned long long)nv);
    atomicAdd(&q.stats64[1], (unsigned long long)nt);
  }
}

// Camera rays of pixel-major batches: each wave takes 64 consecutive rays (one pixel's frames
// when the batch has 64 frames) and walks the tree once for all of them (coherent_walk).
constexpr int kCoherentBlock = 256;
#ifndef NOVA_COHERENT_BLOCKS_PER_CU
#define NOVA_COHERENT_BLOCKS_PER_CU 64  // many more workgroups than fit at once: a static split of the packets balances better
#endif
#ifndef NOVA_COHERENT_WAVES
#define NOVA_COHERENT_WAVES 0  // waves per SIMD the coherent walk is compiled for (0: compiler's choice)
#endif
#if NOVA_COHERENT_WAVES > 0
#define NOVA_COHERENT_OCCUPANCY __attribute__((amdgpu_waves_per_eu(NOVA_COHERENT_WAVES, NOVA_COHERENT_WAVES)))
#else
#define NOVA_COHERENT_OCCUPANCY
#endif
template <bool kCount>
__global__ __launch_bounds__(kCoherentBlock) NOVA_COHERENT_OCCUPANCY void extend_coherent_kernel(TraceCtx tc, Queues q, int which) {
  uint32_t nv = 0, nt = 0;
  const uint32_t n = q.count[which];
  const float *r = q.ray[which];
  const size_t cap = q.cap;
  const uint32_t lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (blockIdx.x == 0 && threadIdx.x == 0) q.count[2] += n;
  for (uint32_t base = (blockIdx.x * (kCoherentBlock / 64) + wv) * 64u; base < n; base += gridDim.x * kCoherentBlock) {
    const uint32_t i = base + lane;
    const bool valid = i < n;
    TraceState ts;
    trace_init(ts, valid ? mk3(r[i], r[cap + i], r[2 * cap + i]) : mk3(0.f, 0.f, 0.f),
               valid ? mk3(r[3 * cap + i], r[4 * cap + i], r[5 * cap + i]) : mk3(0.f, 0.f, 1.f));
    coherent_walk<kCount>(tc, ts, __ballot(valid), 0.000001f, nv, nt);
    if (valid) {
      q.hit[i] = make_float4(__builtin_bit_cast(float, (int32_t)(ts.best)), ts.tbest, ts.bu, ts.bv);
    }
  }
  if (kCount) {
    atomicAdd(&q.stats64[0], (unsigned long long)nv);
    atomicAdd(&q.stats64[1], (unsigned long long)nt);
  }
}

// The adjoint replay reuses the primal pass's hits of its first bounces, in a cache of at
// most this size whatever the job (16 B per path and cached bounce).
#ifndef NOVA_HIT_CACHE_MB
#define NOVA_HIT_CACHE_MB 2048
#endif
constexpr size_t kHitCacheBytes = (size_t)NOVA_HIT_CACHE_MB << 20;

// A finished path's radiance and alpha, by path id.
__device__ __forceinline__ void write_partial(const Queues &q, uint32_t pid, float lx, float ly, float lz, float a) {
  reinterpret_cast<float4 *>(q.partial)[pid] = make_float4(lx, ly, lz, a);
}

#ifndef NOVA_SHADE_WAVES
#define NOVA_SHADE_WAVES 4  // waves per SIMD the shade kernel is compiled for (0: compiler's choice, 135 VGPRs and 3 waves;
                            // 4 waves spill one VGPR and shade 11% faster)
#endif
#if NOVA_SHADE_WAVES > 0
#define NOVA_SHADE_OCCUPANCY __attribute__((amdgpu_waves_per_eu(NOVA_SHADE_WAVES, NOVA_SHADE_WAVES)))
#else
#define NOVA_SHADE_OCCUPANCY
#endif
#ifndef NOVA_SHADE_BLOCKS_PER_CU
#define NOVA_SHADE_BLOCKS_PER_CU 0  // shade workgroups per CU (0: as many as are resident at once)
#endif
constexpr int kShadeBlock = 256;  // 512 and 1024 measured slower (barrier waits)

template <bool kAdj>
__global__ __launch_bounds__(kShadeBlock) NOVA_SHADE_OCCUPANCY void shade_kernel(ShadeParams sp, ShadeCtx sc, const uint32_t *__restrict__ sobol, Queues q, int which,
                                                    AdjointCtx adj) {
  if (q.count[kTailSlot] != 0) return;  // tail_kernel finishes this batch's paths
  // nibble tables of this scatter's sampler dimensions (dim0 .. dim0+4)
  __shared__ uint32_t sob_tab[5 * 128];
  sobol_nibble_fill(sobol, sp.dim0, 5, sob_tab, threadIdx.x, blockDim.x);
  // the replay's texel-gradient table (an array of one element in the primal kernel)
  __shared__ uint32_t gkeys[kAdj ? kGradSlots : 1];
  __shared__ float gvals[kAdj ? 3 * kGradSlots : 1];
  __shared__ uint16_t glist[kAdj ? kGradSlots : 1];
  __shared__ uint32_t gcount[2];  // list lengths of even and odd iterations
  if (kAdj) {
    for (int e = threadIdx.x; e < kGradSlots; e += blockDim.x) {
      gkeys[e] = kGradEmpty;
      gvals[3 * e] = gvals[3 * e + 1] = gvals[3 * e + 2] = 0.f;
    }
    if (threadIdx.x < 2) gcount[threadIdx.x] = 0;
    adj.table_keys = gkeys;
    adj.table_vals = gvals;
    adj.table_list = glist;
  }
  __syncthreads();
  // Iteration it's taken slots go to device memory (one atomic per texel channel) and are
  // cleared; the next iteration's count starts at zero (it was read in iteration it - 1,
  // before that iteration's closing barrier).
  auto flush_table = [&](uint32_t it) {
    const uint32_t used = gcount[it & 1u];
    for (uint32_t e = threadIdx.x; e < used; e += blockDim.x) {
      const uint32_t h = glist[e];
      const uint32_t key = gkeys[h];
      for (int c = 0; c < 3; c++) {
        const float v = gvals[3 * h + c];
        if (v != 0.f) add_grad(adj.grad + key + c, v);
        gvals[3 * h + c] = 0.f;
      }
      gkeys[h] = kGradEmpty;
    }
    if (threadIdx.x == 0) gcount[(it + 1u) & 1u] = 0;
  };
  const SobolNibbles sobn{sob_tab, sp.dim0};
  const uint32_t n = q.count[which];
  const size_t cap = q.cap;
  const float *rin = q.ray[which];
  float *rout = q.ray[which ^ 1];
  int32_t *sout = q.slot[which ^ 1];
  const uint32_t stride = gridDim.x * blockDim.x;
  // every lane of a wave runs the same number of iterations so the ballot below is uniform
  const uint32_t iters = (n + stride - 1) / stride;
  uint32_t escapes = 0;  // wave-uniform count of escaped paths, added to the statistics once
  __shared__ uint32_t wave_cnt[2][kShadeBlock / 64], blk_base[2];
  for (uint32_t it = 0; it < iters; it++) {
    const uint32_t i = it * stride + blockIdx.x * blockDim.x + threadIdx.x;
    bool cont = false, escaped = false;
    f3 no = mk3(0, 0, 0), nd = mk3(0, 0, 0);
    int32_t pid = -1;
    if (i < n) {
      const bool first = sp.bounce == 0;  // camera rays: path id = queue position, fresh path state
      pid = first ? (int32_t)i : q.slot[which][i];
      const f3 o = mk3(rin[i], rin[cap + i], rin[2 * cap + i]);
      const f3 d = mk3(rin[3 * cap + i], rin[4 * cap + i], rin[5 * cap + i]);
      f3 T = first ? mk3(1.f, 1.f, 1.f) : mk3(q.T[pid], q.T[cap + pid], q.T[2 * cap + pid]);
      f3 L = first ? mk3(0.f, 0.f, 0.f) : mk3(q.L[pid], q.L[cap + pid], q.L[2 * cap + pid]);
      float A = first ? 0.f : q.A[pid];
      int prim;
      float ht, hu, hv;
      AdjointCtx padj = adj;
      padj.table_count = &gcount[it & 1u];
      if (kAdj) {  // this path's primal radiance (and hits), left by pass 0 under the same path id
        const size_t pid0 = (uint32_t)pid;
        padj.ltot = adj.ltot + pid0 * 4;
        if (adj.hit_cache && (uint32_t)sp.bounce < adj.cache_bounces) {
          // entries a primal path did not reach hold prim -1 (a miss): always in bounds
          const float4 h = adj.hit_cache[(size_t)sp.bounce * adj.cache_paths + pid0];
          prim = __builtin_bit_cast(int32_t, h.x);
          ht = h.y;
          hu = h.z;
          hv = h.w;
        } else {
          const float4 h = q.hit[i];
          prim = __builtin_bit_cast(int32_t, h.x);
          ht = h.y;
          hu = h.z;
          hv = h.w;
        }
      } else {
        const float4 h = q.hit[i];
        prim = __builtin_bit_cast(int32_t, h.x);
        ht = h.y;
        hu = h.z;
        hv = h.w;
        if (adj.hit_cache && (uint32_t)sp.bounce < adj.cache_bounces)
          adj.hit_cache[(size_t)sp.bounce * adj.cache_paths + pid] = make_float4(__builtin_bit_cast(float, prim), ht, hu, hv);
      }
      cont = shade_path<kAdj>(sp, sc, sobn, path_slot(sp.fp, (uint32_t)pid), o, d, prim, ht, hu, hv, T, L, A, no, nd, escaped, &padj);
      const bool done = !cont;
      q.T[pid] = T.x;
      q.T[cap + pid] = T.y;
      q.T[2 * cap + pid] = T.z;
      if (done) {
        if (!kAdj) write_partial(q, (uint32_t)pid, L.x, L.y, L.z, A);  // a replay reads the primal's
      } else {
        q.L[pid] = L.x;
        q.L[cap + pid] = L.y;
        q.L[2 * cap + pid] = L.z;
        q.A[pid] = A;
      }
    }
    escapes += (uint32_t)__popcll(__ballot(escaped));
    // compact survivors into the next queue
    const uint64_t m = __ballot(cont);
    const uint32_t lane_rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0));
    const uint32_t base = block_reserve<kShadeBlock>(m, &q.count[which ^ 1], it, wave_cnt, blk_base);
    if (kAdj) flush_table(it);  // every lane has splatted this iteration (block_reserve waited)
    if (cont) {
      const uint32_t j = base + lane_rank;
      rout[j] = no.x;
      rout[cap + j] = no.y;
      rout[2 * cap + j] = no.z;
      rout[3 * cap + j] = nd.x;
      rout[4 * cap + j] = nd.y;
      rout[5 * cap + j] = nd.z;
      sout[j] = pid;
    }
    if (kAdj) __syncthreads();  // the table is clear before the next iteration's splats
  }
  if (escapes != 0 && (threadIdx.x & 63) == 0) atomicAdd(&q.stats64[2], (unsigned long long)escapes);
}

// The last bounces of a batch, once few paths remain (the extend launch of bounce `sp.bounce`
// found at most tail_rays rays and recorded it in count[kTailSlot]): one path per lane from
// that bounce to the end, the walk (trace_closest) and shade_path in turn, the radiance
// written when the path ends. Per-bounce launches would each wait for their slowest ray;
// here a path's bounces follow each other at once. Every path takes the same steps with the
// same arithmetic as in the per-bounce launches, so the image is the same. Dynamic LDS: the
// traversal stacks, then the nibble tables of the sampler dimensions from sp.dim0 on.
#ifndef NOVA_TAIL_WAVES
#define NOVA_TAIL_WAVES 0  // waves per SIMD tail_kernel is compiled for (0: the compiler's choice, 3 at its 132 registers)
#endif
#if NOVA_TAIL_WAVES > 0
#define NOVA_TAIL_OCCUPANCY __attribute__((amdgpu_waves_per_eu(NOVA_TAIL_WAVES, NOVA_TAIL_WAVES)))
#else
#define NOVA_TAIL_OCCUPANCY
#endif
template <bool kCount>
__global__ __launch_bounds__(kTraceBlock) NOVA_TAIL_OCCUPANCY void tail_kernel(TraceCtx tc, ShadeParams sp, int max_depth, int rr_start, ShadeCtx sc,
                                                          const uint32_t *__restrict__ sobol, Queues q, int which, int ndims) {
  extern __shared__ uint32_t tail_lds[];
  if ((q.count[kTailSlot] & 0x7FFFFFFFu) != (uint32_t)sp.bounce + 1u) return;
  uint32_t *sob_tab = tail_lds + kLdsStack * kTraceBlock;
  sobol_nibble_fill(sobol, sp.dim0, ndims, sob_tab, threadIdx.x, blockDim.x);
  __syncthreads();
  const SobolNibbles sobn{sob_tab, sp.dim0};
  const uint32_t n = q.count[which];
  const float *rin = q.ray[which];
  const size_t cap = q.cap;
  const uint32_t gtid = blockIdx.x * blockDim.x + threadIdx.x;
  LdsStack stk{tail_lds, tc.gstack, tc.gstride, gtid, threadIdx.x};
  uint32_t nv = 0, nt = 0, segs = 0, escapes = 0;
  for (uint32_t i = gtid; i < n; i += gridDim.x * blockDim.x) {
    const int32_t pid = q.slot[which][i];
    f3 o = mk3(rin[i], rin[cap + i], rin[2 * cap + i]);
    f3 d = mk3(rin[3 * cap + i], rin[4 * cap + i], rin[5 * cap + i]);
    f3 T = mk3(q.T[pid], q.T[cap + pid], q.T[2 * cap + pid]);
    f3 L = mk3(q.L[pid], q.L[cap + pid], q.L[2 * cap + pid]);
    float A = q.A[pid];
    const uint32_t slot = path_slot(sp.fp, (uint32_t)pid);
    ShadeParams s = sp;
    for (int k = sp.bounce; k <= max_depth; k++) {
      s.bounce = k;
      s.last = k == max_depth ? 1 : 0;
      s.rr = (rr_start >= 0 && k >= rr_start) ? 1 : 0;
      float t, u, v;
      const int prim = trace_closest<kCount>(tc, stk, o, d, 0.000001f, t, u, v, nv, nt);
      segs++;
      f3 no = mk3(0, 0, 0), nd = mk3(0, 0, 0);
      bool escaped = false;
      const bool cont = shade_path<false>(s, sc, sobn, slot, o, d, prim, t, u, v, T, L, A, no, nd, escaped);
      escapes += escaped ? 1u : 0u;
      if (!cont) break;
      s.dim0 += 4u + (s.rr ? 1u : 0u);
      o = no;
      d = nd;
    }
    write_partial(q, (uint32_t)pid, L.x, L.y, L.z, A);
  }
  if (segs) {
    atomicAdd(&q.count[2], segs);
    atomicAdd(&q.stats64[16], (unsigned long long)segs);
  }
  if (escapes) atomicAdd(&q.stats64[2], (unsigned long long)escapes);
  if (kCount && (nv | nt)) {
    atomicAdd(&q.stats64[0], (unsigned long long)nv);
    atomicAdd(&q.stats64[1], (unsigned long long)nt);
    atomicAdd(&q.stats64[17], (unsigned long long)nv);
    atomicAdd(&q.stats64[18], (unsigned long long)nt);
  }
}

// Between an extend and a shade launch: the shade's output queue and the next extend's work
// counter start at zero.
__global__ void reset_count_kernel(uint32_t *count, int which) {
  count[which] = 0;
  count[kWorkCounter] = 0;
}

// Engine::writeFramebuffer per frame, in frame order (engine.cpp:351-363); one thread per
// pixel channel, the pixel's frames contiguous in the path-id-ordered partial buffer.
// frames_out (optional) receives every frame's radiance image, frame f at f * npix * 4 floats.
__global__ void accumulate_kernel(const float *__restrict__ partial, FrameParams fp, float *__restrict__ accum, float *__restrict__ final_out,
                                  float *__restrict__ partial_out, float *__restrict__ frames_out) {
  const uint32_t npix4 = (uint32_t)fp.width * (uint32_t)fp.height * 4u;
  const int frames = fp.frames;
  for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < npix4; e += gridDim.x * blockDim.x) {
    const uint32_t p = e >> 2, c = e & 3u;  // p: the pixel's rank in the path numbering
    const uint32_t i = pixel_offset(fp, pixel_of_rank(fp, p)) + c;
    const float *src = partial + (size_t)p * frames * 4 + c;
    if (accum) {  // without accum/final: the caller accumulates (the drop-in's per-frame traverse)
      float a = accum[i];
      for (int f = 0; f < frames; f++) a += src[(size_t)f * 4];  // frame order
      accum[i] = a;
      final_out[i] = a / float(fp.frame0 + frames);  // the last frame's average, the one displayed
    }
    if (partial_out) partial_out[i] = src[(size_t)(frames - 1) * 4];
    if (frames_out)
      for (int f = 0; f < frames; f++) frames_out[(size_t)f * npix4 + i] = src[(size_t)f * 4];
  }
}

// accumulate_kernel for batches of many frames: a workgroup takes 64 pixels and brings their
// frames into LDS 16 at a time with contiguous 16-byte loads (a pixel's frames are adjacent
// in the partial buffer), where one thread per pixel channel reading its own frames strides
// over 1 KiB per pixel. Each channel still adds its frames one by one in frame order, so the
// sums are those of accumulate_kernel bit for bit.
#ifndef NOVA_ACC_TILED_FRAMES
#define NOVA_ACC_TILED_FRAMES 8  // batches of at least this many frames accumulate through LDS tiles
#endif
constexpr int kAccPix = 64, kAccChunk = 16, kAccRow = kAccChunk * 4 + 4;  // row padded against bank conflicts
static_assert(kAccChunk == 16 && kAccRow % 4 == 0, "chunk index split by shifts; rows of whole float4s");
__global__ __launch_bounds__(256) void accumulate_tiled_kernel(const float *__restrict__ partial, FrameParams fp, float *__restrict__ accum,
                                                               float *__restrict__ final_out, float *__restrict__ partial_out,
                                                               float *__restrict__ frames_out) {
  __shared__ __align__(16) float tile[kAccPix * kAccRow];
  const uint32_t npix = (uint32_t)fp.width * (uint32_t)fp.height, npix4 = npix * 4u;
  const int frames = fp.frames;
  const float4 *src4 = reinterpret_cast<const float4 *>(partial);
  const uint32_t q = threadIdx.x >> 2, c = threadIdx.x & 3u;
  for (uint32_t p0 = blockIdx.x * kAccPix; p0 < npix; p0 += gridDim.x * kAccPix) {  // uniform per workgroup
    const uint32_t p = p0 + q;  // this thread's pixel rank
    const bool valid = p < npix;
    const uint32_t i = valid ? pixel_offset(fp, pixel_of_rank(fp, p)) + c : 0u;
    float a = (valid && accum) ? accum[i] : 0.f, last = 0.f;
    for (int f0 = 0; f0 < frames; f0 += kAccChunk) {
      const int nf = min(kAccChunk, frames - f0);
      __syncthreads();  // the previous chunk has been read
      for (int j = threadIdx.x; j < kAccPix * nf; j += blockDim.x) {
        // a full chunk (the common case) splits the index with shifts
        const uint32_t qq = nf == kAccChunk ? (uint32_t)j >> 4 : (uint32_t)j / (uint32_t)nf;
        const uint32_t fl = nf == kAccChunk ? (uint32_t)j & 15u : (uint32_t)j % (uint32_t)nf;
        if (p0 + qq < npix)
          *reinterpret_cast<float4 *>(tile + qq * kAccRow + fl * 4) = src4[(size_t)(p0 + qq) * frames + f0 + fl];
      }
      __syncthreads();
      if (valid) {
        const float *row = tile + q * kAccRow + c;
        for (int fl = 0; fl < nf; fl++) {
          const float v = row[fl * 4];
          a += v;  // frame order
          if (frames_out) frames_out[(size_t)(f0 + fl) * npix4 + i] = v;
        }
        last = row[(nf - 1) * 4];
      }
    }
    if (valid) {
      if (accum) {
        accum[i] = a;
        final_out[i] = a / float(fp.frame0 + frames);
      }
      if (partial_out) partial_out[i] = last;
    }
  }
}

// One kept frame image into the caller's buffers, as accumulate_kernel adds a one-frame
// batch: accum += image, final = accum / (frame + 1), partial = image.
__global__ void frame_add_kernel(const float *__restrict__ image, uint32_t n4, float *__restrict__ accum, float *__restrict__ final_out,
                                 float *__restrict__ partial_out, float frames_so_far) {
  for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < n4; e += gridDim.x * blockDim.x) {
    const float v = image[e];
    if (accum) {
      float a = accum[e];
      a += v;
      accum[e] = a;
      final_out[e] = a / frames_so_far;
    }
    if (partial_out) partial_out[e] = v;
  }
}

// closest-hit entry point: rays n x 6 (AoS)
__global__ __launch_bounds__(kTraceBlock) void closest_hit_kernel(TraceCtx tc, const float *__restrict__ rays, uint64_t n, int32_t *__restrict__ prim,
                                                                  float *__restrict__ tuv) {
  __shared__ uint32_t lds[kLdsStack * kTraceBlock];
  const uint32_t gtid = blockIdx.x * blockDim.x + threadIdx.x;
  for (uint64_t i = gtid; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    const f3 o = mk3(rays[i * 6], rays[i * 6 + 1], rays[i * 6 + 2]);
    const f3 d = mk3(rays[i * 6 + 3], rays[i * 6 + 4], rays[i * 6 + 5]);
    float t, u, v;
    uint32_t nv = 0, nt = 0;
    LdsStack st{lds, tc.gstack, tc.gstride, gtid, threadIdx.x};
    const int p = trace_closest<false>(tc, st, o, d, 0.000001f, t, u, v, nv, nt);
    prim[i] = p;
    tuv[i * 3] = p >= 0 ? t : 0.f;
    tuv[i * 3 + 1] = u;
    tuv[i * 3 + 2] = v;
  }
}

// Render passes (NORMAL / DEPTH / EMISSIVE): one unjittered ray per pixel, the pass value
// (pass_value) accumulated for every frame as the path integrator's frames are.
__global__ __launch_bounds__(kTraceBlock) void pass_kernel(TraceCtx tc, ShadeCtx sc, const uint32_t *__restrict__ sobol, FrameParams fp, int pass,
                                                           float z_near, float z_far, float *__restrict__ accum, float *__restrict__ final_out,
                                                           float *__restrict__ partial_out) {
  __shared__ uint32_t lds[kLdsStack * kTraceBlock];
  const uint32_t npix = (uint32_t)fp.width * (uint32_t)fp.height;
  const uint32_t gtid = blockIdx.x * blockDim.x + threadIdx.x;
  LdsStack stk{lds, tc.gstack, tc.gstride, gtid, threadIdx.x};
  for (uint32_t p = gtid; p < npix; p += gridDim.x * blockDim.x) {
    f3 o, d;
    camera_ray_pass(fp, p, o, d);
    float t, u, v;
    uint32_t nv = 0, nt = 0;
    const int prim = trace_closest<false>(tc, stk, o, d, 0.000001f, t, u, v, nv, nt);
    const float4 val = pass_value(pass, fp, sc, sobol, p, o, d, prim, t, u, v, z_near, z_far);
    const uint32_t i = pixel_offset(fp, p);
    float a[4] = {accum[i], accum[i + 1], accum[i + 2], accum[i + 3]}, fin[4] = {0.f, 0.f, 0.f, 0.f};
    const float vv[4] = {val.x, val.y, val.z, val.w};
    for (int f = 0; f < fp.frames; f++)  // Engine::writeFramebuffer per frame, in frame order
      for (int c = 0; c < 4; c++) {
        a[c] += vv[c];
        fin[c] = a[c] / float(fp.frame0 + f + 1);
      }
    for (int c = 0; c < 4; c++) {
      accum[i + c] = a[c];
      final_out[i + c] = fin[c];
      if (partial_out) partial_out[i + c] = vv[c];
    }
  }
}

#ifndef NOVA_PATHS_PER_BATCH
#define NOVA_PATHS_PER_BATCH (128u << 20)
#endif
constexpr size_t kPathsPerBatch = NOVA_PATHS_PER_BATCH;

static int device_cus(nova_hip_ctx *ctx) {
  int cus = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device) != hipSuccess || cus <= 0) cus = 256;
  return cus;
}

static int ensure_gstack(nova_hip_ctx *ctx, uint32_t threads) {
  const uint32_t bound = traversal_stack_bound(ctx);
  const size_t need = bound > (uint32_t)kLdsStack ? (size_t)(bound - kLdsStack) * threads : 1;
  if (ctx->gstack && ctx->gstack_entries >= need) return NOVA_HIP_SUCCESS;
  if (ctx->gstack) {
    (void)hipDeviceSynchronize();  // queued work of a call that did not wait may still use it
    hipFree(ctx->gstack);
  }
  ctx->gstack = nullptr;
  ctx->gstack_entries = 0;
  NOVA_HIP_CHECK(ctx, hipMalloc(&ctx->gstack, need * sizeof(uint32_t)));
  ctx->gstack_entries = need;
  return NOVA_HIP_SUCCESS;
}

// ------------------------------------------------------------ no-allocation scratch
// Device memory of nova_hip_render_frames_noalloc: queues, hits, path state, radiance and
// traversal overflow stacks of one batch, carved from a fixed array instead of grown with
// hipMalloc. Each device the library runs on holds its own copy until the process ends.
constexpr size_t kNoAllocScratchBytes = (size_t)320 << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mutex;
static void *g_scratch_ptr[kMaxDevices];
static hipError_t g_scratch_err[kMaxDevices];
static bool g_scratch_done[kMaxDevices];

// Once per device, with that device current: the device address of `scratch` (host code
// must not use the array's name as an address).
int noalloc_scratch_setup(int device) {
  if (device < 0 || device >= kMaxDevices) return NOVA_HIP_INVALID_ARGUMENT;
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_done[device]) {
    void *ptr = nullptr;
    g_scratch_err[device] = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    g_scratch_ptr[device] = g_scratch_err[device] == hipSuccess ? ptr : nullptr;
    g_scratch_done[device] = true;
  }
  return g_scratch_ptr[device] ? NOVA_HIP_SUCCESS : NOVA_HIP_DEVICE_ERROR;
}

static void *noalloc_scratch_of(int device) {
  if (device < 0 || device >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  return g_scratch_done[device] ? g_scratch_ptr[device] : nullptr;
}

#ifndef NOVA_TAIL_RAYS
#define NOVA_TAIL_RAYS 262144  // a bounce with at most this many rays hands the batch's paths to tail_kernel (0: never)
#endif
// NOVA_TAIL_RAYS, or the NOVA_HIP_TAIL_RAYS environment variable when set (read once per
// process; a tuning knob).
static uint32_t tail_rays_setting() {
  static const uint32_t v = [] {
    const char *e = std::getenv("NOVA_HIP_TAIL_RAYS");
    return e ? (uint32_t)std::strtoul(e, nullptr, 10) : (uint32_t)NOVA_TAIL_RAYS;
  }();
  return v;
}

#ifndef NOVA_TILE_FRAMES
#define NOVA_TILE_FRAMES 64  // batches of fewer frames take their pixels in 8x8 tiles (pixel_of_rank)
#endif
// NOVA_TILE_FRAMES, or the NOVA_HIP_TILE_FRAMES environment variable when set (read once per
// process; a tuning knob).
static int tile_frames_setting() {
  static const int v = [] {
    const char *e = std::getenv("NOVA_HIP_TILE_FRAMES");
    return e ? (int)std::strtol(e, nullptr, 10) : (int)NOVA_TILE_FRAMES;
  }();
  return v;
}

// Counters start at zero through a kernel, not a memset, so the call can be captured.
__global__ void zero_words_kernel(uint32_t *p, uint32_t n) {
  for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0u;
}

// Bump allocator over the fixed array (256-byte aligned pieces).
struct ScratchCarve {
  unsigned char *base;
  size_t used = 0, limit = 0;
  void *take(size_t bytes) {
    const size_t at = (used + 255) / 256 * 256;
    if (at + bytes > limit) return nullptr;
    used = at + bytes;
    return base + at;
  }
};

// Bytes a batch of `cap` paths needs in the fixed array: the queues and path state (116 B
// per path), the counters and the traversal overflow stacks of `threads` lanes.
static size_t noalloc_bytes(size_t cap, uint32_t threads, uint32_t stack_bound) {
  const size_t per_path[] = {cap * 12 * sizeof(float), cap * 2 * sizeof(int32_t), cap * 4 * sizeof(float), cap * 7 * sizeof(float),
                             cap * 4 * sizeof(float)};
  size_t total = 256 + 256;  // counters
  for (size_t b : per_path) total += (b + 255) / 256 * 256;
  const size_t g = stack_bound > (uint32_t)kLdsStack ? (size_t)(stack_bound - kLdsStack) * threads : 1;
  return total + (g * sizeof(uint32_t) + 255) / 256 * 256;
}

}  // namespace nova_hip

using namespace nova_hip;

extern "C" int nova_hip_closest_hit(nova_hip_ctx *ctx, const float *rays_dev, uint64_t n, int32_t *prim_dev, float *tuv_dev, void *stream) {
  if (!ctx) return NOVA_HIP_INVALID_ARGUMENT;
  if (!ctx->bvh_ready) return set_error(ctx, NOVA_HIP_SCENE_NOT_PROCESSED, "BVH not built");
  if (n == 0) return NOVA_HIP_SUCCESS;
  if (!rays_dev || !prim_dev || !tuv_dev) return set_error(ctx, NOVA_HIP_INVALID_ARGUMENT, "null buffer");
  NOVA_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  const uint32_t blocks = (uint32_t)std::min<uint64_t>((n + kTraceBlock - 1) / kTraceBlock, (uint64_t)device_cus(ctx) * 8);
  const uint32_t threads = blocks * kTraceBlock;
  int st = ensure_gstack(ctx, threads);
  if (st) return st;
  TraceCtx tc{ctx->nodes4, ctx->tri_leaf, ctx->gstack, threads, ctx->tie};
  hipLaunchKernelGGL(closest_hit_kernel, dim3(blocks), dim3(kTraceBlock), 0, (hipStream_t)stream, tc, rays_dev, n, prim_dev, tuv_dev);
  NOVA_HIP_CHECK(ctx, hipGetLastError());
  return NOVA_HIP_SUCCESS;
}

// integrator_dispatch (Integrator.cpp:17-45): 0 for the path integrator, else the PASS_* that
// writes the frame last.
static int resolve_integrator(nova_hip_ctx *ctx, int32_t flags, int *pass) {
  *pass = 0;
  if (flags == 0) return NOVA_HIP_SUCCESS;
  if (!(flags & NOVA_HIP_INTEGRATOR_PATH)) return set_error(ctx, NOVA_HIP_RENDERER_EXCEPTION, "INVALID_INTEGRATOR: integrator_flags without PATH");
  if (flags & NOVA_HIP_INTEGRATOR_EMISSIVE)
    *pass = PASS_EMISSIVE;
  else if (flags & NOVA_HIP_INTEGRATOR_DEPTH)
    *pass = PASS_DEPTH;
  else if (flags & NOVA_HIP_INTEGRATOR_NORMAL)
    *pass = PASS_NORMAL;
  else if (!(flags & NOVA_HIP_INTEGRATOR_COMBINED))
    return set_error(ctx, NOVA_HIP_RENDERER_EXCEPTION, "INVALID_RENDER_MODE: no render pass in integrator_flags");
  return NOVA_HIP_SUCCESS;
}

// The context's start and end events, made on first use.
static int call_events(nova_hip_ctx *ctx, hipEvent_t *ev0, hipEvent_t *ev1) {
  if (!ctx->ev_begin) NOVA_HIP_CHECK(ctx, hipEventCreate(&ctx->ev_begin));
  if (!ctx->ev_end) NOVA_HIP_CHECK(ctx, hipEventCreate(&ctx->ev_end));
  *ev0 = ctx->ev_begin;
  *ev1 = ctx->ev_end;
  return NOVA_HIP_SUCCESS;
}

static int render_pass(nova_hip_ctx *ctx, const nova_hip_render_params *p, int pass, float *accum_dev, float *final_dev, float *partial_dev,
                       hipStream_t s, bool async) {
  const uint32_t npix = (uint32_t)p->width * (uint32_t)p->height;
  const uint32_t blocks = std::min<uint32_t>((npix + kTraceBlock - 1) / kTraceBlock, (uint32_t)device_cus(ctx) * 8);
  const uint32_t threads = blocks * kTraceBlock;
  int st = ensure_gstack(ctx, threads);
  if (st) return st;
  TraceCtx tc{ctx->nodes4, ctx->tri_leaf, ctx->gstack, threads, ctx->tie};
  ShadeCtx sc{ctx->textures, ctx->tex_u32, ctx->tex_f32, ctx->materials, ctx->mesh_xf, ctx->shade, ctx->acos_lut, ctx->envmap, -1};
  FrameParams fp;
  fp.width = p->width;
  fp.height = p->height;
  fp.frame0 = p->frame_start;
  fp.frames = p->frame_count;
  fp.frame_minor = 0;
  fp.tiled = 0;
  fp.offset = 0;
  fp.vinvert = p->vertical_invert;
  fp.policy = NOVA_HIP_INDEX_CPU;
  for (int i = 0; i < 16; i++) {
    fp.inv_P[i] = p->inv_P[i];
    fp.inv_V[i] = p->inv_V[i];
  }
  if (async) {  // queued; statistics untouched
    hipLaunchKernelGGL(pass_kernel, dim3(blocks), dim3(kTraceBlock), 0, s, tc, sc, ctx->sobol_table, fp, pass, p->z_near, p->z_far, accum_dev,
                       final_dev, partial_dev);
    NOVA_HIP_CHECK(ctx, hipGetLastError());
    return NOVA_HIP_SUCCESS;
  }
  hipEvent_t ev0, ev1;
  int est = call_events(ctx, &ev0, &ev1);
  if (est) return est;
  NOVA_HIP_CHECK(ctx, hipEventRecord(ev0, s));
  hipLaunchKernelGGL(pass_kernel, dim3(blocks), dim3(kTraceBlock), 0, s, tc, sc, ctx->sobol_table, fp, pass, p->z_near, p->z_far, accum_dev, final_dev,
                     partial_dev);
  NOVA_HIP_CHECK(ctx, hipGetLastError());
  NOVA_HIP_CHECK(ctx, hipEventRecord(ev1, s));
  NOVA_HIP_CHECK(ctx, hipEventSynchronize(ev1));
  float ms = 0.f;
  NOVA_HIP_CHECK(ctx, hipEventElapsedTime(&ms, ev0, ev1));
  ctx->stats = nova_hip_render_stats{};
  ctx->stats.total_ms = ms;
  ctx->stats.segments = npix;
  ctx->stats.extend_launches = 1;
  return NOVA_HIP_SUCCESS;
}

// Renders (adj == nullptr) or renders and replays for the adjoint, batch by batch.
// With noalloc (the device address of the fixed scratch array) nothing is allocated, no
// event is recorded and nothing waits: the work is ordered on the stream and the render
// statistics are left as they were.
static int render_impl(nova_hip_ctx *ctx, const nova_hip_render_params *p, float *accum_dev, float *final_dev, float *partial_dev, void *stream,
                       const AdjointCtx *adj, unsigned char *noalloc = nullptr, float *frames_out = nullptr) {
  if (!ctx || !p) return NOVA_HIP_INVALID_ARGUMENT;
  if (!ctx->bvh_ready) return set_error(ctx, NOVA_HIP_SCENE_NOT_PROCESSED, "scene/BVH not built");
  if (p->width <= 0 || p->height <= 0 || p->width > 65536 || p->height > 65536)
    return set_error(ctx, NOVA_HIP_INVALID_RENDERBUFFER_DIM, "Wrong buffer width / height dimension.");
  if (p->frame_count <= 0 || p->frame_start < 0 || p->max_depth < 0 || (int64_t)p->frame_start + p->frame_count > INT32_MAX)
    return set_error(ctx, NOVA_HIP_INVALID_ARGUMENT, "bad frame range or depth");
  if (p->index_policy != NOVA_HIP_INDEX_CPU && p->index_policy != NOVA_HIP_INDEX_OPTIX)
    return set_error(ctx, NOVA_HIP_INVALID_ARGUMENT, "unknown index_policy");
  // queued without a wait: nothing is read back, so neither timing nor counting
  const bool async = (p->flags & NOVA_HIP_RENDER_ASYNC) != 0;
  if (async && (adj || (p->flags & (NOVA_HIP_RENDER_TIMING | NOVA_HIP_RENDER_COUNT_TRAVERSAL))))
    return set_error(ctx, NOVA_HIP_INVALID_ARGUMENT, "ASYNC reads nothing back: not with TIMING, COUNT_TRAVERSAL or the adjoint");
  // accum and final together, or neither with a partial buffer (one frame's radiance only)
  // or the per-frame images
  if (!adj && (!accum_dev != !final_dev || (!accum_dev && !partial_dev && !frames_out)))
    return set_error(ctx, NOVA_HIP_INVALID_BUFFER_STATE, "null render buffer");
  if ((((uintptr_t)accum_dev | (uintptr_t)final_dev | (uintptr_t)partial_dev) & 3u) != 0)
    return set_error(ctx, NOVA_HIP_INVALID_BUFFER_STATE, "render buffers must be 4-byte aligned float arrays");
  if (p->sampling_roughness_texture >= 0) {
    if ((size_t)p->sampling_roughness_texture >= ctx->host_textures.size() ||
        ctx->host_textures[p->sampling_roughness_texture].format != NOVA_HIP_TEX_F32)
      return set_error(ctx, NOVA_HIP_INVALID_ARGUMENT, "sampling_roughness_texture is not an FP32 texture of the scene");
  }
  // sampler dimensions: 2 + sum over scatters of (4 + RR roll) must stay inside the table
  {
    uint32_t dims = 2;
    for (int k = 0; k <= p->max_depth; k++) dims += 4 + ((p->rr_start >= 0 && k >= p->rr_start) ? 1 : 0);
    if (dims > 256) return set_error(ctx, NOVA_HIP_INVALID_ARGUMENT, "max_depth needs more than 256 sampler dimensions");
  }
  int pass = 0;
  int st = resolve_integrator(ctx, p->integrator_flags, &pass);
  if (st) return st;
  if (pass != 0 && adj) return set_error(ctx, NOVA_HIP_INVALID_ARGUMENT, "the adjoint differentiates the path integrator only");
  hipStream_t s = (hipStream_t)stream;
  NOVA_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  if (pass != 0) {
    if (!accum_dev) return set_error(ctx, NOVA_HIP_INVALID_BUFFER_STATE, "render passes need accum and final buffers");
    return render_pass(ctx, p, pass, accum_dev, final_dev, partial_dev, s, async);
  }
  const size_t npix = (size_t)p->width * p->height;
  // Frames per batch: up to kPathsPerBatch paths in flight (about 116 B of queue and path
  // state each, 15.6 GB at the limit). Large batches matter because the last bounces hold
  // few rays whose launches cost their slowest ray's latency, a fixed cost per batch.
  int fb = (int)std::max<size_t>(1, std::min<size_t>((size_t)p->frame_count, kPathsPerBatch / npix));
  const int cus = device_cus(ctx);
  auto trace_threads_of = [&](size_t c) {
    return (uint32_t)std::min<size_t>((c + kTraceBlock - 1) / kTraceBlock, (size_t)cus * NOVA_TRACE_BLOCKS_PER_CU) * kTraceBlock;
  };
  if (noalloc) {
    // as many frames per batch as the fixed array holds
    while (fb > 0 && noalloc_bytes(npix * fb, trace_threads_of(npix * fb), traversal_stack_bound(ctx)) > kNoAllocScratchBytes) fb--;
    if (fb == 0)
      return set_error(ctx, NOVA_HIP_INVALID_ARGUMENT,
                       "image of " + std::to_string(npix) + " pixels needs " +
                           std::to_string(noalloc_bytes(npix, trace_threads_of(npix), traversal_stack_bound(ctx))) +
                           " bytes of scratch per frame; the no-allocation entry point holds " + std::to_string(kNoAllocScratchBytes));
  }
  const size_t cap = npix * fb;
  if (cap >= (1u << 31)) return set_error(ctx, NOVA_HIP_INVALID_ARGUMENT, "image too large");
  hipError_t e = hipSuccess;
  ScratchCarve carve{noalloc, 0, kNoAllocScratchBytes};
  auto piece = [&](int slot, size_t bytes) -> void * {
    if (!noalloc) return ensure_scratch(ctx, slot, bytes, &e);
    void *r = carve.take(bytes);
    if (!r) e = hipErrorOutOfMemory;
    return r;
  };
  Queues q;
  q.cap = (uint32_t)cap;
  float *rays = (float *)piece(0, cap * 12 * sizeof(float));
  if (!rays) return check_hip(ctx, e, "queue alloc");
  int32_t *slots = (int32_t *)piece(1, cap * 2 * sizeof(int32_t));
  if (!slots) return check_hip(ctx, e, "queue alloc");
  uint32_t *counts = (uint32_t *)piece(2, 256);
  if (!counts) return check_hip(ctx, e, "counter alloc");
  unsigned long long *stats64 = (unsigned long long *)(counts + 8);
  float *hits = (float *)piece(3, cap * 4 * sizeof(float));
  if (!hits) return check_hip(ctx, e, "hit alloc");
  float *state = (float *)piece(4, cap * 7 * sizeof(float));
  if (!state) return check_hip(ctx, e, "state alloc");
  float *partial = (float *)piece(5, cap * 4 * sizeof(float));
  if (!partial) return check_hip(ctx, e, "partial alloc");
  q.ray[0] = rays;
  q.ray[1] = rays + cap * 6;
  q.slot[0] = slots;
  q.slot[1] = slots + cap;
  q.count = counts;
  q.stats64 = stats64;
  q.hit = reinterpret_cast<float4 *>(hits);
  q.T = state;
  q.L = state + 3 * cap;
  q.A = state + 6 * cap;
  q.partial = partial;
  Queues q0 = q;  // camera-ray launches count their node visits and triangle tests apart
  q0.stats64 = stats64 + 4;

  const uint32_t trace_threads = trace_threads_of(cap);
  const uint32_t trace_blocks = trace_threads / kTraceBlock;
  uint32_t *gstack = nullptr;
  if (noalloc) {
    const uint32_t bound = traversal_stack_bound(ctx);
    gstack = (uint32_t *)carve.take((bound > (uint32_t)kLdsStack ? (size_t)(bound - kLdsStack) * trace_threads : 1) * sizeof(uint32_t));
    if (!gstack) return set_error(ctx, NOVA_HIP_INVALID_ARGUMENT, "traversal stacks do not fit the no-allocation scratch");
  } else {
    st = ensure_gstack(ctx, trace_threads);
    if (st) return st;
    gstack = ctx->gstack;
  }
  TraceCtx tc{ctx->nodes4, ctx->tri_leaf, gstack, trace_threads, ctx->tie, ctx->nodes4f};
  ShadeCtx sc{ctx->textures, ctx->tex_u32, ctx->tex_f32, ctx->materials, ctx->mesh_xf, ctx->shade, ctx->acos_lut, ctx->envmap,
              p->sampling_roughness_texture};
  AdjointCtx replay = adj ? AdjointCtx{partial, adj->dimg, adj->scale, adj->grad, nullptr, 0u, 0u}
                          : AdjointCtx{nullptr, nullptr, 0.f, nullptr, nullptr, 0u, 0u};
  // The replay reuses the primal pass's hits of as many leading bounces as fit in
  // kHitCacheBytes (16 B per path of the batch and bounce); it traces the others again.
  if (adj && !(p->flags & NOVA_HIP_ADJOINT_RETRACE)) {
    const size_t bounces = std::min<size_t>((size_t)p->max_depth + 1, kHitCacheBytes / (cap * sizeof(float4)));
    if (bounces > 0) {
      replay.hit_cache = (float4 *)ensure_scratch(ctx, 6, bounces * cap * sizeof(float4), &e);
      if (!replay.hit_cache) return check_hip(ctx, e, "hit cache alloc");
      replay.cache_paths = (uint32_t)cap;
      replay.cache_bounces = (uint32_t)bounces;
    }
  }
  // The shade loop splits the queue evenly over its grid, so the grid is what fits on the
  // chip at once: a second round of workgroups would run after the first at lower occupancy.
  int shade_per_cu[2] = {NOVA_SHADE_BLOCKS_PER_CU, NOVA_SHADE_BLOCKS_PER_CU};
  for (int a = 0; a < 2; a++) {
    if (shade_per_cu[a] > 0) continue;
    const void *fn = a ? reinterpret_cast<const void *>(shade_kernel<true>) : reinterpret_cast<const void *>(shade_kernel<false>);
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&shade_per_cu[a], fn, kShadeBlock, 0) != hipSuccess || shade_per_cu[a] <= 0) shade_per_cu[a] = 8;
  }
  const uint32_t shade_blocks[2] = {(uint32_t)std::min<size_t>((cap + kShadeBlock - 1) / kShadeBlock, (size_t)cus * shade_per_cu[0]),
                                    (uint32_t)std::min<size_t>((cap + kShadeBlock - 1) / kShadeBlock, (size_t)cus * shade_per_cu[1])};
  if (noalloc)
    hipLaunchKernelGGL(zero_words_kernel, dim3(1), dim3(64), 0, s, counts, 64u);
  else
    NOVA_HIP_CHECK(ctx, hipMemsetAsync(counts, 0, 256, s));

  const bool timing = (p->flags & NOVA_HIP_RENDER_TIMING) != 0;
  const bool counting = (p->flags & NOVA_HIP_RENDER_COUNT_TRAVERSAL) != 0;
  std::vector<hipEvent_t> kev;  // pairs of events around launches; kkind[i]: what pair i timed
  std::vector<int> kkind;       // 0 extend (bounce 0), 1 extend, 2 shade, 3 tail
  auto stamp = [&](void) -> int {
    if (kev.size() == ctx->ev_pool.size()) {  // the pool grows to the most launches one call timed
      hipEvent_t ev;
      NOVA_HIP_CHECK(ctx, hipEventCreate(&ev));
      ctx->ev_pool.push_back(ev);
    }
    const hipEvent_t ev = ctx->ev_pool[kev.size()];
    kev.push_back(ev);
    NOVA_HIP_CHECK(ctx, hipEventRecord(ev, s));
    return NOVA_HIP_SUCCESS;
  };
  // The last bounces go to tail_kernel once at most tail_rays rays remain (the primal
  // integrator only; the adjoint replay keeps per-bounce launches).
  const uint32_t tail_rays = (adj || (p->flags & NOVA_HIP_RENDER_NO_TAIL)) ? 0u : tail_rays_setting();
  // a hint from the previous call: same batch size, depth and RR (and scene: a new upload or
  // BVH clears it); not for the no-allocation entry point, which reads nothing back
  const bool use_hint = tail_rays > 0 && !noalloc && ctx->tail_hint > 0 && ctx->tail_hint_key[0] == (uint64_t)cap &&
                        ctx->tail_hint_key[1] == (uint64_t)p->max_depth && ctx->tail_hint_key[2] == (uint64_t)(int64_t)p->rr_start;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  if (!noalloc && !async) {
    if ((st = call_events(ctx, &ev0, &ev1))) return st;
    NOVA_HIP_CHECK(ctx, hipEventRecord(ev0, s));
  }
  uint64_t launches = 0, camera_rays = 0;
  // 64-bit frame counter: the step past a range that ends at INT32_MAX does not wrap
  const int64_t frame_end = (int64_t)p->frame_start + p->frame_count;
  for (int64_t f0 = p->frame_start; f0 < frame_end; f0 += fb) {
    const int frames = (int)std::min<int64_t>(fb, frame_end - f0);
    FrameParams fp;
    fp.width = p->width;
    fp.height = p->height;
    fp.frame0 = (int32_t)f0;
    fp.frames = frames;
    fp.offset = p->max_samples_offset;
    fp.vinvert = p->vertical_invert;
    fp.policy = p->index_policy;
    for (int i = 0; i < 16; i++) {
      fp.inv_P[i] = p->inv_P[i];
      fp.inv_V[i] = p->inv_V[i];
    }
    const uint32_t active = (uint32_t)(npix * frames);
    camera_rays += (uint64_t)active * (adj ? 2 : 1);
    for (int pass = 0; pass < (adj ? 2 : 1); pass++) {  // pass 1: the adjoint replay of pass 0's paths
    fp.frame_minor = 1;  // the replay numbers paths as pass 0 does
    fp.tiled = frames < tile_frames_setting() ? 1 : 0;
    if (pass == 0 && replay.hit_cache)  // every entry starts as a miss (prim -1)
      NOVA_HIP_CHECK(ctx, hipMemsetAsync(replay.hit_cache, 0xFF, (size_t)replay.cache_bounces * cap * sizeof(float4), s));
    hipLaunchKernelGGL(generate_kernel, dim3(std::min<uint32_t>((active + 255) / 256, (uint32_t)cus * 8)), dim3(256), 0, s, fp, ctx->sobol_table, q,
                       active);
    NOVA_HIP_CHECK(ctx, hipGetLastError());
    uint32_t dim = 2;
    int which = 0;
    // camera rays of a pixel-major batch with at least NOVA_COHERENT_MIN_FRAMES frames: a
    // wave holds 64 / frames pixels' rays, close enough to walk the tree together (a wave of
    // 64 pixels of one row is not)
#ifndef NOVA_COHERENT_MIN_FRAMES
#define NOVA_COHERENT_MIN_FRAMES 8
#endif
    const bool coherent0 = fp.frame_minor && frames >= NOVA_COHERENT_MIN_FRAMES && !(p->flags & NOVA_HIP_RENDER_PER_RAY_TRAVERSAL) &&
                           traversal_stack_bound(ctx) < (uint32_t)kWaveStack;
    for (int k = 0; k <= p->max_depth; k++) {
      // this bounce may start the tail: the sampler dimensions from its scatter to the last
      // (4 per scatter, 5 where it rolls RR) as nibble tables beside the traversal stacks
      int ndims = 0;
      for (int j = k; j <= p->max_depth; j++) ndims += 4 + ((p->rr_start >= 0 && j >= p->rr_start) ? 1 : 0);
      const size_t tail_lds = ((size_t)kLdsStack * kTraceBlock + (size_t)ndims * 128) * sizeof(uint32_t);
      const bool tail_here = tail_rays > 0 && k >= 1 && k < p->max_depth && tail_lds <= 65536;
      // The previous call of the same shape started its tail at bounce tail_hint: start it
      // there at the latest (the extend launch forces it) and launch nothing after it.
      const bool tail_last = tail_here && use_hint && k == ctx->tail_hint;
      if (timing && (st = stamp())) return st;
      if (timing) kkind.push_back(k == 0 ? 0 : 1);
      if (pass == 1 && replay.hit_cache && (uint32_t)k < replay.cache_bounces) {
        // the replay's rays are pass 0's: their hits are in the cache
      } else if (k == 0 && coherent0) {
        const uint32_t cblocks = std::min<uint32_t>((active + kCoherentBlock - 1) / kCoherentBlock, (uint32_t)cus * NOVA_COHERENT_BLOCKS_PER_CU);
        if (counting)
          hipLaunchKernelGGL(extend_coherent_kernel<true>, dim3(cblocks), dim3(kCoherentBlock), 0, s, tc, q0, which);
        else
          hipLaunchKernelGGL(extend_coherent_kernel<false>, dim3(cblocks), dim3(kCoherentBlock), 0, s, tc, q, which);
      } else if (counting) {
        hipLaunchKernelGGL(extend_kernel<true>, dim3(trace_blocks), dim3(kTraceBlock), 0, s, tc, k == 0 ? q0 : q, which, k, tail_here ? tail_rays : 0u,
                           tail_last ? 1 : 0);
      } else {
        hipLaunchKernelGGL(extend_kernel<false>, dim3(trace_blocks), dim3(kTraceBlock), 0, s, tc, q, which, k, tail_here ? tail_rays : 0u,
                           tail_last ? 1 : 0);
      }
      if (timing && (st = stamp())) return st;
      hipLaunchKernelGGL(reset_count_kernel, dim3(1), dim3(1), 0, s, counts, which ^ 1);
      ShadeParams spp;
      spp.fp = fp;
      spp.bounce = k;
      spp.last = (k == p->max_depth) ? 1 : 0;
      spp.rr = (p->rr_start >= 0 && k >= p->rr_start) ? 1 : 0;
      spp.dim0 = dim;
      if (timing && (st = stamp())) return st;
      if (timing) kkind.push_back(2);
      if (pass == 0)
        hipLaunchKernelGGL(shade_kernel<false>, dim3(shade_blocks[0]), dim3(kShadeBlock), 0, s, spp, sc, ctx->sobol_table, q, which, replay);
      else
        hipLaunchKernelGGL(shade_kernel<true>, dim3(shade_blocks[1]), dim3(kShadeBlock), 0, s, spp, sc, ctx->sobol_table, q, which, replay);
      if (timing && (st = stamp())) return st;
      NOVA_HIP_CHECK(ctx, hipGetLastError());
      if (tail_here) {
        // at most trace_threads lanes: the overflow stacks are sized for the extend grid
        const uint32_t tblocks = std::min<uint32_t>({(uint32_t)((std::min<size_t>(tail_last ? cap : tail_rays, cap) + kTraceBlock - 1) / kTraceBlock),
                                                     trace_blocks, (uint32_t)cus * 4});
        if (timing && (st = stamp())) return st;
        if (timing) kkind.push_back(3);
        if (counting)
          hipLaunchKernelGGL(tail_kernel<true>, dim3(tblocks), dim3(kTraceBlock), tail_lds, s, tc, spp, p->max_depth, p->rr_start, sc, ctx->sobol_table, q,
                             which, ndims);
        else
          hipLaunchKernelGGL(tail_kernel<false>, dim3(tblocks), dim3(kTraceBlock), tail_lds, s, tc, spp, p->max_depth, p->rr_start, sc, ctx->sobol_table, q,
                             which, ndims);
        if (timing && (st = stamp())) return st;
        NOVA_HIP_CHECK(ctx, hipGetLastError());
      }
      dim += 4 + (spp.rr ? 1 : 0);
      which ^= 1;
      launches++;
      if (tail_last) break;  // tail_kernel has finished every path
    }
    }
    if (!adj) {
      float *const fr_out = frames_out ? frames_out + (size_t)(f0 - p->frame_start) * npix * 4 : (float *)nullptr;
      float *const last_out = (f0 + frames == frame_end) ? partial_dev : (float *)nullptr;
      if (frames >= NOVA_ACC_TILED_FRAMES)
        hipLaunchKernelGGL(accumulate_tiled_kernel, dim3(std::min<size_t>((npix + kAccPix - 1) / kAccPix, (size_t)cus * 16)), dim3(256), 0, s, partial,
                           fp, accum_dev, final_dev, last_out, fr_out);
      else
        hipLaunchKernelGGL(accumulate_kernel, dim3(std::min<size_t>((npix * 4 + 255) / 256, (size_t)cus * 16)), dim3(256), 0, s, partial, fp,
                           accum_dev, final_dev, last_out, fr_out);
      NOVA_HIP_CHECK(ctx, hipGetLastError());
    }
  }
  if (noalloc || async) return NOVA_HIP_SUCCESS;  // stream-ordered; statistics untouched
  NOVA_HIP_CHECK(ctx, hipEventRecord(ev1, s));
  NOVA_HIP_CHECK(ctx, hipEventSynchronize(ev1));
  float ms = 0.f;
  hipEventElapsedTime(&ms, ev0, ev1);
  // the counters and the 64-bit statistics after them in one copy (the per-frame engine
  // loop pays for every synchronous copy)
  uint32_t words[8 + 2 * 20] = {};
  static_assert(sizeof(words) <= 256, "counters piece");
  NOVA_HIP_CHECK(ctx, hipMemcpy(words, counts, sizeof(words), hipMemcpyDeviceToHost));
  const uint32_t seg = words[2];
  unsigned long long s64[20];
  std::memcpy(s64, words + 8, sizeof(s64));
  if (tail_rays > 0) {  // where the last batch's tail started: the next call's hint
    const uint32_t slot = words[kTailSlot];
    ctx->tail_hint = (slot == 0 || (slot & 0x80000000u)) ? 0 : (int)slot - 1;  // a forced start with many rays: learn again
    ctx->tail_hint_key[0] = (uint64_t)cap;
    ctx->tail_hint_key[1] = (uint64_t)p->max_depth;
    ctx->tail_hint_key[2] = (uint64_t)(int64_t)p->rr_start;
  }
  float ext_ms = 0.f, shade_ms = 0.f, ext0_ms = 0.f, tail_ms = 0.f;
  for (size_t k = 0; k < kkind.size() && 2 * k + 1 < kev.size(); k++) {
    float a = 0.f;
    hipEventElapsedTime(&a, kev[2 * k], kev[2 * k + 1]);
    if (kkind[k] <= 1) ext_ms += a;
    if (kkind[k] == 0) ext0_ms += a;  // bounce 0 of a batch pass
    if (kkind[k] == 2) shade_ms += a;
    if (kkind[k] == 3) tail_ms += a;
  }
  ctx->stats.total_ms = ms;
  ctx->stats.extend_ms = ext_ms;
  ctx->stats.shade_ms = shade_ms;
  ctx->stats.other_ms = timing ? ms - ext_ms - shade_ms - tail_ms : 0.f;
  ctx->stats.tail_ms = timing ? tail_ms : 0.f;
  ctx->stats.tail_segments = s64[16];
  ctx->stats.tail_node_visits = counting ? s64[17] : 0;
  ctx->stats.tail_triangle_tests = counting ? s64[18] : 0;
  ctx->stats.segments = seg;
  ctx->stats.extend_launches = launches;
  ctx->stats.node_visits = counting ? s64[0] + s64[4] : 0;
  ctx->stats.triangle_tests = counting ? s64[1] + s64[5] : 0;
  ctx->stats.escaped = s64[2];
  ctx->stats.extend_first_ms = timing ? ext0_ms : 0.f;
  ctx->stats.camera_rays = camera_rays;
  ctx->stats.node_visits_first = counting ? s64[4] : 0;
  ctx->stats.triangle_tests_first = counting ? s64[5] : 0;
  ctx->stats.walk_wave_turns = counting ? s64[8] + s64[12] : 0;
  ctx->stats.walk_lane_turns = counting ? s64[9] + s64[13] : 0;
  ctx->stats.walk_tri_rounds = counting ? s64[10] + s64[14] : 0;
  return NOVA_HIP_SUCCESS;
}

extern "C" int nova_hip_render_frames(nova_hip_ctx *ctx, const nova_hip_render_params *p, float *accum_dev, float *final_dev, float *partial_dev,
                                      void *stream) {
  return render_impl(ctx, p, accum_dev, final_dev, partial_dev, stream, nullptr);
}

extern "C" int nova_hip_render_frame_lookahead(nova_hip_ctx *ctx, const nova_hip_render_params *p, int32_t lookahead, float *accum_dev,
                                               float *final_dev, float *partial_dev, void *stream) {
  if (!ctx || !p) return NOVA_HIP_INVALID_ARGUMENT;
  if (p->frame_count != 1 || p->frame_start < 0) return set_error(ctx, NOVA_HIP_INVALID_ARGUMENT, "one frame per call (frame_count 1)");
  if (lookahead < 1 || lookahead > 1024) return set_error(ctx, NOVA_HIP_INVALID_ARGUMENT, "lookahead must be in [1, 1024]");
  if (!accum_dev != !final_dev || (!accum_dev && !partial_dev)) return set_error(ctx, NOVA_HIP_INVALID_BUFFER_STATE, "null render buffer");
  if ((((uintptr_t)accum_dev | (uintptr_t)final_dev | (uintptr_t)partial_dev) & 3u) != 0)
    return set_error(ctx, NOVA_HIP_INVALID_BUFFER_STATE, "render buffers must be 4-byte aligned float arrays");
  int pass = 0;
  int st = resolve_integrator(ctx, p->integrator_flags, &pass);
  if (st) return st;
  // frames past INT32_MAX do not exist: at most the frames left below it
  lookahead = (int32_t)std::min<int64_t>(lookahead, (int64_t)INT32_MAX - p->frame_start);
  if (pass != 0 || lookahead <= 1) return render_impl(ctx, p, accum_dev, final_dev, partial_dev, stream, nullptr);
  if (p->width <= 0 || p->height <= 0 || p->width > 65536 || p->height > 65536)
    return set_error(ctx, NOVA_HIP_INVALID_RENDERBUFFER_DIM, "Wrong buffer width / height dimension.");
  const size_t n4 = (size_t)p->width * p->height * 4;
  if (n4 >= (1ull << 32)) return set_error(ctx, NOVA_HIP_INVALID_ARGUMENT, "image too large to keep frames of");
  nova_hip_render_params key = *p;
  key.frame_start = 0;
  const bool hit = ctx->ahead_valid && ctx->ahead_gen == ctx->scene_gen && std::memcmp(&key, &ctx->ahead_key, sizeof(key)) == 0 &&
                   p->frame_start >= ctx->ahead_first && p->frame_start < ctx->ahead_first + ctx->ahead_count;
  hipStream_t s = (hipStream_t)stream;
  NOVA_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  if (!hit) {
    ctx->ahead_valid = false;
    hipError_t e = hipSuccess;
    float *store = (float *)ensure_scratch(ctx, 12, (size_t)lookahead * n4 * sizeof(float), &e);
    if (!store) return check_hip(ctx, e, "lookahead frame store alloc");
    nova_hip_render_params ahead = *p;
    ahead.frame_count = lookahead;
    // camera rays by the per-ray walk, as a one-frame render traces them: the per-wave walk
    // of many frames can settle a grazing ray's near-tie differently (its lanes visit leaves
    // in the first lane's order)
    ahead.flags |= NOVA_HIP_RENDER_PER_RAY_TRAVERSAL;
    st = render_impl(ctx, &ahead, nullptr, nullptr, nullptr, stream, nullptr, nullptr, store);
    if (st) return st;
    ctx->ahead_key = key;
    ctx->ahead_first = p->frame_start;
    ctx->ahead_count = lookahead;
    ctx->ahead_gen = ctx->scene_gen;
    ctx->ahead_valid = true;
  }
  const float *image = (const float *)ctx->scratch[12].ptr + (size_t)(p->frame_start - ctx->ahead_first) * n4;
  hipLaunchKernelGGL(frame_add_kernel, dim3((uint32_t)std::min<size_t>((n4 + 255) / 256, (size_t)device_cus(ctx) * 16)), dim3(256), 0, s, image,
                     (uint32_t)n4, accum_dev, final_dev, partial_dev, float(p->frame_start + 1));
  NOVA_HIP_CHECK(ctx, hipGetLastError());
  return NOVA_HIP_SUCCESS;
}

extern "C" int nova_hip_render_frames_noalloc(nova_hip_ctx *ctx, const nova_hip_render_params *p, float *accum_dev, float *final_dev,
                                              float *partial_dev, void *stream) {
  if (!ctx || !p) return NOVA_HIP_INVALID_ARGUMENT;
  if (p->flags & (NOVA_HIP_RENDER_TIMING | NOVA_HIP_RENDER_COUNT_TRAVERSAL))
    return set_error(ctx, NOVA_HIP_INVALID_ARGUMENT, "timing and traversal counting need events and reads: not offered without allocation");
  int pass = 0;
  int st = resolve_integrator(ctx, p->integrator_flags, &pass);
  if (st) return st;
  if (pass != 0) return set_error(ctx, NOVA_HIP_INVALID_ARGUMENT, "the no-allocation entry point renders the path integrator only");
  unsigned char *base = static_cast<unsigned char *>(noalloc_scratch_of(ctx->device));
  if (!base) return set_error(ctx, NOVA_HIP_DEVICE_ERROR, "no-allocation scratch address unavailable on this device");
  return render_impl(ctx, p, accum_dev, final_dev, partial_dev, stream, nullptr, base);
}

extern "C" int nova_hip_render_adjoint(nova_hip_ctx *ctx, const nova_hip_render_params *p, const float *d_image_dev, float scale,
                                       float *grad_pool_dev, void *stream) {
  if (!ctx || !p) return NOVA_HIP_INVALID_ARGUMENT;
  if (!d_image_dev || !grad_pool_dev) return set_error(ctx, NOVA_HIP_INVALID_BUFFER_STATE, "null adjoint buffer");
  const AdjointCtx adj{nullptr, d_image_dev, scale, grad_pool_dev, nullptr, 0u, 0u};
  return render_impl(ctx, p, nullptr, nullptr, nullptr, stream, &adj);
}

extern "C" int nova_hip_render_stats_get(nova_hip_ctx *ctx, nova_hip_render_stats *out) {
  if (!ctx || !out) return NOVA_HIP_INVALID_ARGUMENT;
  *out = ctx->stats;
  return NOVA_HIP_SUCCESS;
}
