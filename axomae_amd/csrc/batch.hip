// Batched rendering of many scenes in shared launches.
//
// Every path of every scene gets a global slot: scenes back to back (in the order given),
// each scene pixel-major as nova_hip_render_frames numbers its paths (path = pixel * frames +
// frame; the sampler slot of a path is its frame-major slot, path_slot). The slots are cut into
// waves of `queue_capacity` paths; a wave runs generate -> (extend, reset, shade) per depth
// level -> accumulate -> advance, and each of those is one grid-stride launch over the
// wave's combined queue, whatever scenes the paths belong to (a path carries its scene id
// through the compaction). Wave parameters live in device memory, so the sequence of one
// wave is the same for every wave: it is captured once into a HIP graph and replayed.
//
// Each path computes exactly what the single-scene renderer computes for it (same scene
// data, sample index and arithmetic) and the accumulate stage adds a wave's frames to the
// per-scene accumulators in frame order, so a scene's image equals its nova_hip_render_frames
// image bit for bit where both trace the camera rays with the same walk (per ray or per wave;
// the two differ, rarely, on a ray grazing a shared triangle edge).
#include <algorithm>
#include <vector>

#include "device_math.h"
#include "device_path.h"
#include "device_queue.h"
#include "device_shading.h"
#include "device_trace.h"
#include "nova_internal.h"

namespace nova_hip {

struct BatchScene {
  TraceCtx tc;    // gstack fields and tie unused (the launch passes its own stack; tie below)
  TieTables tie;  // the scene's tie tables, in the scene record so a lane computes their address
  ShadeCtx sc;
  FrameParams fp;
  uint64_t first;   // first global slot
  uint64_t slots;   // npix * frames
  uint32_t npix;
  float *accum;
};

struct WaveDesc {
  uint64_t start;
  uint32_t count;
  uint32_t seg_begin, seg_end;
  uint32_t pix_total;  // sum of the segments' pixel counts
};

// The part of one scene inside one wave: local paths [a, b), which cover pixels
// p0 .. p0 + (its pixel count) - 1, with prefix pix_base among the wave's segments.
struct Segment {
  uint32_t scene;
  uint32_t a, b;
  uint32_t pix_base;
  uint32_t p0;
};

struct BatchState {
  uint32_t wave;
  uint32_t count[2];  // queue sizes
  uint32_t work;      // rays of the current queue claimed by the extend launch
  uint32_t do_accum;
  unsigned long long segments;  // rays traced
};

struct BatchQueues {
  float *ray[2];
  int32_t *slot[2];  // wave-relative slot
  int32_t *sid[2];   // scene index
  float4 *hit;        // closest hit per queue entry: (prim id bits, t, u, v), one 16-byte record
  float *T, *L, *A;
  float *partial;  // wave-relative slot x 4
  uint32_t cap;
};

__device__ __forceinline__ int find_scene(const BatchScene *__restrict__ scenes, int n, uint64_t g) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (scenes[mid].first <= g)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

__global__ __launch_bounds__(256) void batch_generate(const BatchScene *__restrict__ scenes, int nscenes, const WaveDesc *__restrict__ waves,
                                                      BatchState *st, const uint32_t *__restrict__ sobol, BatchQueues q) {
  __shared__ uint32_t sob_tab[2 * 128];  // jitter dims 0-1 as per-nibble tables
  sobol_nibble_fill(sobol, 0, 2, sob_tab, threadIdx.x, blockDim.x);
  __syncthreads();
  const SobolNibbles sobn{sob_tab, 0};
  const WaveDesc w = waves[st->wave];
  const size_t n = q.cap;
  for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < w.count; r += gridDim.x * blockDim.x) {
    const uint64_t g = w.start + r;
    const int s = find_scene(scenes, nscenes, g);
    const uint32_t l = (uint32_t)(g - scenes[s].first);
    f3 o, d;
    camera_ray(scenes[s].fp, sobn, path_slot(scenes[s].fp, l), o, d);
    float *ray = q.ray[0];
    ray[r] = o.x;
    ray[n + r] = o.y;
    ray[2 * n + r] = o.z;
    ray[3 * n + r] = d.x;
    ray[4 * n + r] = d.y;
    ray[5 * n + r] = d.z;
    q.slot[0][r] = (int32_t)r;
    q.sid[0][r] = s;
    q.T[r] = q.T[n + r] = q.T[2 * n + r] = 1.f;
    q.L[r] = q.L[n + r] = q.L[2 * n + r] = 0.f;
    q.A[r] = 0.f;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->count[0] = w.count;
    st->count[1] = 0;
    st->work = 0;
  }
}

// Closest hits of the queue with the single-scene kernel's lane refill: a lane whose ray
// has ended takes the next unclaimed ray once kBatchRefill lanes of its wave are idle.
constexpr uint32_t kBatchRefill = 56;
// Tie tables of a lane's scene: an address computed from the scene index, not a pointer
// loaded from the scene record at every leaf.
struct SceneTie {
  const BatchScene *scenes;
  int32_t sid;
  __device__ const TieTables *operator()(const TraceCtx &) const { return &scenes[sid].tie; }
};
__global__ __launch_bounds__(kTraceBlock) __attribute__((amdgpu_waves_per_eu(6, 6))) void batch_extend(
    const BatchScene *__restrict__ scenes, BatchState *st, BatchQueues q, int which, uint32_t *gstack, uint32_t gstride) {
  __shared__ uint32_t lds[kLdsStack * kTraceBlock];
  const uint32_t n = st->count[which];
  const float *r = q.ray[which];
  const size_t cap = q.cap;
  const uint32_t gtid = blockIdx.x * blockDim.x + threadIdx.x;
  if (gtid == 0) st->segments += n;
  LdsStack stk{lds, gstack, gstride, gtid, threadIdx.x};
  // first rays by thread index, refills claimed past them (as extend_kernel)
  const uint32_t lane = threadIdx.x & 63;
  const uint32_t nthreads = gridDim.x * blockDim.x;
  TraceState ts;
  bool have = false, exhausted = nthreads >= n;
  uint32_t ray = 0;
  int32_t sid = 0;
  TraceCtx tc{nullptr, nullptr, nullptr, 0, nullptr};  // the lane's scene, loaded when it takes a ray (not at every step)
  uint32_t nv = 0, nt = 0;
  for (uint32_t turns = 0; turns < (1u << 30); turns++) {
    const uint64_t idle = __ballot(!have);
    uint32_t base = gtid - lane;  // first turn: lane l takes ray gtid
    bool fill = turns == 0;
    if (!fill && !exhausted && (uint32_t)__popcll(idle) >= kBatchRefill) {
      const int leader = __builtin_ffsll((long long)idle) - 1;
      if ((int)lane == leader) base = nthreads + atomicAdd(&st->work, (uint32_t)__popcll(idle));
      base = __shfl(base, leader, 64);
      if (base >= n) exhausted = true;
      fill = true;
    }
    if (fill && !have) {
      const uint32_t i = base + __builtin_amdgcn_mbcnt_hi((uint32_t)(idle >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)idle, 0));
      if (i < n) {
        trace_init(ts, mk3(r[i], r[cap + i], r[2 * cap + i]), mk3(r[3 * cap + i], r[4 * cap + i], r[5 * cap + i]));
        ray = i;
        sid = q.sid[which][i];
        tc = scenes[sid].tc;
        have = true;
      }
    }
    if (__ballot(have) == 0) break;
    if (have && trace_step<false>(tc, stk, ts, 0.000001f, nv, nt, SceneTie{scenes, sid})) {
      q.hit[ray] = make_float4(__builtin_bit_cast(float, (int32_t)(ts.best)), ts.tbest, ts.bu, ts.bv);
      have = false;
    }
  }
}

// Camera rays (queue 0 right after batch_generate, in slot order) walked once per wave of 64
// consecutive slots, as the single-scene renderer walks its pixel-major camera rays. Used
// when every scene's path count is a multiple of 64, so the 64 slots of a packet belong to
// one scene (waves start at multiples of the queue capacity, itself a multiple of 256).
constexpr int kBatchCoherentBlock = 256;
constexpr uint32_t kBatchCoherentBlocksPerCu = 64;  // grid-stride over the packets, as the single-scene walk
__global__ __launch_bounds__(kBatchCoherentBlock) void batch_extend_coherent(const BatchScene *__restrict__ scenes, BatchState *st, BatchQueues q) {
  const uint32_t n = st->count[0];
  const float *r = q.ray[0];
  const size_t cap = q.cap;
  const uint32_t lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (blockIdx.x == 0 && threadIdx.x == 0) st->segments += n;
  uint32_t nv = 0, nt = 0;
  for (uint32_t base = (blockIdx.x * (kBatchCoherentBlock / 64) + wv) * 64u; base < n; base += gridDim.x * kBatchCoherentBlock) {
    const uint32_t i = base + lane;
    const bool valid = i < n;
    const int32_t sid = __builtin_amdgcn_readfirstlane(q.sid[0][base]);
    TraceCtx tc = scenes[sid].tc;
    tc.tie = &scenes[sid].tie;
    TraceState ts;
    trace_init(ts, valid ? mk3(r[i], r[cap + i], r[2 * cap + i]) : mk3(0.f, 0.f, 0.f),
               valid ? mk3(r[3 * cap + i], r[4 * cap + i], r[5 * cap + i]) : mk3(0.f, 0.f, 1.f));
    coherent_walk<false>(tc, ts, __ballot(valid), 0.000001f, nv, nt);
    if (valid) {
      q.hit[i] = make_float4(__builtin_bit_cast(float, (int32_t)(ts.best)), ts.tbest, ts.bu, ts.bv);
    }
  }
}

// Between an extend and a shade launch: the shade's output queue and the next extend's work
// counter start at zero.
__global__ void batch_reset(BatchState *st, int which) {
  st->count[which] = 0;
  st->work = 0;
}

__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4, 4))) void batch_shade(const BatchScene *__restrict__ scenes, const WaveDesc *__restrict__ waves, BatchState *st,
                                                   const uint32_t *__restrict__ sobol, BatchQueues q, int which, int bounce, int last, int rr,
                                                   uint32_t dim0) {
  __shared__ uint32_t sob_tab[5 * 128];  // nibble tables of dims dim0 .. dim0+4
  sobol_nibble_fill(sobol, dim0, 5, sob_tab, threadIdx.x, blockDim.x);
  __syncthreads();
  const SobolNibbles sobn{sob_tab, dim0};
  const uint32_t n = st->count[which];
  const uint64_t wstart = waves[st->wave].start;
  const size_t cap = q.cap;
  const float *rin = q.ray[which];
  float *rout = q.ray[which ^ 1];
  const uint32_t stride = gridDim.x * blockDim.x;
  const uint32_t iters = (n + stride - 1) / stride;  // uniform over the grid: the compaction below needs every thread
  __shared__ uint32_t wave_cnt[2][4], blk_base[2];
  for (uint32_t it = 0; it < iters; it++) {
    const uint32_t i = it * stride + blockIdx.x * blockDim.x + threadIdx.x;
    bool cont = false, escaped = false;
    f3 no = mk3(0, 0, 0), nd = mk3(0, 0, 0);
    int32_t slot = -1, sid = 0;
    if (i < n) {
      slot = q.slot[which][i];
      sid = q.sid[which][i];
      const BatchScene &S = scenes[sid];
      const f3 o = mk3(rin[i], rin[cap + i], rin[2 * cap + i]);
      const f3 d = mk3(rin[3 * cap + i], rin[4 * cap + i], rin[5 * cap + i]);
      f3 T = mk3(q.T[slot], q.T[cap + slot], q.T[2 * cap + slot]);
      f3 L = mk3(q.L[slot], q.L[cap + slot], q.L[2 * cap + slot]);
      float A = q.A[slot];
      ShadeParams sp;
      sp.fp = S.fp;
      sp.bounce = bounce;
      sp.last = last;
      sp.rr = rr;
      sp.dim0 = dim0;
      const uint32_t l = (uint32_t)(wstart + (uint64_t)slot - S.first);
      const float4 h = q.hit[i];
      cont = shade_path(sp, S.sc, sobn, path_slot(S.fp, l), o, d, __builtin_bit_cast(int32_t, h.x), h.y, h.z, h.w, T, L, A, no, nd, escaped);
      q.T[slot] = T.x;
      q.T[cap + slot] = T.y;
      q.T[2 * cap + slot] = T.z;
      if (!cont) {
        reinterpret_cast<float4 *>(q.partial)[slot] = make_float4(L.x, L.y, L.z, A);
      } else {
        q.L[slot] = L.x;
        q.L[cap + slot] = L.y;
        q.L[2 * cap + slot] = L.z;
        q.A[slot] = A;
      }
    }
    const uint64_t m = __ballot(cont);
    const uint32_t lane_rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0));
    const uint32_t base = block_reserve<256>(m, &st->count[which ^ 1], it, wave_cnt, blk_base);
    if (cont) {
      const uint32_t j = base + lane_rank;
      rout[j] = no.x;
      rout[cap + j] = no.y;
      rout[2 * cap + j] = no.z;
      rout[3 * cap + j] = nd.x;
      rout[4 * cap + j] = nd.y;
      rout[5 * cap + j] = nd.z;
      q.slot[which ^ 1][j] = slot;
      q.sid[which ^ 1][j] = sid;
    }
  }
}

// accum += each frame of the wave, in frame order, pixel by pixel (Engine::writeFramebuffer,
// engine.cpp:351-363).
__global__ __launch_bounds__(256) void batch_accumulate(const BatchScene *__restrict__ scenes, const WaveDesc *__restrict__ waves,
                                                        const Segment *__restrict__ segs, const BatchState *st, const float *__restrict__ partial) {
  if (!st->do_accum) return;
  const WaveDesc w = waves[st->wave];
  for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < w.pix_total; k += gridDim.x * blockDim.x) {
    uint32_t lo = w.seg_begin, hi = w.seg_end - 1;
    while (lo < hi) {
      const uint32_t mid = (lo + hi + 1) >> 1;
      if (segs[mid].pix_base <= k)
        lo = mid;
      else
        hi = mid - 1;
    }
    const Segment sg = segs[lo];
    const BatchScene &S = scenes[sg.scene];
    const uint32_t p = sg.p0 + (k - sg.pix_base);
    const uint32_t F = (uint32_t)S.fp.frames;
    const uint32_t fa = sg.a > p * F ? sg.a - p * F : 0u, fb = std::min(F, sg.b - p * F);  // this wave's frames of pixel p
    float4 *acc = reinterpret_cast<float4 *>(S.accum + pixel_offset(S.fp, p));
    float4 a = *acc;
    for (uint32_t f = fa; f < fb; f++) {
      const uint32_t l = p * F + f;
      const float4 v = reinterpret_cast<const float4 *>(partial)[S.first + l - w.start];
      a.x += v.x;
      a.y += v.y;
      a.z += v.z;
      a.w += v.w;
    }
    *acc = a;
  }
}

__global__ void batch_advance(BatchState *st) { st->wave += 1; }

__global__ void batch_set(BatchState *st, uint32_t wave, uint32_t do_accum) {
  st->wave = wave;
  st->do_accum = do_accum;
  st->count[0] = st->count[1] = 0;
  st->work = 0;
  st->segments = 0;
}

__global__ void batch_zero(float *__restrict__ p, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = 0.f;
}

// final = accum / (frame_start + frame_count), the division the last frame's write makes
__global__ void batch_finalize(const float *__restrict__ accum, float *__restrict__ final_out, size_t n, float frames_total) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    final_out[i] = accum[i] / frames_total;
}

struct BatchPlan {
  std::vector<BatchScene> scenes;
  uint64_t total = 0;
  int max_depth = 0, rr_start = -1;
  uint32_t stack_bound = 0;
  bool coherent0 = true;  // bounce 0 walks packets of 64 camera rays (batch_extend_coherent)
};

struct WavePlan {
  std::vector<WaveDesc> waves;
  std::vector<Segment> segs;
  bool ok = true;  // false when a wave would start inside a packet of the per-wave camera walk
};

static WavePlan plan_waves(const BatchPlan &bp, uint64_t start, uint64_t end, uint32_t cap) {
  WavePlan wp;
  for (uint64_t w0 = start; w0 < end; w0 += cap) {
    // the per-wave camera walk reads one scene index per packet of 64 paths
    if (bp.coherent0 && w0 % 64 != 0) wp.ok = false;
    const uint64_t w1 = std::min<uint64_t>(end, w0 + cap);
    WaveDesc wd{w0, (uint32_t)(w1 - w0), (uint32_t)wp.segs.size(), 0, 0};
    for (uint32_t s = 0; s < bp.scenes.size(); s++) {
      const BatchScene &S = bp.scenes[s];
      const uint64_t a = std::max(w0, S.first), b = std::min(w1, S.first + S.slots);
      if (a >= b) continue;
      const uint32_t la = (uint32_t)(a - S.first), lb = (uint32_t)(b - S.first), F = (uint32_t)S.fp.frames;
      const uint32_t p0 = la / F, p1 = (lb - 1) / F;  // pixels of paths la .. lb-1 (pixel-major)
      wp.segs.push_back(Segment{s, la, lb, wd.pix_total, p0});
      wd.pix_total += p1 - p0 + 1;
    }
    wd.seg_end = (uint32_t)wp.segs.size();
    wp.waves.push_back(wd);
  }
  return wp;
}

static int device_cus_of(int device) {
  int cus = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess || cus <= 0) cus = 256;
  return cus;
}

// Device buffers of one batch call (owned by the executor context's scratch slots 6..11).
struct BatchBuffers {
  BatchScene *scenes = nullptr;
  WaveDesc *waves = nullptr;
  Segment *segs = nullptr;
  BatchState *state = nullptr;
  uint32_t *gstack = nullptr;
  BatchQueues q{};
};

static int alloc_buffers(nova_hip_ctx *ex, const BatchPlan &bp, const WavePlan &wp, uint32_t cap, uint32_t trace_threads, BatchBuffers &B,
                         hipStream_t s) {
  hipError_t e;
  char *qbuf = (char *)ensure_scratch(ex, 6, (size_t)cap * (12 + 2 + 2 + 4 + 7 + 4) * 4, &e);
  if (!qbuf) return check_hip(ex, e, "batch queue alloc");
  float *f = (float *)qbuf;
  B.q.cap = cap;
  B.q.ray[0] = f;
  B.q.ray[1] = f + (size_t)cap * 6;
  int32_t *ip = (int32_t *)(f + (size_t)cap * 12);
  B.q.slot[0] = ip;
  B.q.slot[1] = ip + cap;
  B.q.sid[0] = ip + 2 * (size_t)cap;
  B.q.sid[1] = ip + 3 * (size_t)cap;
  B.q.hit = reinterpret_cast<float4 *>(ip + 4 * (size_t)cap);  // 16-byte aligned: cap is a multiple of 256
  float *hf = (float *)(ip + 8 * (size_t)cap);
  B.q.T = hf;
  B.q.L = hf + 3 * (size_t)cap;
  B.q.A = hf + 6 * (size_t)cap;
  B.q.partial = hf + 7 * (size_t)cap;  // 4 floats per slot, 16-byte aligned (cap is a multiple of 256)
  B.scenes = (BatchScene *)ensure_scratch(ex, 7, bp.scenes.size() * sizeof(BatchScene), &e);
  if (!B.scenes) return check_hip(ex, e, "batch scene table alloc");
  B.waves = (WaveDesc *)ensure_scratch(ex, 8, std::max<size_t>(1, wp.waves.size()) * sizeof(WaveDesc), &e);
  if (!B.waves) return check_hip(ex, e, "batch wave table alloc");
  B.segs = (Segment *)ensure_scratch(ex, 9, std::max<size_t>(1, wp.segs.size()) * sizeof(Segment), &e);
  if (!B.segs) return check_hip(ex, e, "batch segment table alloc");
  B.state = (BatchState *)ensure_scratch(ex, 10, sizeof(BatchState), &e);
  if (!B.state) return check_hip(ex, e, "batch state alloc");
  const size_t gneed = bp.stack_bound > (uint32_t)kLdsStack ? (size_t)(bp.stack_bound - kLdsStack) * trace_threads : 1;
  B.gstack = (uint32_t *)ensure_scratch(ex, 11, gneed * sizeof(uint32_t), &e);
  if (!B.gstack) return check_hip(ex, e, "batch stack alloc");
  NOVA_HIP_CHECK(ex, hipMemcpyAsync(B.scenes, bp.scenes.data(), bp.scenes.size() * sizeof(BatchScene), hipMemcpyHostToDevice, s));
  if (!wp.waves.empty())
    NOVA_HIP_CHECK(ex, hipMemcpyAsync(B.waves, wp.waves.data(), wp.waves.size() * sizeof(WaveDesc), hipMemcpyHostToDevice, s));
  if (!wp.segs.empty()) NOVA_HIP_CHECK(ex, hipMemcpyAsync(B.segs, wp.segs.data(), wp.segs.size() * sizeof(Segment), hipMemcpyHostToDevice, s));
  return NOVA_HIP_SUCCESS;
}

// Launches one wave's sequence on s (captured into a graph by the caller, or run directly).
static void launch_wave(const BatchPlan &bp, const BatchBuffers &B, const nova_hip_batch_config &cfg, int cus, const uint32_t *sobol,
                        hipStream_t s) {
  const int ns = (int)bp.scenes.size();
  const uint32_t cap = B.q.cap;
  const uint32_t gen_blocks = std::min<uint32_t>((cap + 255) / 256, (uint32_t)cus * 8);
  const uint32_t trace_blocks = std::min<uint32_t>((cap + kTraceBlock - 1) / kTraceBlock, (uint32_t)cus * cfg.trace_blocks_per_cu);
  const uint32_t shade_blocks = std::min<uint32_t>((cap + 255) / 256, (uint32_t)cus * cfg.shade_blocks_per_cu);
  hipLaunchKernelGGL(batch_generate, dim3(gen_blocks), dim3(256), 0, s, B.scenes, ns, B.waves, B.state, sobol, B.q);
  uint32_t dim = 2;
  int which = 0;
  for (int k = 0; k <= bp.max_depth; k++) {
    if (k == 0 && bp.coherent0)
      hipLaunchKernelGGL(batch_extend_coherent, dim3(std::min<uint32_t>((cap + kBatchCoherentBlock - 1) / kBatchCoherentBlock, (uint32_t)cus * kBatchCoherentBlocksPerCu)),
                         dim3(kBatchCoherentBlock), 0, s, B.scenes, B.state, B.q);
    else
      hipLaunchKernelGGL(batch_extend, dim3(trace_blocks), dim3(kTraceBlock), 0, s, B.scenes, B.state, B.q, which, B.gstack,
                         trace_blocks * (uint32_t)kTraceBlock);
    hipLaunchKernelGGL(batch_reset, dim3(1), dim3(1), 0, s, B.state, which ^ 1);
    const int rr = (bp.rr_start >= 0 && k >= bp.rr_start) ? 1 : 0;
    hipLaunchKernelGGL(batch_shade, dim3(shade_blocks), dim3(256), 0, s, B.scenes, B.waves, B.state, sobol, B.q, which, k,
                       k == bp.max_depth ? 1 : 0, rr, dim);
    dim += 4 + rr;
    which ^= 1;
  }
  hipLaunchKernelGGL(batch_accumulate, dim3(std::min<uint32_t>((cap + 255) / 256, (uint32_t)cus * 8)), dim3(256), 0, s, B.scenes, B.waves,
                     B.segs, B.state, B.q.partial);
  hipLaunchKernelGGL(batch_advance, dim3(1), dim3(1), 0, s, B.state);
}

struct GraphExec {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  ~GraphExec() {
    if (exec) hipGraphExecDestroy(exec);
    if (graph) hipGraphDestroy(graph);
  }
};

static int capture_wave(nova_hip_ctx *ex, const BatchPlan &bp, const BatchBuffers &B, const nova_hip_batch_config &cfg, int cus,
                        const uint32_t *sobol, hipStream_t s, GraphExec &g) {
  NOVA_HIP_CHECK(ex, hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  launch_wave(bp, B, cfg, cus, sobol, s);
  hipError_t e = hipStreamEndCapture(s, &g.graph);
  if (e != hipSuccess) return check_hip(ex, e, "hipStreamEndCapture");
  NOVA_HIP_CHECK(ex, hipGraphInstantiate(&g.exec, g.graph, nullptr, nullptr, 0));
  return NOVA_HIP_SUCCESS;
}

// The executor's non-blocking stream (default priority), made to wait for the caller's
// stream so the batch sees the caller's earlier work.
static int executor_stream(nova_hip_ctx *ex, hipStream_t user, hipStream_t *out) {
  if (!ex->batch_stream) NOVA_HIP_CHECK(ex, hipStreamCreateWithFlags(&ex->batch_stream, hipStreamNonBlocking));
  hipEvent_t ev;
  NOVA_HIP_CHECK(ex, hipEventCreateWithFlags(&ev, hipEventDisableTiming));
  hipError_t e = hipEventRecord(ev, user);
  if (e == hipSuccess) e = hipStreamWaitEvent(ex->batch_stream, ev, 0);
  hipEventDestroy(ev);
  if (e != hipSuccess) return check_hip(ex, e, "batch stream ordering");
  *out = ex->batch_stream;
  return NOVA_HIP_SUCCESS;
}

static int make_plan(nova_hip_ctx *ex, const nova_hip_batch_item *items, uint32_t count, BatchPlan &bp) {
  if (count == 0 || count > 4096) return set_error(ex, NOVA_HIP_INVALID_ARGUMENT, "batch needs 1..4096 items");
  bp.scenes.resize(count);
  uint64_t first = 0;
  for (uint32_t i = 0; i < count; i++) {
    const nova_hip_batch_item &it = items[i];
    nova_hip_ctx *c = it.scene;
    const nova_hip_render_params *p = &it.params;
    if (!c || !c->bvh_ready) return set_error(ex, NOVA_HIP_SCENE_NOT_PROCESSED, "batch item without a built scene");
    if (c->device != ex->device) return set_error(ex, NOVA_HIP_INVALID_ARGUMENT, "batch items must live on the executor's device");
    if (p->width <= 0 || p->height <= 0 || p->width > 65536 || p->height > 65536)
      return set_error(ex, NOVA_HIP_INVALID_RENDERBUFFER_DIM, "Wrong buffer width / height dimension.");
    if (p->frame_count <= 0 || p->frame_start < 0 || p->max_depth < 0 || (int64_t)p->frame_start + p->frame_count > INT32_MAX)
      return set_error(ex, NOVA_HIP_INVALID_ARGUMENT, "bad frame range or depth");
    if (p->index_policy != NOVA_HIP_INDEX_CPU && p->index_policy != NOVA_HIP_INDEX_OPTIX)
      return set_error(ex, NOVA_HIP_INVALID_ARGUMENT, "unknown index_policy");
    const int32_t passes = NOVA_HIP_INTEGRATOR_NORMAL | NOVA_HIP_INTEGRATOR_DEPTH | NOVA_HIP_INTEGRATOR_EMISSIVE;
    if (p->integrator_flags != 0 &&
        (!(p->integrator_flags & NOVA_HIP_INTEGRATOR_PATH) || !(p->integrator_flags & NOVA_HIP_INTEGRATOR_COMBINED) || (p->integrator_flags & passes)))
      return set_error(ex, NOVA_HIP_INVALID_ARGUMENT, "batched rendering runs the path integrator (PATH | COMBINED) only");
    if (i > 0 && (p->max_depth != bp.max_depth || p->rr_start != bp.rr_start))
      return set_error(ex, NOVA_HIP_INVALID_ARGUMENT, "batch items must share max_depth and rr_start");
    if (!it.accum_dev || !it.final_dev) return set_error(ex, NOVA_HIP_INVALID_BUFFER_STATE, "null render buffer");
    // batch_accumulate reads and writes a pixel as one float4
    if ((reinterpret_cast<uintptr_t>(it.accum_dev) | reinterpret_cast<uintptr_t>(it.final_dev)) & 15u)
      return set_error(ex, NOVA_HIP_INVALID_BUFFER_STATE, "batch render buffers must be 16-byte aligned");
    if (p->sampling_roughness_texture >= 0) return set_error(ex, NOVA_HIP_INVALID_ARGUMENT, "sampling_roughness_texture is not batched");
    bp.max_depth = p->max_depth;
    bp.rr_start = p->rr_start;
    const uint64_t npix = (uint64_t)p->width * p->height;
    const uint64_t slots = npix * (uint64_t)p->frame_count;
    if (slots >= (1ull << 31)) return set_error(ex, NOVA_HIP_INVALID_ARGUMENT, "batch item too large");
    BatchScene &S = bp.scenes[i];
    S.tc = TraceCtx{c->nodes4, c->tri_leaf, nullptr, 0, nullptr, c->nodes4f};
    S.tie = TieTables{c->bnodes, c->bparent, c->bdepth, c->leaf_of_pos, c->xf_inv};
    S.sc = ShadeCtx{c->textures, c->tex_u32, c->tex_f32, c->materials, c->mesh_xf, c->shade, c->acos_lut, c->envmap, -1};
    S.fp.width = p->width;
    S.fp.height = p->height;
    S.fp.frame0 = p->frame_start;
    S.fp.frame_minor = 1;
    S.fp.tiled = 0;  // batch_accumulate takes a wave's pixels as one run per scene
    S.fp.frames = p->frame_count;
    S.fp.offset = p->max_samples_offset;
    S.fp.vinvert = p->vertical_invert;
    S.fp.policy = p->index_policy;
    for (int k = 0; k < 16; k++) {
      S.fp.inv_P[k] = p->inv_P[k];
      S.fp.inv_V[k] = p->inv_V[k];
    }
    S.first = first;
    S.slots = slots;
    S.npix = (uint32_t)npix;
    S.accum = it.accum_dev;
    first += slots;
    bp.stack_bound = std::max(bp.stack_bound, traversal_stack_bound(c));
    bp.coherent0 = bp.coherent0 && slots % 64 == 0;
  }
  bp.coherent0 = bp.coherent0 && bp.stack_bound < (uint32_t)kWaveStack;
  uint32_t dims = 2;
  for (int k = 0; k <= bp.max_depth; k++) dims += 4 + ((bp.rr_start >= 0 && k >= bp.rr_start) ? 1 : 0);
  if (dims > 256) return set_error(ex, NOVA_HIP_INVALID_ARGUMENT, "max_depth needs more than 256 sampler dimensions");
  bp.total = first;
  return NOVA_HIP_SUCCESS;
}

static int check_config(nova_hip_ctx *ex, const nova_hip_batch_config &c) {
  if (c.queue_capacity < 256 || c.queue_capacity > (1u << 26) || (c.queue_capacity & 255) != 0)
    return set_error(ex, NOVA_HIP_INVALID_ARGUMENT, "queue_capacity must be a multiple of 256 in [256, 2^26]");
  if (c.trace_blocks_per_cu < 1 || c.trace_blocks_per_cu > 64 || c.shade_blocks_per_cu < 1 || c.shade_blocks_per_cu > 64)
    return set_error(ex, NOVA_HIP_INVALID_ARGUMENT, "blocks per CU must be in [1, 64]");
  return NOVA_HIP_SUCCESS;
}

}  // namespace nova_hip

using namespace nova_hip;

extern "C" int nova_hip_render_batch(nova_hip_ctx *ex, const nova_hip_batch_item *items, uint32_t count, const nova_hip_batch_config *config,
                                     nova_hip_batch_stats *stats, float *completion_ms, void *stream) {
  if (!ex || !items || !config) return NOVA_HIP_INVALID_ARGUMENT;
  int st = check_config(ex, *config);
  if (st) return st;
  NOVA_HIP_CHECK(ex, hipSetDevice(ex->device));
  BatchPlan bp;
  if ((st = make_plan(ex, items, count, bp))) return st;
  // graphs cannot be captured on the null stream: the batch runs on the executor's own
  // stream, ordered after the caller's stream and synchronized before returning
  hipStream_t us = (hipStream_t)stream, s = nullptr;
  if ((st = executor_stream(ex, us, &s))) return st;
  const nova_hip_batch_config cfg = *config;
  const uint32_t cap = (uint32_t)std::min<uint64_t>(cfg.queue_capacity, (bp.total + 255) / 256 * 256);
  const WavePlan wp = plan_waves(bp, 0, bp.total, cap);
  if (!wp.ok) return set_error(ex, NOVA_HIP_INVALID_ARGUMENT, "batch wave start not on a packet boundary");
  const int cus = device_cus_of(ex->device);
  const uint32_t trace_threads = std::min<uint32_t>((cap + kTraceBlock - 1) / kTraceBlock, (uint32_t)cus * cfg.trace_blocks_per_cu) * kTraceBlock;
  BatchBuffers B;
  if ((st = alloc_buffers(ex, bp, wp, cap, trace_threads, B, s))) return st;
  const uint32_t *sobol = ex->sobol_table;
  for (uint32_t i = 0; i < count; i++) {  // accumulators start at zero
    const size_t n = (size_t)items[i].params.width * items[i].params.height * 4;
    hipLaunchKernelGGL(batch_zero, dim3((unsigned)std::min<size_t>((n + 255) / 256, 4096)), dim3(256), 0, s, items[i].accum_dev, n);
  }
  hipLaunchKernelGGL(batch_set, dim3(1), dim3(1), 0, s, B.state, 0u, 1u);
  NOVA_HIP_CHECK(ex, hipGetLastError());
  GraphExec g;
  if (cfg.use_graph && (st = capture_wave(ex, bp, B, cfg, cus, sobol, s, g))) return st;
  std::vector<hipEvent_t> ev(wp.waves.size() + 1);
  for (auto &e : ev) NOVA_HIP_CHECK(ex, hipEventCreate(&e));
  NOVA_HIP_CHECK(ex, hipEventRecord(ev[0], s));
  for (size_t w = 0; w < wp.waves.size(); w++) {
    if (cfg.use_graph)
      NOVA_HIP_CHECK(ex, hipGraphLaunch(g.exec, s));
    else
      launch_wave(bp, B, cfg, cus, sobol, s);
    NOVA_HIP_CHECK(ex, hipEventRecord(ev[w + 1], s));
  }
  for (uint32_t i = 0; i < count; i++) {
    const size_t n = (size_t)items[i].params.width * items[i].params.height * 4;
    hipLaunchKernelGGL(batch_finalize, dim3((unsigned)std::min<size_t>((n + 255) / 256, 4096)), dim3(256), 0, s, items[i].accum_dev,
                       items[i].final_dev, n, (float)(items[i].params.frame_start + items[i].params.frame_count));
  }
  NOVA_HIP_CHECK(ex, hipGetLastError());
  NOVA_HIP_CHECK(ex, hipStreamSynchronize(s));
  float total_ms = 0.f;
  hipEventElapsedTime(&total_ms, ev[0], ev.back());
  if (completion_ms) {
    for (uint32_t i = 0; i < count; i++) {
      const uint64_t last = bp.scenes[i].first + bp.scenes[i].slots - 1;
      const size_t w = (size_t)(last / cap);
      float ms = 0.f;
      hipEventElapsedTime(&ms, ev[0], ev[w + 1]);
      completion_ms[i] = ms;
    }
  }
  for (auto &e : ev) hipEventDestroy(e);
  BatchState hs{};
  NOVA_HIP_CHECK(ex, hipMemcpy(&hs, B.state, sizeof(hs), hipMemcpyDeviceToHost));
  if (stats) {
    stats->total_ms = total_ms;
    stats->paths = bp.total;
    stats->waves = wp.waves.size();
    stats->queue_capacity = cap;
    stats->segments = hs.segments;
  }
  return NOVA_HIP_SUCCESS;
}

extern "C" int nova_hip_batch_autotune(nova_hip_ctx *ex, const nova_hip_batch_item *items, uint32_t count, nova_hip_batch_config *best,
                                       float *best_ns_per_path, void *stream) {
  if (!ex || !items || !best) return NOVA_HIP_INVALID_ARGUMENT;
  NOVA_HIP_CHECK(ex, hipSetDevice(ex->device));
  BatchPlan bp;
  int st = make_plan(ex, items, count, bp);
  if (st) return st;
  hipStream_t s = nullptr;
  if ((st = executor_stream(ex, (hipStream_t)stream, &s))) return st;
  const int cus = device_cus_of(ex->device);
  const uint32_t *sobol = ex->sobol_table;
  // the same limit check_config enforces for nova_hip_render_batch (2^26)
  const uint32_t caps[] = {1u << 22, 1u << 23, 1u << 24, 1u << 25, 1u << 26};
  // workgroups per CU: the extend kernel balances its lanes through a work counter, so it
  // takes what fits at once; the shade loop splits its queue evenly, so a second round of
  // workgroups is a candidate against one round that fits
  int occ_t = 0, occ_s = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ_t, reinterpret_cast<const void *>(batch_extend), kTraceBlock, 0) != hipSuccess || occ_t <= 0)
    occ_t = 8;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ_s, reinterpret_cast<const void *>(batch_shade), 256, 0) != hipSuccess || occ_s <= 0)
    occ_s = 4;
  const uint32_t trace_occ[] = {(uint32_t)occ_t};
  const uint32_t shade_occ[] = {(uint32_t)occ_s, 2u * (uint32_t)occ_s};
  // every candidate runs the same slots: the largest capacity's worth (or the whole batch)
  // from the middle of the batch, as many waves of its own capacity as that takes, so a
  // candidate's time per path includes its waves' fixed costs over the same scenes. A wave
  // must start on a multiple of 64 paths (a packet of the per-wave camera walk never
  // straddles two scenes), so the start is rounded down to a multiple of 256, as capacities are.
  const uint64_t span = std::min<uint64_t>(bp.total, caps[sizeof(caps) / sizeof(caps[0]) - 1]);
  const uint64_t mid = bp.total / 2 / 256 * 256;
  const uint64_t start = bp.total - mid >= span ? mid : (bp.total - span) / 256 * 256;
  const uint64_t end = std::min<uint64_t>(bp.total, start + span);
  float best_t = -1.f;
  hipEvent_t e0, e1;
  NOVA_HIP_CHECK(ex, hipEventCreate(&e0));
  NOVA_HIP_CHECK(ex, hipEventCreate(&e1));
  for (uint32_t cap0 : caps)
    for (uint32_t to : trace_occ)
      for (uint32_t so : shade_occ) {
        const uint32_t cap = (uint32_t)std::min<uint64_t>(cap0, (span + 255) / 256 * 256);
        const WavePlan wp = plan_waves(bp, start, end, cap);
        if (!wp.ok) return set_error(ex, NOVA_HIP_INVALID_ARGUMENT, "batch wave start not on a packet boundary");
        nova_hip_batch_config cfg{cap, to, so, 1};
        const uint32_t trace_threads = std::min<uint32_t>((cap + kTraceBlock - 1) / kTraceBlock, (uint32_t)cus * to) * kTraceBlock;
        BatchBuffers B;
        if ((st = alloc_buffers(ex, bp, wp, cap, trace_threads, B, s))) return st;
        GraphExec g;
        if ((st = capture_wave(ex, bp, B, cfg, cus, sobol, s, g))) return st;
        float t = 0.f;
        for (int rep = 0; rep < 3; rep++) {  // rep 0 warms up
          hipLaunchKernelGGL(batch_set, dim3(1), dim3(1), 0, s, B.state, 0u, 0u);
          NOVA_HIP_CHECK(ex, hipEventRecord(e0, s));
          for (size_t w = 0; w < wp.waves.size(); w++) NOVA_HIP_CHECK(ex, hipGraphLaunch(g.exec, s));
          NOVA_HIP_CHECK(ex, hipEventRecord(e1, s));
          NOVA_HIP_CHECK(ex, hipEventSynchronize(e1));
          float ms = 0.f;
          hipEventElapsedTime(&ms, e0, e1);
          if (rep > 0) t += ms;
        }
        const float ns = t * 1e6f / (2.f * (float)(end - start));
        if (best_t < 0.f || ns < best_t) {
          best_t = ns;
          *best = cfg;
        }
      }
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  if (best_ns_per_path) *best_ns_per_path = best_t;
  return NOVA_HIP_SUCCESS;
}
