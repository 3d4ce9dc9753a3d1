// Queue compaction shared by the wavefront kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace nova_hip {

// Survivor compaction of one loop iteration of a kBlock-thread workgroup: the workgroup's
// waves reserve one range of the output queue with a single atomic on *count, and each
// wave's survivors follow the earlier waves'. Returns the first index of the calling
// wave's part; a survivor's index is that plus its rank among the wave's survivors (mask m).
// Every thread of the workgroup must call it once per iteration `it` (it holds two
// barriers). The tables are double-buffered on `it`, so a wave that runs ahead into the
// next iteration cannot overwrite counts another wave still reads. One atomic per wave on
// a counter shared by the whole chip serialized the kernel: this is 4x fewer at kBlock 256.
template <int kBlock>
__device__ __forceinline__ uint32_t block_reserve(uint64_t m, uint32_t *count, uint32_t it, uint32_t (&wave_cnt)[2][kBlock / 64],
                                                  uint32_t (&blk_base)[2]) {
  const uint32_t wv = threadIdx.x >> 6, buf = it & 1u;
  if ((threadIdx.x & 63) == 0) wave_cnt[buf][wv] = (uint32_t)__popcll(m);
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t tot = 0;
    for (int k = 0; k < kBlock / 64; k++) tot += wave_cnt[buf][k];
    blk_base[buf] = tot ? atomicAdd(count, tot) : 0u;
  }
  __syncthreads();
  uint32_t base = blk_base[buf];
  for (uint32_t k = 0; k < wv; k++) base += wave_cnt[buf][k];
  return base;
}

}  // namespace nova_hip
