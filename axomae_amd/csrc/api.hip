// Context management, error reporting and the Sobol fill entry point.
#include <cstring>
#include <string>

#include "device_math.h"
#include "nova_internal.h"

namespace nova_hip {

static const uint32_t kSobolTable[256][32] = {
#include "sobol_table.inc"
};

int set_error(nova_hip_ctx *ctx, int code, const std::string &msg) {
  if (ctx) ctx->last_error = msg;
  return code;
}

int check_hip(nova_hip_ctx *ctx, hipError_t e, const char *what) {
  std::string m = std::string(what) + ": " + hipGetErrorString(e);
  return set_error(ctx, e == hipErrorOutOfMemory ? NOVA_HIP_OUT_OF_MEMORY : NOVA_HIP_DEVICE_ERROR, m);
}

void *ensure_scratch(nova_hip_ctx *ctx, int slot, size_t bytes, hipError_t *err) {
  DeviceBuffer &b = ctx->scratch[slot];
  *err = hipSuccess;
  if (b.bytes >= bytes && b.ptr) return b.ptr;
  if (b.ptr) {
    // work queued by a call that did not wait (NOVA_HIP_RENDER_ASYNC) may still use it
    (void)hipDeviceSynchronize();
    hipFree(b.ptr);
    b.ptr = nullptr;
    b.bytes = 0;
  }
  size_t want = bytes + bytes / 4 + 256;
  *err = hipMalloc(&b.ptr, want);
  if (*err != hipSuccess) {
    b.ptr = nullptr;
    return nullptr;
  }
  b.bytes = want;
  return b.ptr;
}

// OwenScrambled Sobol for arbitrary (index, dim) pairs: one sample per lane; the
// direction row differs per lane here, so it is read from global memory (L1-resident).
__global__ void sobol_fill_kernel(const uint32_t *__restrict__ table, const uint32_t *__restrict__ idx, const uint32_t *__restrict__ dim,
                                  uint64_t n, float *__restrict__ out, uint32_t *__restrict__ out_bits) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint32_t index = idx[i];
  uint32_t d = dim[i];
  if (d >= 256) {  // outside the direction table: reported as NaN, never read past the table
    out[i] = __builtin_nanf("");
    if (out_bits) out_bits[i] = 0xFFFFFFFFu;
    return;
  }
  uint32_t v = owen_scramble(sobol_xor(table + (size_t)d * 32, index ^ (index >> 1)));
  out[i] = u32_to_unit(v);
  if (out_bits) out_bits[i] = v;
}

// Direction table on the current device for the context-free entry points: one copy per
// device, kept until the process ends; nullptr if it cannot be made.
const uint32_t *device_sobol_table() {
  static uint32_t *tables[64] = {};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return nullptr;
  if (!tables[dev]) {
    uint32_t *t = nullptr;
    if (hipMalloc(&t, sizeof(kSobolTable)) != hipSuccess) return nullptr;
    if (hipMemcpy(t, kSobolTable, sizeof(kSobolTable), hipMemcpyHostToDevice) != hipSuccess) {
      hipFree(t);
      return nullptr;
    }
    tables[dev] = t;
  }
  return tables[dev];
}

}  // namespace nova_hip

using namespace nova_hip;

extern "C" {

int nova_hip_create(int device, nova_hip_ctx **out) {
  if (!out) return NOVA_HIP_INVALID_ARGUMENT;
  *out = nullptr;
  int count = 0;
  hipError_t e = hipGetDeviceCount(&count);
  if (e != hipSuccess || count == 0) return NOVA_HIP_DEVICE_ERROR;
  if (device < 0 || device >= count) return NOVA_HIP_INVALID_ARGUMENT;
  e = hipSetDevice(device);
  if (e != hipSuccess) return NOVA_HIP_DEVICE_ERROR;
  nova_hip_ctx *ctx = new nova_hip_ctx();
  ctx->device = device;
  // the no-allocation entry point's scratch address on this device (its status is reported
  // by that entry point; the others do not use it)
  noalloc_scratch_setup(device);
  e = hipMalloc(&ctx->sobol_table, sizeof(kSobolTable));
  if (e == hipSuccess) e = hipMemcpy(ctx->sobol_table, kSobolTable, sizeof(kSobolTable), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    nova_hip_destroy(ctx);
    return NOVA_HIP_DEVICE_ERROR;
  }
  *out = ctx;
  return NOVA_HIP_SUCCESS;
}

static void free_scene(nova_hip_ctx *ctx) {
  // render work queued without a wait (NOVA_HIP_RENDER_ASYNC) may still read these
  (void)hipDeviceSynchronize();
  void **ptrs[] = {(void **)&ctx->tri_world, (void **)&ctx->shade,     (void **)&ctx->prim_aabb, (void **)&ctx->prim_centroid,
                   (void **)&ctx->mesh_xf,   (void **)&ctx->materials, (void **)&ctx->textures,  (void **)&ctx->tex_u32,
                   (void **)&ctx->tex_f32,   (void **)&ctx->acos_lut,  (void **)&ctx->prim_idx,  (void **)&ctx->bnodes,
                   (void **)&ctx->nodes4,    (void **)&ctx->nodes4f,   (void **)&ctx->tri_leaf,  (void **)&ctx->bparent, (void **)&ctx->bdepth,
                   (void **)&ctx->leaf_of_pos, (void **)&ctx->tie,   (void **)&ctx->xf_inv};
  for (void **p : ptrs)
    if (*p) {
      hipFree(*p);
      *p = nullptr;
    }
  ctx->scene_ready = false;
  ctx->bvh_ready = false;
  ctx->scene_gen++;
}

void nova_hip_free_scene_internal(nova_hip_ctx *ctx) { free_scene(ctx); }

int nova_hip_destroy(nova_hip_ctx *ctx) {
  if (!ctx) return NOVA_HIP_INVALID_ARGUMENT;
  hipSetDevice(ctx->device);
  free_scene(ctx);
  if (ctx->sobol_table) hipFree(ctx->sobol_table);
  for (auto &b : ctx->scratch)
    if (b.ptr) hipFree(b.ptr);
  if (ctx->gstack) hipFree(ctx->gstack);
  if (ctx->batch_stream) hipStreamDestroy(ctx->batch_stream);
  if (ctx->ev_begin) hipEventDestroy(ctx->ev_begin);
  if (ctx->ev_end) hipEventDestroy(ctx->ev_end);
  for (hipEvent_t ev : ctx->ev_pool) hipEventDestroy(ev);
  delete ctx;
  return NOVA_HIP_SUCCESS;
}

const char *nova_hip_last_error(const nova_hip_ctx *ctx) { return ctx ? ctx->last_error.c_str() : "null context"; }

int nova_hip_sobol_fill(const uint32_t *idx_dev, const uint32_t *dim_dev, uint64_t n, float *out_dev, uint32_t *out_bits_dev, void *stream) {
  if (n == 0) return NOVA_HIP_SUCCESS;
  if (!idx_dev || !dim_dev || !out_dev) return NOVA_HIP_INVALID_ARGUMENT;
  const uint32_t *table = device_sobol_table();
  if (!table) return NOVA_HIP_OUT_OF_MEMORY;
  const unsigned block = 256;
  const uint64_t grid = (n + block - 1) / block;
  hipLaunchKernelGGL(sobol_fill_kernel, dim3((unsigned)grid), dim3(block), 0, (hipStream_t)stream, table, idx_dev, dim_dev, n, out_dev,
                     out_bits_dev);
  return hipGetLastError() == hipSuccess ? NOVA_HIP_SUCCESS : NOVA_HIP_DEVICE_ERROR;
}

}  // extern "C"
