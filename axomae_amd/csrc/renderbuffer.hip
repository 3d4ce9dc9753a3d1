// Render buffers on the device: the reference's RenderBufferPack / NvRenderBuffer
// (corelib/nova/api/renderbuffer.cpp:8-230) with every buffer in device memory, and the
// engine's per-frame writeFramebuffer + swapBackBuffer (engine.cpp:351-363, 498-525) done
// by the render itself, so no frame is copied to the host unless the caller reads one.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>
#include <utility>

#include "nova_hip.h"
#include "nova_internal.h"

struct nova_hip_renderbuffer {
  nova_hip_ctx *ctx = nullptr;
  unsigned width = 0, height = 0;
  size_t capacity = 0;  // pixels
  float *front = nullptr, *back = nullptr, *partial = nullptr, *accumulator = nullptr, *depth = nullptr, *normal = nullptr;
};

namespace {

constexpr unsigned kColor = 4, kDepth = 1, kNormal = 3;  // RenderBufferPack::CHANNELS_*

void release(nova_hip_renderbuffer *rb) {
  if (rb->capacity) (void)hipDeviceSynchronize();  // queued render work may still write them
  for (float **p : {&rb->front, &rb->back, &rb->partial, &rb->accumulator, &rb->depth, &rb->normal}) {
    if (*p) (void)hipFree(*p);
    *p = nullptr;
  }
  rb->capacity = 0;
}

// Zeroes `pixels` pixels of every buffer (cleanBuffers over the capacity, cleanCanvas over
// width * height), ordered on `s`.
int clear(nova_hip_renderbuffer *rb, size_t pixels, hipStream_t s) {
  nova_hip_ctx *ctx = rb->ctx;
  if (!pixels) return NOVA_HIP_SUCCESS;
  for (float *p : {rb->front, rb->back, rb->partial, rb->accumulator}) NOVA_HIP_CHECK(ctx, hipMemsetAsync(p, 0, pixels * kColor * sizeof(float), s));
  NOVA_HIP_CHECK(ctx, hipMemsetAsync(rb->depth, 0, pixels * kDepth * sizeof(float), s));
  NOVA_HIP_CHECK(ctx, hipMemsetAsync(rb->normal, 0, pixels * kNormal * sizeof(float), s));
  return NOVA_HIP_SUCCESS;
}

}  // namespace

extern "C" int nova_hip_renderbuffer_create(nova_hip_ctx *ctx, nova_hip_renderbuffer **out) {
  if (!ctx || !out) return NOVA_HIP_INVALID_ARGUMENT;
  *out = new nova_hip_renderbuffer();
  (*out)->ctx = ctx;
  return NOVA_HIP_SUCCESS;
}

extern "C" int nova_hip_renderbuffer_destroy(nova_hip_renderbuffer *rb) {
  if (!rb) return NOVA_HIP_SUCCESS;
  (void)hipSetDevice(rb->ctx->device);
  release(rb);
  delete rb;
  return NOVA_HIP_SUCCESS;
}

extern "C" int nova_hip_renderbuffer_preallocate(nova_hip_renderbuffer *rb, unsigned max_width, unsigned max_height) {
  if (!rb) return NOVA_HIP_INVALID_ARGUMENT;
  nova_hip_ctx *ctx = rb->ctx;
  if (max_width == 0 || max_height == 0) return nova_hip::set_error(ctx, NOVA_HIP_INVALID_RENDERBUFFER_DIM, "render buffer of zero size");
  NOVA_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  release(rb);
  const size_t cap = (size_t)max_width * max_height;
  const std::pair<float **, unsigned> bufs[] = {{&rb->front, kColor},       {&rb->back, kColor},  {&rb->partial, kColor},
                                                {&rb->accumulator, kColor}, {&rb->depth, kDepth}, {&rb->normal, kNormal}};
  for (const auto &b : bufs) {
    if (hipMalloc(b.first, cap * b.second * sizeof(float)) != hipSuccess) {
      (void)hipGetLastError();
      release(rb);
      return nova_hip::set_error(ctx, NOVA_HIP_OUT_OF_MEMORY, "render buffer allocation failed");
    }
  }
  rb->capacity = cap;
  const int rc = clear(rb, cap, nullptr);  // the reference's vectors start zeroed
  if (rc != NOVA_HIP_SUCCESS) return rc;
  NOVA_HIP_CHECK(ctx, hipStreamSynchronize(nullptr));
  return NOVA_HIP_SUCCESS;
}

extern "C" int nova_hip_renderbuffer_resize(nova_hip_renderbuffer *rb, unsigned width, unsigned height) {
  if (!rb) return NOVA_HIP_INVALID_ARGUMENT;
  if (rb->capacity < (size_t)width * height) {
    const int rc = nova_hip_renderbuffer_preallocate(rb, width, height);
    if (rc != NOVA_HIP_SUCCESS) return rc;
  }
  rb->width = width;
  rb->height = height;
  return NOVA_HIP_SUCCESS;
}

extern "C" int nova_hip_renderbuffer_reset(nova_hip_renderbuffer *rb, void *stream) {
  if (!rb) return NOVA_HIP_INVALID_ARGUMENT;
  if (!rb->capacity) return nova_hip::set_error(rb->ctx, NOVA_HIP_INVALID_BUFFER_STATE, "render buffer not allocated");
  NOVA_HIP_CHECK(rb->ctx, hipSetDevice(rb->ctx->device));
  return clear(rb, rb->capacity, (hipStream_t)stream);
}

extern "C" int nova_hip_renderbuffer_clean_canvas(nova_hip_renderbuffer *rb, void *stream) {
  if (!rb) return NOVA_HIP_INVALID_ARGUMENT;
  if (!rb->capacity) return nova_hip::set_error(rb->ctx, NOVA_HIP_INVALID_BUFFER_STATE, "render buffer not allocated");
  NOVA_HIP_CHECK(rb->ctx, hipSetDevice(rb->ctx->device));
  return clear(rb, (size_t)rb->width * rb->height, (hipStream_t)stream);
}

extern "C" int nova_hip_renderbuffer_swap(nova_hip_renderbuffer *rb) {
  if (!rb) return NOVA_HIP_INVALID_ARGUMENT;
  std::swap(rb->front, rb->back);
  return NOVA_HIP_SUCCESS;
}

extern "C" int nova_hip_renderbuffer_get(const nova_hip_renderbuffer *rb, nova_hip_renderbuffer_view *out) {
  if (!rb || !out) return NOVA_HIP_INVALID_ARGUMENT;
  out->front = rb->front;
  out->back = rb->back;
  out->partial = rb->partial;
  out->accumulator = rb->accumulator;
  out->depth = rb->depth;
  out->normal = rb->normal;
  out->width = rb->width;
  out->height = rb->height;
  out->capacity = (uint64_t)rb->capacity;
  return NOVA_HIP_SUCCESS;
}

extern "C" int nova_hip_render_to_renderbuffer(nova_hip_ctx *ctx, const nova_hip_render_params *params, nova_hip_renderbuffer *rb, void *stream) {
  if (!ctx || !params || !rb) return NOVA_HIP_INVALID_ARGUMENT;
  if (rb->ctx->device != ctx->device) return nova_hip::set_error(ctx, NOVA_HIP_INVALID_ARGUMENT, "render buffer belongs to another device");
  if (!rb->capacity) return nova_hip::set_error(ctx, NOVA_HIP_INVALID_BUFFER_STATE, "render buffer not allocated");
  if ((unsigned)params->width != rb->width || (unsigned)params->height != rb->height)
    return nova_hip::set_error(ctx, NOVA_HIP_INVALID_RENDERBUFFER_DIM,
                               "render size " + std::to_string(params->width) + "x" + std::to_string(params->height) + " differs from the render buffer's " +
                                   std::to_string(rb->width) + "x" + std::to_string(rb->height));
  const int rc = nova_hip_render_frames(ctx, params, rb->accumulator, rb->back, rb->partial, stream);
  if (rc != NOVA_HIP_SUCCESS) return rc;
  std::swap(rb->front, rb->back);  // the finished image is displayed; the old front becomes the next back buffer
  return NOVA_HIP_SUCCESS;
}

extern "C" int nova_hip_render_to_renderbuffer_lookahead(nova_hip_ctx *ctx, const nova_hip_render_params *params, int32_t lookahead,
                                                         nova_hip_renderbuffer *rb, void *stream) {
  if (!ctx || !params || !rb) return NOVA_HIP_INVALID_ARGUMENT;
  if (rb->ctx->device != ctx->device) return nova_hip::set_error(ctx, NOVA_HIP_INVALID_ARGUMENT, "render buffer belongs to another device");
  if (!rb->capacity) return nova_hip::set_error(ctx, NOVA_HIP_INVALID_BUFFER_STATE, "render buffer not allocated");
  if ((unsigned)params->width != rb->width || (unsigned)params->height != rb->height)
    return nova_hip::set_error(ctx, NOVA_HIP_INVALID_RENDERBUFFER_DIM,
                               "render size " + std::to_string(params->width) + "x" + std::to_string(params->height) + " differs from the render buffer's " +
                                   std::to_string(rb->width) + "x" + std::to_string(rb->height));
  const int rc = nova_hip_render_frame_lookahead(ctx, params, lookahead, rb->accumulator, rb->back, rb->partial, stream);
  if (rc != NOVA_HIP_SUCCESS) return rc;
  std::swap(rb->front, rb->back);
  return NOVA_HIP_SUCCESS;
}

extern "C" int nova_hip_renderbuffer_read(const nova_hip_renderbuffer *rb, int which, float *host, void *stream) {
  if (!rb || !host) return NOVA_HIP_INVALID_ARGUMENT;
  nova_hip_ctx *ctx = rb->ctx;
  const float *src = nullptr;
  unsigned ch = kColor;
  switch (which) {
    case NOVA_HIP_RB_FRONT: src = rb->front; break;
    case NOVA_HIP_RB_BACK: src = rb->back; break;
    case NOVA_HIP_RB_PARTIAL: src = rb->partial; break;
    case NOVA_HIP_RB_ACCUMULATOR: src = rb->accumulator; break;
    case NOVA_HIP_RB_DEPTH: src = rb->depth; ch = kDepth; break;
    case NOVA_HIP_RB_NORMAL: src = rb->normal; ch = kNormal; break;
    default: return nova_hip::set_error(ctx, NOVA_HIP_INVALID_ARGUMENT, "unknown render buffer " + std::to_string(which));
  }
  if (!src) return nova_hip::set_error(ctx, NOVA_HIP_INVALID_BUFFER_STATE, "render buffer not allocated");
  NOVA_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  const size_t bytes = (size_t)rb->width * rb->height * ch * sizeof(float);
  NOVA_HIP_CHECK(ctx, hipMemcpyAsync(host, src, bytes, hipMemcpyDeviceToHost, (hipStream_t)stream));
  NOVA_HIP_CHECK(ctx, hipStreamSynchronize((hipStream_t)stream));
  return NOVA_HIP_SUCCESS;
}
