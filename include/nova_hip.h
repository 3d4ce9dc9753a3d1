/*
 * nova_hip.h - C ABI of the MI355X (gfx950) Nova hot path.
 *
 * This is the drop-in boundary for the slot the reference fills with its
 * CUDA/OptiX backend: the device accelerator built by Engine::buildAcceleration
 * and launched once per frame through DeviceIntersectorInterface::traverse.
 * Every function returns an int status (NOVA_HIP_SUCCESS = 0); no C++ exception
 * crosses the ABI and nova_hip_last_error() returns the message of the last
 * failure on a context. A context is driven from one host thread; all device
 * work is ordered on the stream passed to it (NULL = the default stream).
 *
 * Citations name files of the reference (HamilcarR/Axomae) relative to its root.
 */
#ifndef NOVA_HIP_H
#define NOVA_HIP_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes: 0..16 keep nova::ERROR_STATE's numbering (corelib/nova/api_common.h:7-27);
 * 100+ are device-layer failures the reference reports as RENDERER_EXCEPTION
 * (corelib/nova/api/engine.cpp:478-486). */
enum nova_hip_status {
  NOVA_HIP_SUCCESS = 0,
  NOVA_HIP_INVALID_ENGINE_STATE = 1,
  NOVA_HIP_INVALID_BUFFER_STATE = 2,
  NOVA_HIP_INVALID_SCENE_TYPE = 4,
  NOVA_HIP_INVALID_TRANSFORM_TYPE = 5,
  NOVA_HIP_SCENE_NOT_PROCESSED = 9,
  NOVA_HIP_OUT_OF_MEMORY = 11,
  NOVA_HIP_INVALID_ARGUMENT = 15,
  NOVA_HIP_RENDERER_EXCEPTION = 16,
  NOVA_HIP_DEVICE_ERROR = 100,
  NOVA_HIP_INVALID_RENDERBUFFER_DIM = 101 /* nova::exception::INVALID_RENDERBUFFER_DIM (utils/DrawEngine.cpp:44-47) */
};

typedef struct nova_hip_ctx nova_hip_ctx;

/* Context on one HIP device. Replaces DeviceAcceleratorInterface::make()
 * (corelib/nova/aggregate/api/device_accelerator_factory.cpp:9-17) and the
 * CUDA context the engine's user callback sets per frame (api_engine.h:31-48). */
int nova_hip_create(int device, nova_hip_ctx **out);
/* Releases every device buffer: DeviceAcceleratorInterface::cleanup() + destructor
 * (corelib/nova/aggregate/api/optix/host.cpp:542-580). */
int nova_hip_destroy(nova_hip_ctx *ctx);
const char *nova_hip_last_error(const nova_hip_ctx *ctx);

/* ------------------------------------------------------------------ scene */

/* One triangle mesh: the Object3D attribute spans (corelib/internal/geometry/Object3D.h:387-394)
 * plus its model matrix (column-major, glm::value_ptr order) and material index.
 * normals/tangents/bitangents are vertex_count*3 floats, uv vertex_count*2. */
typedef struct {
  const float *vertices, *normals, *tangents, *bitangents, *uv;
  const uint32_t *indices;
  uint64_t vertex_count, index_count;
  float transform[16];
  int32_t material;
} nova_hip_mesh_desc;

/* PrincipledMaterial texture pack + refractive indices (corelib/nova/material/NovaMaterials.h:45-138,
 * corelib/nova/api/material.cpp:212-226). Texture fields are indices into the scene's
 * texture list (-1 = absent). albedo_f32 / roughness_f32 and emission_scale are the FP32
 * material-texture and emission extensions (1.0 = reference behaviour). */
typedef struct {
  int32_t albedo, metallic, normal, emissive;
  int32_t albedo_f32, roughness_f32;
  float eta[3], k[3];
  float anisotropy;
  float emission_scale;
} nova_hip_material_desc;

enum nova_hip_texture_format { NOVA_HIP_TEX_U32 = 0, NOVA_HIP_TEX_F32 = 1 };

/* TextureRawData (corelib/nova/texturing/nova_texturing.h:98-112): u32 texels packed
 * (a<<24)|(r<<16)|(g<<8)|b, or `channels` floats per texel. */
typedef struct {
  int32_t width, height, channels;
  int32_t invert_u, invert_v;
  int32_t format;
  const void *data;
} nova_hip_texture_desc;

/* primitive_aggregate_data_s (corelib/nova/aggregate/aggregate_datastructures.h:17-20) with
 * the texture/material holders of NovaResourceManager. Global primitive order is
 * mesh-major, triangle-minor (corelib/nova/api/engine.cpp:159-175). */
typedef struct {
  const nova_hip_mesh_desc *meshes;
  uint32_t mesh_count;
  const nova_hip_material_desc *materials;
  uint32_t material_count;
  const nova_hip_texture_desc *textures;
  uint32_t texture_count;
  int32_t envmap_texture; /* TextureResourcesHolder::getCurrentEnvmap (nova_texturing.h:130-133) */
} nova_hip_scene_desc;

/* Uploads and flattens the scene on the device (host buffers are copied; the caller may
 * free them afterwards). Replaces the OptiX GAS input assembly of OptixAccelerator::build
 * (corelib/nova/aggregate/api/optix/host.cpp:432-528) and the managed-memory views of
 * NovaResourceManager (corelib/nova/manager/NovaResourceManager.h:28-82). */
int nova_hip_scene_upload(nova_hip_ctx *ctx, const nova_hip_scene_desc *scene, void *stream);

/* ------------------------------------------------------------------- BVH */

typedef struct {
  uint32_t primitive_count;
  uint32_t binary_nodes;   /* nodes of the SAH tree, BvhtlBuilder numbering differs (pairs still adjacent) */
  uint32_t binary_depth;
  uint32_t wide_nodes;     /* 4-wide nodes after the collapse */
  uint32_t wide_depth;
  uint32_t leaves;
  float build_ms;          /* device time of SAH build + collapse */
} nova_hip_bvh_info;

/* GPU SAH build with the reference builder's exact split decisions
 * (BvhtlBuilder::buildTriangleBasedScene, PERFORMANCE = 4 candidates per axis, SAH:
 * corelib/nova/aggregate/native/BvhBuilder.cpp:37-253), then a collapse to a 4-wide BVH.
 * `build_candidates` must be 4 (the reference's default build type). The primitive
 * order prim_idx equals the reference builder's bit for bit. */
int nova_hip_bvh_build(nova_hip_ctx *ctx, int build_candidates, nova_hip_bvh_info *info, void *stream);

/* Copies the builder outputs to host memory: prim_idx (primitive_count int32) and the
 * binary nodes as Bvhnl records {float min[3], max[3]; int32 left, primitive_count}
 * (BvhBuilder.h:10-15), binary_nodes entries. Either pointer may be NULL. */
int nova_hip_bvh_read(nova_hip_ctx *ctx, int32_t *prim_idx, void *binary_nodes);

/* Closest hit for a batch of rays in device memory: rays are n x {ox,oy,oz,dx,dy,dz}
 * floats; outputs are the global primitive id (-1 on miss) and {t,u,v} per ray.
 * The same kernel runs in the integrator's extend stage. Replaces
 * GenericHostAccelerator::hit (corelib/nova/aggregate/acceleration_interface.h:13-32,
 * native.cpp:56-114 + Triangle.h:146-213). tnear is 1e-6 as Ray's default (ray/Ray.h:15). */
int nova_hip_closest_hit(nova_hip_ctx *ctx, const float *rays_dev, uint64_t n, int32_t *prim_dev, float *tuv_dev, void *stream);

/* ---------------------------------------------------------------- render */

typedef struct {
  int32_t width, height;
  int32_t max_depth;        /* RenderOptions::getMaxDepth: maxDepth+1 scatters (Integrator.cpp:62-66) */
  int32_t frame_start;      /* first frame index (Engine frame_index) */
  int32_t frame_count;      /* frames rendered, 1 spp each (engine.cpp:498-525) */
  uint32_t max_samples_offset; /* renderer_max_samples added to the sample index (Integrator.h:80) */
  int32_t vertical_invert;  /* EngineResourcesHolder::vertical_invert (Integrator.h:38-42) */
  int32_t rr_start;         /* Russian-roulette extension: first scatter index that rolls, -1 = off */
  float inv_P[16], inv_V[16]; /* CameraResourcesHolder inv_P / inv_V, column-major */
  int32_t flags;            /* NOVA_HIP_RENDER_TIMING | NOVA_HIP_RENDER_COUNT_TRAVERSAL, 0 for none */
  int32_t sampling_roughness_texture; /* -1 for the plain renderer. Gradient-check harness only: an
                             FP32 texture whose channel 0 replaces, for every material with a
                             roughness_f32 texture, the roughness used to sample directions and
                             to weight them by their pdf; the BSDF value keeps roughness_f32. */
  int32_t index_policy;     /* Sobol sample index of pixel (x, y) in frame f:
                             NOVA_HIP_INDEX_CPU: 4*(y*W+x)*f + max_samples_offset, row flipped
                             with vertical_invert (Integrator.h:38-42,80; the parity default);
                             NOVA_HIP_INDEX_OPTIX: (x*H+y)*max_samples_offset + f, the reference's
                             GPU raygen (integrator_optix.cu:43), max_samples_offset = sample_max. */
  int32_t integrator_flags; /* EngineResourcesHolder::integrator_flag (api_common.h:29-49); 0 means
                             PATH | COMBINED. With PATH, the pass bits NORMAL, DEPTH and EMISSIVE
                             select a render pass instead of the path integrator; when several are
                             set the last in integrator_dispatch order wins (COMBINED, NORMAL,
                             DEPTH, EMISSIVE), since each overwrites the frame. */
  float z_near, z_far;      /* CameraResourcesHolder near / far, read by the DEPTH pass */
} nova_hip_render_params;

enum { NOVA_HIP_INDEX_CPU = 0, NOVA_HIP_INDEX_OPTIX = 1 };
/* integrator::TYPE values (api_common.h:29-49) */
enum {
  NOVA_HIP_INTEGRATOR_PATH = 1 << 0,
  NOVA_HIP_INTEGRATOR_COMBINED = 1 << 8,
  NOVA_HIP_INTEGRATOR_NORMAL = 1 << 9,
  NOVA_HIP_INTEGRATOR_DEPTH = 1 << 10,
  NOVA_HIP_INTEGRATOR_EMISSIVE = 1 << 13
};

/* flags: time every extend/shade launch with HIP events on the render stream; count BVH node
 * visits and triangle tests per ray (an instrumented traversal, slower). */
/* NOVA_HIP_RENDER_PER_RAY_TRAVERSAL: trace camera rays one walk per ray even where a wave
 * would walk the tree once for its rays (the result is the same; for tests and tuning). */
/* NOVA_HIP_ADJOINT_RETRACE: the adjoint's replay traces its rays again instead of reusing
 * the primal pass's hits (the same rays, so the same hits; for tests). */
/* NOVA_HIP_RENDER_NO_TAIL: keep one extend and one shade launch per bounce to the end, where
 * the renderer otherwise finishes the last bounces of a batch in one launch, one path per
 * lane, once few paths remain (the result is the same; for tests and tuning). */
/* NOVA_HIP_RENDER_ASYNC: a render call returns once its work is queued on the stream instead
 * of waiting for it to finish (an engine loop can queue its next frame meanwhile). The call
 * reads nothing back: nova_hip_render_stats keeps describing the last call that waited. The
 * context's buffers stay in use until the stream has done the work, so until then the caller
 * makes every other call on the context on the same stream (render calls with the flag or
 * without, texture updates) and synchronizes the stream before reading results on the host.
 * Calls that free or replace device buffers (scene upload, BVH build, render-buffer resize
 * or destroy, a render that needs more scratch, nova_hip_destroy) wait for the device first. Not with TIMING or
 * COUNT_TRAVERSAL, which read back (NOVA_HIP_INVALID_ARGUMENT), nor for the adjoint. No
 * counterpart in the reference, whose traverse waits for its frame. */
enum {
  NOVA_HIP_RENDER_TIMING = 1,
  NOVA_HIP_RENDER_COUNT_TRAVERSAL = 2,
  NOVA_HIP_RENDER_PER_RAY_TRAVERSAL = 4,
  NOVA_HIP_ADJOINT_RETRACE = 8,
  NOVA_HIP_RENDER_NO_TAIL = 16,
  NOVA_HIP_RENDER_ASYNC = 32
};

/* Renders frames [frame_start, frame_start+frame_count) of the path integrator
 * (PathIntegrator::Li, corelib/nova/integrator/Integrator.cpp:59-98, sample indices by
 * params->index_policy) and accumulates them as
 * Engine::writeFramebuffer does (engine.cpp:351-363): accum += partial,
 * final = accum / (frame+1). accum_dev and final_dev are device RGBA32F buffers of
 * width*height*4 floats; accum must hold the previous frames' sum (zero at frame 0).
 * frame_start >= 0, frame_count >= 1 and frame_start + frame_count <= INT32_MAX, else
 * NOVA_HIP_INVALID_ARGUMENT (frame indices are the engine's int sample index).
 * partial_dev (optional) receives the last frame's radiance (HdrBufferStruct::partial_buffer).
 * With accum_dev and final_dev both NULL and partial_dev set, only the last frame's radiance
 * is written, as DeviceIntersectorInterface::traverse leaves it for the engine to accumulate.
 * The render passes (integrator_flags NORMAL / DEPTH / EMISSIVE: NormalIntegrator.cpp,
 * DepthIntegrator.cpp, EmissiveIntegrator.cpp) trace one unjittered ray per pixel, give the
 * same value every frame and accumulate it the same way. A flag word without PATH returns
 * NOVA_HIP_RENDERER_EXCEPTION (nova::exception::INVALID_INTEGRATOR), PATH without a pass
 * bit likewise (INVALID_RENDER_MODE).
 * Replaces gpu_draw -> device_start_integrator -> OptixIntersector::traverse
 * (corelib/nova/utils/DrawEngine_GPU.cpp:14-20, integrator/cuda/job_dispatch.cpp:7-12,
 * aggregate/api/optix/device.cpp:18-21). */
int nova_hip_render_frames(nova_hip_ctx *ctx, const nova_hip_render_params *params, float *accum_dev, float *final_dev,
                           float *partial_dev, void *stream);

/* One frame (params->frame_count must be 1) as nova_hip_render_frames renders it, served from
 * frames rendered ahead, for an engine loop that renders one 1-spp frame per call with the
 * same camera and settings (Engine::renderGpu -> DeviceIntersectorInterface::traverse,
 * engine.cpp:461-494). When an earlier call of this function on the context kept frame
 * frame_start for otherwise identical parameters on the same scene (no scene upload, BVH
 * build or texture update since), that frame's radiance is added to accum_dev (final_dev =
 * accum_dev / (frame_start + 1)) and copied to partial_dev, as a one-frame render would; else
 * frames frame_start .. frame_start + lookahead - 1 are rendered in one call, their images kept
 * in device memory (lookahead * width * height * 16 bytes, grown on demand, freed with the
 * context), and the first is returned. lookahead is in [1, 1024]; frames past INT32_MAX - 1
 * are not rendered ahead. Each frame equals its nova_hip_render_frames frame bit
 * for bit: a path's samples and arithmetic do not depend on the batch it runs in. lookahead 1
 * (or a render pass) renders the frame alone. nova_hip_render_stats describes the last call
 * that rendered. Buffers as for nova_hip_render_frames. No counterpart in the reference. */
int nova_hip_render_frame_lookahead(nova_hip_ctx *ctx, const nova_hip_render_params *params, int32_t lookahead, float *accum_dev,
                                    float *final_dev, float *partial_dev, void *stream);

/* nova_hip_render_frames without any allocation, for callers that cannot allocate device
 * memory (code under stream capture, tight serving loops, hosts that forbid hipMalloc). No
 * counterpart in the reference. The path integrator only (integrator_flags 0 or
 * PATH | COMBINED); NOVA_HIP_RENDER_TIMING and NOVA_HIP_RENDER_COUNT_TRAVERSAL are refused.
 * Queues, path state and traversal stacks come from a fixed 320 MiB device array that the
 * library holds on each device it runs on, sized for one 1920x1080 frame per batch: frames
 * are rendered in batches of as many as fit, and an image whose single frame does not fit is
 * refused with NOVA_HIP_INVALID_ARGUMENT before any work. The call is stream-ordered and
 * does not wait (it records no event and leaves nova_hip_render_stats as it was). Every call
 * on a device shares that device's array: calls that overlap on one device (several
 * streams or threads) race and must be serialized by the caller. The scene and BVH must be
 * built already (those calls allocate); the image equals nova_hip_render_frames' bit for bit. */
int nova_hip_render_frames_noalloc(nova_hip_ctx *ctx, const nova_hip_render_params *params, float *accum_dev, float *final_dev,
                                   float *partial_dev, void *stream);

/* Statistics of the last nova_hip_render_frames call: device time of the whole call and,
 * with NOVA_HIP_RENDER_TIMING, summed extend (traversal) and shade kernel times in
 * milliseconds; path segments traced (rays entering an extend launch) and launches; with
 * NOVA_HIP_RENDER_COUNT_TRAVERSAL, 4-wide node visits and triangle tests summed over rays.
 * The *_first fields and camera_rays cover the camera-ray (bounce 0) launches alone. */
typedef struct {
  float total_ms, extend_ms, shade_ms, other_ms;
  uint64_t segments;
  uint64_t extend_launches;
  uint64_t node_visits;
  uint64_t triangle_tests;
  uint64_t escaped;          /* segments whose ray left the scene (envmap lookups after a miss) */
  float extend_first_ms;     /* with NOVA_HIP_RENDER_TIMING: the camera-ray extend launches */
  uint64_t camera_rays;      /* rays entering the camera-ray launches */
  uint64_t node_visits_first, triangle_tests_first; /* with NOVA_HIP_RENDER_COUNT_TRAVERSAL */
  /* With NOVA_HIP_RENDER_COUNT_TRAVERSAL, the per-ray walk's SIMD use: loop turns of its
   * waves, active lanes summed over those turns (lane_turns / (64 * wave_turns) is the
   * fraction of lanes doing a node step), and wave rounds of the triangle loop (tests /
   * (64 * rounds) is the fraction of lanes testing a triangle in a round). */
  uint64_t walk_wave_turns, walk_lane_turns, walk_tri_rounds;
  /* The last bounces finished one path per lane (see NOVA_HIP_RENDER_NO_TAIL): segments,
   * and with NOVA_HIP_RENDER_COUNT_TRAVERSAL node visits and triangle tests, included in the
   * totals above; with NOVA_HIP_RENDER_TIMING the time of those launches (in neither
   * extend_ms nor shade_ms). */
  uint64_t tail_segments, tail_node_visits, tail_triangle_tests;
  float tail_ms;
} nova_hip_render_stats;
int nova_hip_render_stats_get(nova_hip_ctx *ctx, nova_hip_render_stats *out);

/* --------------------------------------------------------- render buffers */

/* RenderBufferPack + NvRenderBuffer (corelib/nova/api/renderbuffer.cpp:8-230) in device
 * memory: front and back display buffers, partial and accumulator (RGBA32F), depth (one
 * float) and normal (three floats) per pixel. Allocated for a capacity, used at
 * width x height <= capacity; all buffers start zeroed. The depth and normal buffers are
 * allocated and cleared like the reference's, which writes neither on this path. */
typedef struct nova_hip_renderbuffer nova_hip_renderbuffer;
typedef struct {
  float *front, *back, *partial, *accumulator, *depth, *normal;  /* device pointers */
  uint32_t width, height;
  uint64_t capacity;                                              /* pixels */
} nova_hip_renderbuffer_view;
enum { NOVA_HIP_RB_FRONT = 0, NOVA_HIP_RB_BACK = 1, NOVA_HIP_RB_PARTIAL = 2, NOVA_HIP_RB_ACCUMULATOR = 3, NOVA_HIP_RB_DEPTH = 4, NOVA_HIP_RB_NORMAL = 5 };

/* create_renderbuffer (renderbuffer.cpp:232): an empty render buffer on ctx's device;
 * errors of the calls below are reported on ctx (nova_hip_last_error). The render buffer
 * refers to ctx: destroy it before ctx. */
int nova_hip_renderbuffer_create(nova_hip_ctx *ctx, nova_hip_renderbuffer **out);
int nova_hip_renderbuffer_destroy(nova_hip_renderbuffer *rb);
/* RenderBuffer::preallocate (renderbuffer.cpp:126-136, RenderBufferPack::allocate :32-42):
 * capacity max_width * max_height; NOVA_HIP_OUT_OF_MEMORY when the allocation fails. */
int nova_hip_renderbuffer_preallocate(nova_hip_renderbuffer *rb, unsigned max_width, unsigned max_height);
/* RenderBuffer::resize (renderbuffer.cpp:138-146): reallocates (zeroed) only when the
 * capacity is below width * height, then sets the dimensions. */
int nova_hip_renderbuffer_resize(nova_hip_renderbuffer *rb, unsigned width, unsigned height);
/* resetBuffers -> cleanBuffers (renderbuffer.cpp:119-124, 44-57): zeroes the whole capacity;
 * cleanCanvas (:60-73): zeroes width * height pixels. Stream-ordered. */
int nova_hip_renderbuffer_reset(nova_hip_renderbuffer *rb, void *stream);
int nova_hip_renderbuffer_clean_canvas(nova_hip_renderbuffer *rb, void *stream);
/* swapBackBuffer (renderbuffer.cpp:163-166): exchanges the front and back buffers. */
int nova_hip_renderbuffer_swap(nova_hip_renderbuffer *rb);
/* getRenderBuffers / getAccumulator / frontBuffer / backBuffer (renderbuffer.cpp:148-173). */
int nova_hip_renderbuffer_get(const nova_hip_renderbuffer *rb, nova_hip_renderbuffer_view *out);
/* The engine's frame (engine.cpp:498-525): nova_hip_render_frames into the partial buffer,
 * writeFramebuffer (accumulator += partial, back = accumulator / (frame + 1), engine.cpp:351-363)
 * and swapBackBuffer, all on the device. params->width/height must equal the render buffer's
 * dimensions (NOVA_HIP_INVALID_RENDERBUFFER_DIM otherwise). With frame_count 1 a call is one
 * iteration of the engine's loop; with more, the frames are rendered in one batch and the
 * buffers are swapped once, so the front buffer holds the last frame's image and the
 * accumulator all frames' sum. */
int nova_hip_render_to_renderbuffer(nova_hip_ctx *ctx, const nova_hip_render_params *params, nova_hip_renderbuffer *rb, void *stream);
/* One iteration of the engine's loop (params->frame_count 1) as nova_hip_render_to_renderbuffer
 * does it, the frame served by nova_hip_render_frame_lookahead (rendered `lookahead` frames at a
 * time while the parameters stay the same; the same frames bit for bit). No counterpart in the
 * reference. */
int nova_hip_render_to_renderbuffer_lookahead(nova_hip_ctx *ctx, const nova_hip_render_params *params, int32_t lookahead, nova_hip_renderbuffer *rb,
                                              void *stream);
/* Copies one buffer (NOVA_HIP_RB_*; width * height pixels) to host memory, for display:
 * getFramebuffer / sampleColor (renderbuffer.cpp:175-220) without the per-frame copy. */
int nova_hip_renderbuffer_read(const nova_hip_renderbuffer *rb, int which, float *host, void *stream);

/* ---------------------------------------------------------------- batch */

/* One scene of a batched render: a context holding an uploaded scene with its BVH built,
 * the frames to render (as for nova_hip_render_frames; max_depth and rr_start must be the
 * same for every item, sampling_roughness_texture -1; the render flags are not used: a
 * batch always waits and reports nova_hip_batch_stats) and device RGBA32F buffers of
 * width*height*4 floats, 16-byte aligned: accum is zeroed and receives the sum of the
 * frames, final the image accum / (frame_start + frame_count). */
typedef struct {
  nova_hip_ctx *scene;
  nova_hip_render_params params;
  float *accum_dev, *final_dev;
} nova_hip_batch_item;

/* Launch shape of a batched render: paths in flight per wave (a multiple of 256), trace and
 * shade workgroups per CU, and whether the wave sequence is replayed from a HIP graph. */
typedef struct {
  uint32_t queue_capacity;
  uint32_t trace_blocks_per_cu, shade_blocks_per_cu;
  int32_t use_graph;
} nova_hip_batch_config;

typedef struct {
  float total_ms;            /* first wave start to last wave end, device time */
  uint64_t paths, waves, segments;
  uint32_t queue_capacity;
} nova_hip_batch_stats;

/* Renders every item in shared launches: all scenes' paths are numbered back to back and
 * processed in waves of queue_capacity paths, each wave one generate, (extend, shade) per
 * depth level and one accumulate launch over its paths whatever their scene. Each image
 * equals the item's nova_hip_render_frames image bit for bit when both trace its camera rays
 * the same way (see below). completion_ms (optional,
 * count floats) receives the device time at which each item's last wave ended. Items are
 * processed in the order given (sorting them by path count gives shortest-first completion).
 * When every item's path count (width * height * frame_count) is a multiple of 64, camera
 * rays are traced 64 at a time by one walk per wave, which is faster; nova_hip_render_frames
 * does so from 8 frames per batch. The two walks give the same hits except, rarely, on a ray
 * grazing the edge two triangles share (measured: 1 camera ray in 16.6 M), so an item that
 * one of them traces per ray and the other per wave can differ in such a pixel.
 * `executor` is a context that owns the queues and the wave graph (it needs no scene).
 * Replaces a loop of Engine renders, one per scene (engine.cpp:498-551); no counterpart in
 * the reference's API. */
int nova_hip_render_batch(nova_hip_ctx *executor, const nova_hip_batch_item *items, uint32_t count, const nova_hip_batch_config *config,
                          nova_hip_batch_stats *stats, float *completion_ms, void *stream);

/* Autotuner for nova_hip_render_batch: times the captured wave graph of every candidate
 * (queue capacity 2^22 to 2^26 paths x the trace workgroups per CU that are resident at
 * once x one or two rounds of resident shade workgroups per CU) on the same slots, 2^26
 * paths from the middle of the batch (the whole batch when smaller) run as as many waves as
 * the candidate's capacity takes, without accumulating, and returns the fastest per path
 * (best_ns_per_path optional). */
int nova_hip_batch_autotune(nova_hip_ctx *executor, const nova_hip_batch_item *items, uint32_t count, nova_hip_batch_config *best,
                            float *best_ns_per_path, void *stream);

/* ---------------------------------------------------------- differentiation */

/* Element offset of FP32 texture `texture` in the scene's FP32 texel pool and the pool size
 * in floats: a gradient buffer laid out like the pool holds texture t's gradient at
 * grad + offset(t), width*height*channels floats, texel (i,j) channel c at
 * (j*width + i)*channels + c. No counterpart in the reference (it has no gradients). */
int nova_hip_texture_f32_info(nova_hip_ctx *ctx, int32_t texture, uint64_t *offset, uint64_t *pool_floats);

/* Replaces the texels of FP32 texture `texture` with width*height*channels floats from
 * device memory (stream-ordered). Used by the inverse-rendering harness between renders. */
int nova_hip_texture_f32_update(nova_hip_ctx *ctx, int32_t texture, const float *texels_dev, void *stream);

/* Path-replay adjoint (Vicini et al. 2021) of the frames nova_hip_render_frames would render
 * with the same params: for a loss whose gradient with respect to each frame's radiance
 * buffer is scale * d_image (RGBA32F, width*height*4 floats, alpha ignored), adds
 * dLoss/d(texel) to grad_pool_dev (laid out like the FP32 texel pool, see
 * nova_hip_texture_f32_info) for the albedo_f32 (channels 0-2) and roughness_f32
 * (channel 0) textures of every material. Each batch of frames is rendered once to get
 * every path's radiance, then replayed from the same sample indices. The replay takes the
 * primal pass's hits of its leading bounces from a cache of at most 2 GiB (16 B per path
 * and bounce, as many bounces as fit) and traces the deeper bounces again; nothing else is
 * kept per path vertex, so memory stays within the batch's queues plus that cap whatever
 * the depth (NOVA_HIP_ADJOINT_RETRACE: no cache, every bounce traced). Sampling is detached: directions, lobe choices and pdfs are constants, the
 * gradient flows through the BSDF values and Russian-roulette rescaling. With
 * sampling_roughness_texture >= 0 the replay samples with that texture as the renderer
 * does. No counterpart in the reference. */
int nova_hip_render_adjoint(nova_hip_ctx *ctx, const nova_hip_render_params *params, const float *d_image_dev, float scale,
                            float *grad_pool_dev, void *stream);

/* --------------------------------------------------------------- sampler */

/* Owen-scrambled Sobol samples: out[i] = SobolGenerator::generate(idx[i], dim[i]) with the
 * reference's seed 0xDEADBEEF and 256 x 32 direction table
 * (corelib/internal/common/math/math_random.h:29-97). Device pointers; dims must be < 256.
 * out_bits (optional) receives the scrambled 32-bit integers before the float convert. */
int nova_hip_sobol_fill(const uint32_t *idx_dev, const uint32_t *dim_dev, uint64_t n, float *out_dev, uint32_t *out_bits_dev,
                        void *stream);

/* ----------------------------------------------------------------- bakes */

/* Diffuse irradiance of FP32 equirect maps, batched in one launch:
 * TextureOperations<float>::computeDiffuseIrradiance (sources/processing/TextureProcessing.cpp:9-64,
 * TextureProcessing.h:253-301). env_dev[i] is a device pointer to map i (w[i] x h[i] x c[i]
 * floats); out_dev[i] receives out_w[i] x out_h[i] x c[i] floats. The pointer/size arrays
 * are host arrays of length count. */
int nova_hip_bake_irradiance(const float *const *env_dev, const int32_t *w, const int32_t *h, const int32_t *c, uint32_t count,
                             float *const *out_dev, const int32_t *out_w, const int32_t *out_h, int32_t samples, void *stream);

/* GGX-prefiltered specular of FP32 equirect maps at one roughness per output, batched:
 * the GLSL baker envmap_prefilter.frag:1-116 driven by RenderPipeline::preFilterEnvmap
 * (sources/renderer/RenderPipeline.cpp:94-150), on equirect with level-0 lookups. */
int nova_hip_bake_prefilter(const float *const *env_dev, const int32_t *w, const int32_t *h, const int32_t *c, uint32_t count,
                            float *const *out_dev, const int32_t *out_w, const int32_t *out_h, const float *roughness,
                            int32_t samples, void *stream);

/* One job of a batched bake: an FP32 equirect map (device pointer, width x height x channels,
 * channels 3 or 4; a 4-channel map 16-byte aligned) and its output (device pointer,
 * out_width x out_height x channels). */
enum nova_hip_bake_kind {
  NOVA_HIP_BAKE_IRRADIANCE = 0,       /* computeDiffuseIrradiance with its pgc3d hash (the reference's sampler) */
  NOVA_HIP_BAKE_IRRADIANCE_SOBOL = 1, /* the same estimator on Owen-scrambled Sobol points, dims 0-1,
                                         index texel * samples + s (extension) */
  NOVA_HIP_BAKE_PREFILTER = 2,        /* envmap_prefilter.frag at `roughness` */
  NOVA_HIP_BAKE_BRDF_LUT = 3,         /* brdf_lookup_table_baker.frag:74-113: no map (env, width, height
                                         unused), channels 2, out_width x out_height x {A, B} with
                                         NdotV = (i + 0.5) / out_width, roughness = (j + 0.5) / out_height;
                                         1024 samples reproduce the shader */
  NOVA_HIP_BAKE_CUBEMAP = 4,          /* envmap_bake.frag:10-24 under the capture views of
                                         RenderPipeline.cpp:14-26 (90-degree faces): out_width ==
                                         out_height == the face size, out holds 6 faces in capture order
                                         (left, right, down, up, back, front), face-major, rows bottom-up,
                                         `channels` floats per texel; the map is read bilinearly at level 0
                                         with REPEAT wrap (the GL state of the equirect texture) */
  NOVA_HIP_BAKE_IRRADIANCE_RIEMANN = 5 /* irradiance_baker.frag:15-39 (steps of 0.0075 in azimuth and
                                         zenith) with the map as its environment, sampled by the same rule;
                                         out is equirect, out_width x out_height x channels, texel (i, j) at
                                         its centre direction; `samples` is not used */
};
typedef struct {
  const float *env;
  int32_t width, height, channels;
  float *out;
  int32_t out_width, out_height;
  int32_t kind;    /* nova_hip_bake_kind */
  float roughness; /* NOVA_HIP_BAKE_PREFILTER only */
} nova_hip_bake_job;

/* Every texel of every job in one launch (irradiance maps and prefilter mips of many
 * environment maps together), `samples` samples per texel. Replaces the per-map
 * launchAsyncDiffuseIrradianceCompute bands (TextureProcessing.h:155-230) and the per-mip
 * GL passes of RenderPipeline::preFilterEnvmap (RenderPipeline.cpp:94-150). */
int nova_hip_bake_batch(const nova_hip_bake_job *jobs, uint32_t count, int32_t samples, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* NOVA_HIP_H */
