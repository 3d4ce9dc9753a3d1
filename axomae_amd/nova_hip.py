"""ctypes binding of libnova_hip.so (the C ABI declared in include/nova_hip.h).

Host-side mirror of the reference's engine calls for the device path: a
:class:`HipEngine` uploads a scene (Engine::buildScene), builds the acceleration
structure (Engine::buildAcceleration) and renders frames into device RGBA32F
buffers (Engine::renderGpu + writeFramebuffer). torch provides device memory and
streams. The library must be built (`python -m axomae_amd.build`); nothing here
falls back to a CPU path: a missing library or device raises.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libnova_hip.so")

ERRORS = {0: "SUCCESS", 1: "INVALID_ENGINE_STATE", 2: "INVALID_BUFFER_STATE", 4: "INVALID_SCENE_TYPE", 5: "INVALID_TRANSFORM_TYPE",
          9: "SCENE_NOT_PROCESSED", 11: "OUT_OF_MEMORY", 15: "INVALID_ARGUMENT", 16: "RENDERER_EXCEPTION", 100: "DEVICE_ERROR",
          101: "INVALID_RENDERBUFFER_DIM"}

EXPORTS = ["nova_hip_create", "nova_hip_destroy", "nova_hip_last_error", "nova_hip_scene_upload", "nova_hip_bvh_build", "nova_hip_bvh_read",
           "nova_hip_closest_hit", "nova_hip_render_frames", "nova_hip_render_stats_get", "nova_hip_sobol_fill", "nova_hip_bake_irradiance",
           "nova_hip_bake_prefilter", "nova_hip_texture_f32_info", "nova_hip_texture_f32_update", "nova_hip_render_adjoint",
           "nova_hip_bake_batch", "nova_hip_render_batch", "nova_hip_batch_autotune", "nova_hip_renderbuffer_create",
           "nova_hip_renderbuffer_destroy", "nova_hip_renderbuffer_preallocate", "nova_hip_renderbuffer_resize", "nova_hip_renderbuffer_reset",
           "nova_hip_renderbuffer_clean_canvas", "nova_hip_renderbuffer_swap", "nova_hip_renderbuffer_get", "nova_hip_render_to_renderbuffer",
           "nova_hip_renderbuffer_read", "nova_hip_render_frames_noalloc", "nova_hip_render_frame_lookahead",
           "nova_hip_render_to_renderbuffer_lookahead"]

RB_FRONT, RB_BACK, RB_PARTIAL, RB_ACCUMULATOR, RB_DEPTH, RB_NORMAL = range(6)
RB_CHANNELS = {RB_FRONT: 4, RB_BACK: 4, RB_PARTIAL: 4, RB_ACCUMULATOR: 4, RB_DEPTH: 1, RB_NORMAL: 3}


class RenderBufferView(C.Structure):
    _fields_ = [("front", C.c_void_p), ("back", C.c_void_p), ("partial", C.c_void_p), ("accumulator", C.c_void_p), ("depth", C.c_void_p),
                ("normal", C.c_void_p), ("width", C.c_uint32), ("height", C.c_uint32), ("capacity", C.c_uint64)]


class NovaHipError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class MeshDesc(C.Structure):
    _fields_ = [("vertices", C.c_void_p), ("normals", C.c_void_p), ("tangents", C.c_void_p), ("bitangents", C.c_void_p), ("uv", C.c_void_p),
                ("indices", C.c_void_p), ("vertex_count", C.c_uint64), ("index_count", C.c_uint64), ("transform", C.c_float * 16),
                ("material", C.c_int32)]


class MaterialDesc(C.Structure):
    _fields_ = [("albedo", C.c_int32), ("metallic", C.c_int32), ("normal", C.c_int32), ("emissive", C.c_int32), ("albedo_f32", C.c_int32),
                ("roughness_f32", C.c_int32), ("eta", C.c_float * 3), ("k", C.c_float * 3), ("anisotropy", C.c_float),
                ("emission_scale", C.c_float)]


class TextureDesc(C.Structure):
    _fields_ = [("width", C.c_int32), ("height", C.c_int32), ("channels", C.c_int32), ("invert_u", C.c_int32), ("invert_v", C.c_int32),
                ("format", C.c_int32), ("data", C.c_void_p)]


class SceneDesc(C.Structure):
    _fields_ = [("meshes", C.POINTER(MeshDesc)), ("mesh_count", C.c_uint32), ("materials", C.POINTER(MaterialDesc)),
                ("material_count", C.c_uint32), ("textures", C.POINTER(TextureDesc)), ("texture_count", C.c_uint32),
                ("envmap_texture", C.c_int32)]


class BvhInfo(C.Structure):
    _fields_ = [("primitive_count", C.c_uint32), ("binary_nodes", C.c_uint32), ("binary_depth", C.c_uint32), ("wide_nodes", C.c_uint32),
                ("wide_depth", C.c_uint32), ("leaves", C.c_uint32), ("build_ms", C.c_float)]


class RenderParams(C.Structure):
    _fields_ = [("width", C.c_int32), ("height", C.c_int32), ("max_depth", C.c_int32), ("frame_start", C.c_int32), ("frame_count", C.c_int32),
                ("max_samples_offset", C.c_uint32), ("vertical_invert", C.c_int32), ("rr_start", C.c_int32), ("inv_P", C.c_float * 16),
                ("inv_V", C.c_float * 16), ("flags", C.c_int32), ("sampling_roughness_texture", C.c_int32), ("index_policy", C.c_int32),
                ("integrator_flags", C.c_int32), ("z_near", C.c_float), ("z_far", C.c_float)]


INDEX_CPU, INDEX_OPTIX = 0, 1
# integrator::TYPE (api_common.h:29-49)
INTEGRATOR_PATH, INTEGRATOR_COMBINED, INTEGRATOR_NORMAL, INTEGRATOR_DEPTH, INTEGRATOR_EMISSIVE = 1, 1 << 8, 1 << 9, 1 << 10, 1 << 13


class BakeJob(C.Structure):
    _fields_ = [("env", C.c_void_p), ("width", C.c_int32), ("height", C.c_int32), ("channels", C.c_int32), ("out", C.c_void_p),
                ("out_width", C.c_int32), ("out_height", C.c_int32), ("kind", C.c_int32), ("roughness", C.c_float)]


BAKE_IRRADIANCE, BAKE_IRRADIANCE_SOBOL, BAKE_PREFILTER, BAKE_BRDF_LUT, BAKE_CUBEMAP, BAKE_IRRADIANCE_RIEMANN = 0, 1, 2, 3, 4, 5

class BatchConfig(C.Structure):
    _fields_ = [("queue_capacity", C.c_uint32), ("trace_blocks_per_cu", C.c_uint32), ("shade_blocks_per_cu", C.c_uint32), ("use_graph", C.c_int32)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class BatchStats(C.Structure):
    _fields_ = [("total_ms", C.c_float), ("paths", C.c_uint64), ("waves", C.c_uint64), ("segments", C.c_uint64), ("queue_capacity", C.c_uint32)]


RENDER_TIMING = 1
RENDER_COUNT_TRAVERSAL = 2
RENDER_PER_RAY_TRAVERSAL = 4  # camera rays one walk per ray (testing and tuning)
ADJOINT_RETRACE = 8  # the adjoint replay traces again instead of reusing the primal's hits (testing)
RENDER_NO_TAIL = 16  # one extend and one shade launch per bounce to the end (testing and tuning)
RENDER_ASYNC = 32  # return once the work is queued; statistics are not read back


class RenderStats(C.Structure):
    _fields_ = [("total_ms", C.c_float), ("extend_ms", C.c_float), ("shade_ms", C.c_float), ("other_ms", C.c_float), ("segments", C.c_uint64),
                ("extend_launches", C.c_uint64), ("node_visits", C.c_uint64), ("triangle_tests", C.c_uint64), ("escaped", C.c_uint64),
                ("extend_first_ms", C.c_float), ("camera_rays", C.c_uint64), ("node_visits_first", C.c_uint64),
                ("triangle_tests_first", C.c_uint64), ("walk_wave_turns", C.c_uint64), ("walk_lane_turns", C.c_uint64),
                ("walk_tri_rounds", C.c_uint64), ("tail_segments", C.c_uint64), ("tail_node_visits", C.c_uint64),
                ("tail_triangle_tests", C.c_uint64), ("tail_ms", C.c_float)]


BNODE_DTYPE = np.dtype([("min", np.float32, 3), ("max", np.float32, 3), ("left", np.int32), ("count", np.int32)])

_LIB = None


def load():
    """Load libnova_hip.so; raises if it has not been built."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with `python -m axomae_amd.build`")
    # torch's HIP runtime and the one this library links share the soname libamdhip64.so.7:
    # the first loaded serves the process, and torch must have its own, so it loads first
    import torch  # noqa: F401
    L = C.CDLL(LIB_PATH)
    P = C.c_void_p
    L.nova_hip_create.argtypes = [C.c_int, C.POINTER(P)]
    L.nova_hip_destroy.argtypes = [P]
    L.nova_hip_last_error.argtypes = [P]
    L.nova_hip_last_error.restype = C.c_char_p
    L.nova_hip_scene_upload.argtypes = [P, C.POINTER(SceneDesc), P]
    L.nova_hip_bvh_build.argtypes = [P, C.c_int, C.POINTER(BvhInfo), P]
    L.nova_hip_bvh_read.argtypes = [P, P, P]
    L.nova_hip_closest_hit.argtypes = [P, P, C.c_uint64, P, P, P]
    L.nova_hip_render_frames.argtypes = [P, C.POINTER(RenderParams), P, P, P, P]
    L.nova_hip_render_frames_noalloc.argtypes = [P, C.POINTER(RenderParams), P, P, P, P]
    L.nova_hip_render_frame_lookahead.argtypes = [P, C.POINTER(RenderParams), C.c_int32, P, P, P, P]
    L.nova_hip_render_stats_get.argtypes = [P, C.POINTER(RenderStats)]
    L.nova_hip_sobol_fill.argtypes = [P, P, C.c_uint64, P, P, P]
    L.nova_hip_render_batch.argtypes = [P, P, C.c_uint32, C.POINTER(BatchConfig), C.POINTER(BatchStats), P, P]
    L.nova_hip_batch_autotune.argtypes = [P, P, C.c_uint32, C.POINTER(BatchConfig), C.POINTER(C.c_float), P]
    L.nova_hip_bake_batch.argtypes = [C.POINTER(BakeJob), C.c_uint32, C.c_int32, P]
    L.nova_hip_texture_f32_info.argtypes = [P, C.c_int32, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]
    L.nova_hip_texture_f32_update.argtypes = [P, C.c_int32, P, P]
    L.nova_hip_render_adjoint.argtypes = [P, C.POINTER(RenderParams), P, C.c_float, P, P]
    L.nova_hip_bake_irradiance.argtypes = [C.POINTER(P), P, P, P, C.c_uint32, C.POINTER(P), P, P, C.c_int32, P]
    L.nova_hip_bake_prefilter.argtypes = [C.POINTER(P), P, P, P, C.c_uint32, C.POINTER(P), P, P, P, C.c_int32, P]
    L.nova_hip_renderbuffer_create.argtypes = [P, C.POINTER(P)]
    L.nova_hip_renderbuffer_destroy.argtypes = [P]
    L.nova_hip_renderbuffer_preallocate.argtypes = [P, C.c_uint, C.c_uint]
    L.nova_hip_renderbuffer_resize.argtypes = [P, C.c_uint, C.c_uint]
    L.nova_hip_renderbuffer_reset.argtypes = [P, P]
    L.nova_hip_renderbuffer_clean_canvas.argtypes = [P, P]
    L.nova_hip_renderbuffer_swap.argtypes = [P]
    L.nova_hip_renderbuffer_get.argtypes = [P, C.POINTER(RenderBufferView)]
    L.nova_hip_render_to_renderbuffer.argtypes = [P, C.POINTER(RenderParams), P, P]
    L.nova_hip_render_to_renderbuffer_lookahead.argtypes = [P, C.POINTER(RenderParams), C.c_int32, P, P]
    L.nova_hip_renderbuffer_read.argtypes = [P, C.c_int, P, P]
    _LIB = L
    return L


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("no HIP device visible: the Nova HIP path needs an MI355X")
    return torch


def _stream(torch):
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def sobol_fill(idx, dims, return_bits=False):
    """Device Owen-scrambled Sobol samples for (index, dim) pairs (torch uint32/int32 tensors on cuda)."""
    torch = _torch()
    L = load()
    idx = idx.to(device="cuda", dtype=torch.int32).contiguous()
    dims = dims.to(device="cuda", dtype=torch.int32).contiguous()
    out = torch.empty(idx.shape, dtype=torch.float32, device="cuda")
    bits = torch.empty(idx.shape, dtype=torch.int32, device="cuda") if return_bits else None
    rc = L.nova_hip_sobol_fill(idx.data_ptr(), dims.data_ptr(), idx.numel(), out.data_ptr(), bits.data_ptr() if bits is not None else None,
                               _stream(torch))
    if rc:
        raise NovaHipError(rc, "sobol_fill")
    return (out, bits) if return_bits else out


def scene_desc(scene):
    """nova_hip_scene_desc for a scenes.Scene; returns (descriptor, arrays that must outlive it)."""
    keep = []

    def ptr(a, dtype):
        a = np.ascontiguousarray(a, dtype=dtype)
        keep.append(a)
        return a.ctypes.data

    meshes = (MeshDesc * max(1, len(scene.meshes)))()
    for i, m in enumerate(scene.meshes):
        d = meshes[i]
        d.vertices = ptr(m.vertices, np.float32)
        d.normals = ptr(m.normals, np.float32)
        d.tangents = ptr(m.tangents, np.float32)
        d.bitangents = ptr(m.bitangents, np.float32)
        d.uv = ptr(m.uv, np.float32)
        d.indices = ptr(m.indices, np.uint32)
        d.vertex_count = len(m.vertices) // 3
        d.index_count = len(m.indices)
        d.transform[:] = [float(x) for x in np.asarray(m.transform, dtype=np.float32)]
        d.material = m.material
    mats = (MaterialDesc * max(1, len(scene.materials)))()
    for i, m in enumerate(scene.materials):
        d = mats[i]
        d.albedo, d.metallic, d.normal, d.emissive = m.albedo, m.metallic, m.normal, m.emissive
        d.albedo_f32, d.roughness_f32 = m.albedo_f32, m.roughness_f32
        d.eta[:] = [float(x) for x in m.eta]
        d.k[:] = [float(x) for x in m.k]
        d.anisotropy = m.anisotropy
        d.emission_scale = m.emission_scale
    texs = (TextureDesc * max(1, len(scene.textures)))()
    for i, t in enumerate(scene.textures):
        d = texs[i]
        d.width, d.height, d.channels = t.width, t.height, t.channels
        d.invert_u, d.invert_v = int(t.invert_u), int(t.invert_v)
        d.format = 1 if t.is_f32 else 0
        d.data = ptr(t.data, np.float32 if t.is_f32 else np.uint32)
    keep += [meshes, mats, texs]
    sd = SceneDesc(meshes, len(scene.meshes), mats, len(scene.materials), texs, len(scene.textures), scene.envmap)
    return sd, keep


def render_params(scene, width, height, frames, max_depth, frame_start=0, rr_start=-1, max_samples_offset=0, vertical_invert=False,
                  inv_P=None, inv_V=None, flags=0, sampling_roughness_texture=-1, index_policy=INDEX_CPU, integrator_flags=0, z_near=0.1,
                  z_far=1000.0):
    p = RenderParams()
    p.sampling_roughness_texture = sampling_roughness_texture
    p.index_policy = index_policy
    p.integrator_flags = integrator_flags
    p.z_near, p.z_far = z_near, z_far
    p.width, p.height, p.max_depth = width, height, max_depth
    p.frame_start, p.frame_count = frame_start, frames
    p.max_samples_offset = max_samples_offset
    p.vertical_invert = int(vertical_invert)
    p.rr_start = rr_start
    p.flags = flags
    p.inv_P[:] = [float(x) for x in (scene.inv_P if inv_P is None else inv_P)]
    p.inv_V[:] = [float(x) for x in (scene.inv_V if inv_V is None else inv_V)]
    return p


class HipEngine:
    """One device context holding a scene, its BVH and render scratch."""

    def __init__(self, device=0):
        self.torch = _torch()
        self.L = load()
        self.ctx = C.c_void_p()
        rc = self.L.nova_hip_create(device, C.byref(self.ctx))
        if rc:
            raise NovaHipError(rc, "nova_hip_create")
        self.info = None
        self.scene = None

    def close(self):
        if self.ctx:
            self.L.nova_hip_destroy(self.ctx)
            self.ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc, what):
        if rc:
            msg = self.L.nova_hip_last_error(self.ctx)
            raise NovaHipError(rc, f"{what}: {msg.decode() if msg else ''}")

    # ---------------------------------------------------------------- scene
    def upload(self, scene):
        sd, keep = scene_desc(scene)
        self._check(self.L.nova_hip_scene_upload(self.ctx, C.byref(sd), _stream(self.torch)), "scene_upload")
        self.scene = scene
        self.prim_count = scene.triangle_count

    def build_bvh(self):
        info = BvhInfo()
        self._check(self.L.nova_hip_bvh_build(self.ctx, 4, C.byref(info), _stream(self.torch)), "bvh_build")
        self.info = info
        return info

    def bvh_arrays(self):
        prim = np.empty(self.info.primitive_count, dtype=np.int32)
        nodes = np.empty(self.info.binary_nodes, dtype=BNODE_DTYPE)
        self._check(self.L.nova_hip_bvh_read(self.ctx, prim.ctypes.data, nodes.ctypes.data), "bvh_read")
        return prim, nodes

    def closest_hit(self, rays):
        """rays: (n, 6) float32 array/tensor of origin+direction; returns (prim int32, tuv float32) tensors."""
        torch = self.torch
        r = torch.as_tensor(np.asarray(rays, dtype=np.float32)).reshape(-1, 6).to("cuda") if not torch.is_tensor(rays) else rays.reshape(-1, 6).to(
            device="cuda", dtype=torch.float32).contiguous()
        n = r.shape[0]
        prim = torch.empty(n, dtype=torch.int32, device="cuda")
        tuv = torch.empty((n, 3), dtype=torch.float32, device="cuda")
        self._check(self.L.nova_hip_closest_hit(self.ctx, r.data_ptr(), n, prim.data_ptr(), tuv.data_ptr(), _stream(torch)), "closest_hit")
        return prim, tuv

    # --------------------------------------------------------------- render
    def render(self, width, height, frames, max_depth, frame_start=0, rr_start=-1, max_samples_offset=0, vertical_invert=False, accum=None,
               inv_P=None, inv_V=None, flags=0, sampling_roughness_texture=-1, index_policy=INDEX_CPU, integrator_flags=0, z_near=0.1,
               z_far=1000.0):
        """Renders frames [frame_start, frame_start+frames); returns (final, accum) device tensors (H, W, 4).
        integrator_flags selects a render pass (INTEGRATOR_PATH | INTEGRATOR_NORMAL, ...); 0 is the path integrator."""
        torch = self.torch
        p = render_params(self.scene, width, height, frames, max_depth, frame_start, rr_start, max_samples_offset, vertical_invert, inv_P, inv_V,
                          flags, sampling_roughness_texture, index_policy, integrator_flags, z_near, z_far)
        if accum is None:
            accum = torch.zeros((height, width, 4), dtype=torch.float32, device="cuda")
        final = torch.empty((height, width, 4), dtype=torch.float32, device="cuda")
        self._check(self.L.nova_hip_render_frames(self.ctx, C.byref(p), accum.data_ptr(), final.data_ptr(), None, _stream(torch)), "render_frames")
        return final, accum

    def render_frame_lookahead(self, width, height, frame, max_depth, lookahead, accum, final=None, partial=None, rr_start=-1, inv_P=None,
                               inv_V=None, vertical_invert=False, flags=0, index_policy=INDEX_CPU, max_samples_offset=0):
        """One frame as :meth:`render` renders it (accum += frame, final = accum / (frame + 1), partial = the
        frame), rendering `lookahead` frames ahead on a miss and serving later frames of the same
        parameters from them (nova_hip_render_frame_lookahead). accum, final and partial are device
        tensors (H, W, 4); final and partial may be None."""
        torch = self.torch
        p = render_params(self.scene, width, height, 1, max_depth, frame, rr_start, max_samples_offset, vertical_invert, inv_P, inv_V, flags,
                          index_policy=index_policy)
        ptr = lambda t: None if t is None else t.data_ptr()
        if final is None and accum is not None:
            final = torch.empty_like(accum)
        self._check(self.L.nova_hip_render_frame_lookahead(self.ctx, C.byref(p), lookahead, ptr(accum), ptr(final), ptr(partial), _stream(torch)),
                    "render_frame_lookahead")
        return final

    def render_to(self, rb, frames, max_depth, frame_start=0, rr_start=-1, max_samples_offset=0, vertical_invert=False, inv_P=None,
                  inv_V=None, integrator_flags=0, index_policy=INDEX_CPU, flags=0):
        """The engine's frame loop into a :class:`RenderBuffer` (render, accumulate, swap) at its dimensions."""
        p = render_params(self.scene, rb.width, rb.height, frames, max_depth, frame_start, rr_start, max_samples_offset, vertical_invert, inv_P,
                          inv_V, flags, -1, index_policy, integrator_flags, 0.1, 1000.0)
        self._check(self.L.nova_hip_render_to_renderbuffer(self.ctx, C.byref(p), rb.rb, _stream(self.torch)), "render_to_renderbuffer")

    def render_to_lookahead(self, rb, frame, max_depth, lookahead, rr_start=-1):
        """One engine-loop frame into a :class:`RenderBuffer`, served from frames rendered `lookahead` at a time."""
        p = render_params(self.scene, rb.width, rb.height, 1, max_depth, frame, rr_start)
        self._check(self.L.nova_hip_render_to_renderbuffer_lookahead(self.ctx, C.byref(p), lookahead, rb.rb, _stream(self.torch)),
                    "render_to_renderbuffer_lookahead")

    # ------------------------------------------------------ differentiation
    def texture_info(self, tex):
        """(offset, pool_floats) of FP32 texture `tex` in the device texel pool."""
        off, pool = C.c_uint64(), C.c_uint64()
        self._check(self.L.nova_hip_texture_f32_info(self.ctx, tex, C.byref(off), C.byref(pool)), "texture_f32_info")
        return off.value, pool.value

    def update_texture(self, tex, texels):
        """Replaces FP32 texture `tex` with a device tensor of its width*height*channels floats."""
        torch = self.torch
        t = self.scene.textures[tex]
        x = texels.to(device="cuda", dtype=torch.float32).contiguous()
        if x.numel() != t.width * t.height * t.channels:
            raise ValueError("texel count does not match the texture")
        self._check(self.L.nova_hip_texture_f32_update(self.ctx, tex, x.data_ptr(), _stream(torch)), "texture_f32_update")
        self._keep_texels = x  # the copy is stream-ordered: keep the source alive until the next call

    def adjoint(self, width, height, frames, max_depth, d_image, scale, frame_start=0, rr_start=-1, max_samples_offset=0, vertical_invert=False,
                sampling_roughness_texture=-1, grad=None, flags=0):
        """Path-replay adjoint: adds dLoss/d(texel) for every FP32 albedo/roughness texture to `grad` (FP32 pool layout,
        allocated zeroed if None), where scale * d_image (H, W, 4) is the loss gradient of each frame's radiance.
        flags=ADJOINT_RETRACE makes the replay trace its rays again instead of reusing the primal pass's hits."""
        torch = self.torch
        _, pool = self.texture_info(next(i for i, t in enumerate(self.scene.textures) if t.is_f32))
        if grad is None:
            grad = torch.zeros(pool, dtype=torch.float32, device="cuda")
        d = d_image.to(device="cuda", dtype=torch.float32).contiguous()
        if d.numel() != width * height * 4:
            raise ValueError("d_image must hold width*height*4 floats")
        p = render_params(self.scene, width, height, frames, max_depth, frame_start, rr_start, max_samples_offset, vertical_invert,
                          flags=flags, sampling_roughness_texture=sampling_roughness_texture)
        self._check(self.L.nova_hip_render_adjoint(self.ctx, C.byref(p), d.data_ptr(), float(scale), grad.data_ptr(), _stream(torch)),
                    "render_adjoint")
        return grad

    def texture_grad(self, grad, tex):
        """View of texture `tex`'s gradient inside a pool-layout gradient buffer, shaped (H, W, C)."""
        off, _ = self.texture_info(tex)
        t = self.scene.textures[tex]
        return grad[off:off + t.width * t.height * t.channels].view(t.height, t.width, t.channels)

    def render_stats(self):
        s = RenderStats()
        self._check(self.L.nova_hip_render_stats_get(self.ctx, C.byref(s)), "render_stats")
        return s


def bake(kind, envmaps, out_sizes, samples, roughness=None):
    """Batched irradiance ('irradiance') or prefilter ('prefilter') bake of FP32 envmaps; returns numpy arrays."""
    torch = _torch()
    L = load()
    n = len(envmaps)
    ins = [torch.as_tensor(np.ascontiguousarray(e.data, dtype=np.float32)).to("cuda") for e in envmaps]
    outs = [torch.zeros(oh * ow * e.channels, dtype=torch.float32, device="cuda") for e, (ow, oh) in zip(envmaps, out_sizes)]
    P = C.c_void_p
    env_ptrs = (P * n)(*[t.data_ptr() for t in ins])
    out_ptrs = (P * n)(*[t.data_ptr() for t in outs])
    w = np.array([e.width for e in envmaps], dtype=np.int32)
    h = np.array([e.height for e in envmaps], dtype=np.int32)
    c = np.array([e.channels for e in envmaps], dtype=np.int32)
    ow = np.array([s[0] for s in out_sizes], dtype=np.int32)
    oh = np.array([s[1] for s in out_sizes], dtype=np.int32)
    st = _stream(torch)
    if kind == "irradiance":
        rc = L.nova_hip_bake_irradiance(env_ptrs, w.ctypes.data, h.ctypes.data, c.ctypes.data, n, out_ptrs, ow.ctypes.data, oh.ctypes.data, samples, st)
    else:
        r = np.asarray(roughness, dtype=np.float32)
        rc = L.nova_hip_bake_prefilter(env_ptrs, w.ctypes.data, h.ctypes.data, c.ctypes.data, n, out_ptrs, ow.ctypes.data, oh.ctypes.data,
                                       r.ctypes.data, samples, st)
    if rc:
        raise NovaHipError(rc, f"bake {kind}")
    return [o.cpu().numpy().reshape(s[1], s[0], e.channels) for o, s, e in zip(outs, out_sizes, envmaps)]


def bake_batch(jobs, samples, keep_on_device=False):
    """One launch for every job: jobs are (envmap Texture or device tensor with (w, h, c), out_w, out_h, kind, roughness).
    Returns the outputs (numpy, or device tensors with keep_on_device) in job order. The outputs
    are views of one device allocation (each job's output starts on a 256-byte boundary); the
    kernels write every element, so it is not cleared."""
    torch = _torch()
    L = load()
    arr = (BakeJob * len(jobs))()
    keep, layout, offset = [], [], 0
    for env, ow, oh, kind, rough in jobs:
        c = 2 if kind == BAKE_BRDF_LUT else (env[1][2] if isinstance(env, tuple) else env.channels)
        faces = 6 if kind == BAKE_CUBEMAP else 1  # a cube map: six ow x ow faces
        layout.append((offset, ow, oh, c, faces))
        offset += -(-faces * oh * ow * c // 64) * 64
    pool = torch.empty(max(offset, 1), dtype=torch.float32, device="cuda")
    base = pool.data_ptr()
    for k, (env, ow, oh, kind, rough) in enumerate(jobs):
        off = layout[k][0]
        if kind == BAKE_BRDF_LUT:  # no map: two output channels (A, B)
            arr[k] = BakeJob(None, 0, 0, 2, base + 4 * off, ow, oh, kind, 0.0)
            continue
        if isinstance(env, tuple):
            src, (w, h, c) = env
        else:
            src = torch.as_tensor(np.ascontiguousarray(env.data, dtype=np.float32)).to("cuda")
            w, h, c = env.width, env.height, env.channels
        keep.append(src)
        arr[k] = BakeJob(src.data_ptr(), w, h, c, base + 4 * off, ow, oh, kind, float(rough))
    rc = L.nova_hip_bake_batch(arr, len(jobs), samples, _stream(torch))
    if rc:
        raise NovaHipError(rc, "bake_batch")
    shape = lambda ow, oh, c, faces: (faces, oh, ow, c) if faces > 1 else (oh, ow, c)
    outs = [pool[off:off + f * oh * ow * c].view(*shape(ow, oh, c, f)) for off, ow, oh, c, f in layout]
    if keep_on_device:
        return outs
    return [o.cpu().numpy() for o in outs]


class BatchItem(C.Structure):
    _fields_ = [("scene", C.c_void_p), ("params", RenderParams), ("accum_dev", C.c_void_p), ("final_dev", C.c_void_p)]


class BatchRenderer:
    """Renders many scenes (one HipEngine each, scene uploaded and BVH built) in shared
    launches through nova_hip_render_batch; owns the executor context."""

    def __init__(self, device=0):
        self.exec = HipEngine(device)

    def close(self):
        self.exec.close()

    def _items(self, jobs):
        torch = self.exec.torch
        arr = (BatchItem * len(jobs))()
        bufs = []
        for k, (eng, width, height, frames, max_depth, kw) in enumerate(jobs):
            p = render_params(eng.scene, width, height, frames, max_depth, **kw)
            acc = torch.empty((height, width, 4), dtype=torch.float32, device="cuda")
            fin = torch.empty((height, width, 4), dtype=torch.float32, device="cuda")
            bufs.append((acc, fin))
            arr[k] = BatchItem(eng.ctx.value, p, acc.data_ptr(), fin.data_ptr())
        return arr, bufs

    def autotune(self, jobs):
        """jobs: (engine, width, height, frames, max_depth, extra render_params kwargs). Returns (BatchConfig, ns per path)."""
        arr, _ = self._items(jobs)
        cfg, ns = BatchConfig(), C.c_float()
        self.exec._check(self.exec.L.nova_hip_batch_autotune(self.exec.ctx, arr, len(jobs), C.byref(cfg), C.byref(ns), _stream(self.exec.torch)),
                         "batch_autotune")
        return cfg, ns.value

    def render(self, jobs, config=None):
        """Returns (final images, BatchStats, per-job completion ms)."""
        if config is None:
            config, _ = self.autotune(jobs)
        arr, bufs = self._items(jobs)
        st = BatchStats()
        comp = np.zeros(len(jobs), dtype=np.float32)
        self.exec._check(self.exec.L.nova_hip_render_batch(self.exec.ctx, arr, len(jobs), C.byref(config), C.byref(st), comp.ctypes.data,
                                                           _stream(self.exec.torch)), "render_batch")
        return [f for _, f in bufs], st, comp


class RenderBuffer:
    """Device render buffers of an engine's context (nova_hip_renderbuffer_*): front/back
    display buffers, partial, accumulator, depth and normal."""

    def __init__(self, engine, width, height, capacity=None):
        self.eng = engine
        self.rb = C.c_void_p()
        engine._check(engine.L.nova_hip_renderbuffer_create(engine.ctx, C.byref(self.rb)), "renderbuffer_create")
        if capacity is not None:
            engine._check(engine.L.nova_hip_renderbuffer_preallocate(self.rb, capacity[0], capacity[1]), "renderbuffer_preallocate")
        self.resize(width, height)

    def resize(self, width, height):
        self.eng._check(self.eng.L.nova_hip_renderbuffer_resize(self.rb, width, height), "renderbuffer_resize")

    def view(self):
        v = RenderBufferView()
        self.eng._check(self.eng.L.nova_hip_renderbuffer_get(self.rb, C.byref(v)), "renderbuffer_get")
        return v

    @property
    def width(self):
        return self.view().width

    @property
    def height(self):
        return self.view().height

    def reset(self):
        self.eng._check(self.eng.L.nova_hip_renderbuffer_reset(self.rb, _stream(self.eng.torch)), "renderbuffer_reset")

    def clean_canvas(self):
        self.eng._check(self.eng.L.nova_hip_renderbuffer_clean_canvas(self.rb, _stream(self.eng.torch)), "renderbuffer_clean_canvas")

    def swap(self):
        self.eng._check(self.eng.L.nova_hip_renderbuffer_swap(self.rb), "renderbuffer_swap")

    def read(self, which=RB_FRONT):
        """One buffer copied to the host as (H, W, channels) float32."""
        v = self.view()
        out = np.empty((v.height, v.width, RB_CHANNELS[which]), dtype=np.float32)
        self.eng._check(self.eng.L.nova_hip_renderbuffer_read(self.rb, which, out.ctypes.data, _stream(self.eng.torch)), "renderbuffer_read")
        return out

    def close(self):
        if self.rb:
            self.eng.L.nova_hip_renderbuffer_destroy(self.rb)
            self.rb = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
