"""The device path's own stages, compiled for the host CPU (tests/emu/host_emu.hip), against
the oracle: camera ray, Moller-Trumbore, hit record, Principled scatter, Russian roulette and
the front-to-back estimator. Runs without a GPU; the GPU tests repeat the comparison with
the real kernels."""
import ctypes as C
import os

import numpy as np
import pytest

from axomae_amd import nova_hip, scenes
from conftest import rel_rmse
from scene_cases import BVH_SCENES, hit_test_rays, near_axis_test_rays

HERE = os.path.dirname(os.path.abspath(__file__))
EMU_DIR = os.path.join(HERE, "emu")
CSRC = os.path.join(HERE, "..", "axomae_amd", "csrc")


@pytest.fixture(scope="module")
def emu():
    import sys
    sys.path.insert(0, EMU_DIR)
    import build_emu
    srcs = [os.path.join(EMU_DIR, "host_emu.hip")] + [os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".h")]
    if not os.path.exists(build_emu.LIB) or os.path.getmtime(build_emu.LIB) < max(os.path.getmtime(f) for f in srcs):
        build_emu.build()
    L = C.CDLL(build_emu.LIB)
    L.emu_render.argtypes = [C.POINTER(nova_hip.SceneDesc), C.POINTER(nova_hip.RenderParams), C.c_void_p, C.c_void_p]
    L.emu_closest_hit.argtypes = [C.POINTER(nova_hip.SceneDesc), C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p,
                                  C.c_void_p, C.c_void_p]
    L.emu_adjoint.argtypes = [C.POINTER(nova_hip.SceneDesc), C.POINTER(nova_hip.RenderParams), C.c_void_p, C.c_float, C.c_void_p]
    return L


def emu_closest_hit(L, sc, oscene, rays):
    """Closest hit through the device traversal code (host build) on the oracle's binary BVH."""
    sd, keep = nova_hip.scene_desc(sc)
    prim_idx, nodes = oscene.bvh_arrays()
    rays = np.ascontiguousarray(rays, dtype=np.float32)
    prim = np.empty(len(rays), dtype=np.int32)
    tuv = np.empty((len(rays), 3), dtype=np.float32)
    counts = np.zeros(2, dtype=np.uint64)
    assert L.emu_closest_hit(C.byref(sd), nodes.ctypes.data, len(nodes), prim_idx.ctypes.data, rays.ctypes.data, len(rays), prim.ctypes.data,
                             tuv.ctypes.data, counts.ctypes.data) == 0
    return prim, tuv, counts


def emu_render(L, sc, w, h, frames, depth, **kw):
    sd, keep = nova_hip.scene_desc(sc)
    p = nova_hip.render_params(sc, w, h, frames, depth, **kw)
    acc = np.zeros(w * h * 4, dtype=np.float32)
    out = np.zeros(w * h * 4, dtype=np.float32)
    assert L.emu_render(C.byref(sd), C.byref(p), acc.ctypes.data, out.ctypes.data) == 0
    return out.reshape(h, w, 4)


def test_box_room_device_code_matches_oracle(oracle, emu):
    sc = scenes.box_room()
    img = emu_render(emu, sc, 48, 48, 8, 5)
    ref, _ = oracle.OracleScene(sc).render(48, 48, 8, 5)
    err = rel_rmse(img, ref)
    exact = np.mean(np.all(img[..., :3] == ref[..., :3], axis=-1))
    print(f"box room: rel_rmse {err:.3e}, bit-identical pixels {exact:.4f}")
    assert err <= 1e-3


def test_terrain_spheres_rr_device_code_matches_oracle(oracle, emu):
    sc = scenes.scaled_million(scale_cells=16, spheres=4, sphere_level=2, width=40, height=24)
    img = emu_render(emu, sc, 40, 24, 3, 8, rr_start=3)
    ref, _ = oracle.OracleScene(sc).render(40, 24, 3, 8, rr_start=3)
    err = rel_rmse(img, ref)
    exact = np.mean(np.all(img[..., :3] == ref[..., :3], axis=-1))
    print(f"terrain+spheres: rel_rmse {err:.3e}, bit-identical pixels {exact:.4f}")
    assert err <= 1e-3


def test_offsets_and_vertical_invert_device_code_matches_oracle(oracle, emu):
    sc = scenes.box_room()
    img = emu_render(emu, sc, 24, 20, 3, 3, frame_start=5, max_samples_offset=17, vertical_invert=True)
    ref, _ = oracle.OracleScene(sc).render(24, 20, 3, 3, frame_start=5, max_samples_offset=17, vertical_invert=True)
    assert rel_rmse(img, ref) <= 1e-3


def test_gpu_raygen_index_policy_device_code_matches_oracle(oracle, emu):
    from axomae_amd.nova_hip import INDEX_OPTIX
    sc = scenes.box_room()
    img = emu_render(emu, sc, 24, 20, 3, 3, frame_start=2, max_samples_offset=16, index_policy=INDEX_OPTIX)
    oracle.set_index_policy(1)
    try:
        ref, _ = oracle.OracleScene(sc).render(24, 20, 3, 3, frame_start=2, max_samples_offset=16)
    finally:
        oracle.set_index_policy(0)
    assert rel_rmse(img, ref) <= 1e-3
    cpu = emu_render(emu, sc, 24, 20, 3, 3, frame_start=2, max_samples_offset=16)
    assert rel_rmse(img, cpu) > 1e-2  # the policy changes the samples


def fp32_textured_scene(seed=0xA70AE004, tex=64):
    """Terrain + spheres whose materials read FP32 albedo (RGB) and roughness textures."""
    rng = np.random.default_rng(seed)
    sc = scenes.scaled_million(scale_cells=12, spheres=3, sphere_level=2, width=32, height=20)
    alb = sc.add_texture(scenes.Texture(tex, tex, 3, rng.uniform(0.1, 0.9, tex * tex * 3).astype(np.float32)))
    rough = sc.add_texture(scenes.Texture(tex, tex, 1, rng.uniform(0.05, 0.9, tex * tex).astype(np.float32)))
    for m in sc.materials:
        m.albedo_f32 = alb
        m.roughness_f32 = rough
    return sc


def test_fp32_material_textures_device_code_matches_oracle(oracle, emu):
    sc = fp32_textured_scene()
    img = emu_render(emu, sc, 32, 20, 2, 4)
    ref, _ = oracle.OracleScene(sc).render(32, 20, 2, 4)
    assert rel_rmse(img, ref) <= 1e-3


@pytest.mark.parametrize("name", ["box_room", "terrain_spheres", "soup_5000", "soup_duplicates", "soup_int_max_fallback", "single_triangle",
                                  "transformed"])
def test_device_traversal_matches_reference_traversal(oracle, emu, name):
    """The 4-wide traversal (widened slabs, parallel-axis containment, near-to-far order,
    exact-tie rule) returns the reference walk's primitive, t, u and v on every ray,
    including exact equal-t ties at shared edges hit by pixel-centre rays and rays running
    inside box face planes."""
    sc = BVH_SCENES[name]()
    rays = hit_test_rays(sc, name, 20000)
    os_ = oracle.OracleScene(sc)
    oprim, otuv = os_.closest_hit(rays)
    prim, tuv, counts = emu_closest_hit(emu, sc, os_, rays)
    bad = np.where(prim != oprim)[0]
    print(f"[emu-hit] {name}: {len(rays)} rays, {len(bad)} mismatches, nodes/ray {counts[0] / len(rays):.1f}", flush=True)
    assert len(bad) == 0, [(int(k), int(prim[k]), int(oprim[k]), tuv[k].tolist(), otuv[k].tolist()) for k in bad[:8]]
    np.testing.assert_array_equal(tuv, otuv)


def check_never_misses(name, prim, tuv, oprim, otuv, max_rate):
    """The walk's widened boxes are supersets of the reference's: it never misses a hit the
    reference walk finds and never returns a farther one. Where it differs, it found a closer
    hit in a box the reference's division-form slab test rounds away (or a hit where the
    reference found none); the rate of those is bounded."""
    missed = np.where((oprim >= 0) & ((prim < 0) | (tuv[:, 0] > otuv[:, 0])))[0]
    assert len(missed) == 0, [(int(k), int(prim[k]), int(oprim[k]), tuv[k].tolist(), otuv[k].tolist()) for k in missed[:8]]
    same = prim == oprim
    np.testing.assert_array_equal(tuv[same], otuv[same])
    rate = 1.0 - same.mean()
    print(f"[near-axis] {name}: {len(prim)} rays, {int((~same).sum())} closer hits than the reference walk ({rate:.2e})", flush=True)
    assert rate <= max_rate


@pytest.mark.parametrize("name", ["box_room", "terrain_spheres", "soup_5000", "soup_duplicates", "soup_int_max_fallback", "transformed"])
def test_near_axis_rays_never_miss_a_reference_hit(oracle, emu, name):
    """Rays whose direction has components of 1e-3 .. 1e-37 (finite reciprocals up to 1e37),
    half of them from origins on box face planes."""
    sc = BVH_SCENES[name]()
    rays = near_axis_test_rays(name, 8000)
    os_ = oracle.OracleScene(sc)
    oprim, otuv = os_.closest_hit(rays)
    prim, tuv, _ = emu_closest_hit(emu, sc, os_, rays)
    check_never_misses(name, prim, tuv, oprim, otuv, 1e-3)


def test_transformed_meshes_render_matches_oracle(oracle, emu):
    """Rotated, non-uniformly scaled and translated meshes: rays go to object space before
    the triangle test, as in Triangle::hit (Triangle.h:150-153)."""
    sc = scenes.transformed_scene()
    img = emu_render(emu, sc, 48, 48, 8, 5)
    ref, _ = oracle.OracleScene(sc).render(48, 48, 8, 5)
    err = rel_rmse(img, ref)
    print(f"[emu-render] transformed 48x48x8 d5 rel_rmse={err:.3e}", flush=True)
    assert err <= 1e-3


def test_sobol_nibble_tables_match_direction_table(emu, oracle):
    """The shade kernels read Sobol samples from per-nibble tables of the launch's five
    dimensions; they give the direction-table (and oracle) samples bit for bit."""
    rng = np.random.default_rng(9)
    idx = np.concatenate([[0, 1, 2, 3, 0xFFFFFFFF, 0x80000000], rng.integers(0, 2 ** 32, 4000, dtype=np.uint64)]).astype(np.uint32)
    emu.emu_sobol_nibble.argtypes = [C.c_uint32, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]
    for dim0 in (2, 6, 33, 120, 251, 253):
        a = np.full((len(idx), 5), np.nan, dtype=np.float32)
        b = np.full((len(idx), 5), np.nan, dtype=np.float32)
        assert emu.emu_sobol_nibble(dim0, idx.ctypes.data, len(idx), a.ctypes.data, b.ctypes.data) == 0
        np.testing.assert_array_equal(a, b)
        dims = [d for d in range(dim0, dim0 + 5) if d < 256]
        I, Dm = np.meshgrid(idx, np.array(dims, dtype=np.uint32), indexing="ij")
        ref = oracle.sobol_float(I.ravel(), Dm.ravel()).reshape(len(idx), len(dims))
        np.testing.assert_array_equal(b[:, :len(dims)], ref)


@pytest.mark.parametrize("w,h,frames", [(1, 1, 1), (7, 5, 3), (20, 20, 3), (37, 29, 1), (64, 16, 2), (1920, 1080, 1), (50, 37, 8)])
def test_tiled_path_numbering_is_a_permutation(emu, w, h, frames):
    """Path id -> sampler slot with pixels in 8x8 tiles (renders of fewer than 64 frames):
    every slot frame * npix + pixel is reached exactly once, a path keeps its frame
    (id % frames), the first 64 pixel ranks of a band of full tiles cover one 8x8 block,
    and the row-by-row numbering is pixel * frames + frame."""
    emu.emu_path_slots.argtypes = [C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_void_p]
    npix, n = w * h, w * h * frames
    tiled = np.empty(n, np.uint32)
    plain = np.empty(n, np.uint32)
    assert emu.emu_path_slots(w, h, frames, 1, tiled.ctypes.data) == 0
    assert emu.emu_path_slots(w, h, frames, 0, plain.ctypes.data) == 0
    np.testing.assert_array_equal(np.sort(tiled), np.arange(n, dtype=np.uint32))
    pid = np.arange(n, dtype=np.uint64)
    np.testing.assert_array_equal(tiled // npix, pid % frames)
    np.testing.assert_array_equal(plain, (pid % frames) * npix + pid // frames)
    if w >= 8 and h >= 8:
        px = tiled[: 64 * frames: frames] % npix
        xs, ys = px % w, px // w
        assert sorted(zip(ys.tolist(), xs.tolist())) == [(y, x) for y in range(8) for x in range(8)]


@pytest.mark.parametrize("pass_", [1, 2, 3])
def test_render_passes_device_code_match_oracle(oracle, emu, pass_):
    """Normal / depth / emissive passes (NormalIntegrator, DepthIntegrator, EmissiveIntegrator):
    the device functions on the host against the oracle, pixel by pixel. The host build
    intersects by brute force in primitive order, which settles exact ties at shared edges
    differently from the reference walk; those pixels are counted, not compared."""
    from axomae_amd.nova_hip import INTEGRATOR_PATH, INTEGRATOR_NORMAL, INTEGRATOR_DEPTH, INTEGRATOR_EMISSIVE
    flag = {1: INTEGRATOR_NORMAL, 2: INTEGRATOR_DEPTH, 3: INTEGRATOR_EMISSIVE}[pass_]
    sc = scenes.box_room()
    w, h = 40, 32
    img = emu_render(emu, sc, w, h, 3, 2, integrator_flags=INTEGRATOR_PATH | flag, z_near=0.1, z_far=20.0).reshape(h, w, 4)
    ref, _ = oracle.OracleScene(sc).render_pass(w, h, 3, pass_, z_near=0.1, z_far=20.0)
    same = np.all(img == ref, axis=2)
    print(f"[pass {pass_}] identical pixels {same.mean():.4f}", flush=True)
    assert same.mean() >= 0.99
    assert np.abs(ref).max() > 0 and np.isfinite(img).all()
