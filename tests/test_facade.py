"""The C++ facade (HipAccelerator, flatten) driven through its own API: flatten() of the
engine's holders on the CPU here; the engine's per-frame GPU loop against the oracle's
engine loop on the MI355X."""
import ctypes as C

import numpy as np
import pytest

from axomae_amd import scenes
from conftest import rel_rmse
from facade_case import Packed, load

gpu = pytest.mark.gpu


@pytest.fixture(scope="module")
def driver():
    return load()


def _flatten(L, sc, omit=False):
    pk = Packed(sc, omit)
    counts = np.zeros(4, dtype=np.int32)
    mm = np.zeros(len(sc.meshes), dtype=np.int32)
    xf = np.zeros(len(sc.meshes) * 16, dtype=np.float32)
    b0 = np.zeros(len(sc.meshes[0].vertices), dtype=np.float32)
    err = C.create_string_buffer(256)
    rc = L.facade_flatten(*pk.args(), counts.ctypes.data, mm.ctypes.data, xf.ctypes.data, b0.ctypes.data, err, 256)
    return rc, counts, mm, xf, b0, err.value


def test_flatten_maps_the_engine_holders(driver):
    """Every mesh keeps its material (from its primitives) and its model matrix (the `m`
    block of its 57-float transform record); materials, textures and the envmap index pass
    through; given bitangents are kept."""
    sc = scenes.transformed_scene()
    rc, counts, mm, xf, b0, err = _flatten(driver, sc)
    assert rc == 0, err
    assert counts.tolist() == [len(sc.meshes), len(sc.materials), len(sc.textures), sc.envmap]
    np.testing.assert_array_equal(mm, [m.material for m in sc.meshes])
    np.testing.assert_array_equal(xf, np.concatenate([m.transform for m in sc.meshes]))
    np.testing.assert_array_equal(b0, sc.meshes[0].bitangents)


def test_flatten_completes_missing_bitangents(driver):
    """A mesh without bitangents gets normalize(cross(n, t)) per vertex, as
    Object3D::compute_missing_vectors derives them (Object3D.h:208-222); the scenes'
    frames are right-handed, so this is the bitangent they store, to rounding."""
    sc = scenes.box_room()
    rc, counts, mm, xf, b0, err = _flatten(driver, sc, omit=True)
    assert rc == 0, err
    n = sc.meshes[0].normals.reshape(-1, 3).astype(np.float64)
    t = sc.meshes[0].tangents.reshape(-1, 3).astype(np.float64)
    c = np.cross(n, t)
    c /= np.linalg.norm(c, axis=1, keepdims=True)
    np.testing.assert_allclose(b0.reshape(-1, 3), c, atol=2e-7)
    np.testing.assert_allclose(b0, sc.meshes[0].bitangents, atol=2e-7)


def test_flatten_rejects_a_texture_without_texels(driver):
    """Inconsistent holders raise inside flatten() and come back as an error message, not a
    crash: here a texture whose texel pointer is null."""
    sc = scenes.box_room()
    pk = Packed(sc)
    pk.tex_ptrs[0] = None
    counts = np.zeros(4, dtype=np.int32)
    mm = np.zeros(len(sc.meshes), dtype=np.int32)
    xf = np.zeros(len(sc.meshes) * 16, dtype=np.float32)
    b0 = np.zeros(len(sc.meshes[0].vertices), dtype=np.float32)
    err = C.create_string_buffer(256)
    assert driver.facade_flatten(*pk.args(), counts.ctypes.data, mm.ctypes.data, xf.ctypes.data, b0.ctypes.data, err, 256) == 1
    assert b"texture without texels" in err.value


@gpu
@pytest.mark.parametrize("name,omit", [("box_room", False), ("transformed", False), ("transformed", True)])
def test_engine_frame_loop_through_the_facade_matches_the_oracle(oracle, driver, name, omit):
    """DeviceAcceleratorInterface::make -> build -> per frame copyParamsToDevice and
    getIntersectorObject()->traverse into HdrBufferStruct::partial_buffer, accumulated on the
    host as Engine::writeFramebuffer does (engine.cpp:351-363, 461-494), against the oracle's
    engine loop on the same scene, frames and depth."""
    sc = scenes.box_room() if name == "box_room" else scenes.transformed_scene()
    W, H, F, D = 96, 80, 6, 5
    pk = Packed(sc, omit)
    img = np.zeros(W * H * 4, dtype=np.float32)
    acc = np.zeros(W * H * 4, dtype=np.float32)
    err = C.create_string_buffer(256)
    inv_P = np.ascontiguousarray(sc.inv_P, dtype=np.float32)
    inv_V = np.ascontiguousarray(sc.inv_V, dtype=np.float32)
    rc = driver.facade_engine_loop(*pk.args(), inv_P.ctypes.data, inv_V.ctypes.data, W, H, F, D, -1, 1, img.ctypes.data, acc.ctypes.data, err,
                                   256)
    assert rc == 0, err.value
    ref, racc = oracle.OracleScene(sc).render(W, H, F, D, threads=8)
    e1, e2 = rel_rmse(img.reshape(H, W, 4), ref), rel_rmse(acc.reshape(H, W, 4), racc.reshape(H, W, 4))
    print(f"[facade] {name} bitangents {'completed' if omit else 'given'}: image rel_rmse {e1:.3e}, accumulator {e2:.3e}", flush=True)
    assert e1 <= 1e-3 and e2 <= 1e-3


@gpu
def test_engine_frame_loop_with_frames_rendered_ahead_is_unchanged(driver):
    """setFrameLookahead(4): traverse renders four frames in one call and serves the next three
    from them; every frame's radiance, so the engine's accumulator and image, equal the
    frame-by-frame loop's bit for bit (10 frames: two full lookahead batches and a cut one)."""
    sc = scenes.box_room()
    W, H, F, D = 96, 80, 10, 5
    pk = Packed(sc)
    inv_P = np.ascontiguousarray(sc.inv_P, dtype=np.float32)
    inv_V = np.ascontiguousarray(sc.inv_V, dtype=np.float32)
    out = {}
    for la in (1, 4):
        img = np.zeros(W * H * 4, dtype=np.float32)
        acc = np.zeros(W * H * 4, dtype=np.float32)
        err = C.create_string_buffer(256)
        rc = driver.facade_engine_loop(*pk.args(), inv_P.ctypes.data, inv_V.ctypes.data, W, H, F, D, 3, la, img.ctypes.data, acc.ctypes.data,
                                       err, 256)
        assert rc == 0, err.value
        out[la] = (img, acc)
    print("[facade] lookahead 4 vs 1: 10 frames", flush=True)
    np.testing.assert_array_equal(out[4][1], out[1][1])
    np.testing.assert_array_equal(out[4][0], out[1][0])
