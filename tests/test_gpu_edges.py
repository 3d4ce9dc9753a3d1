"""Edge cases of the render entry point on the device: one-pixel images, depth 0, a view in
which every ray escapes to the sky, renders that need more than one batch of paths, and the
argument errors the C ABI reports (needs an MI355X)."""
import numpy as np
import pytest

from axomae_amd import scenes
from conftest import rel_rmse

gpu = pytest.mark.gpu


@pytest.fixture(scope="module")
def engine():
    from axomae_amd.nova_hip import HipEngine
    e = HipEngine(0)
    yield e
    e.close()


@gpu
@pytest.mark.parametrize("depth", [0, 1, 8])
def test_one_pixel_image_matches_oracle(oracle, engine, depth):
    """A 1x1 image (one path per frame) at depth 0, 1 and 8, against the oracle."""
    sc = scenes.scaled_million(scale_cells=32, spheres=4, sphere_level=2, width=1, height=1)
    engine.upload(sc)
    engine.build_bvh()
    img, _ = engine.render(1, 1, 16, depth, rr_start=-1)
    ref, _ = oracle.OracleScene(sc).render(1, 1, 16, depth)
    err = rel_rmse(img.cpu().numpy(), ref)
    print(f"[edge] 1x1 x16 depth {depth}: rel_rmse={err:.3e}", flush=True)
    assert err <= 1e-3


@gpu
def test_view_of_the_sky_only(oracle, engine):
    """Every camera ray escapes: the image is the environment map seen through the pixels
    (the integrator's miss branch at depth 0), as in the oracle."""
    sc = scenes.scaled_million(scale_cells=32, spheres=4, sphere_level=2, width=40, height=30)
    scenes.set_camera(sc, eye=(0.0, 50.0, 0.0), center=(0.3, 100.0, 0.2), fovy=40.0, aspect=40 / 30)
    engine.upload(sc)
    engine.build_bvh()
    img, _ = engine.render(40, 30, 8, 5)
    st = engine.render_stats()
    ref, _ = oracle.OracleScene(sc).render(40, 30, 8, 5)
    err = rel_rmse(img.cpu().numpy(), ref)
    print(f"[edge] sky only 40x30 x8: escaped {st.escaped} of {st.segments} segments, rel_rmse={err:.3e}", flush=True)
    assert st.escaped == st.segments == 40 * 30 * 8
    assert err <= 1e-3


@gpu
@pytest.mark.parametrize("first", [0, 2**31 - 10])
def test_render_larger_than_one_batch_equals_frame_by_frame(engine, first):
    """4096 x 4096 x 9 frames (151 M paths) needs two batches of at most 2^27 paths; its
    accumulator and image equal those of nine one-frame calls accumulating in turn, bit for
    bit (frames are accumulated in frame order either way). The second case's range ends at
    frame INT32_MAX - 1, where the step past the second batch would leave 32-bit range."""
    W = H = 4096
    sc = scenes.box_room(width=W, height=H)
    engine.upload(sc)
    engine.build_bvh()
    img, acc = engine.render(W, H, 9, 2, frame_start=first)
    acc_seq = None
    for f in range(9):
        img_seq, acc_seq = engine.render(W, H, 1, 2, frame_start=first + f, accum=acc_seq)
    print(f"[edge] 4096^2 x 9 from frame {first}: one call vs nine", flush=True)
    np.testing.assert_array_equal(acc.cpu().numpy(), acc_seq.cpu().numpy())
    np.testing.assert_array_equal(img.cpu().numpy(), img_seq.cpu().numpy())


@gpu
def test_render_argument_errors(engine):
    """Bad image sizes are refused with INVALID_RENDERBUFFER_DIM (the reference's exception
    for a render buffer of the wrong size, DrawEngine.cpp:44-47), bad frame ranges and depths
    with INVALID_ARGUMENT, before any work; an empty scene with INVALID_SCENE_TYPE, the
    reference's code for an empty primitive list."""
    from axomae_amd.nova_hip import NovaHipError
    INVALID_SCENE_TYPE, INVALID_ARGUMENT, INVALID_RENDERBUFFER_DIM = 4, 15, 101  # ERROR_STATE numbering (api_common.h:7-27)
    sc = scenes.box_room()
    engine.upload(sc)
    engine.build_bvh()
    for kw, code in ((dict(width=0, height=8, frames=1, max_depth=2), INVALID_RENDERBUFFER_DIM),
                     (dict(width=70000, height=8, frames=1, max_depth=2), INVALID_RENDERBUFFER_DIM),
                     (dict(width=8, height=8, frames=0, max_depth=2), INVALID_ARGUMENT),
                     (dict(width=8, height=8, frames=1, max_depth=-1), INVALID_ARGUMENT),
                     (dict(width=8, height=8, frames=2, max_depth=2, frame_start=2**31 - 2), INVALID_ARGUMENT)):
        with pytest.raises(NovaHipError) as e:
            engine.render(**kw)
        assert e.value.code == code, kw
    empty = scenes.Scene(name="empty")
    empty.materials = list(sc.materials)
    with pytest.raises(NovaHipError) as e:
        engine.upload(empty)
    assert e.value.code == INVALID_SCENE_TYPE
