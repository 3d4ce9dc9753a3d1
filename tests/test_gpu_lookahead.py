"""Frames rendered ahead (nova_hip_render_frame_lookahead) on the MI355X: one frame per call as
an engine loop renders it, every frame's accumulator, image and radiance equal to one-frame
renders bit for bit, and frames kept from an earlier call are not served after the camera or
a texture changed."""
import numpy as np
import pytest
import torch

from axomae_amd import scenes

gpu = pytest.mark.gpu


@pytest.fixture(scope="module")
def diff_scene():
    from axomae_amd.nova_hip import HipEngine
    sc, alb, rough = scenes.differentiable_scene(width=64, height=48, tex=32, terrain_cells=12, spheres=2, sphere_level=1, sky=(64, 32))
    e = HipEngine(0)
    e.upload(sc)
    e.build_bvh()
    yield e, sc, alb
    e.close()


def _moved_camera(sc):
    v = np.array(sc.inv_V, dtype=np.float32).reshape(4, 4).copy()
    v[3, 0] += 0.05  # column-major: translation in the last column
    return v.reshape(-1)


@gpu
@pytest.mark.parametrize("lookahead", [3, 8])
def test_lookahead_frames_equal_one_frame_renders(diff_scene, lookahead):
    """13 frames, one per call; from frame 6 the camera moves and at frame 10 a texture changes.
    After every call the accumulator, image and frame radiance equal those of one-frame
    renders with the same sequence of camera and texture."""
    e, sc, alb = diff_scene
    W, H, D = 64, 48, 4
    base_tex = torch.from_numpy(sc.textures[alb].data.copy())
    other_tex = torch.from_numpy(np.ascontiguousarray(sc.textures[alb].data[::-1]).copy())
    runs = {}
    for mode in ("plain", "ahead"):
        e.update_texture(alb, base_tex)
        acc = torch.zeros((H, W, 4), dtype=torch.float32, device="cuda")
        fin = torch.empty_like(acc)
        part = torch.empty_like(acc)
        seq = []
        for f in range(13):
            inv_V = _moved_camera(sc) if f >= 6 else None
            if f == 10:
                e.update_texture(alb, other_tex)
            if mode == "plain":
                fin, acc = e.render(W, H, 1, D, frame_start=f, accum=acc, rr_start=2, inv_V=inv_V)
                seq.append((acc.cpu().numpy().copy(), fin.cpu().numpy().copy()))
            else:
                e.render_frame_lookahead(W, H, f, D, lookahead, acc, fin, part, rr_start=2, inv_V=inv_V)
                seq.append((acc.cpu().numpy().copy(), fin.cpu().numpy().copy(), part.cpu().numpy().copy()))
        runs[mode] = seq
    e.update_texture(alb, base_tex)
    print(f"[lookahead] {lookahead}: 13 frames, camera moved at 6, texture changed at 10", flush=True)
    for f in range(13):
        np.testing.assert_array_equal(runs["ahead"][f][0], runs["plain"][f][0], err_msg=f"accumulator after frame {f}")
        np.testing.assert_array_equal(runs["ahead"][f][1], runs["plain"][f][1], err_msg=f"image after frame {f}")
        prev = runs["plain"][f - 1][0] if f else np.zeros_like(runs["plain"][0][0])
        np.testing.assert_array_equal(prev + runs["ahead"][f][2], runs["plain"][f][0], err_msg=f"radiance of frame {f}")


@gpu
def test_lookahead_with_the_gpu_index_policy_and_vertical_invert(diff_scene):
    """The reference GPU raygen's sample index ((x*H + y)*sample_max + frame) and flipped rows:
    9 frames, 4 ahead, equal to one-frame renders after every call."""
    from axomae_amd.nova_hip import INDEX_OPTIX
    e, sc, _ = diff_scene
    W, H, D = 48, 40, 3
    acc = torch.zeros((H, W, 4), dtype=torch.float32, device="cuda")
    fin, part = torch.empty_like(acc), torch.empty_like(acc)
    acc_p = None
    for f in range(9):
        e.render_frame_lookahead(W, H, f, D, 4, acc, fin, part, vertical_invert=True, index_policy=INDEX_OPTIX, max_samples_offset=9)
        fin_p, acc_p = e.render(W, H, 1, D, frame_start=f, accum=acc_p, vertical_invert=True, index_policy=INDEX_OPTIX, max_samples_offset=9)
        assert torch.equal(acc, acc_p), f
        assert torch.equal(fin, fin_p), f
    print("[lookahead] OPTIX policy, vertical invert: 9 frames, 4 ahead", flush=True)


@gpu
def test_lookahead_arguments(diff_scene):
    """More than one frame per call and a lookahead outside [1, 1024] are refused."""
    from axomae_amd.nova_hip import render_params
    import ctypes as C
    e, sc, _ = diff_scene
    acc = torch.zeros((48, 64, 4), dtype=torch.float32, device="cuda")
    fin = torch.empty_like(acc)
    for frames, la in ((2, 4), (1, 0), (1, 2000)):
        p = render_params(sc, 64, 48, frames, 4)
        rc = e.L.nova_hip_render_frame_lookahead(e.ctx, C.byref(p), la, acc.data_ptr(), fin.data_ptr(), None, None)
        assert rc == 15, (frames, la, rc)  # INVALID_ARGUMENT


@gpu
def test_lookahead_at_the_end_of_the_frame_range(diff_scene):
    """Frames INT32_MAX - 4 .. INT32_MAX - 1 with 8 ahead: a call renders only the frames left
    below INT32_MAX, and every call equals a one-frame render."""
    e, sc, _ = diff_scene
    W, H, D = 32, 24, 3
    acc = torch.zeros((H, W, 4), dtype=torch.float32, device="cuda")
    fin, part = torch.empty_like(acc), torch.empty_like(acc)
    acc_p = None
    for f in range(2**31 - 5, 2**31 - 1):
        e.render_frame_lookahead(W, H, f, D, 8, acc, fin, part)
        fin_p, acc_p = e.render(W, H, 1, D, frame_start=f, accum=acc_p)
        assert torch.equal(acc, acc_p), f
        assert torch.equal(fin, fin_p), f
    print("[lookahead] frames INT32_MAX-4 .. INT32_MAX-1, 8 ahead", flush=True)


@gpu
def test_lookahead_calls_that_do_not_wait(diff_scene):
    """Frames rendered 4 ahead by calls that return once queued (NOVA_HIP_RENDER_ASYNC): a call
    that misses queues the batch and the first frame's accumulation without waiting, later
    calls queue their copies; after 10 calls and one synchronization the accumulator and image
    equal those of waiting one-frame renders."""
    from axomae_amd.nova_hip import RENDER_ASYNC
    e, sc, _ = diff_scene
    W, H, D = 64, 48, 4
    acc = torch.zeros((H, W, 4), dtype=torch.float32, device="cuda")
    fin, part = torch.empty_like(acc), torch.empty_like(acc)
    for f in range(10):
        e.render_frame_lookahead(W, H, f, D, 4, acc, fin, part, rr_start=2, flags=RENDER_ASYNC)
    torch.cuda.synchronize()
    acc_p = None
    for f in range(10):
        fin_p, acc_p = e.render(W, H, 1, D, frame_start=f, accum=acc_p, rr_start=2)
    print("[lookahead] 10 frames, 4 ahead, calls that do not wait", flush=True)
    assert torch.equal(acc, acc_p)
    assert torch.equal(fin, fin_p)


@gpu
def test_lookahead_longer_than_one_batch():
    """4096 x 4096 with 9 frames ahead: the renderer splits them into batches of at most 2^27
    paths (8 frames, then 1), and each batch's frames land at their own place among the kept
    images. Frames 7 (end of the first batch) and 8 (the second) equal one-frame renders."""
    from axomae_amd.nova_hip import HipEngine
    W = H = 4096
    sc = scenes.box_room(width=W, height=H)
    e = HipEngine(0)
    e.upload(sc)
    e.build_bvh()
    acc = torch.zeros((H, W, 4), dtype=torch.float32, device="cuda")
    fin, part = torch.empty_like(acc), torch.empty_like(acc)
    acc_p = None
    for f in range(9):
        e.render_frame_lookahead(W, H, f, 2, 9, acc, fin, part)
        fin_p, acc_p = e.render(W, H, 1, 2, frame_start=f, accum=acc_p)
        if f in (7, 8):
            print(f"[lookahead] 4096^2, 9 ahead: frame {f}", flush=True)
            assert torch.equal(acc, acc_p), f
            assert torch.equal(fin, fin_p), f
    e.close()


@gpu
def test_renderbuffer_loop_with_frames_rendered_ahead(diff_scene):
    """The device engine loop (render, accumulate, swap) served from frames rendered 4 ahead:
    after each of 10 frames the accumulator and the displayed (front) image equal the plain
    loop's."""
    from axomae_amd.nova_hip import RB_ACCUMULATOR, RB_FRONT, RenderBuffer
    e, sc, _ = diff_scene
    W, H, D = 64, 48, 4
    plain, ahead = RenderBuffer(e, W, H), RenderBuffer(e, W, H)
    plain.reset()
    ahead.reset()
    for f in range(10):
        e.render_to(plain, 1, D, frame_start=f, rr_start=2)
        e.render_to_lookahead(ahead, f, D, 4, rr_start=2)
        np.testing.assert_array_equal(ahead.read(RB_ACCUMULATOR), plain.read(RB_ACCUMULATOR), err_msg=f"accumulator after frame {f}")
        np.testing.assert_array_equal(ahead.read(RB_FRONT), plain.read(RB_FRONT), err_msg=f"front after frame {f}")
    print("[lookahead] render buffer loop: 10 frames, 4 ahead", flush=True)
    plain.close()
    ahead.close()
