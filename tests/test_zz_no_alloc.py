"""nova_hip_render_frames_noalloc (the render entry point that allocates nothing: its queues,
path state and traversal stacks come from a fixed device array) against
nova_hip_render_frames through the C ABI: final image, accumulator and last-frame radiance
equal bit for bit on two calls in a row, the second smaller than the first, and sizes
beyond the fixed array refused before any work."""
import ctypes as C

import numpy as np
import pytest

from axomae_amd import scenes

gpu = pytest.mark.gpu


def _call(engine, fn, sc, W, H, frames, depth, rr_start):
    import torch
    from axomae_amd import nova_hip
    p = nova_hip.render_params(sc, W, H, frames, depth, rr_start=rr_start)
    acc = torch.zeros((H, W, 4), dtype=torch.float32, device="cuda")
    fin = torch.empty_like(acc)
    part = torch.empty_like(acc)
    rc = fn(engine.ctx, C.byref(p), acc.data_ptr(), fin.data_ptr(), part.data_ptr(), nova_hip._stream(torch))
    torch.cuda.synchronize()
    return rc, fin.cpu().numpy(), acc.cpu().numpy(), part.cpu().numpy()


@gpu
def test_render_frames_noalloc_equals_render_frames():
    from axomae_amd.nova_hip import HipEngine
    sc = scenes.million_triangles()
    e = HipEngine(0)
    e.upload(sc)
    e.build_bvh()
    # a 1080p frame fills most of the array (one frame per batch); then a smaller image
    for W, H, frames in ((1920, 1080, 3), (640, 360, 5)):
        print(f"[noalloc] {W}x{H} x {frames} frames, depth 8, RR from scatter 3", flush=True)
        rc, fin, acc, part = _call(e, e.L.nova_hip_render_frames_noalloc, sc, W, H, frames, 8, 3)
        assert rc == 0, e.L.nova_hip_last_error(e.ctx)
        rc, rfin, racc, rpart = _call(e, e.L.nova_hip_render_frames, sc, W, H, frames, 8, 3)
        assert rc == 0
        np.testing.assert_array_equal(fin, rfin)
        np.testing.assert_array_equal(acc, racc)
        np.testing.assert_array_equal(part, rpart)
        assert np.isfinite(fin).all() and fin[..., :3].max() > 0
    e.close()


@gpu
def test_render_frames_noalloc_refuses_what_does_not_fit():
    import torch
    from axomae_amd import nova_hip
    from axomae_amd.nova_hip import HipEngine, RENDER_TIMING
    sc = scenes.box_room()
    e = HipEngine(0)
    e.upload(sc)
    e.build_bvh()
    W, H = 4096, 4096  # 16.8 M paths per frame: about 1.9 GB of queues and state
    p = nova_hip.render_params(sc, W, H, 1, 5)
    rc = e.L.nova_hip_render_frames_noalloc(e.ctx, C.byref(p), 256, 256, None, nova_hip._stream(torch))
    assert rc == 15 and b"no-allocation" in e.L.nova_hip_last_error(e.ctx)
    p = nova_hip.render_params(sc, 64, 64, 1, 5, flags=RENDER_TIMING)
    rc = e.L.nova_hip_render_frames_noalloc(e.ctx, C.byref(p), 256, 256, None, nova_hip._stream(torch))
    assert rc == 15
    e.close()
