"""Parity at the benchmark configurations' own sizes (MI355X): the box room at 256x256,
16 spp, depth 5 in full; the 1,048,576-triangle 1080p workload on a band of rows of one
frame (the CPU renders only that band); one scene of the 64-scene batch; full-size maps of
the IBL bake batch. Radiance and bakes within relative RMSE 1e-3 of the oracle."""
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


def _threads():
    import os
    try:
        return max(1, min(len(os.sched_getaffinity(0)), 16))
    except AttributeError:
        return 8


@gpu
def test_box_room_full_config(oracle, engine):
    sc = scenes.box_room()
    engine.upload(sc)
    engine.build_bvh()
    img = engine.render(256, 256, 16, 5)[0].cpu().numpy()
    ref, _ = oracle.OracleScene(sc).render(256, 256, 16, 5, threads=_threads())
    err = rel_rmse(img, ref)
    print(f"[config] box room 256x256x16 d5 rel_rmse={err:.3e}", flush=True)
    assert err <= 1e-3


@gpu
def test_million_triangles_1080p_band(oracle, engine):
    sc = scenes.million_triangles()
    engine.upload(sc)
    engine.build_bvh()
    W, H, frame = 1920, 1080, 1
    import torch
    partial = torch.empty((H, W, 4), dtype=torch.float32, device="cuda")
    from axomae_amd import nova_hip
    import ctypes as C
    p = nova_hip.render_params(sc, W, H, 1, 8, frame_start=frame, rr_start=3)
    acc = torch.zeros((H, W, 4), dtype=torch.float32, device="cuda")
    fin = torch.empty_like(acc)
    engine._check(engine.L.nova_hip_render_frames(engine.ctx, C.byref(p), acc.data_ptr(), fin.data_ptr(), partial.data_ptr(),
                                                  nova_hip._stream(torch)), "render_frames")
    gpu_img = partial.cpu().numpy()
    y0, y1 = 500, 524  # rows of the reference's y (row 0 at the bottom of the image buffer order)
    osc = oracle.OracleScene(sc)
    ref = oracle.render_rows(osc, W, H, frame, y0, y1, 8, rr_start=3, threads=_threads()).reshape(H, W, 4)
    err = rel_rmse(gpu_img[y0:y1], ref[y0:y1])
    print(f"[config] 1M triangles 1080p frame {frame} rows {y0}-{y1 - 1} d8 rr3 rel_rmse={err:.3e}", flush=True)
    assert err <= 1e-3


@gpu
def test_batch_scene_matches_oracle(oracle):
    from axomae_amd.nova_hip import BatchConfig, BatchRenderer, HipEngine
    specs = scenes.batch_scene_specs()
    seed, tris, res, spp = min(specs, key=lambda s: s[1] * s[2] * s[2] * s[3])
    small = scenes.batch_scene_specs(seed=1, count=3)
    engines, jobs = [], []
    for sd, t, r, s in [(seed, tris, res, spp)] + [(a, min(b, 20000), 64, 4) for a, b, _, _ in small]:
        e = HipEngine(0)
        e.upload(scenes.batch_scene(sd, t, r))
        e.build_bvh()
        engines.append(e)
        jobs.append((e, r, r, s, 6, {}))
    br = BatchRenderer(0)
    finals, st, _ = br.render(jobs, BatchConfig(1 << 20, 8, 8, 1))
    sc = engines[0].scene
    ref, _ = oracle.OracleScene(sc).render(res, res, spp, 6, threads=_threads())
    err = rel_rmse(finals[0].cpu().numpy(), ref)
    print(f"[config] batch scene {tris} triangles {res}^2 x {spp} d6 rel_rmse={err:.3e}", flush=True)
    assert err <= 1e-3
    br.close()
    for e in engines:
        e.close()


@gpu
def test_bake_batch_full_size_maps(oracle):
    from axomae_amd import nova_hip
    specs = scenes.bake_envmap_specs()
    picks = [max(range(len(specs)), key=lambda i: specs[i][0]), min(range(len(specs)), key=lambda i: specs[i][0])]
    jobs, refs = [], []
    for i in picks:
        w, h, d, si, ss = specs[i]
        env = scenes.sky_envmap(w, h, d, si, ss)
        jobs.append((env, 64, 32, nova_hip.BAKE_IRRADIANCE_SOBOL, 0.0))
        refs.append((env, 64, 32, None))
        for lvl in (1, 3, 5):
            jobs.append((env, 256 >> lvl, 128 >> lvl, nova_hip.BAKE_PREFILTER, lvl / 5.0))
            refs.append((env, 256 >> lvl, 128 >> lvl, lvl / 5.0))
    outs = nova_hip.bake_batch(jobs, 1024)
    for (env, w, h, r), o in zip(refs, outs):
        ref = oracle.bake_irradiance(env, w, h, 1024, threads=_threads(), sobol=True) if r is None else \
            oracle.bake_prefilter(env, w, h, r, 1024, threads=_threads())
        err = rel_rmse(o, ref)
        print(f"[config] bake {env.width}x{env.height} -> {w}x{h} {'irradiance' if r is None else f'prefilter r={r}'} rel_rmse={err:.3e}",
              flush=True)
        assert err <= 1e-3
