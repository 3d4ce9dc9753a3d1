"""Batched rendering on the MI355X: every scene of a batch equals its own
nova_hip_render_frames image bit for bit, with the wave sequence replayed from a HIP graph
or launched directly, and with queue capacities small enough that waves split frames and
mix scenes; the autotuner returns one of its candidates."""
import numpy as np
import pytest

from axomae_amd import scenes

gpu = pytest.mark.gpu

SPECS = [(11, 3000, 40, 3), (12, 800, 24, 5), (13, 12000, 56, 2), (14, 1500, 32, 4), (15, 5000, 20, 6)]
DEPTH = 6


# every path count a multiple of 64: bounce 0 walks packets of 64 camera rays
SPECS64 = [(11, 3000, 40, 8), (13, 12000, 56, 2), (14, 1500, 32, 4), (16, 2000, 24, 16)]


@pytest.fixture(scope="module", params=["mixed", "packets64"])
def batch_jobs(request):
    from axomae_amd.nova_hip import HipEngine
    specs = SPECS if request.param == "mixed" else SPECS64
    engines, jobs, refs, segments = [], [], [], 0
    for seed, tris, res, spp in specs:
        e = HipEngine(0)
        e.upload(scenes.batch_scene(seed, tris, res))
        e.build_bvh()
        engines.append(e)
        ref, _ = e.render(res, res, spp, DEPTH, rr_start=3)
        segments += e.render_stats().segments
        refs.append(ref.cpu().numpy())
        jobs.append((e, res, res, spp, DEPTH, {"rr_start": 3}))
    yield jobs, refs, segments
    for e in engines:
        e.close()


@gpu
@pytest.mark.parametrize("use_graph", [1, 0])
@pytest.mark.parametrize("cap", [2048, 1 << 20])
def test_batch_equals_single_scene_renders(batch_jobs, use_graph, cap):
    from axomae_amd.nova_hip import BatchConfig, BatchRenderer
    jobs, refs, segments = batch_jobs
    br = BatchRenderer(0)
    finals, st, comp = br.render(jobs, BatchConfig(cap, 8, 8, use_graph))
    print(f"[batch] cap {cap} graph {use_graph}: {st.waves} waves, {st.paths} paths, {st.segments} segments, {st.total_ms:.2f} ms", flush=True)
    assert st.paths == sum(w * h * s for _, w, h, s, _, _ in jobs)
    assert st.segments == segments  # rays traced, as the single-scene renders count them
    for k, (f, ref) in enumerate(zip(finals, refs)):
        np.testing.assert_array_equal(f.cpu().numpy(), ref, err_msg=f"scene {k}")
    assert np.all(np.diff(comp) >= 0) and comp[-1] <= st.total_ms + 1e-3
    br.close()


@gpu
def test_batch_autotune_picks_a_candidate(batch_jobs):
    from axomae_amd.nova_hip import BatchRenderer
    jobs, refs, _ = batch_jobs
    br = BatchRenderer(0)
    cfg, ns = br.autotune(jobs)
    print(f"[batch] autotune: {cfg.as_dict()} {ns:.3f} ns/path", flush=True)
    assert 1 <= cfg.trace_blocks_per_cu <= 64 and 1 <= cfg.shade_blocks_per_cu <= 64 and cfg.use_graph == 1
    assert cfg.queue_capacity % 256 == 0 and cfg.queue_capacity <= 1 << 26  # what nova_hip_render_batch accepts
    assert ns > 0
    finals, st, _ = br.render(jobs, cfg)
    for f, ref in zip(finals, refs):
        np.testing.assert_array_equal(f.cpu().numpy(), ref)
    br.close()


@gpu
def test_batch_autotune_wave_starts_on_a_packet_boundary():
    """Two 1000x1000 scenes at 5 and 4 spp: 9,000,000 paths, 64 modulo 128, every scene's
    path count a multiple of 64 (camera rays go in packets of 64). The autotuner's trial
    waves (here over the whole batch, at every candidate capacity) must each start on a
    packet boundary so that no packet mixes two scenes. The tuned configuration is accepted
    by the renderer and every image equals its single-scene render."""
    from axomae_amd.nova_hip import BatchRenderer, HipEngine
    engines, jobs, refs = [], [], []
    for seed, spp in ((21, 5), (22, 4)):
        e = HipEngine(0)
        e.upload(scenes.batch_scene(seed, 2000, 1000))
        e.build_bvh()
        engines.append(e)
        refs.append(e.render(1000, 1000, spp, DEPTH, rr_start=3)[0].cpu().numpy())
        jobs.append((e, 1000, 1000, spp, DEPTH, {"rr_start": 3}))
    assert sum(w * h * s for _, w, h, s, _, _ in jobs) % 128 == 64
    br = BatchRenderer(0)
    cfg, ns = br.autotune(jobs)
    print(f"[batch] autotune 2 x 1000^2: {cfg.as_dict()} {ns:.3f} ns/path", flush=True)
    finals, st, _ = br.render(jobs, cfg)
    for k, (f, ref) in enumerate(zip(finals, refs)):
        np.testing.assert_array_equal(f.cpu().numpy(), ref, err_msg=f"scene {k}")
    br.close()
    for e in engines:
        e.close()


@gpu
def test_batch_refuses_a_frame_range_past_int32(batch_jobs):
    """A job whose frames would run past INT32_MAX is refused with INVALID_ARGUMENT before any work."""
    from axomae_amd.nova_hip import BatchConfig, BatchRenderer, NovaHipError
    jobs, _, _ = batch_jobs
    e, w, h, _, d, _ = jobs[0]
    br = BatchRenderer(0)
    with pytest.raises(NovaHipError) as err:
        br.render([(e, w, h, 2, d, {"frame_start": 2**31 - 2})], BatchConfig(1 << 20, 8, 8, 1))
    assert err.value.code == 15  # INVALID_ARGUMENT
    br.close()
