"""The GL viewer's IBL bakes on the MI355X against the oracle: the equirect -> cube map
capture (envmap_bake.frag under RenderPipeline's capture views) and the Riemann-sum
irradiance (irradiance_baker.frag), as bake kinds of nova_hip_bake_batch, within relative
RMSE 1e-3 (the GPU regroups the Riemann sum by azimuth; the cube capture is the same
arithmetic)."""
import numpy as np
import pytest

from axomae_amd import scenes
from conftest import rel_rmse

gpu = pytest.mark.gpu


def _threads():
    import os
    try:
        return max(1, min(len(os.sched_getaffinity(0)), 16))
    except AttributeError:
        return 8


@gpu
@pytest.mark.parametrize("spec", [0, 5])
def test_cubemap_and_riemann_irradiance_match_the_oracle(oracle, spec):
    from axomae_amd import nova_hip
    w, h, d, si, ss = scenes.bake_envmap_specs()[spec]
    env = scenes.sky_envmap(w, h, d, si, ss)
    S, IW, IH = 64, 32, 16
    cube, irr = nova_hip.bake_batch([(env, S, S, nova_hip.BAKE_CUBEMAP, 0.0), (env, IW, IH, nova_hip.BAKE_IRRADIANCE_RIEMANN, 0.0)], 1)
    ref_cube = oracle.bake_cubemap(env, S, threads=_threads())
    ref_irr = oracle.bake_irradiance_riemann(env, IW, IH, threads=_threads())
    e_cube, e_irr = rel_rmse(cube, ref_cube), rel_rmse(irr, ref_irr)
    print(f"[gl-bake] {w}x{h}: cube map 6x{S}^2 rel_rmse={e_cube:.3e}, Riemann irradiance {IW}x{IH} rel_rmse={e_irr:.3e}", flush=True)
    assert cube.shape == (6, S, S, 4) and irr.shape == (IH, IW, 4)
    assert e_cube <= 1e-3 and e_irr <= 1e-3
    assert (cube[..., 3] == 0).all() and (irr[..., 3] == 0).all()


@gpu
def test_gl_bakes_share_a_batch_with_the_sampled_bakes(oracle):
    """Cube captures, Riemann irradiance, Sobol irradiance and prefilter jobs in one call:
    each output equals the one of its own single-kind call."""
    from axomae_amd import nova_hip
    env = scenes.sky_envmap(256, 128, (0.3, 0.8, 0.2), 200.0, 0.01)
    jobs = [(env, 16, 16, nova_hip.BAKE_CUBEMAP, 0.0), (env, 16, 8, nova_hip.BAKE_IRRADIANCE_SOBOL, 0.0),
            (env, 8, 4, nova_hip.BAKE_IRRADIANCE_RIEMANN, 0.0), (env, 32, 16, nova_hip.BAKE_PREFILTER, 0.4)]
    together = nova_hip.bake_batch(jobs, 256)
    for job, out in zip(jobs, together):
        alone = nova_hip.bake_batch([job], 256)[0]
        np.testing.assert_array_equal(out, alone)
