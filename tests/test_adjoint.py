"""Path-replay adjoint on the CPU: the oracle's reverse-mode gradient against central finite
differences of the oracle's own renders, and the device replay code (host build) against
the oracle's gradient.

Loss = sum over pixels and RGB of (I - target)^2 with I = the mean of the frames, so
dLoss/d(frame radiance) = (1/F) * 2 (I - target). Finite differences use common random
numbers (the same sample indices) and take the loss difference in float64 as
sum (I+ - I-)(I+ + I- - 2 target). The roughness gradient is the detached-sampling one:
directions, lobe choices and pdfs are held fixed, which the FD renders reproduce by drawing
the sampling roughness from an unperturbed copy of the texture.
"""
import ctypes as C
import copy

import numpy as np
import pytest

from axomae_amd import nova_hip, scenes
from test_device_code_on_cpu import emu, fp32_textured_scene  # noqa: F401  (fixture)

W, H, F, D = 32, 20, 2, 3


def adjoint_case():
    sc = fp32_textured_scene()
    alb, rough = sc.materials[0].albedo_f32, sc.materials[0].roughness_f32
    t = sc.textures[rough]
    samp = sc.add_texture(scenes.Texture(t.width, t.height, 1, t.data.copy()))
    return sc, alb, rough, samp


def loss_and_dimg(img, target):
    d = img[..., :3].astype(np.float64) - target[..., :3]
    dimg = np.zeros(img.shape, dtype=np.float32)
    dimg[..., :3] = 2 * d
    return float((d * d).sum()), dimg


@pytest.fixture(scope="module")
def case(oracle):
    sc, alb, rough, samp = adjoint_case()
    os_ = oracle.OracleScene(sc)
    img, _ = os_.render(W, H, F, D, sampling_roughness_texture=samp)
    target = img.astype(np.float64) + np.random.default_rng(3).normal(0, 0.05, img.shape)
    _, dimg = loss_and_dimg(img, target)
    grad = os_.render_adjoint(W, H, F, D, dimg, 1.0 / F, sampling_roughness_texture=samp)
    return sc, alb, rough, samp, target, dimg, grad, os_.f32_pool()[0]


def fd_loss_delta(oracle, sc, tex, k, h, target, samp):
    imgs = []
    for sgn in (1, -1):
        s2 = copy.deepcopy(sc)
        s2.textures[tex].data = s2.textures[tex].data.copy()
        s2.textures[tex].data[k] += sgn * h
        im, _ = oracle.OracleScene(s2).render(W, H, F, D, sampling_roughness_texture=samp)
        imgs.append(im[..., :3].astype(np.float64))
    return float(((imgs[0] - imgs[1]) * (imgs[0] + imgs[1] - 2 * target[..., :3])).sum()) / (2 * h)


@pytest.mark.parametrize("which", ["albedo", "roughness"])
def test_oracle_adjoint_matches_central_differences(oracle, case, which):
    sc, alb, rough, samp, target, dimg, grad, offs = case
    tex = alb if which == "albedo" else rough
    t = sc.textures[tex]
    g = grad[offs[tex]:offs[tex] + t.width * t.height * t.channels]
    rng = np.random.default_rng(5)
    # texels whose gradient stands above the FD noise floor of float32 renders at h = 1e-3
    # (a few 1e-4 of the largest gradient here); the 1% bar applies to those
    nz = np.flatnonzero(np.abs(g) > 1e-2 * np.abs(g).max())
    picks = np.concatenate([np.argsort(-np.abs(g))[:4], rng.choice(nz, 4, replace=False)])
    for k in picks:
        fd = fd_loss_delta(oracle, sc, tex, int(k), 1e-3, target, samp)
        print(f"[fd] {which} elem {k}: adjoint {g[k]:.6e} fd {fd:.6e}", flush=True)
        assert abs(g[k] - fd) <= 0.01 * abs(fd) + 2e-4 * np.abs(g).max()


@pytest.mark.parametrize("rr_start", [-1, 1])
def test_device_replay_code_matches_oracle_adjoint(oracle, emu, rr_start):  # noqa: F811
    """Hand-derived closed forms and front-to-back replay (device code, host build) against
    dual numbers over the recursion (oracle), with metal and non-metal lobes, with and
    without Russian roulette."""
    sc, alb, rough, samp = adjoint_case()
    os_ = oracle.OracleScene(sc)
    img, _ = os_.render(W, H, F, D, rr_start=rr_start)
    target = img + np.random.default_rng(4).normal(0, 0.05, img.shape).astype(np.float32)
    _, dimg = loss_and_dimg(img, target)
    g_ref = os_.render_adjoint(W, H, F, D, dimg, 1.0 / F, rr_start=rr_start, sampling_roughness_texture=samp)
    sd, keep = nova_hip.scene_desc(sc)
    p = nova_hip.render_params(sc, W, H, F, D, rr_start=rr_start, sampling_roughness_texture=samp)
    g = np.zeros(len(g_ref), dtype=np.float32)
    assert emu.emu_adjoint(C.byref(sd), C.byref(p), dimg.ctypes.data, 1.0 / F, g.ctypes.data) == 0
    err = np.linalg.norm(g - g_ref) / np.linalg.norm(g_ref)
    print(f"[adjoint] host replay vs oracle rr_start={rr_start}: rel L2 {err:.2e}", flush=True)
    assert np.count_nonzero(g_ref) > 1000
    assert err <= 1e-5
