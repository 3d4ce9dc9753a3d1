"""Path-replay adjoint on the MI355X: against the oracle's CPU gradient, against central
finite differences of the HIP renderer (common random numbers, float64 loss differences),
and through the torch.autograd.Function of the inverse-rendering harness. The
differentiation workload is checked at its own size: 1280x720, 32 spp, depth 4, about
100k triangles, 512x512 FP32 albedo and roughness textures, L2 loss against a render of
textures perturbed by Gaussian noise (sigma 0.05)."""
import numpy as np
import pytest

from axomae_amd import scenes
from test_adjoint import D, F, H, W, adjoint_case, loss_and_dimg

gpu = pytest.mark.gpu


@pytest.fixture(scope="module")
def engine():
    from axomae_amd.nova_hip import HipEngine
    e = HipEngine(0)
    yield e
    e.close()


@gpu
@pytest.mark.parametrize("rr_start", [-1, 1])
def test_gpu_adjoint_matches_oracle_adjoint(oracle, engine, rr_start):
    import torch
    sc, alb, rough, samp = adjoint_case()
    engine.upload(sc)
    engine.build_bvh()
    img = engine.render(W, H, F, D, rr_start=rr_start)[0].cpu().numpy()
    os_ = oracle.OracleScene(sc)
    ref, _ = os_.render(W, H, F, D, rr_start=rr_start)
    target = ref + np.random.default_rng(4).normal(0, 0.05, ref.shape).astype(np.float32)
    _, dimg = loss_and_dimg(ref, target)
    g_ref = os_.render_adjoint(W, H, F, D, dimg, 1.0 / F, rr_start=rr_start, sampling_roughness_texture=samp)
    g = engine.adjoint(W, H, F, D, torch.from_numpy(dimg), 1.0 / F, rr_start=rr_start, sampling_roughness_texture=samp)
    g = g.cpu().numpy()
    err = np.linalg.norm(g - g_ref) / np.linalg.norm(g_ref)
    print(f"[adjoint] gpu vs oracle rr_start={rr_start}: rel L2 {err:.2e}, image rel {np.abs(img - ref).max():.2e}", flush=True)
    assert err <= 1e-4


def _fd_setup(engine, width, height, frames, depth):
    import torch
    sc, alb, rough = scenes.differentiable_scene(width=width, height=height, terrain_cells=120, spheres=6, sphere_level=3, tex=256,
                                                 sky=(256, 128))
    t = sc.textures[rough]
    samp = sc.add_texture(scenes.Texture(t.width, t.height, 1, t.data.copy()))
    engine.upload(sc)
    engine.build_bvh()
    rng = np.random.default_rng(7)
    tgt_alb = np.clip(sc.textures[alb].data + rng.normal(0, 0.05, sc.textures[alb].data.shape), 0.0, 1.0).astype(np.float32)
    tgt_rough = np.clip(sc.textures[rough].data + rng.normal(0, 0.05, sc.textures[rough].data.shape), 0.01, 1.0).astype(np.float32)
    engine.update_texture(alb, torch.from_numpy(tgt_alb))
    engine.update_texture(rough, torch.from_numpy(tgt_rough))
    target = engine.render(width, height, frames, depth, sampling_roughness_texture=samp)[0].double()
    engine.update_texture(alb, torch.from_numpy(sc.textures[alb].data))
    engine.update_texture(rough, torch.from_numpy(sc.textures[rough].data))
    img = engine.render(width, height, frames, depth, sampling_roughness_texture=samp)[0].double()
    return sc, alb, rough, samp, target, img


@gpu
def test_gpu_adjoint_matches_finite_differences(engine):
    """64 texels per texture, drawn among those whose gradient stands above the float32
    FD noise floor; albedo on the plain renderer (sampling does not depend on albedo),
    roughness with detached sampling."""
    import torch
    Wd, Hd, Fd, Dd = 192, 108, 8, 4
    sc, alb, rough, samp, target, img = _fd_setup(engine, Wd, Hd, Fd, Dd)
    dimg = (2 * (img - target)).float()
    dimg[..., 3] = 0
    grad = engine.adjoint(Wd, Hd, Fd, Dd, dimg, 1.0 / Fd, sampling_roughness_texture=samp)
    torch.cuda.synchronize()
    h = 1e-3
    rng = np.random.default_rng(11)
    for tex, samp_tex in ((alb, -1), (rough, samp)):
        base = torch.from_numpy(sc.textures[tex].data).cuda()
        g = engine.texture_grad(grad, tex).reshape(-1).cpu().numpy().astype(np.float64)
        picks = rng.choice(np.flatnonzero(np.abs(g) > 1e-2 * np.abs(g).max()), 64, replace=False)
        worst = 0.0
        for k in picks:
            ims = []
            for sgn in (1, -1):
                x = base.clone()
                x[int(k)] += sgn * h
                engine.update_texture(tex, x)
                ims.append(engine.render(Wd, Hd, Fd, Dd, sampling_roughness_texture=samp_tex)[0][..., :3].double())
            engine.update_texture(tex, base)
            fd = float(((ims[0] - ims[1]) * (ims[0] + ims[1] - 2 * target[..., :3])).sum()) / (2 * h)
            err = abs(g[k] - fd) / (abs(fd) + 2e-2 * np.abs(g).max())
            worst = max(worst, err)
            assert abs(g[k] - fd) <= 0.01 * abs(fd) + 2e-4 * np.abs(g).max(), (tex, int(k), g[k], fd)
        print(f"[fd-gpu] texture {tex}: 64 texels, worst scaled error {worst:.2e}", flush=True)


@gpu
def test_autograd_function_backward(engine):
    """loss.backward() through TextureRender returns the adjoint's texel gradients, and the
    directional derivative along a random texel direction matches a central difference."""
    import torch
    from axomae_amd import diff
    Wd, Hd, Fd, Dd = 96, 54, 4, 3
    sc, alb, rough, samp, target, _ = _fd_setup(engine, Wd, Hd, Fd, Dd)
    t_alb, t_rough = sc.textures[alb], sc.textures[rough]
    a = torch.from_numpy(t_alb.data).cuda().reshape(t_alb.height, t_alb.width, 3).requires_grad_(True)
    r = torch.from_numpy(t_rough.data).cuda().reshape(t_rough.height, t_rough.width, 1).requires_grad_(True)
    cfg = dict(width=Wd, height=Hd, frames=Fd, max_depth=Dd)
    img = diff.render(a, r, engine, alb, rough, **cfg)
    loss = ((img[..., :3].double() - target[..., :3]) ** 2).sum()
    loss.backward()
    assert a.grad is not None and torch.isfinite(a.grad).all() and a.grad.abs().max() > 0
    # directional check on albedo (sampling does not depend on it, so the plain renderer is exact)
    v = torch.randn_like(a) * (a.grad.abs() > 0)
    eps = 1e-3
    with torch.no_grad():
        lp = ((diff.render(a + eps * v, r, engine, alb, rough, **cfg)[..., :3].double() - target[..., :3]) ** 2).sum()
        lm = ((diff.render(a - eps * v, r, engine, alb, rough, **cfg)[..., :3].double() - target[..., :3]) ** 2).sum()
    fd = float((lp - lm) / (2 * eps))
    an = float((a.grad.double() * v.double()).sum())
    print(f"[autograd] directional albedo derivative: adjoint {an:.6e} fd {fd:.6e}", flush=True)
    assert abs(an - fd) <= 0.01 * abs(fd)


@gpu
@pytest.mark.parametrize("rr_start", [-1, 1])
def test_gpu_adjoint_replay_reuses_the_primal_hits(engine, rr_start):
    """The replay reads the primal pass's hits instead of tracing its rays again (the same
    rays): the gradient equals the retracing replay's up to the order of the float atomics."""
    import torch
    from axomae_amd.nova_hip import ADJOINT_RETRACE

    sc, alb, rough, samp = adjoint_case()
    engine.upload(sc)
    engine.build_bvh()
    img = engine.render(W, H, F, D, rr_start=rr_start)[0]
    gen = torch.Generator(device="cpu").manual_seed(3)
    dimg = torch.randn(img.shape, generator=gen).cuda()
    dimg[..., 3] = 0
    g_cache = engine.adjoint(W, H, F, D, dimg, 1.0 / F, rr_start=rr_start, sampling_roughness_texture=samp).cpu().numpy()
    g_trace = engine.adjoint(W, H, F, D, dimg, 1.0 / F, rr_start=rr_start, sampling_roughness_texture=samp, flags=ADJOINT_RETRACE).cpu().numpy()
    err = np.linalg.norm(g_cache.astype(np.float64) - g_trace) / np.linalg.norm(g_trace.astype(np.float64))
    print(f"[adjoint] cached hits vs retrace rr_start={rr_start}: rel L2 {err:.2e}", flush=True)
    assert np.count_nonzero(g_trace) > 1000 and err <= 1e-5


C4 = dict(W=1280, H=720, F=32, D=4)


@pytest.fixture(scope="module")
def c4_engine():
    from axomae_amd.nova_hip import HipEngine
    e = HipEngine(0)
    yield e
    e.close()


@pytest.fixture(scope="module")
def c4(c4_engine):
    """The differentiation workload at full size with an unperturbed copy of the roughness
    texture for detached sampling; target rendered from perturbed textures; the image, the
    loss gradient per frame and the GPU adjoint's texel gradients."""
    import torch
    engine = c4_engine
    W, H, F, D = C4["W"], C4["H"], C4["F"], C4["D"]
    sc, alb, rough = scenes.differentiable_scene(width=W, height=H)
    t = sc.textures[rough]
    samp = sc.add_texture(scenes.Texture(t.width, t.height, 1, t.data.copy()))
    engine.upload(sc)
    engine.build_bvh()
    rng = np.random.default_rng(0xA70AE004)
    ta = np.clip(sc.textures[alb].data + rng.normal(0, 0.05, sc.textures[alb].data.shape), 0, 1).astype(np.float32)
    tr = np.clip(sc.textures[rough].data + rng.normal(0, 0.05, sc.textures[rough].data.shape), 0.01, 1).astype(np.float32)
    engine.update_texture(alb, torch.from_numpy(ta))
    engine.update_texture(rough, torch.from_numpy(tr))
    target = engine.render(W, H, F, D, sampling_roughness_texture=samp)[0].double()
    engine.update_texture(alb, torch.from_numpy(sc.textures[alb].data))
    engine.update_texture(rough, torch.from_numpy(sc.textures[rough].data))
    img = engine.render(W, H, F, D, sampling_roughness_texture=samp)[0]
    dimg = (2 * (img.double() - target)).float()
    dimg[..., 3] = 0
    grad = engine.adjoint(W, H, F, D, dimg, 1.0 / F, sampling_roughness_texture=samp)
    torch.cuda.synchronize()
    return dict(sc=sc, alb=alb, rough=rough, samp=samp, target=target, img=img, dimg=dimg, grad=grad)


@gpu
def test_c4_adjoint_matches_oracle_adjoint_full_size(oracle, c4):
    """Every texel gradient of the full-size workload against the oracle's CPU reverse-mode
    gradient (float64) of the same 29.5 M paths."""
    W, H, F, D = C4["W"], C4["H"], C4["F"], C4["D"]
    osc = oracle.OracleScene(c4["sc"])
    import os
    threads = max(1, min(len(os.sched_getaffinity(0)), 16))
    g_ref = osc.render_adjoint(W, H, F, D, c4["dimg"].cpu().numpy(), 1.0 / F, sampling_roughness_texture=c4["samp"], threads=threads)
    g = c4["grad"].cpu().numpy().astype(np.float64)
    err = np.linalg.norm(g - g_ref) / np.linalg.norm(g_ref)
    print(f"[c4] adjoint vs oracle 1280x720x32 d4: rel L2 {err:.2e}, nonzero texels {np.count_nonzero(g_ref)}", flush=True)
    assert np.count_nonzero(g_ref) > 100000 and err <= 1e-4


def _frame_partials(engine, W, H, F, D, samp_tex):
    """The F frames of a render one call each (frame_count = 1), their radiance buffers
    stacked as (F, H, W, 3) float32: the per-path values the image averages."""
    import ctypes as C
    import torch
    from axomae_amd import nova_hip
    parts = torch.empty((F, H, W, 4), dtype=torch.float32, device="cuda")
    acc = torch.zeros((H, W, 4), dtype=torch.float32, device="cuda")
    fin = torch.empty_like(acc)
    for f in range(F):
        p = nova_hip.render_params(engine.scene, W, H, 1, D, frame_start=f, sampling_roughness_texture=samp_tex)
        engine._check(engine.L.nova_hip_render_frames(engine.ctx, C.byref(p), acc.data_ptr(), fin.data_ptr(), parts[f].data_ptr(),
                                                      nova_hip._stream(torch)), "render_frames")
    return parts[..., :3]


@gpu
def test_c4_adjoint_matches_central_differences_on_random_texels(c4_engine, c4):
    """Central differences (h = 1e-3, float64 loss differences, common random numbers) on
    256 uniformly drawn entries of each texture, gated at 1% of the difference.

    Albedo is differenced on the plain renderer (sampling does not read it); roughness with
    the sampling texture held fixed, as the detached estimator defines its derivative.
    The image differenced is the float64 mean of the 32 frames' float32 radiance buffers.
    (Differencing the renderer's float32 accumulation instead loses up to a few percent of
    the signal on texels seen only through indirect light: one texel's step moves each
    pixel by about an ulp, and the 32 rounded additions then bias the difference.)
    Noise floor: what is left is the float32 rounding of each path's radiance; for the
    changed (frame, pixel) entries E it gives the loss difference a standard deviation of
    about sigma = sqrt(sum_E (3 ulp(x) s_p)^2) / (32 * 2h), s = img+ + img- - 2 target.
    Entries with |fd| under 300 sigma (3 sigma above 1% of the difference) are below the
    floor: they are counted and must agree to 1% plus 5 sigma. Entries no path reaches
    (identical renders, zero gradient) are counted apart."""
    import torch
    engine = c4_engine
    W, H, F, D = C4["W"], C4["H"], C4["F"], C4["D"]
    sc, target = c4["sc"], c4["target"][..., :3]
    h = 1e-3
    rng = np.random.default_rng(0xFD04)
    for tex, samp_tex in ((c4["alb"], -1), (c4["rough"], c4["samp"])):
        base = torch.from_numpy(sc.textures[tex].data).cuda()
        g = engine.texture_grad(c4["grad"], tex).reshape(-1).double()
        picks = rng.choice(base.numel(), 256, replace=False)
        above = below = unseen = 0
        worst = 0.0
        for k in picks.tolist():
            parts = []
            for sgn in (1, -1):
                x = base.clone()
                x[k] += sgn * h
                engine.update_texture(tex, x)
                parts.append(_frame_partials(engine, W, H, F, D, samp_tex))
            engine.update_texture(tex, base)
            p, m = parts[0].double().mean(0), parts[1].double().mean(0)
            s = p + m - 2 * target
            fd = float(((p - m) * s).sum()) / (2 * h)
            changed = parts[0] != parts[1]
            gk = float(g[k])
            if not bool(changed.any()) and gk == 0.0:
                unseen += 1
                continue
            a = parts[0].abs()
            ulp = (torch.nextafter(a, torch.full_like(a, float("inf"))) - a).double()
            sigma = float(torch.sqrt((((3 * ulp * s.unsqueeze(0)) ** 2) * changed).sum())) / (F * 2 * h)
            if abs(fd) >= 300 * sigma and abs(fd) > 0:
                above += 1
                worst = max(worst, abs(gk - fd) / abs(fd))
                assert abs(gk - fd) <= 0.01 * abs(fd), (tex, k, gk, fd, sigma)
            else:
                below += 1
                assert abs(gk - fd) <= 0.01 * abs(fd) + 5 * sigma, (tex, k, gk, fd, sigma)
        print(f"[c4-fd] texture {tex}: 256 entries, {above} above the noise floor (worst relative error {worst:.2e}), "
              f"{below} below it, {unseen} not reached by any path", flush=True)
        assert above >= 32
