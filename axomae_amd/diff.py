"""Inverse-rendering harness: the HIP path-replay adjoint behind a torch.autograd.Function.

Used by the tests and the gradient benchmark only. The forward pass writes the albedo and
roughness texels into the scene's FP32 texture pool and renders frames [0, frames); the
backward pass calls nova_hip_render_adjoint, which re-renders each batch of frames to get
every path's radiance and replays the paths from the same sample indices, so memory stays
at one batch of path state whatever the number of frames.
"""
import torch

from .nova_hip import HipEngine


class TextureRender(torch.autograd.Function):
    """image = render(albedo, roughness); albedo (H, W, 3) and roughness (H, W, 1) float32 cuda tensors."""

    @staticmethod
    def forward(ctx, albedo, roughness, engine: HipEngine, albedo_tex: int, roughness_tex: int, cfg: dict):
        engine.update_texture(albedo_tex, albedo.detach())
        engine.update_texture(roughness_tex, roughness.detach())
        final, _ = engine.render(cfg["width"], cfg["height"], cfg["frames"], cfg["max_depth"], rr_start=cfg.get("rr_start", -1))
        ctx.save_for_backward(albedo.detach(), roughness.detach())
        ctx.engine, ctx.tex, ctx.cfg = engine, (albedo_tex, roughness_tex), cfg
        return final

    @staticmethod
    def backward(ctx, grad_image):
        albedo, roughness = ctx.saved_tensors
        engine, (ta, tr), cfg = ctx.engine, ctx.tex, ctx.cfg
        # the engine may have rendered other texels since the forward pass
        engine.update_texture(ta, albedo)
        engine.update_texture(tr, roughness)
        # final = (1/frames) * sum of the frames' radiance, so each frame gets grad / frames
        grad = engine.adjoint(cfg["width"], cfg["height"], cfg["frames"], cfg["max_depth"], grad_image, 1.0 / cfg["frames"],
                              rr_start=cfg.get("rr_start", -1))
        ga = engine.texture_grad(grad, ta).reshape(albedo.shape).clone() if ctx.needs_input_grad[0] else None
        gr = engine.texture_grad(grad, tr).reshape(roughness.shape).clone() if ctx.needs_input_grad[1] else None
        return ga, gr, None, None, None, None


def render(albedo, roughness, engine, albedo_tex, roughness_tex, **cfg):
    return TextureRender.apply(albedo, roughness, engine, albedo_tex, roughness_tex, cfg)
