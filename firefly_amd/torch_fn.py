"""torch.autograd.Function over the path tracer and its path-replay adjoint.

A thin wrapper for tests and inverse-rendering scripts: the forward renders
(storing every path's radiance), the backward calls the adjoint with
d loss / d image and returns d loss / d texture and d loss / d environment.
Only the rendering is on the GPU path; torch holds the device buffers.
"""
from __future__ import annotations

import torch

from . import abi


class Renderer:
    """A fixed request (scene, environment, camera, size, spp, depth, seed)
    whose albedo texture `texture` and environment texels are parameters."""

    def __init__(self, ctx: abi.Context, scene: abi.Scene, env: abi.EnvMap, cam: abi.Camera, width: int,
                 height: int, spp: int, max_depth: int, seed: int = 0, texture: int = 0, spp_batch: int = 1,
                 deterministic: bool = False):
        self.ctx, self.scene, self.env, self.cam = ctx, scene, env, cam
        self.width, self.height, self.spp, self.max_depth, self.seed = width, height, spp, max_depth, seed
        self.texture, self.spp_batch = texture, spp_batch
        self.flags = abi.FF_ADJOINT_DETERMINISTIC if deterministic else 0


class PathTrace(torch.autograd.Function):
    """image [h][w][3] = mean of `spp` samples, as a function of
    (albedo texture [th][tw][3], environment [eh][ew][3])."""

    @staticmethod
    def forward(fctx, texture: torch.Tensor, env_rgb: torch.Tensor, R: Renderer):
        tex = texture.detach().contiguous().float()
        env = env_rgb.detach().contiguous().float()
        R.scene.set_texture(R.texture, tex)
        R.env.update(env)
        film = torch.zeros((R.height, R.width, 4), dtype=torch.float32, device=tex.device)
        paths = torch.empty((R.spp, R.height, R.width, 4), dtype=torch.float32, device=tex.device)
        abi.render_paths(R.ctx, R.scene, R.env, R.cam, R.width, R.height, R.spp, R.max_depth, R.seed, film, paths,
                         spp_batch=R.spp_batch)
        fctx.R = R
        fctx.save_for_backward(paths)
        fctx.shapes = (texture.shape, env_rgb.shape)
        return film[..., :3] / float(R.spp)

    @staticmethod
    def backward(fctx, grad_image: torch.Tensor):
        R = fctx.R
        (paths,) = fctx.saved_tensors
        d = torch.zeros((R.height, R.width, 4), dtype=torch.float32, device=paths.device)
        d[..., :3] = grad_image
        g_tex = torch.zeros(fctx.shapes[0], dtype=torch.float32, device=paths.device)
        g_env = torch.zeros(fctx.shapes[1], dtype=torch.float32, device=paths.device)
        tex_grads = [None] * (R.texture + 1)
        tex_grads[R.texture] = g_tex
        abi.adjoint(R.ctx, R.scene, R.env, R.cam, R.width, R.height, R.spp, R.max_depth, R.seed, paths, d,
                    env_grad=g_env, tex_grads=tex_grads, spp_batch=R.spp_batch, flags=R.flags)
        return g_tex, g_env, None


def render(texture: torch.Tensor, env_rgb: torch.Tensor, R: Renderer) -> torch.Tensor:
    return PathTrace.apply(texture, env_rgb, R)
