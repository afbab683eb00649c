"""torch.autograd binding of the path-replay adjoint (tpt_render_adjoint,
tpt_render_adjoint_texels).

``render_albedo(albedo, scene, ...)`` renders the scene with the
NormalizedLambert albedos taken from ``albedo`` (one RGB row per material;
rows of other materials are ignored) and returns the image as a tensor whose
backward pass runs the adjoint kernel: ``loss(render_albedo(a, ...)).backward()``
fills ``a.grad`` with dLoss/d(albedo).  ``render_texture(texels, index, ...)``
does the same for the level-0 texels of one OpenPBR base-colour texture.
``render_params(params, scene, ...)`` takes one row of five OpenPBR
parameters per material (base colour RGB, specular_roughness,
base_metalness; rows of other materials are ignored) and differentiates
through tpt_render_adjoint_params.
"""
from __future__ import annotations

import numpy as np
import torch

from .material import NormalizedLambert


def with_albedo(materials, albedo: np.ndarray) -> list:
    """The material list with every NormalizedLambert albedo replaced by the
    matching row of ``albedo`` (shape (len(materials), 3))."""
    out = []
    for i, m in enumerate(materials):
        out.append(NormalizedLambert(tuple(float(x) for x in albedo[i])) if isinstance(m, NormalizedLambert) else m)
    return out


class AlbedoRender(torch.autograd.Function):
    @staticmethod
    def forward(ctx, albedo, device_scene, materials, camera, width, height, spp, max_depth):
        a = albedo.detach().to("cpu", torch.float32).numpy()
        device_scene.set_materials(with_albedo(materials, a))
        img = device_scene.render(camera, width, height, spp, max_depth)
        ctx.call = (device_scene, materials, a, camera, width, height, spp, max_depth)
        return torch.from_numpy(img).to(albedo.device)

    @staticmethod
    def backward(ctx, grad_image):
        device_scene, materials, a, camera, width, height, spp, max_depth = ctx.call
        device_scene.set_materials(with_albedo(materials, a))  # the forward's parameters
        g = device_scene.render_adjoint(camera, width, height, spp, max_depth,
                                        grad_image.detach().to("cpu", torch.float32).numpy())
        return torch.from_numpy(g).to(grad_image.device), None, None, None, None, None, None, None


def render_albedo(albedo: torch.Tensor, device_scene, materials, camera, width: int, height: int, spp: int,
                  max_depth: int) -> torch.Tensor:
    """Differentiable render (see the module docstring)."""
    return AlbedoRender.apply(albedo, device_scene, materials, camera, width, height, spp, max_depth)


class TextureRender(torch.autograd.Function):
    @staticmethod
    def forward(ctx, texels, texture, device_scene, camera, width, height, spp, max_depth):
        t = texels.detach().to("cpu", torch.float32).numpy()
        device_scene.set_texture(texture, t)
        img = device_scene.render(camera, width, height, spp, max_depth)
        ctx.call = (device_scene, texture, t, camera, width, height, spp, max_depth)
        return torch.from_numpy(img).to(texels.device)

    @staticmethod
    def backward(ctx, grad_image):
        device_scene, texture, t, camera, width, height, spp, max_depth = ctx.call
        device_scene.set_texture(texture, t)  # the forward's texels
        _, grads = device_scene.render_adjoint_texels(camera, width, height, spp, max_depth,
                                                      grad_image.detach().to("cpu", torch.float32).numpy())
        return torch.from_numpy(grads[texture]).to(grad_image.device), None, None, None, None, None, None, None


def render_texture(texels: torch.Tensor, texture: int, device_scene, camera, width: int, height: int, spp: int,
                   max_depth: int) -> torch.Tensor:
    """Differentiable render in the level-0 texels of one base-colour texture
    of the scene (index ``texture``): the forward uploads ``texels``
    (tpt_scene_set_texture) and renders; the backward runs the path-replay
    adjoint and returns dLoss/d(texels) (tpt_render_adjoint_texels)."""
    return TextureRender.apply(texels, texture, device_scene, camera, width, height, spp, max_depth)


def with_params(materials, params: np.ndarray) -> list:
    """The material list with every OpenPBR material's base colour, specular
    roughness and metalness replaced by the matching row of ``params``
    (shape (len(materials), 5))."""
    from .material import OpenPbr

    out = []
    for i, m in enumerate(materials):
        if isinstance(m, OpenPbr):
            r = params[i]
            m = m.with_(base_color=(float(r[0]), float(r[1]), float(r[2])), specular_roughness=float(r[3]),
                        base_metalness=float(r[4]))
        out.append(m)
    return out


class ParamsRender(torch.autograd.Function):
    @staticmethod
    def forward(ctx, params, device_scene, materials, camera, width, height, spp, max_depth):
        p = params.detach().to("cpu", torch.float32).numpy()
        device_scene.set_materials(with_params(materials, p))
        img = device_scene.render(camera, width, height, spp, max_depth)
        ctx.call = (device_scene, materials, p, camera, width, height, spp, max_depth)
        return torch.from_numpy(img).to(params.device)

    @staticmethod
    def backward(ctx, grad_image):
        device_scene, materials, p, camera, width, height, spp, max_depth = ctx.call
        device_scene.set_materials(with_params(materials, p))  # the forward's parameters
        g, _ = device_scene.render_adjoint_params(camera, width, height, spp, max_depth,
                                                  grad_image.detach().to("cpu", torch.float32).numpy())
        return torch.from_numpy(g).to(grad_image.device), None, None, None, None, None, None, None


def render_params(params: torch.Tensor, device_scene, materials, camera, width: int, height: int, spp: int,
                  max_depth: int) -> torch.Tensor:
    """Differentiable render with respect to OpenPBR base colour, roughness
    and metalness (see the module docstring)."""
    return ParamsRender.apply(params, device_scene, materials, camera, width, height, spp, max_depth)
