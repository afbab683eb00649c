"""torch.autograd binding of the path-replay adjoint (tpt_render_adjoint,
tpt_render_adjoint_texels, tpt_render_adjoint_params).

``render_albedo(albedo, scene, ...)`` renders the scene with the
NormalizedLambert albedos taken from ``albedo`` (one RGB row per material;
rows of other materials are ignored and get zero gradient) and returns the
image as a tensor whose backward pass runs the adjoint kernel:
``loss(render_albedo(a, ...)).backward()`` fills ``a.grad`` with
dLoss/d(albedo).  ``render_texture(texels, index, ...)`` does the same for the
level-0 texels of one OpenPBR base-colour texture.  ``render_params(params,
scene, ...)`` takes one row of five OpenPBR parameters per material (base
colour RGB, specular_roughness, base_metalness; rows of other materials are
ignored and get zero gradient) and differentiates through
tpt_render_adjoint_params.  ``render_views(params, texels, scene, materials,
views)`` renders many views of uneven resolution in one wavefront batch and
differentiates all of them in one batched adjoint launch
(tpt_render_adjoint_batch), with respect to the parameter rows and every
texture's texels.

The image, the incoming image gradient and the parameter / texel gradients
stay in device memory: the library's calls are enqueued on torch's current
stream with device pointers (no host copy of an image or a gradient).  Only
the material table itself (a few hundred bytes per material) goes through
host memory, because tpt_scene_set_materials takes host records.
"""
from __future__ import annotations

import numpy as np
import torch

from .material import NormalizedLambert, OpenPbr


def _dev(device_scene) -> torch.device:
    return torch.device("cuda", torch.cuda.current_device())


def _bind_stream(device_scene) -> None:
    """Enqueue the library's work on torch's current stream, so the device
    buffers it reads and writes are ordered with torch's own kernels."""
    device_scene.device.set_stream(torch.cuda.current_stream().cuda_stream)


def _grad_image(grad_image: torch.Tensor) -> torch.Tensor:
    return grad_image.detach().to(_dev(None), torch.float32).contiguous()


def with_albedo(materials, albedo: np.ndarray) -> list:
    """The material list with every NormalizedLambert albedo replaced by the
    matching row of ``albedo`` (shape (len(materials), 3))."""
    out = []
    for i, m in enumerate(materials):
        out.append(NormalizedLambert(tuple(float(x) for x in albedo[i])) if isinstance(m, NormalizedLambert) else m)
    return out


def _row_mask(materials, kind, cols: int, device) -> torch.Tensor:
    """1 for the rows of materials of the given class, 0 elsewhere."""
    m = torch.tensor([1.0 if isinstance(x, kind) else 0.0 for x in materials], dtype=torch.float32)
    return m.to(device)[:, None].expand(len(materials), cols)


def _render(device_scene, camera, width, height, spp, max_depth) -> torch.Tensor:
    _bind_stream(device_scene)
    img = torch.empty((height, width, 3), dtype=torch.float32, device=_dev(device_scene))
    device_scene.render_device(camera, width, height, spp, max_depth, img)
    return img


class AlbedoRender(torch.autograd.Function):
    @staticmethod
    def forward(ctx, albedo, device_scene, materials, camera, width, height, spp, max_depth):
        a = albedo.detach().to("cpu", torch.float32).numpy()
        _bind_stream(device_scene)
        device_scene.set_materials(with_albedo(materials, a))
        img = _render(device_scene, camera, width, height, spp, max_depth)
        ctx.call = (device_scene, materials, a, camera, width, height, spp, max_depth)
        return img.to(albedo.device)

    @staticmethod
    def backward(ctx, grad_image):
        device_scene, materials, a, camera, width, height, spp, max_depth = ctx.call
        _bind_stream(device_scene)
        device_scene.set_materials(with_albedo(materials, a))  # the forward's parameters
        g_img = _grad_image(grad_image)
        g = torch.zeros((len(materials), 3), dtype=torch.float32, device=g_img.device)
        device_scene.render_adjoint_device(camera, width, height, spp, max_depth, g_img, g)
        # the forward reads only NormalizedLambert rows (tpt_render_adjoint
        # also fills OpenPBR base-colour rows)
        g = g * _row_mask(materials, NormalizedLambert, 3, g.device)
        return g.to(grad_image.device), None, None, None, None, None, None, None


def render_albedo(albedo: torch.Tensor, device_scene, materials, camera, width: int, height: int, spp: int,
                  max_depth: int) -> torch.Tensor:
    """Differentiable render (see the module docstring)."""
    return AlbedoRender.apply(albedo, device_scene, materials, camera, width, height, spp, max_depth)


def _texel_offsets(device_scene) -> list:
    offs, o = [], 0
    for shape in device_scene.texture_shapes:
        offs.append(o)
        o += int(np.prod(shape))
    return offs + [o]


class TextureRender(torch.autograd.Function):
    @staticmethod
    def forward(ctx, texels, texture, device_scene, camera, width, height, spp, max_depth):
        t = texels.detach().to(_dev(device_scene), torch.float32).contiguous()
        _bind_stream(device_scene)
        device_scene.set_texture_device(texture, t)
        img = _render(device_scene, camera, width, height, spp, max_depth)
        ctx.call = (device_scene, texture, t, camera, width, height, spp, max_depth)
        return img.to(texels.device)

    @staticmethod
    def backward(ctx, grad_image):
        device_scene, texture, t, camera, width, height, spp, max_depth = ctx.call
        _bind_stream(device_scene)
        device_scene.set_texture_device(texture, t)  # the forward's texels
        g_img = _grad_image(grad_image)
        offs = _texel_offsets(device_scene)
        d_alb = torch.zeros((device_scene.material_count, 3), dtype=torch.float32, device=g_img.device)
        d_tex = torch.zeros(max(1, offs[-1]), dtype=torch.float32, device=g_img.device)
        device_scene.render_adjoint_texels_device(camera, width, height, spp, max_depth, g_img, d_alb, d_tex)
        g = d_tex[offs[texture]:offs[texture + 1]].reshape(device_scene.texture_shapes[texture])
        return g.to(grad_image.device), None, None, None, None, None, None, None


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
        _bind_stream(device_scene)
        device_scene.set_materials(with_params(materials, p))
        img = _render(device_scene, camera, width, height, spp, max_depth)
        ctx.call = (device_scene, materials, p, camera, width, height, spp, max_depth)
        return img.to(params.device)

    @staticmethod
    def backward(ctx, grad_image):
        device_scene, materials, p, camera, width, height, spp, max_depth = ctx.call
        _bind_stream(device_scene)
        device_scene.set_materials(with_params(materials, p))  # the forward's parameters
        g_img = _grad_image(grad_image)
        g = torch.zeros((len(materials), 5), dtype=torch.float32, device=g_img.device)
        device_scene.render_adjoint_params_device(camera, width, height, spp, max_depth, g_img, g)
        # the forward reads only OpenPBR rows (tpt_render_adjoint_params also
        # fills NormalizedLambert albedo rows)
        g = g * _row_mask(materials, OpenPbr, 5, g.device)
        return g.to(grad_image.device), None, None, None, None, None, None, None


def render_params(params: torch.Tensor, device_scene, materials, camera, width: int, height: int, spp: int,
                  max_depth: int) -> torch.Tensor:
    """Differentiable render with respect to OpenPBR base colour, roughness
    and metalness (see the module docstring)."""
    return ParamsRender.apply(params, device_scene, materials, camera, width, height, spp, max_depth)


class ViewsRender(torch.autograd.Function):
    @staticmethod
    def forward(ctx, params, texels, device_scene, materials, views):
        p = params.detach().to("cpu", torch.float32).numpy()
        _bind_stream(device_scene)
        device_scene.set_materials(with_params(materials, p))
        offs = _texel_offsets(device_scene)
        tx = None
        if texels is not None:
            tx = texels.detach().to(_dev(device_scene), torch.float32).contiguous()
            for t in range(len(device_scene.texture_shapes)):
                device_scene.set_texture_device(t, tx[offs[t]:offs[t + 1]].view(device_scene.texture_shapes[t]))
        outs = [torch.empty((h, w, 3), dtype=torch.float32, device=_dev(device_scene)) for _, w, h, _, _ in views]
        device_scene.render_batch_device(list(views), outs)
        ctx.call = (device_scene, materials, p, tx, views, texels is not None)
        return tuple(outs)

    @staticmethod
    def backward(ctx, *grad_images):
        device_scene, materials, p, tx, views, has_tex = ctx.call
        _bind_stream(device_scene)
        device_scene.set_materials(with_params(materials, p))  # the forward's parameters
        offs = _texel_offsets(device_scene)
        if tx is not None:
            for t in range(len(device_scene.texture_shapes)):
                device_scene.set_texture_device(t, tx[offs[t]:offs[t + 1]].view(device_scene.texture_shapes[t]))
        dev = _dev(device_scene)
        gi = [(g if g is not None else torch.zeros((h, w, 3), device=dev)).detach().to(dev, torch.float32).contiguous()
              for g, (_, w, h, _, _) in zip(grad_images, views)]
        adj_views = [(cam, w, h, spp, depth, g) for (cam, w, h, spp, depth), g in zip(views, gi)]
        d_p = torch.zeros((len(materials), 5), dtype=torch.float32, device=dev)
        d_t = torch.zeros(max(1, offs[-1]), dtype=torch.float32, device=dev) if has_tex else None
        device_scene.render_adjoint_batch_device(adj_views, d_p, d_t)
        d_p = d_p * _row_mask(materials, OpenPbr, 5, dev)
        return d_p, (d_t[:offs[-1]] if has_tex else None), None, None, None


def render_views(params: torch.Tensor, texels, device_scene, materials, views):
    """Differentiable batched render of many views in one launch: ``views``
    are (camera, width, height, spp, max_depth) tuples; ``params`` one row of
    five OpenPBR parameters per material (as render_params); ``texels`` the
    level-0 texels of every texture of the scene concatenated (flat, in the
    scene's texture order) or None.  The forward is one wavefront batch
    (tpt_render_batch); the backward one batched path-replay adjoint over
    every view (tpt_render_adjoint_batch).  Returns one (height, width, 3)
    image per view."""
    return ViewsRender.apply(params, texels, device_scene, materials, tuple(views))
