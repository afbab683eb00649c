"""Batched multi-view adjoint (tpt_render_adjoint_batch) on a small
instance of the batched inverse-rendering city (scenes.city_inverse: 64
learnable OpenPBR materials, four base-colour textures, the sky and the
lamps; here a 10 x 10 tower grid, 64^2 textures and four views of uneven
size):

* the batch's gradient is the sum over the views of each view's
  tpt_render_adjoint_params gradient (FP64 accumulation on both sides);
* the FP32 accumulation (per-sample LDS epochs folded into compensated
  per-block sums) agrees with the FP64 one (the error is stated);
* against central differences of the summed L2 loss over all views (common
  random numbers): base-colour and metalness entries at h = 1e-2 and texel
  patches at h = 1e-3 within 5 % relative L2, roughness entries to the
  adjoint's sign (a parameter change moves the lobe-choice boundaries and
  needs a step at which the flipped samples are resolved; roughness also
  moves the sampled directions across visibility edges; see
  test_gpu_inverse_cornell);
* the torch binding (tpt.autograd.render_views) returns the same
  gradients."""
import numpy as np
import pytest
import torch

from tpt import camera, scenes
from tpt.autograd import render_views, with_params

pytestmark = pytest.mark.gpu

DEPTH = 8


@pytest.fixture(scope="module")
def city(device):
    c = scenes.city_inverse(grid=10, rocks=200, lights=32, tex_size=64, views=4, view_sizes=(24, 48),
                            sky_size=(64, 32))
    ds = c.scene.upload(device)
    yield c, ds
    ds.close()


def theta_of(mats, n):
    t = np.zeros((len(mats), 5))
    for i, m in enumerate(mats[:n]):
        t[i, :3] = m.base_color
        t[i, 3] = m.specular_roughness
        t[i, 4] = m.base_metalness
    return t


def frames(c, spp):
    return [(camera.pinhole(**v, width=w, height=h), w, h, spp, DEPTH) for v, w, h in c.views]


def test_batch_equals_sum_of_views_and_fp32_matches_fp64(city):
    c, ds = city
    fr = frames(c, 16)
    imgs = ds.render_batch(fr)
    rng = np.random.default_rng(0)
    dls = [rng.normal(size=im.shape).astype(np.float32) for im in imgs]
    views = [f + (dl,) for f, dl in zip(fr, dls)]
    g64, t64 = ds.render_adjoint_batch(views, fp64=True)
    sum_p = np.zeros_like(g64, dtype=np.float64)
    sum_t = [np.zeros(t.shape) for t in t64]
    for (cam, w, h, spp, depth), dl in zip(fr, dls):
        g, tg = ds.render_adjoint_params(cam, w, h, spp, depth, dl, fp64=True)
        sum_p += g
        for a, b in zip(sum_t, tg):
            a += b
    scale = np.abs(sum_p).max()
    assert scale > 0 and np.all(g64[-1] == 0)  # the emitter row
    assert np.allclose(g64, sum_p, rtol=0, atol=1e-5 * scale)
    for a, b in zip(t64, sum_t):
        assert np.allclose(a, b, rtol=0, atol=1e-5 * max(np.abs(b).max(), 1e-12))
    g32, t32 = ds.render_adjoint_batch(views)
    err = np.abs(g32 - g64).max() / scale
    terr = max(np.abs(a - b).max() / max(np.abs(b).max(), 1e-12) for a, b in zip(t32, t64))
    print(f"FP32 compensated vs FP64: parameters {err:.2e}, texels {terr:.2e} of the largest gradient")
    assert err < 1e-5 and terr < 1e-4


def test_batch_matches_central_differences(city):
    c, ds = city
    spp = 2048
    fr = frames(c, spp)
    theta = theta_of(c.materials, c.learnable)
    tgt = theta.copy()
    tgt[:c.learnable, :3] = np.clip(tgt[:c.learnable, :3] + 0.1, 0, 1)
    ds.set_materials(with_params(c.materials, tgt))
    targets = [im.astype(np.float64) for im in ds.render_batch(fr)]

    def loss(t, texels=None):
        ds.set_materials(with_params(c.materials, t))
        if texels is not None:
            ds.set_texture(*texels)
        return sum(float(((im.astype(np.float64) - T) ** 2).sum()) for im, T in zip(ds.render_batch(fr), targets))

    ds.set_materials(with_params(c.materials, theta))
    imgs = ds.render_batch(fr)
    views = [f + ((2.0 * (im - T)).astype(np.float32),) for f, im, T in zip(fr, imgs, targets)]
    g, tg = ds.render_adjoint_batch(views)
    # the base-colour and metalness entries with the largest gradients, and
    # the largest roughness entries (roughness moves the sampled directions:
    # its differences jump at visibility edges, so those are held to the
    # adjoint's sign, as in test_gpu_inverse_cornell)
    gl = g[:c.learnable].copy()
    gl[:, 3] = 0.0
    picks = [tuple(int(v) for v in np.unravel_index(k, gl.shape)) for k in np.argsort(-np.abs(gl).ravel())[:5]]
    picks += [(int(i), 3) for i in np.argsort(-np.abs(g[:c.learnable, 3]))[:2]]
    h = 1e-2
    fd, ad = [], []
    for i, col in picks:
        t = theta.copy()
        t[i, col] += h
        lp = loss(t)
        t[i, col] -= 2 * h
        lm = loss(t)
        fd.append((lp - lm) / (2 * h))
        ad.append(float(g[i, col]))
        print(f"  parameter ({i}, {col}): adjoint {ad[-1]:.5g} fd {fd[-1]:.5g}", flush=True)
    ds.set_materials(with_params(c.materials, theta))
    # texel patches of the texture with the largest gradient
    ti = int(np.argmax([np.abs(t).max() for t in tg]))
    base = c.scene.textures[ti].pixels.copy()
    y, x, ch = np.unravel_index(int(np.argmax(np.abs(tg[ti]))), tg[ti].shape)
    fdt, adt = [], []
    for dy, dx in ((0, 0), (8, 0), (0, 8)):
        ys = slice(max(0, y + dy - 4), y + dy + 4)
        xs = slice(max(0, x + dx - 4), x + dx + 4)
        p = base.copy()
        p[ys, xs, ch] += 1e-3
        lp = loss(theta, (ti, p))
        p[ys, xs, ch] -= 2e-3
        lm = loss(theta, (ti, p))
        fdt.append((lp - lm) / 2e-3)
        adt.append(float(tg[ti][ys, xs, ch].sum()))
        print(f"  texel patch ({y + dy}, {x + dx}, {ch}): adjoint {adt[-1]:.5g} fd {fdt[-1]:.5g}", flush=True)
    ds.set_texture(ti, base)
    rel = lambda a, b: float(np.linalg.norm(np.array(a) - np.array(b)) / np.linalg.norm(b))
    print(f"batched adjoint vs central differences: base colour + metalness {rel(ad[:5], fd[:5]):.4f}, "
          f"roughness {rel(ad[5:], fd[5:]):.4f}, texel patches {rel(adt, fdt):.4f}")
    assert rel(ad[:5], fd[:5]) < 0.05 and rel(adt, fdt) < 0.05
    assert np.all(np.sign(ad[5:]) == np.sign(fd[5:]))


def test_torch_render_views(city, device):
    c, ds = city
    fr = frames(c, 8)
    theta = theta_of(c.materials, c.learnable)
    texels = torch.cat([torch.from_numpy(t.pixels.reshape(-1)) for t in c.scene.textures]).cuda()
    params = torch.tensor(theta, dtype=torch.float32, device="cuda", requires_grad=True)
    tx = texels.clone().requires_grad_(True)
    imgs = render_views(params, tx, ds, c.materials, fr)
    assert [tuple(i.shape) for i in imgs] == [(h, w, 3) for _, w, h, _, _ in fr]
    loss = sum((im ** 2).sum() for im in imgs)
    loss.backward()
    torch.cuda.synchronize()
    ds.set_materials(with_params(c.materials, theta.astype(np.float32)))
    views = [f + ((2.0 * im.detach().cpu().numpy()).astype(np.float32),) for f, im in zip(fr, imgs)]
    want, want_t = ds.render_adjoint_batch(views)
    got = params.grad.cpu().numpy()
    # every OpenPBR row, the emitter's too (its metalness scales a metal lobe
    # whose weight is linear in it, so it reaches the image)
    assert np.allclose(got, want, rtol=1e-5, atol=1e-6 * np.abs(want).max())
    wt = np.concatenate([t.ravel() for t in want_t])
    assert np.allclose(tx.grad.cpu().numpy(), wt, rtol=1e-5, atol=1e-6 * np.abs(wt).max())
