"""Path-replay adjoint with respect to OpenPBR base colour, specular
roughness and metalness (tpt_render_adjoint_params), checked against central
finite differences of the rendered L2 loss with common random numbers.  The
walls are glossy metal / dielectric blends, so the lobe probabilities move
with every parameter and lobe choices flip under the finite-difference step
(the adjoint holds them fixed: its estimator stays unbiased); depth 3 keeps
Russian roulette out, and many samples keep the flips' noise at a few
percent.  Also: rows of the albedo-only entry points are unchanged, and a
textured roughness slot gets no gradient."""
import numpy as np
import pytest

from tpt import camera, scenes

pytestmark = pytest.mark.gpu

W = H = 64
PARAMS = [(0, 3), (1, 3), (2, 3), (0, 4), (1, 4), (2, 4), (2, 0), (2, 1), (2, 2)]


def materials(theta):
    """theta: (3, 5) rows of base colour RGB, roughness, metalness for white, red, green."""
    cs = scenes.cornell(specular_weight=1.0)
    mats = list(cs.materials)
    for i in range(3):
        mats[i] = mats[i].with_(base_color=tuple(float(x) for x in theta[i, :3]),
                                specular_roughness=float(theta[i, 3]), base_metalness=float(theta[i, 4]))
    return cs, mats


BASE = np.array([[0.8, 0.8, 0.8, 0.5, 0.4], [0.8, 0.1, 0.1, 0.35, 0.2], [0.1, 0.8, 0.1, 0.6, 0.7]], np.float64)
TARGET = np.array([[0.7, 0.75, 0.8, 0.35, 0.6], [0.9, 0.2, 0.1, 0.5, 0.1], [0.2, 0.7, 0.3, 0.45, 0.5]], np.float64)


@pytest.fixture(scope="module")
def setup(device):
    cs, mats = materials(BASE)
    ds = cs.scene.upload(device)
    cam = camera.pinhole(**cs.camera, width=W, height=H)
    yield cs, ds, cam
    ds.close()


def render(ds, theta, cam, spp, depth):
    _, mats = materials(theta)
    ds.set_materials(mats)
    return ds.render_batch([(cam, W, H, spp, depth)])[0].astype(np.float64)


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def test_roughness_metalness_and_base_colour_match_fd(setup):
    cs, ds, cam = setup
    # 16,384 samples per pixel: the finite differences' lobe-flip noise
    # falls to about 1.5% relative L2 here (2,048 samples: about 12%)
    spp, depth, h = 16384, 3, 2e-2
    target = render(ds, TARGET, cam, spp, depth)
    img = render(ds, BASE, cam, spp, depth)
    g, _ = ds.render_adjoint_params(cam, W, H, spp, depth, (2.0 * (img - target)).astype(np.float32))
    assert g.shape == (len(cs.materials), 5)
    fd = np.zeros(len(PARAMS))
    for k, (i, c) in enumerate(PARAMS):
        t = BASE.copy()
        t[i, c] += h
        lp = ((render(ds, t, cam, spp, depth) - target) ** 2).sum()
        t[i, c] -= 2 * h
        lm = ((render(ds, t, cam, spp, depth) - target) ** 2).sum()
        fd[k] = (lp - lm) / (2 * h)
    ga = np.array([g[i, c] for i, c in PARAMS])
    print("adjoint", ga, "\nfd     ", fd, "\nrel L2", rel_l2(ga, fd))
    assert np.all(g[3] == 0)  # the light
    assert rel_l2(ga, fd) < 0.05


def test_albedo_entry_point_rows_unchanged(setup):
    """The base-colour rows of materials whose BSDF is affine in the base
    colour (metalness 0) are the exact ratios of tpt_render_adjoint."""
    cs, ds, cam = setup
    theta = BASE.copy()
    theta[:, 4] = 0.0
    img = render(ds, theta, cam, 64, 3)
    dl = (2.0 * (img - 0.5 * img)).astype(np.float32)
    g5, _ = ds.render_adjoint_params(cam, W, H, 64, 3, dl)
    g3 = ds.render_adjoint(cam, W, H, 64, 3, dl)
    assert np.allclose(g5[:, :3], g3, rtol=1e-5, atol=1e-7 * np.abs(g3).max())
    assert np.any(g5[:3, 3] != 0) and np.any(g5[:3, 4] != 0)


def test_torch_autograd_params(setup):
    import torch

    from tpt.autograd import render_params

    cs, ds, cam = setup
    _, mats = materials(BASE)
    rows = np.zeros((len(mats), 5), np.float32)
    rows[:3] = BASE
    target = torch.from_numpy(render(ds, TARGET, cam, 64, 3).astype(np.float32))
    p = torch.tensor(rows, requires_grad=True)
    img = render_params(p, ds, mats, cam, W, H, 64, 3)
    loss = ((img - target) ** 2).sum()
    loss.backward()
    ds.set_materials(mats)
    want, _ = ds.render_adjoint_params(cam, W, H, 64, 3, (2.0 * (img.detach().numpy() - target.numpy())))
    assert np.allclose(p.grad.numpy(), want, rtol=1e-5, atol=1e-6 * np.abs(want).max())
    assert np.all(p.grad.numpy()[3] == 0)


def test_smooth_lobes_match_fd(device):
    """Roughness-0 walls: a smooth metal (metalness 0.6) and a smooth
    dielectric (metalness 0.3), whose BSDF samples are mirror (delta)
    samples.  Such a sample's weight is the lobe's layer weight times its
    Fresnel term over the lobe probability, so the metalness and base
    colour reach everything collected after the mirror bounce; the adjoint
    differentiates that numerator with the lobe probability held fixed.
    Checked against central differences (common random numbers)."""
    cs = scenes.cornell(specular_weight=1.0)
    mats = list(cs.materials)
    theta = np.array([[0.8, 0.7, 0.6, 0.0, 0.6], [0.8, 0.2, 0.1, 0.0, 0.3], [0.1, 0.8, 0.1, 0.4, 0.2]], np.float64)

    def with_theta(t):
        out = list(mats)
        for i in range(3):
            out[i] = out[i].with_(base_color=tuple(float(x) for x in t[i, :3]), specular_roughness=float(t[i, 3]),
                                  base_metalness=float(t[i, 4]))
        return out

    ds = cs.scene.upload(device)
    try:
        cam = camera.pinhole(**cs.camera, width=W, height=H)
        spp, depth, h = 16384, 3, 1e-2

        def img(t):
            ds.set_materials(with_theta(t))
            return ds.render_batch([(cam, W, H, spp, depth)])[0].astype(np.float64)

        target = img(theta + np.array([[0.05, -0.05, 0.05, 0.0, -0.1], [-0.1, 0.1, 0.05, 0.0, 0.1],
                                       [0.05, 0.0, 0.05, 0.1, 0.1]]))
        base = img(theta)
        ds.set_materials(with_theta(theta))
        g, _ = ds.render_adjoint_params(cam, W, H, spp, depth, (2.0 * (base - target)).astype(np.float32))
        picks = [(0, 4), (0, 0), (0, 1), (0, 2), (1, 4), (1, 0), (1, 1)]
        fd = []
        for i, c in picks:
            t = theta.copy()
            t[i, c] += h
            lp = ((img(t) - target) ** 2).sum()
            t[i, c] -= 2 * h
            lm = ((img(t) - target) ** 2).sum()
            fd.append((lp - lm) / (2 * h))
        ga = np.array([g[i, c] for i, c in picks])
        fd = np.array(fd)
        print("smooth lobes: adjoint", ga, "\n  fd", fd, "\n  rel L2", rel_l2(ga, fd), flush=True)
        assert rel_l2(ga, fd) < 0.05
    finally:
        ds.close()
