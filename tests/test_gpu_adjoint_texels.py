"""Path-replay adjoint with respect to OpenPBR base colours and base-colour
texels (tpt_render_adjoint_texels), checked against central finite
differences with common random numbers: exactly (to float rounding) where
the estimator is linear in the parameters (depth 1 with diffuse-only OpenPBR
walls: the lobe choice and pdf do not depend on the base colour, the lookup
is linear in the texels, and the quadratic loss makes central differences
exact), closely where it is a polynomial (depth 3, no Russian roulette),
within 5% relative L2 at depth 6 with diffuse walls, where roulette decisions
move with the base colour but are held fixed in the adjoint, and within 10%
with glossy walls, whose lobe choices move too; materials outside the linear class get zero rows; texel updates
through tpt_scene_set_texture; and the torch.autograd binding."""
import numpy as np
import pytest
import torch

from tpt import camera, scenes
from tpt.autograd import render_texture
from tpt.material import OpenPbr
from tpt.texture import Texture

pytestmark = pytest.mark.gpu

W = H = 32


def textured_box(spec_color, spec_weight=0.0):
    """Cornell geometry; OpenPBR walls (white floor / back / ceiling with
    base-colour textures on the back wall (8 x 8) and floor (6 x 5), red,
    green) with the given specular colour and weight; OpenPBR boxes; the
    OpenPBR emitter."""
    cs = scenes.cornell(build=False)
    s = cs.scene
    rng = np.random.default_rng(3)
    t_back = s.add_texture(Texture(rng.uniform(0.3, 1.0, (8, 8, 3)).astype(np.float32)))
    t_floor = s.add_texture(Texture(rng.uniform(0.3, 1.0, (5, 6, 3)).astype(np.float32)))
    # the white mesh holds floor, ceiling and back wall: give the back wall
    # its own mesh and material (3 quads -> split)
    white = s.meshes[0]
    back_tris = np.array([4, 5])  # quad 2 of the white mesh (floor, ceiling, back)
    from tpt.scene import Mesh
    keep = np.array([0, 1, 2, 3])
    back = Mesh(white.positions, white.indices[back_tris], normals=white.normals, uvs=white.uvs)
    s.meshes[0] = Mesh(white.positions, white.indices[keep], normals=white.normals, uvs=white.uvs)
    bm = s.add_mesh(back)
    s.add_instance(bm, 4, np.eye(4))
    s.build_qbvh()
    mk = lambda c, **kw: OpenPbr.new(c).with_(specular_color=spec_color, specular_weight=spec_weight, **kw)
    mats = [mk((0.8, 0.8, 0.8), base_color_texture=t_floor), mk((0.8, 0.1, 0.1)), mk((0.1, 0.8, 0.1)),
            cs.materials[3], mk((0.9, 0.9, 0.9), base_color_texture=t_back)]
    s.materials[:] = mats
    return cs, mats, [t_back, t_floor]


def with_base(materials, base):
    return [m.with_(base_color=tuple(float(x) for x in base[i])) if i != 3 else m for i, m in enumerate(materials)]


@pytest.fixture(scope="module", params=["diffuse", "glossy"])
def setup(device, request):
    if request.param == "diffuse":
        cs, mats, tex = textured_box((0.0, 0.0, 0.0), 0.0)
    else:
        cs, mats, tex = textured_box((1.0, 1.0, 1.0), 1.0)
    ds = cs.scene.upload(device)
    cam = camera.pinhole(**cs.camera, width=W, height=H)
    base = np.array([m.base_color for m in mats], np.float32)
    yield request.param, cs, mats, ds, cam, base, tex
    ds.close()


def frame(ds, cam, spp, depth):
    """One image through the wavefront path tracer: identical to
    ds.render (checked in test_gpu_wavefront.py) and, at 32 x 32 pixels with
    thousands of samples each, far faster than one thread per pixel."""
    return ds.render_batch([(cam, W, H, spp, depth)])[0].astype(np.float64)


def render(ds, mats, cam, base, spp, depth):
    ds.set_materials(with_base(mats, base))
    return frame(ds, cam, spp, depth)


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


PARAMS = [(i, c) for i in (0, 1, 4) for c in range(3)]


def check(setup, depth, spp, h, tol):
    kind, cs, mats, ds, cam, base, tex = setup
    tgt = base.copy()
    tgt[0] = (0.5, 0.6, 0.7)
    tgt[4] = (0.7, 0.9, 0.4)
    target = render(ds, mats, cam, tgt, spp, depth)
    img = render(ds, mats, cam, base, spp, depth)
    g, tg = ds.render_adjoint_texels(cam, W, H, spp, depth, (2.0 * (img - target)).astype(np.float32))
    fd = np.zeros(len(PARAMS))
    for k, (i, c) in enumerate(PARAMS):
        a = base.copy()
        a[i, c] += h
        lp = ((render(ds, mats, cam, a, spp, depth) - target) ** 2).sum()
        a[i, c] -= 2 * h
        lm = ((render(ds, mats, cam, a, spp, depth) - target) ** 2).sum()
        fd[k] = (lp - lm) / (2 * h)
    ga = np.array([g[i, c] for i, c in PARAMS])
    print(kind, "depth", depth, "params: adjoint", ga, "\n  fd", fd, "rel", rel_l2(ga, fd))
    assert rel_l2(ga, fd) < tol
    assert np.all(g[3] == 0)  # the emitter's base weight is 0
    # texels: the largest adjoint gradients of the back wall's texture
    ds.set_materials(with_base(mats, base))
    px0 = cs.scene.textures[tex[0]].pixels.copy()
    gt = tg[tex[0]]
    flat = np.argsort(-np.abs(gt).ravel())[:8]
    fdt, adt = [], []
    for f in flat:
        y, x, c = np.unravel_index(f, gt.shape)
        px = px0.copy()
        px[y, x, c] += h
        ds.set_texture(tex[0], px)
        lp = ((frame(ds, cam, spp, depth) - target) ** 2).sum()
        px[y, x, c] -= 2 * h
        ds.set_texture(tex[0], px)
        lm = ((frame(ds, cam, spp, depth) - target) ** 2).sum()
        fdt.append((lp - lm) / (2 * h))
        adt.append(gt[y, x, c])
    ds.set_texture(tex[0], px0)
    fdt, adt = np.array(fdt), np.array(adt)
    print("  texels: adjoint", adt, "\n  fd", fdt, "rel", rel_l2(adt, fdt))
    assert rel_l2(adt, fdt) < tol


def test_base_color_and_texel_adjoint_exact_where_linear(setup):
    if setup[0] != "diffuse":
        pytest.skip("linear only for diffuse-only walls")
    check(setup, 1, 64, 0.05, 2e-3)


def test_base_color_and_texel_adjoint_without_roulette(setup):
    """Depth 3 (Russian roulette starts at the fourth vertex): with
    diffuse-only walls the estimator is a polynomial in the parameters
    under common random numbers, so central differences agree closely."""
    if setup[0] != "diffuse":
        pytest.skip("the glossy walls' lobe choice moves with the base colour")
    check(setup, 3, 1024, 1e-2, 5e-3)


def test_base_color_and_texel_adjoint_matches_fd_depth6(setup):
    """Depth 6 with diffuse walls: roulette decisions flip under the
    finite-difference step, which the adjoint holds fixed; a 1e-2 step keeps
    the flips' zero-mean noise in the differences small."""
    if setup[0] != "diffuse":
        pytest.skip("glossy walls: see the depth-3 test")
    check(setup, 6, 4096, 1e-2, 0.05)


def test_glossy_base_color_and_texel_adjoint_matches_fd(setup):
    """Glossy walls (specular weight 1): the lobe probabilities move with the
    base colour, so lobe choices flip under the finite-difference step (the
    adjoint holds them fixed; the estimator of the derivative stays
    unbiased); depth 3 keeps roulette out, and 16,384 samples per pixel with
    a 2e-2 step keep the flips' noise at a few percent."""
    if setup[0] != "glossy":
        pytest.skip("diffuse walls: see the tests above")
    check(setup, 3, 16384, 2e-2, 0.10)


def test_nonlinear_materials_get_zero_rows(device):
    cs, mats, tex = textured_box((1.0, 1.0, 1.0), 1.0)
    mats[1] = mats[1].with_(base_metalness=0.5)
    mats[2] = mats[2].with_(coat_weight=0.5, coat_darkening=1.0)
    cs.scene.materials[:] = mats
    ds = cs.scene.upload(device)
    try:
        cam = camera.pinhole(**cs.camera, width=16, height=16)
        img = ds.render(cam, 16, 16, 16, 4)
        g, _ = ds.render_adjoint_texels(cam, 16, 16, 16, 4, (2.0 * img).astype(np.float32))
        assert np.all(g[1] == 0) and np.all(g[2] == 0)
        assert np.any(g[0] != 0) and np.any(g[4] != 0)
    finally:
        ds.close()


def test_set_texture_rebuilds_the_pyramid_and_rejects_emission(device):
    cs = scenes.cornell_textured()
    ds = cs.scene.upload(device)
    try:
        t = cs.scene.textures[0].pixels
        new = (t * 0.5).astype(np.float32)
        ds.set_texture(0, new)
        assert np.array_equal(ds.texture_level(0, 0), new)
        assert np.allclose(ds.texture_level(0, 1), 0.5 * np.asarray(cs.scene.textures[0].pixels).reshape(
            32, 2, 32, 2, 3).mean(axis=(1, 3)), atol=1e-6)
        import tpt
        with pytest.raises(tpt._lib.TptError, match="emission"):
            ds.set_texture(4, cs.scene.textures[4].pixels)
    finally:
        ds.close()


def test_torch_autograd_texels(setup):
    kind, cs, mats, ds, cam, base, tex = setup
    if kind != "diffuse":
        pytest.skip("one material setup is enough")
    ds.set_materials(with_base(mats, base))
    px0 = cs.scene.textures[tex[0]].pixels
    target = torch.from_numpy(ds.render(cam, W, H, 64, 3) * np.float32(0.8))
    t = torch.tensor(px0, requires_grad=True)
    img = render_texture(t, tex[0], ds, cam, W, H, 64, 3)
    loss = ((img - target) ** 2).sum()
    loss.backward()
    _, ref = ds.render_adjoint_texels(cam, W, H, 64, 3, (2.0 * (img.detach() - target)).numpy())
    assert torch.allclose(t.grad, torch.from_numpy(ref[tex[0]]), rtol=1e-5, atol=1e-7)
    ds.set_texture(tex[0], px0)
