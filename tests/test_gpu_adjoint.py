"""Path-replay adjoint (tpt_render_adjoint) of an L2 image loss with respect
to the NormalizedLambert albedos, checked against central finite differences
with common random numbers (the same sample sequence at theta +- h): exactly
(to float rounding) where the estimator is linear in the albedos (depth 1,
the loss is then quadratic and central differences are exact), within 5%
relative L2 at depth 6 where Russian roulette decisions are frozen in the
adjoint; against a CPU finite-difference gradient of the oracle renderer
with the loss summed in float64; and through torch.autograd."""
import numpy as np
import pytest
import torch

import oracle
from tpt import camera, scenes
from tpt.autograd import render_albedo, with_albedo
from tpt.material import NormalizedLambert

pytestmark = pytest.mark.gpu

W = H = 32


@pytest.fixture(scope="module")
def setup(device):
    cs = scenes.cornell_lambert()
    ds = cs.scene.upload(device)
    cam = camera.pinhole(**cs.camera, width=W, height=H)
    base = np.array([m.rho if isinstance(m, NormalizedLambert) else (0, 0, 0) for m in cs.materials], np.float32)
    target_albedo = base.copy()
    target_albedo[0] = (0.5, 0.6, 0.7)
    target_albedo[1] = (0.9, 0.2, 0.3)
    yield cs, ds, cam, base, target_albedo
    ds.close()


def render(ds, cs, cam, albedo, spp, depth):
    ds.set_materials(with_albedo(cs.materials, albedo))
    return ds.render(cam, W, H, spp, depth).astype(np.float64)


def loss_and_grad(ds, cs, cam, albedo, target, spp, depth, fp64=False):
    img = render(ds, cs, cam, albedo, spp, depth)
    r = img - target
    g = ds.render_adjoint(cam, W, H, spp, depth, (2.0 * r).astype(np.float32), fp64=fp64)
    return float((r * r).sum()), g


def fd_grad(ds, cs, cam, albedo, target, spp, depth, h, params):
    out = np.zeros_like(albedo, dtype=np.float64)
    for (i, c) in params:
        a = albedo.copy()
        a[i, c] += h
        lp = ((render(ds, cs, cam, a, spp, depth) - target) ** 2).sum()
        a[i, c] -= 2 * h
        lm = ((render(ds, cs, cam, a, spp, depth) - target) ** 2).sum()
        out[i, c] = (lp - lm) / (2 * h)
    return out


PARAMS = [(i, c) for i in range(3) for c in range(3)]


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def test_adjoint_exact_where_linear(setup):
    cs, ds, cam, base, tgt = setup
    target = render(ds, cs, cam, tgt, 64, 1)
    _, g = loss_and_grad(ds, cs, cam, base, target, 64, 1)
    fd = fd_grad(ds, cs, cam, base, target, 64, 1, 0.05, PARAMS)
    sel = np.array([g[i, c] for i, c in PARAMS]), np.array([fd[i, c] for i, c in PARAMS])
    print("depth 1: adjoint", sel[0], "\n  fd", sel[1], "rel", rel_l2(*sel))
    assert rel_l2(*sel) < 2e-3
    assert np.all(g[3] == 0)  # the Emissive light has no albedo


def test_adjoint_matches_finite_differences_depth6(setup):
    cs, ds, cam, base, tgt = setup
    spp, depth = 4096, 6
    target = render(ds, cs, cam, tgt, 1024, depth)
    _, g = loss_and_grad(ds, cs, cam, base, target, spp, depth)
    fd = fd_grad(ds, cs, cam, base, target, spp, depth, 1e-3, PARAMS)
    a, b = np.array([g[i, c] for i, c in PARAMS]), np.array([fd[i, c] for i, c in PARAMS])
    print("depth 6: adjoint", a, "\n  fd", b, "rel L2", rel_l2(a, b))
    assert rel_l2(a, b) < 0.05


def test_fp32_block_sums_against_fp64(setup):
    cs, ds, cam, base, tgt = setup
    target = render(ds, cs, cam, tgt, 64, 6)
    _, g32 = loss_and_grad(ds, cs, cam, base, target, 256, 6)
    _, g64 = loss_and_grad(ds, cs, cam, base, target, 256, 6, fp64=True)
    err = np.abs(g32 - g64).max() / np.abs(g64).max()
    print("FP32 vs FP64 per-block sums: max error relative to the largest gradient", err)
    assert err < 1e-4


def test_adjoint_against_cpu_fd_reference(setup):
    """Central differences of the oracle renderer (CPU, loss summed in
    float64) on a smaller frame, against the GPU adjoint of the same frame."""
    cs, ds, cam, base, tgt = setup
    w = h = 16
    spp, depth = 1024, 6
    cam16 = camera.pinhole(**cs.camera, width=w, height=h)

    def cpu_render(albedo):
        mats = with_albedo(cs.materials, albedo)
        cs.scene.materials[:] = mats
        return oracle.OracleRenderer(cs.scene, mats).render(cam16, w, h, spp, depth, threads=16).astype(np.float64)

    target = cpu_render(tgt)
    ds.set_materials(with_albedo(cs.materials, base))
    img = ds.render(cam16, w, h, spp, depth).astype(np.float64)
    g = ds.render_adjoint(cam16, w, h, spp, depth, (2.0 * (img - target)).astype(np.float32))
    cpu = np.zeros(9)
    for k, (i, c) in enumerate(PARAMS):
        a = base.copy()
        a[i, c] += 1e-3
        lp = ((cpu_render(a) - target) ** 2).sum()
        a[i, c] -= 2e-3
        lm = ((cpu_render(a) - target) ** 2).sum()
        cpu[k] = (lp - lm) / 2e-3
    cs.scene.materials[:] = cs.materials
    gg = np.array([g[i, c] for i, c in PARAMS])
    print("adjoint", gg, "\n  cpu fd", cpu, "rel L2", rel_l2(gg, cpu))
    assert rel_l2(gg, cpu) < 0.05


def test_torch_autograd(setup):
    cs, ds, cam, base, tgt = setup
    target = torch.from_numpy(render(ds, cs, cam, tgt, 64, 4).astype(np.float32))
    a = torch.tensor(base, requires_grad=True)
    img = render_albedo(a, ds, cs.materials, cam, W, H, 64, 4)
    loss = ((img - target) ** 2).sum()
    loss.backward()
    ds.set_materials(with_albedo(cs.materials, base))
    ref = ds.render_adjoint(cam, W, H, 64, 4, (2.0 * (img.detach() - target)).numpy())
    assert torch.allclose(a.grad, torch.from_numpy(ref), rtol=1e-6, atol=0)
