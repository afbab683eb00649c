"""Environment light on the GPU (R/light/environment.rs): the outdoor scene
(boxes on a floor under a procedural sky with an analytic sun) with OpenPBR
and with NormalizedLambert materials, against the oracle (same estimator; the
direction <-> lat-long mapping goes through acos / atan2 / sin / cos, so
values agree to float rounding), relMSE against a high-sample image, the
wavefront path tracer against the megakernel, and the adjoint against
finite differences with the sky lighting the scene."""
import numpy as np
import pytest

import golden_images
import oracle
from tpt import camera, scenes
from tpt.autograd import with_albedo

pytestmark = pytest.mark.gpu


def relmse(x, ref):
    return float(np.mean((x - ref) ** 2 / (ref * ref + 1e-2)))


@pytest.fixture(scope="module", params=["openpbr", "lambert"])
def outdoor(request, device):
    cs = scenes.outdoor(request.param)
    ds = cs.scene.upload(device)
    yield cs, ds
    ds.close()


def test_env_render_matches_oracle(outdoor):
    cs, ds = outdoor
    w = h = 32
    cam = camera.pinhole(**cs.camera, width=w, height=h)
    gpu = ds.render(cam, w, h, 8, 6)
    cpu = oracle.OracleRenderer(cs.scene, cs.materials).render(cam, w, h, 8, 6)
    rel = np.abs(gpu - cpu) / np.maximum(np.abs(cpu), 1e-3)
    print("mean rel diff", rel.mean(), "identical", np.mean(gpu == cpu))
    assert np.all(np.isfinite(gpu))
    assert rel.mean() < 2e-3 and abs(gpu.mean() - cpu.mean()) <= 2e-3 * cpu.mean()


def test_env_relmse_no_worse_than_cpu(outdoor):
    cs, ds = outdoor
    name = "outdoor_openpbr" if cs.materials[0].__class__.__name__ == "OpenPbr" else "outdoor_lambert"
    e_gpu, e_cpu = golden_images.gpu_and_cpu_relmse(name, cs, ds, 32)
    print(f"{name} relMSE at 32 spp against the CPU reference: gpu {e_gpu:.5f} cpu {e_cpu:.5f}")
    assert e_gpu <= 1.1 * e_cpu + 1e-6


def test_env_wavefront_matches_megakernel(outdoor):
    cs, ds = outdoor
    w, h = 40, 28
    cam = camera.pinhole(**cs.camera, width=w, height=h)
    mega = ds.render(cam, w, h, 16, 6)
    (wave,) = ds.render_batch([(cam, w, h, 16, 6)], exact_shadow=True)
    assert np.array_equal(mega.view(np.uint32), wave.view(np.uint32))


def test_env_adjoint_matches_finite_differences(device):
    cs = scenes.outdoor("lambert")
    ds = cs.scene.upload(device)
    w = h = 24
    spp, depth = 2048, 4
    cam = camera.pinhole(**cs.camera, width=w, height=h)
    base = np.array([m.rho for m in cs.materials], np.float32)
    tgt = base.copy()
    tgt[1] = (0.3, 0.6, 0.2)

    def render(a, n=spp):
        ds.set_materials(with_albedo(cs.materials, a))
        return ds.render(cam, w, h, n, depth).astype(np.float64)

    target = render(tgt, 512)
    img = render(base)
    g = ds.render_adjoint(cam, w, h, spp, depth, (2.0 * (img - target)).astype(np.float32))
    fd = np.zeros_like(g, dtype=np.float64)
    for i in range(3):
        for c in range(3):
            a = base.copy()
            a[i, c] += 1e-3
            lp = ((render(a) - target) ** 2).sum()
            a[i, c] -= 2e-3
            lm = ((render(a) - target) ** 2).sum()
            fd[i, c] = (lp - lm) / 2e-3
    rel = np.linalg.norm(g - fd) / np.linalg.norm(fd)
    print("env adjoint", g.ravel(), "\n  fd", fd.ravel(), "rel L2", rel)
    ds.close()
    assert rel < 0.05
