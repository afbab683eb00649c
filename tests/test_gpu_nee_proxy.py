"""The OpenPBR proxy NEE policy (tpt_scene_set_nee_policy,
TPT_NEE_OPENPBR_PROXY) on the device:

* the GPU image under the proxy policy against the oracle's under the same
  policy (same estimator and sample sequence; the node importances go
  through exp / erf / sqrt of the device maths library vs glibc, so the
  images agree to float rounding, not bit for bit), on the glossy Cornell box
  and on the textured one (the query reads the textured base colour);
* the wavefront path tracer (which stores the query of the previous vertex
  for the reverse pmf) against the megakernel (which keeps it in registers):
  bit for bit with the exact shadow rule, on the glossy and textured boxes
  and on a small city;
* unbiasedness: against the oracle's 16,384-spp image of the reference
  policy the GPU proxy image has the CPU proxy image's relMSE (the reference
  policy's image is the same expectation), and at 4,096 spp its mean is the
  reference's;
* the variance gain at equal samples on the glossy box and on a night city
  lit by its lamps alone (every city material is OpenPBR, so under the
  reference policy the lamps are reached by BSDF sampling only);
* the batched adjoint under the proxy policy against central differences of
  the loss rendered under the same policy.
"""
import numpy as np
import pytest

import golden_images
import oracle
from tpt import camera, scenes
from tpt.autograd import with_params
from tpt.imageio import relmse

pytestmark = pytest.mark.gpu


def bits_equal(a, b):
    return np.array_equal(np.ascontiguousarray(a).view(np.uint32), np.ascontiguousarray(b).view(np.uint32))


@pytest.fixture(scope="module", params=["glossy", "textured"])
def box(device, request):
    cs = scenes.cornell(specular_weight=1.0) if request.param == "glossy" else scenes.cornell_textured()
    ds = cs.scene.upload(device)
    ds.set_nee_policy("openpbr_proxy")
    yield request.param, cs, ds
    ds.close()


def test_proxy_render_matches_oracle(box):
    name, cs, ds = box
    w = h = 32
    cam = camera.pinhole(**cs.camera, width=w, height=h)
    gpu = ds.render(cam, w, h, 8, 8)
    cpu = oracle.OracleRenderer(cs.scene, cs.materials, nee_policy="openpbr_proxy").render(cam, w, h, 8, 8)
    assert np.all(np.isfinite(gpu))
    rel = np.abs(gpu - cpu) / np.maximum(np.abs(cpu), 1e-3)
    print(f"{name}: mean rel diff {rel.mean():.2e} max {rel.max():.2e} bit-identical {np.mean(gpu == cpu):.3f}")
    assert rel.mean() < 1e-3
    assert abs(gpu.mean() - cpu.mean()) <= 1e-4 * cpu.mean()


def test_wavefront_matches_megakernel_under_proxy(box):
    name, cs, ds = box
    w, h, spp, depth = 48, 40, 8, 8
    cam = camera.pinhole(**cs.camera, width=w, height=h)
    mega = ds.render(cam, w, h, spp, depth)
    (wave,) = ds.render_batch([(cam, w, h, spp, depth)], exact_shadow=True)
    print(f"{name}: identical pixels {np.mean(np.all(mega == wave, axis=2)):.4f}")
    assert bits_equal(mega, wave)


def test_proxy_is_unbiased_against_the_cpu_reference(device):
    """cornell_glossy: the reference policy's 16,384-spp CPU image is the
    expectation of either policy.  GPU proxy relMSE at 64 spp no worse than
    the CPU proxy's (x 1.1); at 4,096 spp the GPU proxy image's mean within
    0.3 % of the reference's."""
    cs = scenes.cornell(specular_weight=1.0)
    ref, size, depth, _ = golden_images.reference("cornell_glossy", cs.scene)
    ds = cs.scene.upload(device)
    try:
        ds.set_nee_policy("openpbr_proxy")
        cam = camera.pinhole(**cs.camera, width=size, height=size)
        gpu = ds.render(cam, size, size, 64, depth)
        cpu = oracle.OracleRenderer(cs.scene, cs.materials, nee_policy="openpbr_proxy").render(
            cam, size, size, 64, depth, threads=16)
        e_gpu, e_cpu = relmse(gpu, ref), relmse(cpu, ref)
        hi = ds.render(cam, size, size, 4096, depth)
        print(f"proxy relMSE at 64 spp: gpu {e_gpu:.5f} cpu {e_cpu:.5f}; 4096-spp mean {hi.mean():.5f} "
              f"reference {ref.mean():.5f} (relMSE {relmse(hi, ref):.2e})")
        assert e_gpu <= 1.1 * e_cpu + 1e-6
        assert abs(hi.mean() - ref.mean()) <= 0.003 * ref.mean()
    finally:
        ds.close()


def test_proxy_lowers_variance_on_the_glossy_box(device):
    cs = scenes.cornell(specular_weight=1.0)
    ref, size, depth, _ = golden_images.reference("cornell_glossy", cs.scene)
    ds = cs.scene.upload(device)
    try:
        cam = camera.pinhole(**cs.camera, width=size, height=size)
        base = ds.render(cam, size, size, 16, depth)
        ds.set_nee_policy("openpbr_proxy")
        prox = ds.render(cam, size, size, 16, depth)
        e_base, e_prox = relmse(base, ref), relmse(prox, ref)
        print(f"glossy box relMSE at 16 spp: reference policy {e_base:.4f}, proxy {e_prox:.4f}")
        assert e_prox < 0.2 * e_base
    finally:
        ds.close()


@pytest.fixture(scope="module")
def night_city(device):
    c = scenes.city(grid=12, rocks=300, lights=64, sky_size=(256, 128), tex_size=64)
    c.scene.environment = None
    ds = c.scene.upload(device)
    # from above the towers, straight down: streets and walls under the lamps
    cam = camera.pinhole(eye=(0.0, 80.0, 0.0), look_at=(0.0, 0.0, 0.0), up=(0.0, 0.0, 1.0), fov_y_deg=40.0,
                         width=64, height=64)
    yield c, ds, cam
    ds.close()


def test_lamps_get_nee_and_lower_the_variance_at_night(night_city):
    """The night city (lamps only): both policies converge to the same image
    (means of 8,192-spp images within 1 %), and at 16 spp the proxy's relMSE
    against the proxy's 8,192-spp image is below half of the reference
    policy's."""
    c, ds, cam = night_city
    try:
        ds.set_nee_policy("reference")
        base_hi = ds.render(cam, 64, 64, 8192, 4)
        base = ds.render(cam, 64, 64, 16, 4)
        ds.set_nee_policy("openpbr_proxy")
        prox_hi = ds.render(cam, 64, 64, 8192, 4)
        prox = ds.render(cam, 64, 64, 16, 4)
    finally:
        ds.set_nee_policy("reference")
    e_base, e_prox = relmse(base, prox_hi), relmse(prox, prox_hi)
    print(f"night city: means at 8192 spp reference policy {base_hi.mean():.5f} proxy {prox_hi.mean():.5f}; "
          f"relMSE at 16 spp reference policy {e_base:.4f} proxy {e_prox:.4f}")
    assert prox_hi.mean() > 0
    assert abs(base_hi.mean() - prox_hi.mean()) <= 0.01 * prox_hi.mean()
    assert e_prox < 0.5 * e_base


def test_night_city_wavefront_matches_megakernel_under_proxy(night_city):
    c, ds, cam = night_city
    try:
        ds.set_nee_policy("openpbr_proxy")
        mega = ds.render(cam, 64, 64, 4, 4)
        (wave,) = ds.render_batch([(cam, 64, 64, 4, 4)], exact_shadow=True)
    finally:
        ds.set_nee_policy("reference")
    assert bits_equal(mega, wave)


def test_batched_adjoint_under_proxy_matches_central_differences(device):
    """The path-replay adjoint holds sampling pdfs and MIS weights fixed;
    under the proxy the light-tree pmf depends on the parameters too, and the
    weights still sum to one, so the gradient stays unbiased: base-colour
    entries against central differences (h = 1e-2, common random numbers) of
    the summed L2 loss over two views of the small inverse-rendering city,
    within 5 % relative L2."""
    c = scenes.city_inverse(grid=10, rocks=200, lights=32, tex_size=64, views=2, view_sizes=(24, 32),
                            sky_size=(64, 32))
    ds = c.scene.upload(device)
    try:
        ds.set_nee_policy("openpbr_proxy")
        fr = [(camera.pinhole(**v, width=w, height=h), w, h, 1024, 8) for v, w, h in c.views]
        theta = np.zeros((len(c.materials), 5))
        for i, m in enumerate(c.materials[:c.learnable]):
            theta[i, :3] = m.base_color
            theta[i, 3] = m.specular_roughness
            theta[i, 4] = m.base_metalness
        tgt = theta.copy()
        tgt[:c.learnable, :3] = np.clip(tgt[:c.learnable, :3] + 0.1, 0, 1)
        ds.set_materials(with_params(c.materials, tgt))
        targets = [im.astype(np.float64) for im in ds.render_batch(fr)]

        def loss(t):
            ds.set_materials(with_params(c.materials, t))
            return sum(float(((im.astype(np.float64) - T) ** 2).sum()) for im, T in zip(ds.render_batch(fr), targets))

        ds.set_materials(with_params(c.materials, theta))
        imgs = ds.render_batch(fr)
        views = [f + ((2.0 * (im - T)).astype(np.float32),) for f, im, T in zip(fr, imgs, targets)]
        g, _ = ds.render_adjoint_batch(views)
        gl = g[:c.learnable, :3]
        picks = [tuple(int(v) for v in np.unravel_index(k, gl.shape)) for k in np.argsort(-np.abs(gl).ravel())[:5]]
        fd, ad = [], []
        for i, col in picks:
            t = theta.copy()
            t[i, col] += 1e-2
            lp = loss(t)
            t[i, col] -= 2e-2
            lm = loss(t)
            fd.append((lp - lm) / 2e-2)
            ad.append(float(g[i, col]))
            print(f"  ({i}, {col}): adjoint {ad[-1]:.5g} fd {fd[-1]:.5g}", flush=True)
        rel = float(np.linalg.norm(np.array(ad) - np.array(fd)) / np.linalg.norm(fd))
        print(f"proxy-policy adjoint vs central differences: {rel:.4f}")
        assert rel < 0.05
    finally:
        ds.close()
