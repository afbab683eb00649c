"""GPU path tracer vs the oracle's restatement of the reference renderer
(R/main.rs:98-126 + R/integrator/mis.rs on OpenPBR: the configuration's
diffuse walls, specular weight 0, and the reference's defaults, specular
weight 1, roughness 0.3, IOR 1.5).

The two run the same estimator on the same sample sequence: the sampler,
camera, traversal, shading frame and most BSDF arithmetic agree bit for bit,
and the transcendental functions (cos / sin of direction sampling, atan2 of
the albedo-table lookup, pow, exp of the device maths library vs the host
libm) may differ in the last place, after which a path can take a different
branch.  So most pixels are bit-identical
and the images agree to the Monte Carlo noise level; and against the CPU
oracle's 16,384-spp image (tests/golden/reference_images.npz) the GPU image's
relMSE (mean over pixels and channels of (x - r)^2 / (r^2 + 1e-2)) is no
worse than the CPU image's at equal samples per pixel.
"""
import numpy as np
import pytest

import golden_images
import oracle
import tpt
from tpt import camera, scenes
from tpt.material import OpenPbr
from tpt.imageio import relmse

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", params=[0.0, 1.0], ids=["diffuse", "glossy"])
def cornell(device, request):
    # the configuration's diffuse OpenPBR walls, and the reference's
    # OpenPbrMaterial::new defaults (specular weight 1: the GGX tables)
    cs = scenes.cornell(specular_weight=request.param)
    ds = cs.scene.upload(device)
    yield cs, ds
    ds.close()


def test_render_matches_oracle_pixel_by_pixel(cornell):
    cs, ds = cornell
    w = h = 48
    cam = camera.pinhole(**cs.camera, width=w, height=h)
    gpu = ds.render(cam, w, h, 8, 8)
    cpu = oracle.OracleRenderer(cs.scene, cs.materials).render(cam, w, h, 8, 8)
    assert np.all(np.isfinite(gpu))
    same = np.all(gpu.view(np.uint32) == cpu.view(np.uint32), axis=2)
    print("bit-identical pixels:", same.mean())
    assert same.mean() > 0.8
    # where they differ, the difference is at the Monte Carlo noise level
    assert abs(gpu.mean() - cpu.mean()) <= 0.05 * cpu.mean() + 1e-3


def test_render_relmse_no_worse_than_cpu(cornell):
    """Against the CPU oracle's 16,384-spp image (a committed fixture, not a
    GPU render), the GPU image at 64 spp is no worse than the CPU's."""
    cs, ds = cornell
    name = "cornell_diffuse" if cs.materials[0].specular_weight == 0.0 else "cornell_glossy"
    e_gpu, e_cpu = golden_images.gpu_and_cpu_relmse(name, cs, ds, 64)
    print(f"{name} relMSE at 64 spp against the CPU reference: gpu {e_gpu:.5f} cpu {e_cpu:.5f}")
    assert e_gpu <= 1.1 * e_cpu + 1e-6


def test_render_is_deterministic_and_running_mean(cornell):
    cs, ds = cornell
    w = h = 16
    cam = camera.pinhole(**cs.camera, width=w, height=h)
    a = ds.render(cam, w, h, 4, 8)
    b = ds.render(cam, w, h, 4, 8)
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    # exposure is a final multiply (R/main.rs:120)
    cam2 = camera.pinhole(**cs.camera, width=w, height=h, exposure=2.0)
    c = ds.render(cam2, w, h, 4, 8)
    assert np.array_equal(c, (a * np.float32(2.0)).astype(np.float32))


def test_unsupported_material_is_rejected(device):
    cs = scenes.cornell(build=False)
    # dispersion (the spectral BTDF) is not carried
    cs.scene.materials[0] = OpenPbr.new((0.8, 0.8, 0.8)).with_(transmission_weight=0.5,
                                                                 transmission_dispersion_scale=0.5)
    cs.scene.build_qbvh()
    ds = cs.scene.upload(device)
    cam = camera.pinhole(**cs.camera, width=8, height=8)
    with pytest.raises(tpt._lib.TptError, match="transmission_dispersion_scale"):
        ds.render(cam, 8, 8, 1, 8)
    ds.close()


# ---- next-event estimation through the light tree (NormalizedLambert walls) ----
@pytest.fixture(scope="module")
def cornell_nee(device):
    cs = scenes.cornell_lambert()
    ds = cs.scene.upload(device)
    yield cs, ds
    ds.close()


def test_light_tree_on_device_matches_host_build(cornell_nee):
    cs, ds = cornell_nee
    from tpt.scene import build_light_tree
    a, b = ds.light_tree(), build_light_tree(cs.scene)
    assert len(a) == 3 and np.array_equal(a.view(np.uint8), b.view(np.uint8))


def test_nee_render_matches_oracle(cornell_nee):
    """Same estimator and sample sequence; the light-tree importances go
    through exp / expm1 / exp2 (device maths library vs glibc), so values
    agree to float rounding, not bit for bit: per-channel relative error of
    the image below 1e-3 on average and the image means equal to 1e-4."""
    cs, ds = cornell_nee
    w = h = 32
    cam = camera.pinhole(**cs.camera, width=w, height=h)
    gpu = ds.render(cam, w, h, 8, 8)
    cpu = oracle.OracleRenderer(cs.scene, cs.materials).render(cam, w, h, 8, 8)
    assert np.all(np.isfinite(gpu))
    rel = np.abs(gpu - cpu) / np.maximum(np.abs(cpu), 1e-3)
    print("mean rel diff", rel.mean(), "max", rel.max(), "bit-identical", np.mean(gpu == cpu))
    assert rel.mean() < 1e-3
    assert abs(gpu.mean() - cpu.mean()) <= 1e-4 * cpu.mean()


def test_nee_render_relmse_no_worse_than_cpu(cornell_nee):
    cs, ds = cornell_nee
    e_gpu, e_cpu = golden_images.gpu_and_cpu_relmse("cornell_lambert_nee", cs, ds, 64)
    print(f"NEE relMSE at 64 spp against the CPU reference: gpu {e_gpu:.5f} cpu {e_cpu:.5f}")
    assert e_gpu <= 1.1 * e_cpu + 1e-6


def test_nee_lowers_variance(cornell_nee, cornell):
    """Light-tree NEE with MIS beats BSDF sampling alone on the same
    geometry: the Lambert scene's relMSE at 16 spp is below the OpenPBR
    (BSDF-sampling-only) scene's, each against its own 16384-spp image."""
    errs = []
    for cs, ds in (cornell_nee, cornell):
        cam = camera.pinhole(**cs.camera, width=32, height=32)
        ref = ds.render(cam, 32, 32, 16384, 8)
        errs.append(relmse(ds.render(cam, 32, 32, 16, 8), ref))
    print("relMSE at 16 spp: NEE", errs[0], "BSDF only", errs[1])
    assert errs[0] < errs[1]
