"""Directional lights on the GPU (the light sampler's Directional category,
R/light/mod.rs:133-191, 330-337; sample_li, R/light/directional.rs:34-43;
delta-light NEE without MIS, R/integrator/mis.rs:177-206): the closed form of
a Lambert floor at depth 1, the sun-and-sky scene against the oracle pixel by
pixel, and the wavefront path tracer against the megakernel bit for bit."""
import numpy as np
import pytest

import oracle
from tpt import camera, scenes
from test_oracle_directional import floor_scene

pytestmark = pytest.mark.gpu

SUN = [((0.3, -1.0, 0.2), (1.0, 0.9, 0.7), 3.0), ((-0.5, -0.6, 0.4), (0.4, 0.5, 1.0), 0.5)]


def test_lambert_floor_closed_form_on_device(device):
    direction, color, intensity, rho = (0.3, -1.0, 0.2), (1.0, 0.5, 0.25), 2.0, (0.7, 0.5, 0.3)
    s, mats, cam = floor_scene(direction, color, intensity, rho)
    ds = s.upload(device)
    try:
        w = h = 16
        c = camera.pinhole(**cam, width=w, height=h)
        img = ds.render(c, w, h, 4, 1)
        d = np.asarray(direction, np.float64)
        d /= np.linalg.norm(d)
        want = np.asarray(rho) / np.pi * np.asarray(color) * intensity * abs(d[1])
        assert np.allclose(img.reshape(-1, 3), want, rtol=2e-6)
        (wave,) = ds.render_batch([(c, w, h, 4, 1)], exact_shadow=True)
        assert np.array_equal(wave.view(np.uint32), img.view(np.uint32))
    finally:
        ds.close()


@pytest.fixture(scope="module", params=["lambert", "openpbr"])
def sun_scene(device, request):
    cs = scenes.outdoor(request.param, sun=SUN)
    ds = cs.scene.upload(device)
    yield cs, ds
    ds.close()


def test_sun_and_sky_render_matches_oracle(sun_scene):
    cs, ds = sun_scene
    w = h = 40
    cam = camera.pinhole(**cs.camera, width=w, height=h)
    gpu = ds.render(cam, w, h, 8, 6)
    cpu = oracle.OracleRenderer(cs.scene, cs.materials).render(cam, w, h, 8, 6)
    same = np.all(gpu.view(np.uint32) == cpu.view(np.uint32), axis=2)
    close = np.all(np.abs(gpu - cpu) <= 1e-3 * (np.abs(cpu) + 1e-3), axis=2)
    print("bit-identical pixels", same.mean(), "within 1e-3", close.mean())
    assert close.mean() > 0.8
    assert abs(gpu.mean() - cpu.mean()) <= 0.05 * cpu.mean() + 1e-3


def test_sun_and_sky_wavefront_equals_megakernel(sun_scene):
    cs, ds = sun_scene
    w, h = 48, 40
    cam = camera.pinhole(**cs.camera, width=w, height=h)
    mega = ds.render(cam, w, h, 8, 6)
    (wave,) = ds.render_batch([(cam, w, h, 8, 6)], exact_shadow=True)
    assert np.array_equal(mega.view(np.uint32), wave.view(np.uint32))
