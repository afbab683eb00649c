"""Pins the oracle's environment light against the reference's tests
(R/light/environment.rs:536-846: uv round trip, uniform pdf 1/4pi, sample
pdf == pdf query, zero maps, MIS-compensated weights), float64 closed forms
(the pdf integrates to 1 over the sphere; total power = integral of L), and
the MIS closed forms of R/integrator/mis.rs:460-506 (environment emission
weights, a direct camera hit)."""
import math

import numpy as np
import pytest

import oracle
import tpt
from tpt.material import NormalizedLambert


def env_renderer(pixels, scale=1.0, rotate_y=0.0, meshes=(), instances=(), materials=(NormalizedLambert(),)):
    s = tpt.Scene()
    for m in meshes:
        s.add_mesh(m)
    for m in materials:
        s.add_material(m)
    for mesh, mat, l2w in instances:
        s.add_instance(mesh, mat, l2w)
    s.build_qbvh()
    s.set_environment_light(tpt.Environment(np.asarray(pixels, np.float32), scale, rotate_y))
    return oracle.OracleRenderer(s, s.materials)


def uniform(w, h, v):
    return np.full((h, w, 3), v, np.float32)


@pytest.mark.parametrize("d", [(0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1), (1, 0, 0), (-1, 0, 0)])
def test_uv_round_trip_on_axes(d):  # environment.rs:536-552
    uv = oracle.env_uv(d)
    assert np.allclose(oracle.env_dir(*uv), d, atol=1e-5)


def test_uv_to_direction_is_unit():
    assert abs(np.linalg.norm(oracle.env_dir(0.37, 0.62)) - 1.0) < 1e-5


def test_uniform_pdf_and_power():  # environment.rs:560-571
    r = env_renderer(uniform(128, 64, 1.0))
    assert abs(r.env_pdf(oracle.env_dir(0.25, 0.5)) - 1 / (4 * math.pi)) < 5e-3
    assert abs(r.env_total_power() - 4 * math.pi) / (4 * math.pi) < 2e-3  # integral of L = 1 over the sphere


def test_pdf_integrates_to_one_and_matches_samples():  # environment.rs:573-593 plus quadrature
    px = np.full((8, 16, 3), 0.05, np.float32)
    px[3, 9] = 10.0
    px[5, 2] = (4.0, 1.0, 0.5)
    r = env_renderer(px)
    n = 256  # midpoint rule, 16 x 16 points per texel
    th = (np.arange(n) + 0.5) * math.pi / n
    ph = (np.arange(2 * n) + 0.5) * math.pi / n
    total = 0.0
    for t in th:
        for p in ph:
            d = (math.sin(t) * math.sin(p), math.cos(t), math.sin(t) * math.cos(p))
            total += r.env_pdf(d) * math.sin(t)
    total *= (math.pi / n) ** 2
    assert abs(total - 1.0) < 0.01
    for mis in (False, True):
        for (u, v) in [(0.1, 0.1), (0.5, 0.5), (0.8, 0.3), (0.95, 0.85), (0.33, 0.77)]:
            d, rad, pdf = r.env_sample(u, v, mis)
            assert abs(np.linalg.norm(d) - 1) < 1e-5 and pdf > 0
            assert abs(pdf - r.env_pdf(d, mis)) / pdf < 1e-3
            assert np.allclose(rad, r.env_radiance(d))


def test_zero_and_uniform_maps():  # environment.rs:595-610
    r = env_renderer(uniform(8, 4, 0.0))
    assert r.env_sample(0.5, 0.5) is None and r.env_pdf((0, 1, 0)) == 0.0
    r = env_renderer(uniform(8, 4, 1.0))
    assert r.env_sample(0.5, 0.5, mis=True) is None and r.env_pdf((0, 1, 0), mis=True) == 0.0


def test_peaked_map_concentrates_samples():  # environment.rs:612-637
    px = np.full((16, 32, 3), 1e-4, np.float32)
    px[6, 10] = 1e6
    r = env_renderer(px)
    for (u, v) in [(0.1, 0.1), (0.4, 0.4), (0.7, 0.7), (0.95, 0.05)]:
        d, _, _ = r.env_sample(u, v)
        uv = oracle.env_uv(d)
        assert int(uv[0] * 32) == 10 and int(uv[1] * 16) == 6


def test_mis_compensated_weights():  # environment.rs:639-693
    px = np.ones((2, 4, 3), np.float32)
    px[1, 2] = 9.0
    r = env_renderer(px)
    assert r.env_pdf(oracle.env_dir(0.125, 0.25), mis=True) == 0.0
    assert r.env_pdf(oracle.env_dir(2.5 / 4, 1.5 / 2), mis=True) > 0.0
    px = np.zeros((4, 1, 3), np.float32)
    px[0] = 4.0
    px[1] = 1.2
    r = env_renderer(px)
    assert r.env_pdf(oracle.env_dir(0.5, 0.375), mis=True) > 0.0


def test_rotation_moves_the_map():
    px = np.full((8, 16, 3), 0.01, np.float32)
    px[4, 0] = 50.0  # near phi = 0 (+z) at the horizon
    r0, r1 = env_renderer(px), env_renderer(px, rotate_y=math.pi / 2)
    d0, _, _ = r0.env_sample(0.5, 0.5)
    d1, _, _ = r1.env_sample(0.5, 0.5)
    # Mat3::from_rotation_y(pi/2) maps +z to +x
    rot = np.array([[0, 0, 1], [0, 1, 0], [-1, 0, 0]], np.float64)
    assert np.allclose(rot @ d0, d1, atol=1e-5)


def test_environment_emission_mis():  # mis.rs:460-487
    r = env_renderer(uniform(32, 16, 2.0))
    assert np.all(r.emitted_infinite((1, 1, 1), 0.0, False, (0, 1, 0)) == 0)
    assert np.allclose(r.emitted_infinite((1, 1, 1), 1.0, True, (0, 1, 0)), 2.0, atol=1e-5)
    assert r.emitted_infinite((1, 1, 1), 1000.0, False, (0, 1, 0))[0] > 2.0 * 0.9


def test_camera_ray_reads_the_environment():  # mis.rs:489-504
    c = np.array([0.1, 0.4, 0.9], np.float32)
    r = env_renderer(np.tile(c, (8, 16, 1)))
    out = r.trace((0, 0, 0), (0, 1, 0), (1, 1), 1, (0, 0), 0, 4)
    assert np.allclose(out, c, atol=1e-5)


def test_env_nee_plus_bsdf_is_unbiased_for_a_lambert_floor():
    """A Lambert floor under a sky: the depth-1 estimate (env NEE + BSDF
    sample, balance heuristic) converges to rho/pi * float64 integral of
    L cos over the upper hemisphere."""
    h, w = 32, 64
    v = (np.arange(h) + 0.5) / h
    sky = np.zeros((h, w, 3), np.float32)
    sky[:, :, :] = (0.5 + 1.5 * np.maximum(np.cos(v * math.pi), 0))[:, None, None]
    sky[6, 20] = 200.0
    floor = tpt.Mesh(np.array([(-10, 0, -10), (10, 0, -10), (-10, 0, 10), (10, 0, 10)], np.float32),
                     np.array([[0, 2, 1], [1, 2, 3]], np.uint32))
    r = env_renderer(sky, meshes=[floor], instances=[(0, 0, np.eye(4))], materials=[NormalizedLambert((0.8,) * 3)])
    origin = np.array([0.0, 1.0, -1.0])
    d = (np.array([0.0, 0.0, 0.0]) - origin) / math.sqrt(2.0)
    n = 4096
    acc = np.zeros(3)
    for si in range(n):
        acc += r.trace(origin, d, (1, 1), n, (0, 0), si, 1)
    est = acc / n
    # E = sum over texels of L * cos(theta) * solid angle, upper hemisphere (y > 0)
    e = 0.0
    for j in range(h):
        th0, th1 = j * math.pi / h, (j + 1) * math.pi / h
        # integral over the row of cos(theta) sin(theta) d theta, times dphi
        if math.cos(th0) <= 0:
            continue
        a, b = th0, min(th1, math.pi / 2)
        ring = 0.5 * (math.sin(b) ** 2 - math.sin(a) ** 2) * (2 * math.pi / w)
        e += sky[j, :, 0].astype(np.float64).sum() * ring
    expected = 0.8 / math.pi * e
    assert np.allclose(est, expected, rtol=0.02), (est, expected)
