"""Directional lights in the oracle (R/light/directional.rs, the
Directional category of R/light/mod.rs): the reference's own tests
(direction normalised, sample radiance color * intensity, one light-sampler
entry per light with pmf by luminance * intensity), category pmfs by hand
with the environment and the light tree present, and a closed form: a
NormalizedLambert floor lit only by a directional light renders at depth 1
to rho / pi * color * intensity * cos(theta) exactly (next-event estimation
of a delta light is deterministic, and nothing else lights the floor)."""
import numpy as np

import oracle
from tpt import camera, scenes
from tpt.material import NormalizedLambert
from tpt.scene import DirectionalLight, Scene


def test_directional_light_normalizes_direction():
    d = DirectionalLight.new((0.0, 0.0, -4.0), (1, 1, 1), 1.0)
    assert abs(np.linalg.norm(d.direction) - 1.0) < 1e-6
    assert d.direction == (0.0, 0.0, -1.0)


def test_directional_lights_each_get_their_own_entry():
    cs = scenes.outdoor("lambert", sky=False, sun=[((0, 0, -1), (1, 1, 1), 1.0), ((-1, 0, 0), (1, 1, 1), 2.0)])
    cats = oracle.OracleRenderer(cs.scene, cs.materials).light_categories()
    assert [k for k, _ in cats] == [2, 3]
    p0, p1 = cats[0][1], cats[1][1]
    assert p1 > p0 and abs(p0 - 1 / 3) < 1e-6 and abs(p1 - 2 / 3) < 1e-6


def test_category_pmfs_with_environment_and_tree():
    cs = scenes.cornell_lambert()
    cs.scene.add_directional_light(DirectionalLight.new((0.3, -1.0, 0.2), (1.0, 0.5, 0.25), 2.0))
    r = oracle.OracleRenderer(cs.scene, cs.materials)
    cats = r.light_categories()
    flux = float(r.light_tree()["flux"][0])
    w_dir = (0.2126 * 1.0 + 0.7152 * 0.5 + 0.0722 * 0.25) * 2.0
    assert [k for k, _ in cats] == [0, 2]
    assert abs(cats[0][1] - flux / (flux + w_dir)) < 1e-6
    assert abs(cats[1][1] - w_dir / (flux + w_dir)) < 1e-6
    # with the sky too: tree, environment, directional
    cs2 = scenes.outdoor("lambert", sun=[((0.3, -1.0, 0.2), (1.0, 0.5, 0.25), 2.0)])
    r2 = oracle.OracleRenderer(cs2.scene, cs2.materials)
    cats2 = r2.light_categories()
    p_env = r2.env_total_power()
    assert [k for k, _ in cats2] == [1, 2]
    assert abs(cats2[0][1] - p_env / (p_env + w_dir)) < 1e-6


def floor_scene(direction, color, intensity, rho=(0.7, 0.5, 0.3)):
    s = Scene()
    m = s.add_mesh(scenes.flat_quads([[(-10, 0, -10), (-10, 0, 10), (10, 0, 10), (10, 0, -10)]]))
    s.add_instance(m, 0, np.eye(4))
    s.build_qbvh()
    mats = [NormalizedLambert(rho)]
    s.add_material(mats[0])
    s.add_directional_light(DirectionalLight.new(direction, color, intensity))
    cam = dict(eye=(0.0, 2.0, -3.0), look_at=(0.0, 0.0, 0.0), up=(0.0, 1.0, 0.0), fov_y_deg=40.0)
    return s, mats, cam


def test_lambert_floor_under_a_directional_light_matches_the_closed_form():
    direction, color, intensity, rho = (0.3, -1.0, 0.2), (1.0, 0.5, 0.25), 2.0, (0.7, 0.5, 0.3)
    s, mats, cam = floor_scene(direction, color, intensity, rho)
    w = h = 16
    img = oracle.OracleRenderer(s, mats).render(camera.pinhole(**cam, width=w, height=h), w, h, 4, 1)
    d = np.asarray(direction, np.float64)
    d /= np.linalg.norm(d)
    cos = abs(d[1])  # |n . (-d)| with n = +y
    want = np.asarray(rho) / np.pi * np.asarray(color) * intensity * cos
    assert np.allclose(img.reshape(-1, 3), want, rtol=2e-6)
