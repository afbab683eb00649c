"""The oracle against the committed golden vectors (tests/golden/*.npz,
written by tests/golden/make_golden.py): the oracle, and the seeded scene
generators both sides share, still give the recorded answers bit for bit."""
import ctypes
import sys
from pathlib import Path

import numpy as np
import pytest

import oracle
from tpt import camera, scenes

GOLDEN = Path(__file__).resolve().parent / "golden"
sys.path.insert(0, str(GOLDEN))
import make_golden as mg  # noqa: E402

L = oracle.lib()
L.orc_path_sampler.argtypes = [ctypes.c_uint32] * 6 + [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
L.orc_camera_rays.argtypes = [ctypes.c_void_p] + [ctypes.c_uint32] * 4 + [ctypes.c_void_p]


def bits(a):
    return np.ascontiguousarray(a).view(np.uint8)


def test_cornell_queries():
    z = np.load(GOLDEN / "cornell_queries.npz")
    cs = scenes.cornell()
    assert str(z["digest"]) == mg.mesh_digest(cs.scene)
    assert np.array_equal(bits(mg.cornell_rays()), bits(z["rays"]))
    o = oracle.oracle_scene_from(cs.scene)
    assert np.array_equal(bits(o.closest_hit(z["rays"])), bits(z["hits"]))
    assert np.array_equal(o.unoccluded(z["shadow_rays"], z["shadow_target"]), z["unoccluded"])


def test_sphere_queries():
    z = np.load(GOLDEN / "spheres_queries.npz")
    sc, rays = mg.sphere_case()
    assert str(z["digest"]) == mg.mesh_digest(sc)
    assert np.array_equal(bits(rays), bits(z["rays"]))
    assert np.array_equal(bits(oracle.oracle_scene_from(sc).closest_hit(rays)), bits(z["hits"]))


@pytest.mark.parametrize("case", mg.SAMPLER_CASES)
def test_sampler_values(case):
    w, h, spp, si = case
    z = np.load(GOLDEN / "sampler_camera.npz")
    vals = np.zeros((w * h, 2 + 14 * mg.SAMPLER_DEPTH), np.float32)
    seeds = np.zeros((w * h, 1 + mg.SAMPLER_DEPTH), np.uint32)
    for i in range(w * h):
        L.orc_path_sampler(w, h, spp, i % w, i // w, si, mg.SAMPLER_DEPTH, vals[i].ctypes.data, seeds[i].ctypes.data)
    assert np.array_equal(bits(vals), bits(z[f"vals_{w}x{h}_{spp}_{si}"]))
    assert np.array_equal(seeds, z[f"seeds_{w}x{h}_{spp}_{si}"])
    assert np.all((vals >= 0) & (vals < 1))


def test_camera_rays():
    z = np.load(GOLDEN / "sampler_camera.npz")
    w, h, spp, si = mg.CAMERA_CASE
    cam = camera.pinhole(**scenes.cornell(build=False).camera, width=w, height=h)
    assert np.array_equal(bits(camera.camera_array(cam)), bits(z["camera_array"]))
    ref = np.zeros((w * h, 13), np.float32)
    L.orc_camera_rays(z["camera_array"].ctypes.data, w, h, spp, si, ref.ctypes.data)
    assert np.array_equal(bits(ref), bits(z["camera_rays"]))


def test_cornell_render():
    z = np.load(GOLDEN / "cornell_render.npz")
    w, h, spp, depth = (int(v) for v in z["case"])
    cs = scenes.cornell()
    cam = camera.pinhole(**cs.camera, width=w, height=h)
    img = oracle.OracleRenderer(cs.scene, cs.materials).render(cam, w, h, spp, depth, threads=3)
    assert np.array_equal(bits(img), bits(z["image"]))


def test_reference_images_match_the_scenes_and_the_oracle_converges_to_them():
    """The CPU 16,384-spp reference images (tests/golden/reference_images.npz)
    belong to the current scene generators, and the oracle's own renders
    converge to them: relMSE falls roughly as 1/spp from 16 to 256 spp."""
    import golden_images
    from tpt import camera
    from make_reference_images import CASES

    for name in ("cornell_diffuse", "cornell_lambert_nee"):
        make, size, depth = CASES[name]
        cs = make()
        ref, size, depth, spp = golden_images.reference(name, cs.scene)
        assert spp == 16384 and ref.shape == (size, size, 3) and np.all(np.isfinite(ref)) and ref.mean() > 0
        cam = camera.pinhole(**cs.camera, width=size, height=size)
        r = oracle.OracleRenderer(cs.scene, cs.materials)
        e16 = golden_images.relmse(r.render(cam, size, size, 16, depth), ref)
        e256 = golden_images.relmse(r.render(cam, size, size, 256, depth), ref)
        assert e256 < e16 / 6.0, (name, e16, e256)
