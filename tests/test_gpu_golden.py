"""libtpt against the committed golden vectors (tests/golden/*.npz): the
same answers bit for bit for the ray queries, the sampler and the camera,
and the same Monte Carlo image up to the transcendental-function paths."""
import sys
from pathlib import Path

import numpy as np
import pytest

from tpt import camera, scenes

pytestmark = pytest.mark.gpu

GOLDEN = Path(__file__).resolve().parent / "golden"
sys.path.insert(0, str(GOLDEN))
import make_golden as mg  # noqa: E402


def bits(a):
    return np.ascontiguousarray(a).view(np.uint8)


@pytest.mark.parametrize("quantized", [False, True])
def test_cornell_queries(device, quantized):
    z = np.load(GOLDEN / "cornell_queries.npz")
    cs = scenes.cornell()
    ds = cs.scene.upload(device)
    got = ds.closest_hit(z["rays"], quantized=quantized)
    ref = z["hits"]
    if not quantized:
        assert np.array_equal(bits(got), bits(ref))
    else:  # quantised boxes change only the visit order, so only exact ties may differ
        diff = (got["triangle"] != ref["triangle"]) | (got["instance"] != ref["instance"])
        assert np.all(got["t"][diff] == ref["t"][diff]) and np.all(ref["triangle"][diff] != 0xFFFFFFFF)
        assert np.array_equal(bits(got["t"][~diff]), bits(ref["t"][~diff]))
        assert diff.mean() < 0.01
    for exact in (False, True):
        occ = ds.unoccluded(z["shadow_rays"], z["shadow_target"], quantized=quantized, exact=exact)
        assert np.array_equal(occ, z["unoccluded"])
    ds.close()


def test_sphere_queries(device):
    z = np.load(GOLDEN / "spheres_queries.npz")
    sc, _ = mg.sphere_case()
    ds = sc.upload(device)
    assert np.array_equal(bits(ds.closest_hit(z["rays"])), bits(z["hits"]))
    ds.close()


@pytest.mark.parametrize("case", mg.SAMPLER_CASES)
def test_sampler_values(device, case):
    w, h, spp, si = case
    z = np.load(GOLDEN / "sampler_camera.npz")
    vals, seeds = camera.path_randoms(device, w, h, spp, si, mg.SAMPLER_DEPTH)
    assert np.array_equal(bits(vals), bits(z[f"vals_{w}x{h}_{spp}_{si}"]))
    assert np.array_equal(seeds, z[f"seeds_{w}x{h}_{spp}_{si}"])


def test_camera_rays(device):
    z = np.load(GOLDEN / "sampler_camera.npz")
    w, h, spp, si = mg.CAMERA_CASE
    cam = camera.pinhole(**scenes.cornell(build=False).camera, width=w, height=h)
    rays, diff = camera.camera_rays(device, cam, w, h, spp, si, differentials=True)
    ref = z["camera_rays"]
    assert np.array_equal(bits(rays["o"]), bits(ref[:, 0:3]))
    assert np.array_equal(bits(rays["d"]), bits(ref[:, 3:6]))
    assert np.array_equal(bits(diff), bits(ref[:, 6:13]))


def test_cornell_render(device):
    z = np.load(GOLDEN / "cornell_render.npz")
    w, h, spp, depth = (int(v) for v in z["case"])
    cs = scenes.cornell()
    ds = cs.scene.upload(device)
    img = ds.render(camera.pinhole(**cs.camera, width=w, height=h), w, h, spp, depth)
    ds.close()
    ref = z["image"]
    same = np.all(img.view(np.uint32) == ref.view(np.uint32), axis=2)
    print("bit-identical pixels:", same.mean())
    assert same.mean() > 0.8
