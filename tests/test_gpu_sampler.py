"""GPU parity of the sampler and camera: bit-exact with the oracle's
restatement of PathSampler (R/sampler/path_sampler.rs) and
PinholeCamera::generate_ray_differential (R/scene/camera.rs:48-96)."""
import ctypes

import numpy as np
import pytest

import oracle
from tpt import camera

pytestmark = pytest.mark.gpu

L = oracle.lib()
L.orc_path_sampler.argtypes = [ctypes.c_uint32] * 6 + [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
L.orc_camera_rays.argtypes = [ctypes.c_void_p] + [ctypes.c_uint32] * 4 + [ctypes.c_void_p]


@pytest.mark.parametrize("w,h,spp,si", [(37, 23, 64, 5), (64, 64, 1, 0), (20, 11, 256, 255), (9, 17, 7, 6)])
def test_path_randoms_match_oracle(device, w, h, spp, si):
    depth = 9
    vals, seeds = camera.path_randoms(device, w, h, spp, si, depth)
    ref_v = np.zeros_like(vals)
    ref_s = np.zeros_like(seeds)
    for i in range(w * h):
        L.orc_path_sampler(w, h, spp, i % w, i // w, si, depth, ref_v[i].ctypes.data, ref_s[i].ctypes.data)
    assert np.array_equal(vals.view(np.uint32), ref_v.view(np.uint32))
    assert np.array_equal(seeds, ref_s)


@pytest.mark.parametrize("w,h,spp,si", [(64, 48, 64, 3), (33, 20, 1, 0), (128, 128, 16, 15)])
def test_camera_rays_match_oracle(device, w, h, spp, si):
    cam = camera.pinhole((0.5, 0.5, -1.4), (0.5, 0.5, 0.5), (0, 1, 0), 40.0, w, h)
    rays, diff = camera.camera_rays(device, cam, w, h, spp, si, differentials=True)
    ref = np.zeros((w * h, 13), np.float32)
    L.orc_camera_rays(camera.camera_array(cam).ctypes.data, w, h, spp, si, ref.ctypes.data)
    assert np.array_equal(rays["o"].view(np.uint32), ref[:, 0:3].view(np.uint32))
    assert np.array_equal(rays["d"].view(np.uint32), ref[:, 3:6].view(np.uint32))
    assert np.array_equal(diff.view(np.uint32), ref[:, 6:13].view(np.uint32))
    assert np.all(np.isinf(rays["tmax"]))
