"""A scene uploaded from its on-disk cache (DeviceScene.from_cache:
tpt_scene_desc_load + tpt_scene_create) renders and traces exactly as the
scene uploaded from the host build, and differentiates to the adjoint's
run-to-run rounding."""
import numpy as np
import pytest

from tpt import camera, scenes
from tpt.scene import DeviceScene

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("which", ["cornell_textured", "outdoor", "cornell_inverse"])
def test_cached_upload_renders_identically(device, tmp_path, which):
    if which == "outdoor":
        cs = scenes.outdoor(sky_size=(128, 64))
    elif which == "cornell_inverse":
        cs = scenes.cornell_inverse(glossy=True, seed=0)
    else:
        cs = scenes.cornell_textured()
    path = tmp_path / "scene.tpts"
    cs.scene.save_cache(path)
    direct = cs.scene.upload(device)
    cached = DeviceScene.from_cache(device, path)
    try:
        assert cached.device_bytes == direct.device_bytes
        assert cached.texture_shapes == direct.texture_shapes
        cam = camera.pinhole(**cs.camera, width=96, height=64)
        a = direct.render_batch([(cam, 96, 64, 8, 6)])[0]
        b = cached.render_batch([(cam, 96, 64, 8, 6)])[0]
        assert a.any()
        np.testing.assert_array_equal(a, b)
        rays = scenes.primary_rays(96, 64, cs.camera["eye"], cs.camera["look_at"], cs.camera["up"],
                                   cs.camera["fov_y_deg"])
        np.testing.assert_array_equal(direct.closest_hit(rays), cached.closest_hit(rays))
        if which == "cornell_inverse":
            dimg = np.ones((64, 96, 3), np.float32)
            pa, ta = direct.render_adjoint_params(cam, 96, 64, 4, 6, dimg)
            pb, tb = cached.render_adjoint_params(cam, 96, 64, 4, 6, dimg)
            # the adjoint's in-block float atomics order additions differently
            # from run to run: equal to the last bits
            np.testing.assert_allclose(pa, pb, rtol=1e-6, atol=1e-6 * np.abs(pa).max())
            for x, y in zip(ta, tb):
                np.testing.assert_allclose(x, y, rtol=1e-5, atol=1e-6 * np.abs(x).max())
    finally:
        direct.close()
        cached.close()
