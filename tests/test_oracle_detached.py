"""The oracle's detached evaluation (OracleRenderer.render_rect_detached_f64):
one renderer's paths (sampling decisions, densities, MIS weights, roulette)
carrying another renderer's BSDF values.  The FP64 reference gradient of the
adjoint (tests/golden/fp64_gradient.npz) is made of its central
differences, so it is pinned here:

- with the evaluation materials equal to the path's, it is the plain render,
  bit for bit, in both builds, also through delta lobes (smooth glass);
- with other materials it is an unbiased estimator of the other materials'
  render (importance sampling by the first material's densities), checked
  against that render at many samples, including delta-lobe numerators
  (a tinted smooth glass), a rough thin-walled pane and texels;
- its central differences are linear in a Lambert albedo's step where the
  albedo enters every path once (the differences at two steps agree)."""
import numpy as np
import pytest

import oracle
from tpt import camera, scenes


def _crop(r, ev, cs, size, spp, depth, rect):
    cam = camera.pinhole(**cs.camera, width=size, height=size)
    x0, x1, y0, y1 = rect
    return r.render_rect_detached_f64(ev, cam, size, size, spp, depth, x0, x1, y0, y1)


@pytest.mark.parametrize("fp64", [False, True])
@pytest.mark.parametrize("which", ["inverse", "glass"])
def test_equal_materials_reproduce_the_render(which, fp64):
    cs = scenes.cornell_inverse(glossy=True, seed=0) if which == "inverse" else scenes.cornell_glass()
    a = oracle.OracleRenderer(cs.scene, cs.materials, fp64=fp64)
    b = oracle.OracleRenderer(cs.scene, cs.materials, fp64=fp64)
    cam = camera.pinhole(**cs.camera, width=64, height=64)
    plain = a.render_rect_f64(cam, 64, 64, 16, 6, 20, 44, 16, 48)
    detached = a.render_rect_detached_f64(b, cam, 64, 64, 16, 6, 20, 44, 16, 48)
    assert plain.any()
    np.testing.assert_array_equal(detached, plain)


def test_builds_cannot_be_mixed():
    cs = scenes.cornell_inverse(glossy=True, seed=0)
    a = oracle.OracleRenderer(cs.scene, cs.materials, fp64=False)
    b = oracle.OracleRenderer(cs.scene, cs.materials, fp64=True)
    cam = camera.pinhole(**cs.camera, width=8, height=8)
    with pytest.raises(ValueError):
        a.render_rect_detached_f64(b, cam, 8, 8, 1, 2, 0, 8, 0, 8)


def _unbiased(cs, mats_b, texels_b=None, rect=(8, 24, 8, 24), size=32, spp=2048, depth=4):
    """Mean over the crop of the detached estimator (paths of cs.materials,
    values of mats_b) and of mats_b's plain render, at spp samples."""
    a = oracle.OracleRenderer(cs.scene, cs.materials)
    if texels_b is not None:
        keep = cs.scene.textures[0].pixels
        cs.scene.textures[0].pixels = texels_b
    b = oracle.OracleRenderer(cs.scene, mats_b)
    if texels_b is not None:
        cs.scene.textures[0].pixels = keep
    cam = camera.pinhole(**cs.camera, width=size, height=size)
    x0, x1, y0, y1 = rect
    det = a.render_rect_detached_f64(b, cam, size, size, spp, depth, x0, x1, y0, y1)
    plain = b.render_rect_f64(cam, size, size, spp, depth, x0, x1, y0, y1)
    assert np.isfinite(det).all()
    return det.mean(axis=(0, 1)), plain.mean(axis=(0, 1))


def test_detached_estimator_is_unbiased_for_openpbr_parameters_and_texels():
    cs = scenes.cornell_inverse(glossy=True, seed=0)
    mats = list(cs.materials)
    for i in range(16):  # every glossy material: other colour, roughness, metalness
        m = mats[i]
        mats[i] = m.with_(base_color=tuple(0.8 * c + 0.1 for c in m.base_color),
                          specular_roughness=min(0.9, m.specular_roughness * 1.3),
                          base_metalness=0.7 * m.base_metalness + 0.1)
    texels = np.ascontiguousarray(cs.scene.textures[0].pixels[:, :, ::-1] * 0.7, np.float32)
    det, plain = _unbiased(cs, mats, texels)
    # the two are independent estimates over 256 pixels x 2048 samples
    print("detached", det, "plain", plain)
    np.testing.assert_allclose(det, plain, rtol=0.03)
    # and the detached estimate differs from the unperturbed image by more
    # than the tolerance (the test sees the change)
    base = oracle.OracleRenderer(cs.scene, cs.materials)
    cam = camera.pinhole(**cs.camera, width=32, height=32)
    b0 = base.render_rect_f64(cam, 32, 32, 2048, 4, 8, 24, 8, 24).mean(axis=(0, 1))
    print("unperturbed", b0)
    assert np.abs(plain / b0 - 1).max() > 0.15


def test_detached_estimator_is_unbiased_through_delta_lobes():
    cs = scenes.cornell_glass()
    mats = list(cs.materials)
    mats[3] = mats[3].with_(transmission_color=(0.5, 0.8, 1.0))  # the smooth glass sphere: delta lobes only
    mats[1] = mats[1].with_(base_color=(0.3, 0.6, 0.3))
    det, plain = _unbiased(cs, mats, rect=(8, 24, 10, 26), spp=4096)
    print("detached", det, "plain", plain)
    np.testing.assert_allclose(det, plain, rtol=0.03)


def test_differences_of_a_lambert_albedo_are_step_independent():
    # the NormalizedLambert Cornell box at depth 1 (direct light only): every
    # path carries the albedo of the surface it sees once, so the detached
    # loss-free image is linear in it and its central differences do not
    # depend on the step
    import dataclasses
    cs = scenes.cornell_lambert()
    a = oracle.OracleRenderer(cs.scene, cs.materials, fp64=True)
    idx = next(i for i, m in enumerate(cs.materials) if type(m).__name__ == "NormalizedLambert")

    rho = np.array(cs.materials[idx].rho, np.float32)

    def image(delta):
        mats = list(cs.materials)
        mats[idx] = dataclasses.replace(mats[idx], rho=tuple(float(v) for v in rho + np.float32(delta)))
        return _crop(a, oracle.OracleRenderer(cs.scene, mats, fp64=True), cs, 32, 16, 1, (0, 32, 0, 32))

    def differences(h):  # the records hold f32: divide by the steps they hold
        steps = (rho + np.float32(h)).astype(np.float64) - (rho - np.float32(h)).astype(np.float64)
        return (image(h) - image(-h)) / steps

    d1, d2 = differences(1e-3), differences(1e-1)
    assert np.abs(d1).max() > 0
    np.testing.assert_allclose(d1, d2, rtol=1e-7, atol=1e-9)
