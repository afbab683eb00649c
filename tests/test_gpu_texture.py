"""Textures on the GPU against the oracle's restatement of the reference's
Texture (oracle/texture.h, R/material/texture.rs): the device-built mip
pyramid bit for bit, every lookup mode query by query, and the textured path
tracer (megakernel and wavefront) against the oracle's renderer.

Bilinear and nearest lookups and the pyramid use only +, *, floor and fmod,
so they are bit-exact.  The EWA weights call expf and the level choice
log2f, whose device implementations differ from the host libm in the last
place: filtered values agree to a relative 1e-5, except where a level
boundary falls between the two log2f results (at most 0.5% of random
queries, each still a convex combination of texels)."""
import numpy as np
import pytest

import golden_images
import oracle
import tpt
from tpt import _lib, camera, scenes
from tpt.material import OpenPbr
from tpt.scene import Scene
from tpt.texture import Texture
from tpt.imageio import relmse

pytestmark = pytest.mark.gpu

SHAPES = [(64, 64, 3), (40, 48, 3), (3, 5), (7, 1), (9, 13, 3), (1, 1, 3), (32, 32)]


def _texture_scene(textures):
    s = Scene()
    m = s.add_mesh(scenes.flat_quads([[(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0)]]))
    s.add_instance(m, 0, np.eye(4))
    s.add_material(OpenPbr.new((0.5, 0.5, 0.5)))
    for t in textures:
        s.add_texture(t)
    s.build_qbvh()
    return s


@pytest.fixture(scope="module")
def textured(device):
    rng = np.random.default_rng(7)
    texs = [Texture(rng.uniform(-0.2, 1.5, shp).astype(np.float32)) for shp in SHAPES]
    ds = _texture_scene(texs).upload(device)
    yield ds, texs, [oracle.OracleTexture(t.pixels) for t in texs]
    ds.close()


def test_pyramid_matches_oracle_bit_for_bit(textured):
    ds, texs, orcs = textured
    for i, o in enumerate(orcs):
        info = ds.texture_info(i)
        assert info["levels"] == o.levels
        assert info["channels"] == o.channels
        assert info["max_value"] == np.float32(o.max_value)
        for lv in range(o.levels):
            g, c = ds.texture_level(i, lv), o.level(lv)
            assert g.shape == c.shape, (i, lv)
            assert np.array_equal(g.view(np.uint32), c.view(np.uint32)), (i, lv)


def _queries(rng, n):
    uv = rng.uniform(-2.0, 3.0, (n, 2)).astype(np.float32)
    mag = 10.0 ** rng.uniform(-4.5, 0.3, (n, 1))
    d0 = (rng.normal(size=(n, 2)) * mag).astype(np.float32)
    aniso = 10.0 ** rng.uniform(-1.5, 0.0, (n, 1))
    d1 = (rng.normal(size=(n, 2)) * mag * aniso).astype(np.float32)
    d0[: n // 50] = 0.0  # zero footprints: bilinear level 0
    d1[n // 50: n // 25, 0] = np.inf  # non-finite: bilinear level 0
    return uv, d0, d1


@pytest.mark.parametrize("mode", [_lib.TEXTURE_BILINEAR, _lib.TEXTURE_NEAREST])
def test_exact_lookups_match_oracle(textured, mode):
    ds, texs, orcs = textured
    rng = np.random.default_rng(100 + mode)
    for i, o in enumerate(orcs):
        uv, d0, d1 = _queries(rng, 4000)
        g = ds.texture_sample(i, uv, d0, d1, mode=mode)
        c = o.sample(uv, d0, d1, mode=mode)
        assert np.array_equal(g.view(np.uint32), c.reshape(g.shape).view(np.uint32)), (i, mode)


@pytest.mark.parametrize("mode", [_lib.TEXTURE_FILTERED, _lib.TEXTURE_TRILINEAR, _lib.TEXTURE_MIP_BILINEAR])
def test_filtered_lookups_match_oracle(textured, mode):
    ds, texs, orcs = textured
    rng = np.random.default_rng(200 + mode)
    for i, o in enumerate(orcs):
        uv, d0, d1 = _queries(rng, 4000)
        g = ds.texture_sample(i, uv, d0, d1, mode=mode)
        c = o.sample(uv, d0, d1, mode=mode).reshape(g.shape)
        assert np.all(np.isfinite(g))
        close = np.all(np.abs(g - c) <= 1e-5 * np.abs(c) + 2e-6, axis=1)
        assert close.mean() >= 0.995, (i, mode, close.mean())
        lo, hi = texs[i].pixels.min(), texs[i].pixels.max()
        assert np.all(g >= lo - 1e-5) and np.all(g <= hi + 1e-5)


def test_reference_texture_tests_on_device(device):
    """The reference's own texture tests (R/material/texture.rs:621-723),
    run through the device lookups."""
    t2 = Texture(np.array([[[1, 0, 0], [0, 1, 0]], [[0, 0, 1], [1, 1, 1]]], np.float32))
    ds = _texture_scene([t2, Texture(np.array([[0.0, 1.0], [1.0, 0.0]], np.float32))]).upload(device)
    try:
        B, F, T = _lib.TEXTURE_BILINEAR, _lib.TEXTURE_FILTERED, _lib.TEXTURE_TRILINEAR
        assert np.array_equal(ds.texture_sample(0, [[0.25, 0.25]], mode=B)[0], [1, 0, 0])
        assert np.array_equal(ds.texture_sample(0, [[0.75, 0.75]], mode=B)[0], [1, 1, 1])
        assert np.allclose(ds.texture_sample(0, [[0.5, 0.5]], mode=B)[0], 0.5, atol=1e-6)
        assert np.array_equal(ds.texture_sample(0, [[1.25, -0.75]], mode=B)[0], [1, 0, 0])
        assert abs(ds.texture_sample(1, [[0.5, 0.5]], mode=B)[0, 0] - 0.5) < 1e-6
        assert np.allclose(ds.texture_level(0, 1)[0, 0], 0.5, atol=1e-6)
        assert np.allclose(ds.texture_sample(0, [[0.25, 0.25]], [[1, 0]], [[0, 1]], mode=T)[0], 0.5, atol=1e-6)
        assert np.allclose(ds.texture_sample(0, [[0.25, 0.25]], [[1, 0]], [[0, 1]], mode=F)[0], 0.5, atol=1e-6)
    finally:
        ds.close()


@pytest.mark.parametrize("bad,what", [(dict(base_color_texture=5), "not a texture"),
                                      (dict(base_color_texture=1), "RGB"),
                                      (dict(specular_roughness_texture=0), "scalar")])
def test_invalid_texture_slots_are_rejected(device, bad, what):
    s = _texture_scene([Texture(np.ones((4, 4, 3), np.float32)), Texture(np.ones((4, 4), np.float32))])
    s.materials[0] = s.materials[0].with_(**bad)
    with pytest.raises(tpt._lib.TptError, match=what):
        s.upload(device)


# ---- the textured path tracer ------------------------------------------------------------------


@pytest.fixture(scope="module")
def cornell_tex(device):
    cs = scenes.cornell_textured()
    ds = cs.scene.upload(device)
    yield cs, ds
    ds.close()


def test_textured_render_matches_oracle(cornell_tex):
    cs, ds = cornell_tex
    w = h = 48
    cam = camera.pinhole(**cs.camera, width=w, height=h)
    gpu = ds.render(cam, w, h, 8, 8)
    cpu = oracle.OracleRenderer(cs.scene, cs.materials).render(cam, w, h, 8, 8)
    assert np.all(np.isfinite(gpu))
    rel = np.max(np.abs(gpu - cpu) / (np.abs(cpu) + 1e-3), axis=2)
    close = rel < 1e-3
    print("pixels within 1e-3:", close.mean(), "bit-identical:",
          np.all(gpu.view(np.uint32) == cpu.view(np.uint32), axis=2).mean())
    assert close.mean() > 0.8
    assert abs(gpu.mean() - cpu.mean()) <= 0.05 * cpu.mean() + 1e-3


def test_textures_change_the_image(cornell_tex, device):
    cs, ds = cornell_tex
    w = h = 24
    cam = camera.pinhole(**cs.camera, width=w, height=h)
    a = ds.render(cam, w, h, 4, 8)
    plain = scenes.cornell_textured()
    plain.scene.materials[:] = [m.with_(base_color_texture=None) if isinstance(m, OpenPbr) else m
                                for m in plain.scene.materials]
    dp = plain.scene.upload(device)
    try:
        b = dp.render(cam, w, h, 4, 8)
    finally:
        dp.close()
    assert not np.array_equal(a, b)


def test_textured_wavefront_equals_megakernel(cornell_tex):
    cs, ds = cornell_tex
    w, h = 40, 32
    cam = camera.pinhole(**cs.camera, width=w, height=h)
    mega = ds.render(cam, w, h, 4, 8)
    (wave,) = ds.render_batch([(cam, w, h, 4, 8)])
    assert np.array_equal(mega.view(np.uint32), wave.view(np.uint32))


def test_textured_render_relmse_no_worse_than_cpu(cornell_tex):
    cs, ds = cornell_tex
    e_gpu, e_cpu = golden_images.gpu_and_cpu_relmse("cornell_textured", cs, ds, 64)
    print(f"textured relMSE at 64 spp against the CPU reference: gpu {e_gpu:.5f} cpu {e_cpu:.5f}")
    assert e_gpu <= 1.1 * e_cpu + 1e-6
