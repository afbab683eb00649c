"""The procedural city on the GPU at a reduced size (12 x 12 towers, 300
rocks, 64 lamps, a 256 x 128 sky with the sun): three frames of uneven
resolution at max depth 16 through the megakernel and through one batched
wavefront launch with FP32 and with 8-bit quantised nodes.  With the exact
shadow rule the wavefront images are the megakernel's bit for bit; every
light category (lamps through the light tree, the sky, the sun in it)
reaches the image.  8-bit nodes: hits equal the FP32 hits except on
exact ties, which the images may show in a few pixels."""
import numpy as np
import pytest

from tpt import camera, scenes

pytestmark = pytest.mark.gpu

SIZES = [(96, 54), (64, 64), (40, 24)]


@pytest.fixture(scope="module")
def city(device):
    c = scenes.city(grid=12, rocks=300, lights=64, sky_size=(256, 128), tex_size=64)
    ds = c.scene.upload(device)
    extent = 0.5 * 12 * 4.0
    views = [dict(eye=(-0.9 * extent, 0.5 * extent, -0.9 * extent), look_at=(0.0, 0.0, 0.0), up=(0.0, 1.0, 0.0),
                  fov_y_deg=50.0),
             # street level: x = 0 is a street line between two rows of plots
             dict(eye=(0.0, 1.7, -2.0), look_at=(0.0, 4.0, 20.0), up=(0.0, 1.0, 0.0), fov_y_deg=70.0),
             dict(eye=(0.0, 3.0 * extent, 0.1), look_at=(0.0, 0.0, 0.0), up=(0.0, 0.0, 1.0), fov_y_deg=60.0)]
    frames = [(camera.pinhole(**v, width=w, height=h), w, h, 4, 16) for v, (w, h) in zip(views, SIZES)]
    yield c, ds, frames
    ds.close()


def test_wavefront_batch_matches_megakernel(city):
    c, ds, frames = city
    mega = [ds.render(*f) for f in frames]
    wave = ds.render_batch(frames, exact_shadow=True)
    for (_, w, h, _, _), a, b in zip(frames, mega, wave):
        assert a.shape == b.shape == (h, w, 3)
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), f"frame {w}x{h}"
    for img in mega:
        assert np.all(np.isfinite(img)) and np.all(img >= 0) and img.mean() > 0


def test_material_sorted_shading_keeps_the_image(city, device):
    """wf_sort 1: each bounce's hits are shaded grouped by material (a
    counting sort of the hit queue); a path's result does not depend on the
    order paths are shaded in, so the image is the megakernel's bit for bit."""
    c, ds, frames = city
    mega = [ds.render(*f) for f in frames]
    plain_q8 = ds.render_batch(frames, quantized=True, exact_shadow=True)
    device.set_param("wf_sort", 1)
    try:
        for q8, want in ((False, mega), (True, plain_q8)):
            got = ds.render_batch(frames, quantized=q8, exact_shadow=True)
            for a, b in zip(want, got):
                assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), f"q8={q8}"
    finally:
        device.set_param("wf_sort", 0)


def test_q8_batch_differs_only_through_ties(city):
    """8-bit nodes change the order children are visited in, so where two
    triangles are hit at the same distance (a shared edge; t within 2 ulp)
    the first one found can differ, and that path continues from the other
    triangle.  Anywhere else the images are the FP32 images: few pixels
    differ, and the frame means agree."""
    c, ds, frames = city
    mega = [ds.render(*f) for f in frames]
    wave = ds.render_batch(frames, quantized=True, exact_shadow=True)
    for a, b in zip(mega, wave):
        differ = np.any(a != b, axis=-1).mean()
        assert differ <= 0.02, differ
        assert abs(float(a.mean()) - float(b.mean())) <= 2e-2 * float(a.mean())


def test_q8_hits_equal_fp32_hits_except_ties(city):
    c, ds, frames = city
    rng = np.random.default_rng(5)
    extent = 0.5 * 12 * 4.0
    o = np.stack([rng.uniform(-extent, extent, 20000), rng.uniform(0.5, 30.0, 20000),
                  rng.uniform(-extent, extent, 20000)], 1)
    d = rng.normal(size=(20000, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = scenes.make_rays(o.astype(np.float32), d.astype(np.float32))
    got = ds.closest_hit(rays)
    q = ds.closest_hit(rays, quantized=True)
    assert (got["triangle"] != 0xFFFFFFFF).mean() > 0.3
    diff = np.nonzero((q["instance"] != got["instance"]) | (q["triangle"] != got["triangle"]))[0]
    ai = q["t"][diff].astype(np.float32).view(np.int32).astype(np.int64)
    bi = got["t"][diff].astype(np.float32).view(np.int32).astype(np.int64)
    assert np.all(np.abs(ai - bi) <= 2), "q8 differs beyond a tie"
    same = np.setdiff1d(np.arange(len(rays)), diff)
    assert np.array_equal(q["t"][same].view(np.uint32), got["t"][same].view(np.uint32))


def test_early_exit_shadow_batch_is_close(city):
    """The timed schedule (early-exit shadow walk) differs from the exact
    rule only where another surface is hit at exactly the light sample's
    distance: the image mean agrees to float noise."""
    c, ds, frames = city
    mega = [ds.render(*f) for f in frames]
    wave = ds.render_batch(frames, quantized=True)
    for a, b in zip(mega, wave):
        assert abs(float(a.mean()) - float(b.mean())) <= 1e-3 * max(float(a.mean()), 1e-6)


def test_lamps_light_the_street_at_night(device):
    """Without the sky, a frame over the streets is lit by the lamps alone.
    Every city material is OpenPBR, which under the reference's policy takes
    no light-tree sample (OpenPbrMaterial has no light-tree precompute,
    R/material/open_pbr.rs:592-606), so the lamps reach the image through
    BSDF-sampled paths that hit an emitter (emission weight 1)."""
    c = scenes.city(grid=12, rocks=300, lights=64, sky_size=(256, 128), tex_size=64)
    c.scene.environment = None
    ds = c.scene.upload(device)
    try:
        # from above the towers, straight down: streets and walls under the lamps
        cam = camera.pinhole(eye=(0.0, 80.0, 0.0), look_at=(0.0, 0.0, 0.0), up=(0.0, 0.0, 1.0), fov_y_deg=40.0,
                             width=32, height=32)
        img = ds.render(cam, 32, 32, 16, 4)
        assert np.all(np.isfinite(img)) and img.mean() > 0
    finally:
        ds.close()
