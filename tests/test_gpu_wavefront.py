"""The wavefront path tracer (tpt_render_batch) against the megakernel
(tpt_render): the same estimator, sample sequence and arithmetic per path, so
with the exact shadow rule every frame of a batch is the megakernel's image
bit for bit with FP32 nodes, whatever the batch composition, paths in flight
or graph setting; the default early-exit shadow rule differs only on exact
t ties."""
import numpy as np
import pytest

import tpt
from tpt import camera, scenes

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", params=["openpbr", "lambert_nee"])
def scene(request, device):
    cs = scenes.cornell() if request.param == "openpbr" else scenes.cornell_lambert()
    ds = cs.scene.upload(device)
    yield cs, ds
    ds.close()


def bits_equal(a, b):
    return np.array_equal(np.ascontiguousarray(a).view(np.uint32), np.ascontiguousarray(b).view(np.uint32))


def test_single_frame_matches_megakernel(scene):
    cs, ds = scene
    w, h, spp, depth = 40, 30, 16, 8
    cam = camera.pinhole(**cs.camera, width=w, height=h)
    mega = ds.render(cam, w, h, spp, depth)
    (wave,) = ds.render_batch([(cam, w, h, spp, depth)], exact_shadow=True)
    print("identical pixels", np.mean(np.all(mega == wave, axis=2)))
    assert bits_equal(mega, wave)
    (fast,) = ds.render_batch([(cam, w, h, spp, depth)])
    same = np.mean(np.all(mega == fast, axis=2))
    print("early-exit shadows: identical pixels", same)
    assert same > 0.999 and abs(fast.mean() - mega.mean()) <= 1e-4 * mega.mean()


def test_more_paths_than_resident_threads_match_megakernel(scene):
    """262,144 paths in one pass: more than the shade kernel has resident
    threads, so the launch shape matters (a grid capped at the resident
    capacity once gave wrong OpenPBR paths here)."""
    cs, ds = scene
    w, h, spp, depth = 128, 128, 16, 8
    cam = camera.pinhole(**cs.camera, width=w, height=h)
    mega = ds.render(cam, w, h, spp, depth)
    (wave,) = ds.render_batch([(cam, w, h, spp, depth)], exact_shadow=True)
    print("identical pixels", np.mean(np.all(mega == wave, axis=2)))
    assert bits_equal(mega, wave)


def test_uneven_batch_matches_per_frame_renders(scene):
    cs, ds = scene
    frames = []
    for (w, h, spp, depth) in [(48, 32, 8, 8), (17, 29, 5, 6), (64, 64, 3, 8), (9, 7, 33, 4)]:
        frames.append((camera.pinhole(**cs.camera, width=w, height=h), w, h, spp, depth))
    outs = ds.render_batch(frames, exact_shadow=True)
    for f, img in zip(frames, outs):
        assert bits_equal(img, ds.render(*f))


@pytest.mark.parametrize("paths,graph", [(65536, 1), (65536, 0), (1 << 21, 0)])
def test_paths_in_flight_and_graph_do_not_change_the_image(device, scene, paths, graph):
    cs, ds = scene
    w = h = 96
    cam = camera.pinhole(**cs.camera, width=w, height=h)
    ref = ds.render(cam, w, h, 12, 8)
    device.set_param("wf_paths", paths)
    device.set_param("wf_graph", graph)
    try:
        (img,) = ds.render_batch([(cam, w, h, 12, 8)], exact_shadow=True)
    finally:
        device.set_param("wf_paths", 1 << 21)
        device.set_param("wf_graph", 1)
    assert bits_equal(img, ref)


def test_quantized_nodes_render_the_same_image(scene):
    """8-bit nodes change the traversal order only: hits differ only on exact
    ties, so the image is the FP32 one up to such paths."""
    cs, ds = scene
    w = h = 48
    cam = camera.pinhole(**cs.camera, width=w, height=h)
    (a,) = ds.render_batch([(cam, w, h, 16, 8)])
    (b,) = ds.render_batch([(cam, w, h, 16, 8)], quantized=True)
    same = np.mean(np.all(a == b, axis=2))
    print("q8 identical pixels", same)
    assert same > 0.99 and abs(a.mean() - b.mean()) <= 1e-3 * a.mean()


def test_bad_frames_are_rejected(scene):
    cs, ds = scene
    cam = camera.pinhole(**cs.camera, width=8, height=8)
    with pytest.raises(tpt._lib.TptError):
        ds.render_batch([(cam, 8, 8, 0, 8)])
    assert ds.render_batch([]) == []


@pytest.mark.parametrize("make", ["cornell_textured", "cornell_lambert", "glossy"])
def test_material_sorted_shading_matches_megakernel(device, make):
    cs = scenes.cornell(specular_weight=1.0) if make == "glossy" else getattr(scenes, make)()
    ds = cs.scene.upload(device)
    try:
        w, h = 96, 64
        cam = camera.pinhole(**cs.camera, width=w, height=h)
        ref = ds.render(cam, w, h, 16, 8)
        device.set_param("wf_sort", 1)
        (img,) = ds.render_batch([(cam, w, h, 16, 8)], exact_shadow=True)
        assert bits_equal(img, ref)
    finally:
        device.set_param("wf_sort", 0)
        ds.close()


def test_autotuner_picks_a_setting_no_slower_and_keeps_the_image(device, scene):
    cs, ds = scene
    w = h = 128
    cam = camera.pinhole(**cs.camera, width=w, height=h)
    ref = ds.render(cam, w, h, 8, 8)
    try:
        res = ds.autotune([(cam, w, h, 64, 8)], tune_spp=64)  # 1 Mi paths: two paths-in-flight candidates fit
        print("autotune", res)
        assert res["ms_best"] <= res["ms_default"] and res["trials"] >= 10
        assert res["wf_paths"] in (1 << 19, 1 << 20, 1 << 21, 1 << 22, 1 << 23, 1 << 24)
        assert res["blocks_per_cu"] in (0, 2, 3, 4, 6, 8) and res["refill_below"] in (8, 16, 24, 32, 48)
        assert res["wf_sort"] in (0, 1, 2, 4, 8)
        assert res["lds_depth"] in (8, 12, 16, 24, 32) and res["wf_tile"] in (128, 256, 512)
        (img,) = ds.render_batch([(cam, w, h, 8, 8)], exact_shadow=True)
        assert bits_equal(img, ref)
    finally:
        for name, value in DEFAULTS.items():
            device.set_param(name, value)


DEFAULTS = dict(wf_paths=1 << 21, blocks_per_cu=0, refill_below=40, wf_sort=0, lds_depth=16, wf_tile=256)


@pytest.mark.parametrize("name,value", [("lds_depth", 4), ("lds_depth", 32), ("wf_tile", 128), ("wf_tile", 512),
                                        ("wf_sort", 2), ("wf_sort", 8)])
def test_tuning_knobs_do_not_change_the_image(device, scene, name, value):
    """The traversal stack's LDS depth, the shade queue's tile and the sort's
    granularity change only how the work is scheduled: the image stays the
    megakernel's bit for bit."""
    cs, ds = scene
    w = h = 80
    cam = camera.pinhole(**cs.camera, width=w, height=h)
    ref = ds.render(cam, w, h, 12, 8)
    device.set_param(name, value)
    try:
        (img,) = ds.render_batch([(cam, w, h, 12, 8)], exact_shadow=True)
    finally:
        device.set_param(name, DEFAULTS[name])
    assert bits_equal(img, ref)


def test_tuning_knob_ranges_are_checked(device):
    for name, value in [("lds_depth", 3), ("lds_depth", 33), ("wf_tile", 64), ("wf_tile", 300), ("wf_sort", 65)]:
        with pytest.raises(tpt._lib.TptError, match="out of range"):
            device.set_param(name, value)
