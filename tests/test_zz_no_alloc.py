"""tpt_trace_noalloc (the allocation-free trace entry point) against
tpt_trace_closest / tpt_trace_any on the same device buffers: bit-identical
hits and visibility, on two calls in a row, the second with fewer rays (the
shared scratch keeps the first call's contents and must be cleared)."""
import ctypes

import numpy as np
import pytest
import torch

import tpt
from tpt import _lib, scenes

pytestmark = pytest.mark.gpu


def _dev(a: np.ndarray) -> torch.Tensor:
    return torch.from_numpy(np.ascontiguousarray(a).view(np.uint8).copy()).cuda()


def _noalloc(ds, query, rays_t, n, target_t, dist_t, hits_t, out_t, flags):
    L = _lib.lib()
    st = L.tpt_trace_noalloc(ds.handle, query, _lib.ptr(rays_t), n, _lib.ptr(target_t), _lib.ptr(dist_t),
                             _lib.ptr(hits_t), _lib.ptr(out_t), flags | _lib.TPT_FLAG_DEVICE_POINTERS)
    assert st == _lib.TPT_OK, _lib.last_error()


@pytest.fixture(scope="module")
def field(device):
    sf = scenes.sphere_field(n_spheres=60, level=3, seed=5)
    scene = scenes.single_mesh_scene(sf.mesh)
    rays = scenes.primary_rays(320, 240, **scenes.TRAVERSAL_CAMERA)
    ds = scene.upload(device)
    hits = ds.closest_hit(rays)
    sh = scenes.shadow_rays(rays, hits, sf.mesh, 60_000, n_lights=16)
    yield ds, rays, sh
    ds.close()


@pytest.mark.parametrize("q8", [False, True])
def test_noalloc_closest_matches_trace_closest(device, field, q8):
    ds, rays, _ = field
    device.set_stream(torch.cuda.current_stream().cuda_stream)
    flags = _lib.TPT_FLAG_QUANTIZED if q8 else 0
    for n in (len(rays), len(rays) // 3 + 17):  # second call: fewer rays
        r = _dev(rays[:n])
        ref = torch.zeros(n * 24, dtype=torch.uint8, device="cuda")
        got = torch.full((n * 24,), 0xAB, dtype=torch.uint8, device="cuda")
        ds.closest_hit_device(r, ref, n, quantized=q8)
        _noalloc(ds, 0, r, n, None, None, got, None, flags)
        torch.cuda.synchronize()
        assert torch.equal(ref, got), f"closest hits differ (n={n}, q8={q8})"
        h = got.cpu().numpy().view(tpt.HIT_DTYPE)
        assert (h["triangle"] != tpt.INVALID_INDEX).mean() > 0.1


@pytest.mark.parametrize("exact", [False, True])
def test_noalloc_shadow_matches_trace_any(device, field, exact):
    ds, _, sh = field
    device.set_stream(torch.cuda.current_stream().cuda_stream)
    flags = _lib.TPT_FLAG_ANY_EXACT if exact else 0
    for n in (len(sh.rays), 20_011):
        r = _dev(sh.rays[:n])
        tg = _dev(sh.target[:n])
        ref = torch.zeros(n, dtype=torch.uint8, device="cuda")
        got = torch.full((n,), 7, dtype=torch.uint8, device="cuda")
        ds.unoccluded_device(r, tg, None, ref, n, exact=exact)
        _noalloc(ds, 1, r, n, tg, None, None, got, flags)
        torch.cuda.synchronize()
        assert torch.equal(ref, got), f"visibility differs (n={n}, exact={exact})"
        frac = got.float().mean().item()
        assert 0.05 < frac < 0.95


def test_noalloc_rejects_host_pointers_and_bad_queries(device, field):
    ds, rays, _ = field
    L = _lib.lib()
    hits = np.zeros(4, tpt.HIT_DTYPE)
    assert L.tpt_trace_noalloc(ds.handle, 0, rays.ctypes.data, 4, None, None, hits.ctypes.data, None, 0) \
        == _lib.TPT_ERROR_BAD_ARGUMENT
    assert b"DEVICE_POINTERS" in L.tpt_last_error()
    assert L.tpt_trace_noalloc(ds.handle, 2, None, 4, None, None, None, None, _lib.TPT_FLAG_DEVICE_POINTERS) \
        == _lib.TPT_ERROR_BAD_ARGUMENT
    assert L.tpt_trace_noalloc(ds.handle, 0, ctypes.c_void_p(16), 1 << 31, None, None, ctypes.c_void_p(16), None,
                               _lib.TPT_FLAG_DEVICE_POINTERS) == _lib.TPT_ERROR_BAD_ARGUMENT
    assert b"2^31" in L.tpt_last_error()
