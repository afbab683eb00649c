"""The QBVH build on the GPU (tpt_qbvh_build_device) against the host build
(tpt_qbvh_build, the reference's build_qbvh restated: R/qbvh.rs:144-516):
the same node array, byte for byte, and the same primitive order.

Inputs: seeded random boxes around the sizes where the GPU build changes
phase (ranges of more than 2048 primitives are split level by level over
chunks, smaller ones by one block), clustered boxes (deep, unbalanced
splits), coincident boxes (no split: leaves up to 16 primitives, then the
forced median split), boxes on a plane (a zero-extent centroid axis), and
the triangle bounds and instance bounds of the Cornell box, the traversal
sphere field and the city miniature."""
import time

import numpy as np
import pytest

from tpt import scenes
from tpt.scene import build_qbvh

pytestmark = pytest.mark.gpu


def _same(lo, hi, device, label):
    t = time.perf_counter()
    host = build_qbvh(lo, hi)
    t_host = time.perf_counter() - t
    t = time.perf_counter()
    gpu = build_qbvh(lo, hi, device=device)
    t_gpu = time.perf_counter() - t
    print(f"{label}: {len(lo)} primitives, {len(host.nodes)} nodes; host {t_host * 1e3:.1f} ms, "
          f"GPU {t_gpu * 1e3:.1f} ms", flush=True)
    assert len(gpu.nodes) == len(host.nodes), label
    assert gpu.nodes.tobytes() == host.nodes.tobytes(), label
    np.testing.assert_array_equal(gpu.primitive_indices, host.primitive_indices, err_msg=label)


def _random_boxes(n, seed, spread=100.0, size=1.0):
    rng = np.random.default_rng(seed)
    c = rng.uniform(-spread, spread, (n, 3)).astype(np.float32)
    e = rng.uniform(0.0, size, (n, 3)).astype(np.float32)
    return c - e, c + e


@pytest.mark.parametrize("n", [1, 3, 4, 5, 16, 17, 100, 2047, 2048, 2049, 5000, 100_000, 1_000_000])
def test_random_boxes(device, n):
    lo, hi = _random_boxes(n, seed=n)
    _same(lo, hi, device, f"random {n}")


def test_clustered_boxes(device):
    rng = np.random.default_rng(7)
    centres = rng.uniform(-1000, 1000, (40, 3))
    pts = []
    for k, c in enumerate(centres):  # clusters of very different sizes and spreads
        m = int(rng.integers(1, 20000))
        pts.append(c + rng.normal(0.0, 10.0 ** rng.uniform(-3, 1), (m, 3)))
    p = np.concatenate(pts).astype(np.float32)
    e = rng.uniform(0.0, 0.05, p.shape).astype(np.float32)
    _same(p - e, p + e, device, "clustered")


@pytest.mark.parametrize("n", [12, 16, 40, 3000])
def test_coincident_boxes(device, n):
    lo = np.tile(np.float32([1.0, 2.0, 3.0]), (n, 1))
    hi = lo + np.float32(0.5)
    _same(lo, hi, device, f"coincident {n}")


def test_mixed_coincident_and_spread(device):
    # coincident groups among spread boxes: forced splits inside a larger tree
    lo, hi = _random_boxes(20000, seed=3)
    lo[:5000] = lo[0]
    hi[:5000] = hi[0]
    lo[9000:9100] = lo[9000]
    hi[9000:9100] = hi[9000]
    _same(lo, hi, device, "mixed coincident")


def test_planar_boxes(device):
    lo, hi = _random_boxes(30000, seed=5)
    lo[:, 1] = 0.0
    hi[:, 1] = 0.0
    _same(lo, hi, device, "planar")


def test_scene_meshes_and_instances(device):
    field = scenes.single_mesh_scene(scenes.sphere_field(n_spheres=60, level=3, seed=2).mesh, build=False)
    for name, sc in [("cornell", scenes.cornell(build=False).scene),
                     ("sphere field", field),
                     ("city miniature", scenes.city(grid=12, rocks=300, lights=64, sky_size=(64, 32),
                                                   tex_size=64, build=False).scene)]:
        for i, m in enumerate(sc.meshes):
            lo, hi = m.triangle_bounds()
            _same(lo, hi, device, f"{name} mesh {i}")
        if len(sc.instances) > 1:
            lo = np.stack([x.world_bounds[0] for x in sc.instances])
            hi = np.stack([x.world_bounds[1] for x in sc.instances])
            _same(lo, hi, device, f"{name} instances")


def test_device_built_scene_renders_identically(device):
    cs = scenes.cornell_textured(build=False)
    cs.scene.build_qbvh(device=device)
    ds = cs.scene.upload(device)
    ref = scenes.cornell_textured()
    rs = ref.scene.upload(device)
    from tpt import camera
    cam = camera.pinhole(**cs.camera, width=64, height=64)
    a = ds.render_batch([(cam, 64, 64, 4, 6)])[0]
    b = rs.render_batch([(cam, 64, 64, 4, 6)])[0]
    ds.close()
    rs.close()
    np.testing.assert_array_equal(a, b)


def test_bad_input(device):
    from tpt import _lib
    from tpt._lib import lib
    assert not lib().tpt_qbvh_build_device(device.handle, None, None, 4, 0)
    assert "null bounds" in _lib.last_error()
    assert build_qbvh(np.zeros((0, 3)), np.zeros((0, 3)), device=device) is None


def _scene_q8(ds, tlas):
    import ctypes
    from tpt import _lib
    from tpt._lib import lib, ptr
    n = ctypes.c_size_t(0)
    assert lib().tpt_scene_q8_nodes(ds.handle, tlas, None, 0, ctypes.byref(n)) == 0
    out = np.zeros(n.value, _lib.Q8_NODE_DTYPE)
    if n.value:
        assert lib().tpt_scene_q8_nodes(ds.handle, tlas, ptr(out), n.value, ctypes.byref(n)) == 0, _lib.last_error()
    return out


def _host_q8(nodes, base):
    from tpt import _lib
    from tpt._lib import lib, ptr
    out = np.zeros(len(nodes), _lib.Q8_NODE_DTYPE)
    if len(nodes):
        assert lib().tpt_qbvh_quantize(ptr(nodes), len(nodes), ptr(out)) == 0
    interior = (out["child"] & np.uint32(0x80000000)) == 0
    out["child"] = np.where(interior, out["child"] + np.uint32(base), out["child"])
    return out


def test_scene_q8_nodes_are_encoded_on_the_device_as_on_the_host(device):
    """tpt_scene_create encodes the 8-bit nodes on the device; they are the
    bytes tpt_qbvh_quantize writes on the host (interior children rebased
    into the scene's node array)."""
    field = scenes.single_mesh_scene(scenes.sphere_field(n_spheres=60, level=3, seed=2).mesh)
    for name, sc in [("cornell", scenes.cornell().scene), ("sphere field", field),
                     ("city miniature", scenes.city(grid=12, rocks=300, lights=64, sky_size=(64, 32),
                                                   tex_size=64).scene)]:
        ds = sc.upload(device)
        try:
            blas = _scene_q8(ds, 0)
            want, base = [], 0
            for m in sc.meshes:
                want.append(_host_q8(m.qbvh.nodes, base))
                base += len(m.qbvh.nodes)
            want = np.concatenate(want)
            assert len(blas) == len(want), name
            assert blas.tobytes() == want.tobytes(), name
            tl = _scene_q8(ds, 1)
            tw = _host_q8(sc.qbvh.nodes, 0) if sc.qbvh is not None else np.zeros(0)
            assert tl.tobytes() == tw.tobytes(), name
            print(f"{name}: {len(blas)} BLAS and {len(tl)} TLAS 8-bit nodes identical", flush=True)
        finally:
            ds.close()
