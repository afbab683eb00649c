"""The on-disk scene cache (tpt_scene_desc_save / tpt_scene_desc_load /
tpt_scene_desc_free; no reference counterpart, host code only): a saved
built scene loads back into a descriptor that holds the same arrays, a
second save of the loaded descriptor writes the same bytes, and a
truncated, corrupt, foreign or other-version file is refused with a reason."""
import ctypes

import numpy as np
import pytest

from tpt import _lib, scenes
from tpt._lib import lib
from tpt.scene import ClosestHitError, Scene, scene_desc

HEADER = 56  # magic, version, six record sizes, pad, payload bytes, checksum


def checksum(payload: bytes) -> int:
    """The cache's payload checksum: FNV-1a 64 over 64-bit little-endian
    words in four interleaved lanes, folded in order with the word count."""
    basis, prime, m = 1469598103934665603, 1099511628211, (1 << 64) - 1
    lanes = [basis, basis ^ 1, basis ^ 2, basis ^ 3]
    words = np.frombuffer(payload, "<u8")
    for i, w in enumerate(words.tolist()):
        lanes[i & 3] = ((lanes[i & 3] ^ w) * prime) & m
    h = basis
    for v in lanes:
        h = ((h ^ v) * prime) & m
    return ((h ^ len(words)) * prime) & m


def _array(addr, dtype, count, shape=None):
    if count == 0:
        return np.zeros((0,), dtype)
    buf = (ctypes.c_char * (np.dtype(dtype).itemsize * count)).from_address(addr)
    a = np.frombuffer(buf, dtype, count).copy()
    return a.reshape(shape) if shape else a


@pytest.fixture(scope="module", params=["cornell_textured", "outdoor"])
def scene(request):
    if request.param == "outdoor":
        return scenes.outdoor(sky_size=(64, 32)).scene
    return scenes.cornell_textured().scene


def _load(path):
    d = lib().tpt_scene_desc_load(str(path).encode())
    assert d, _lib.last_error()
    return d


def test_round_trip_holds_every_array(scene, tmp_path):
    p = tmp_path / "scene.tpts"
    scene.save_cache(p)
    d = _load(p)
    try:
        desc = _lib.SceneDesc.from_address(d)
        assert desc.mesh_count == len(scene.meshes)
        assert desc.instance_count == len(scene.instances)
        assert desc.material_count == len(scene.materials)
        assert desc.texture_count == len(scene.textures)
        assert desc.directional_count == len(scene.directional_lights)
        meshes = (_lib.MeshDesc * desc.mesh_count).from_address(desc.meshes)
        for m, md in zip(scene.meshes, meshes):
            np.testing.assert_array_equal(_array(md.positions, np.float32, 3 * md.vertex_count, (-1, 3)),
                                          m.positions)
            np.testing.assert_array_equal(_array(md.indices, np.uint32, 3 * md.triangle_count, (-1, 3)), m.indices)
            assert md.node_count == len(m.qbvh.nodes)
            assert _array(md.nodes, _lib.NODE_DTYPE, md.node_count).tobytes() == m.qbvh.nodes.tobytes()
            np.testing.assert_array_equal(_array(md.primitive_indices, np.uint32, md.triangle_count),
                                          m.qbvh.primitive_indices)
            assert (md.normals is not None) == (m.normals is not None)
            if m.normals is not None:
                np.testing.assert_array_equal(_array(md.normals, np.float32, 3 * md.vertex_count, (-1, 3)),
                                              np.asarray(m.normals, np.float32).reshape(-1, 3))
            if m.uvs is not None:
                np.testing.assert_array_equal(_array(md.uvs, np.float32, 2 * md.vertex_count, (-1, 2)),
                                              np.asarray(m.uvs, np.float32).reshape(-1, 2))
        assert _array(desc.tlas_nodes, _lib.NODE_DTYPE, desc.tlas_node_count).tobytes() == scene.qbvh.nodes.tobytes()
        if scene.environment is not None:
            env = scene.environment.pixels
            assert (desc.env_height, desc.env_width) == env.shape[:2]
            np.testing.assert_array_equal(_array(desc.env_pixels, np.float32, env.size, env.shape), env)
            assert desc.env_scale == np.float32(scene.environment.scale)
        for t, td in zip(scene.textures, (_lib.TextureDesc * desc.texture_count).from_address(desc.textures or 0)
                         if desc.texture_count else []):
            px = np.asarray(t.pixels, np.float32)
            np.testing.assert_array_equal(_array(td.pixels, np.float32, px.size, px.shape), px)
    finally:
        lib().tpt_scene_desc_free(d)


def test_checksum_is_the_stated_one(tmp_path):
    scene = scenes.cornell_textured().scene
    p = tmp_path / "s.tpts"
    scene.save_cache(p)
    raw = p.read_bytes()
    assert raw[:8] == b"TPTSCENE"
    assert int.from_bytes(raw[40:48], "little") == len(raw) - HEADER
    assert int.from_bytes(raw[48:56], "little") == checksum(raw[HEADER:])


def test_a_loaded_descriptor_saves_the_same_bytes(scene, tmp_path):
    a, b = tmp_path / "a.tpts", tmp_path / "b.tpts"
    scene.save_cache(a)
    d = _load(a)
    try:
        assert lib().tpt_scene_desc_save(d, str(b).encode()) == 0, _lib.last_error()
    finally:
        lib().tpt_scene_desc_free(d)
    assert a.read_bytes() == b.read_bytes()
    # and the saved descriptor is the one the upload path builds
    desc, keep = scene_desc(scene)
    c = tmp_path / "c.tpts"
    assert lib().tpt_scene_desc_save(ctypes.byref(desc), str(c).encode()) == 0
    assert c.read_bytes() == a.read_bytes()


def _refused(path, reason):
    assert not lib().tpt_scene_desc_load(str(path).encode())
    assert reason in _lib.last_error(), _lib.last_error()


def test_bad_files_are_refused(tmp_path):
    scene = scenes.cornell_textured().scene
    p = tmp_path / "s.tpts"
    scene.save_cache(p)
    raw = bytearray(p.read_bytes())
    cases = {
        "missing": None,
        "magic": raw[:0] + b"NOTSCENE" + raw[8:],
        "version": raw[:8] + (7).to_bytes(4, "little") + raw[12:],
        "layout": raw[:12] + (999).to_bytes(4, "little") + raw[16:],
        "truncated": raw[:-9],
        "longer": raw + b"\0" * 8,
        "flipped": raw[:HEADER + 100] + bytes([raw[HEADER + 100] ^ 1]) + raw[HEADER + 101:],
        "header only": raw[:HEADER - 4],
    }
    reasons = {"missing": "cannot open", "magic": "not a scene cache", "version": "format version 7",
               "layout": "other record layouts", "truncated": "payload size", "longer": "payload size",
               "flipped": "checksum", "header only": "not a scene cache"}
    for name, data in cases.items():
        q = tmp_path / f"{name.replace(' ', '_')}.tpts"
        if data is not None:
            q.write_bytes(bytes(data))
        _refused(q, reasons[name])


def test_payload_sizes_are_checked_after_the_checksum(tmp_path):
    # a payload whose counts promise more arrays than it holds (checksum
    # recomputed, so only the array bounds can refuse it)
    scene = scenes.cornell_textured().scene
    p = tmp_path / "s.tpts"
    scene.save_cache(p)
    raw = bytearray(p.read_bytes())
    payload = bytearray(raw[HEADER:])
    payload[0:4] = (1000).to_bytes(4, "little")  # mesh_count
    raw[48:56] = checksum(bytes(payload)).to_bytes(8, "little")
    q = tmp_path / "counts.tpts"
    q.write_bytes(bytes(raw[:HEADER]) + bytes(payload))
    _refused(q, "arrays do not match the payload")


def test_an_unbuilt_scene_is_not_saved(tmp_path):
    with pytest.raises(ClosestHitError):
        Scene().save_cache(tmp_path / "x.tpts")
    assert not (tmp_path / "x.tpts").exists()
