"""Host side of the textures (no GPU): the Python Texture mirror, the ABI
records that carry textures, the null-argument checks of the texture entry
points, and the host light-tree build reading an emission texture's maximum
(Texture::max_value into OpenPbrMaterial::max_emission,
R/material/open_pbr.rs:381-391) against the oracle's build."""
import ctypes

import numpy as np
import pytest

import oracle
from tpt import _lib, scenes
from tpt.material import MaterialRecord, NO_TEXTURE, OpenPbr, OpenPbrRecord
from tpt.scene import build_light_tree
from tpt.texture import Texture, checkerboard


def test_texture_from_pixels_shapes_and_errors():
    rgb = Texture.from_pixels(3, 2, np.arange(18))
    assert rgb.pixels.shape == (2, 3, 3) and rgb.channels == 3 and rgb.width == 3 and rgb.height == 2
    sc = Texture.from_pixels(3, 2, np.arange(6))
    assert sc.pixels.shape == (2, 3) and sc.channels == 1
    with pytest.raises(ValueError):
        Texture.from_pixels(3, 2, np.arange(7))
    with pytest.raises(ValueError):
        Texture.from_pixels(0, 2, [])


def test_from_rgba_channel_and_max_value():
    t = Texture.from_rgba_channel(2, 1, bytes([10, 20, 30, 40, 110, 120, 130, 140]), 3)
    assert np.allclose(t.pixels, [[40 / 255, 140 / 255]])
    with pytest.raises(ValueError):
        Texture.from_rgba_channel(2, 1, bytes(8), 4)
    assert Texture(np.array([[-1.0, 0.5], [2.5, np.nan]], np.float32)).max_value() == 2.5
    assert Texture(np.full((2, 2), -3.0, np.float32)).max_value() == 0.0


def test_checkerboard_cells():
    t = checkerboard(8, 8, 2, (1, 1, 1), (0, 0, 0))
    assert np.all(t.pixels[:4, :4] == 1) and np.all(t.pixels[:4, 4:] == 0) and np.all(t.pixels[4:, 4:] == 1)


def test_material_record_carries_texture_slots():
    r = OpenPbr(base_color_texture=2, emission_color_texture=0).open_pbr_record()
    assert ctypes.sizeof(r) == 192 and ctypes.sizeof(MaterialRecord) == 200
    assert r.base_color_texture == 2 and r.emission_color_texture == 0
    assert r.base_metalness_texture == NO_TEXTURE and r.fuzz_roughness_texture == NO_TEXTURE
    # the slot fields sit where include/tpt.h puts them: after fuzz_roughness
    assert OpenPbrRecord.base_color_texture.offset == OpenPbrRecord.fuzz_roughness.offset + 4
    assert ctypes.sizeof(_lib.TextureDesc) == 24 and _lib.TEXTURE_QUERY_DTYPE.itemsize == 24


def test_texture_entry_points_reject_null_arguments():
    L = _lib.lib()
    assert L.tpt_texture_sample(None, 0, 0, None, 1, None, 0) == _lib.TPT_ERROR_BAD_ARGUMENT
    assert L.tpt_texture_info(None, 0, None, None, None) == _lib.TPT_ERROR_BAD_ARGUMENT
    assert L.tpt_texture_level(None, 0, 0, None, 0, None, None) == _lib.TPT_ERROR_BAD_ARGUMENT
    assert L.tpt_scene_set_texture(None, 0, None, 0) == _lib.TPT_ERROR_BAD_ARGUMENT
    assert L.tpt_render_adjoint_texels(None, None, None, None, None, 0) == _lib.TPT_ERROR_BAD_ARGUMENT


def test_light_tree_reads_the_emission_texture_maximum():
    cs = scenes.cornell_textured()
    host = build_light_tree(cs.scene)
    orc = oracle.OracleRenderer(cs.scene, cs.materials).light_tree()
    assert len(host) == len(orc["flux"])
    assert np.array_equal(host["flux"], orc["flux"])
    # 2 triangles of area 0.02 each, emission 15 x texture maximum (0.8)
    tmax = cs.scene.textures[cs.materials[7].emission_color_texture].max_value()
    assert abs(tmax - 0.8) < 1e-6
    assert np.allclose(host["flux"][0], 2 * 0.02 * 15.0 * tmax, rtol=1e-5)


def test_light_tree_rejects_a_missing_emission_texture():
    cs = scenes.cornell(build=True)
    cs.scene.materials[3] = cs.materials[3].with_(emission_color_texture=5)
    with pytest.raises(_lib.TptError, match="emission_color_texture"):
        build_light_tree(cs.scene)
