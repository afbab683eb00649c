"""The oracle's texture restatement (oracle/texture.h) pinned against the
reference's own texture tests (R/material/texture.rs:621-723) and against what
the mathematics fixes: the box-filtered pyramid preserves the mean, bilinear
lookup reproduces an affine function of the texel coordinates exactly, EWA
of a constant is that constant, EWA centred on a texel of a linear ramp with an
axis-aligned footprint returns the ramp there (symmetric lattice, symmetric
weights), every filtered value is a convex combination of texels, and the
level choice follows log2 of the footprint (a one-texel checkerboard is grey
from level 1 on)."""
import numpy as np
import pytest

from oracle import OracleTexture

FILTERED, BILINEAR, NEAREST, TRILINEAR, MIP_BILINEAR = 0, 1, 2, 3, 4


def tex_2x2():
    """test_texture() of the reference: 2x2 red, green / blue, white."""
    return OracleTexture(np.array([[[1, 0, 0], [0, 1, 0]], [[0, 0, 1], [1, 1, 1]]], np.float32))


# ---- the reference's tests ---------------------------------------------------------------------
def test_samples_texel_centers():
    t = tex_2x2()
    assert np.array_equal(t.sample([0.25, 0.25], mode=BILINEAR)[0], [1, 0, 0])
    assert np.array_equal(t.sample([0.75, 0.75], mode=BILINEAR)[0], [1, 1, 1])


def test_bilinearly_interpolates_between_texels():
    assert np.allclose(tex_2x2().sample([0.5, 0.5], mode=BILINEAR)[0], 0.5, atol=1e-6)


def test_wraps_uvs():
    assert np.array_equal(tex_2x2().sample([1.25, -0.75], mode=BILINEAR)[0], [1, 0, 0])


def test_scalar_texture_samples_single_channel():
    assert OracleTexture(np.array([[0.25]], np.float32)).sample([0.0, 0.0], mode=BILINEAR)[0, 0] == 0.25


def test_scalar_texture_bilerp_returns_average():
    t = OracleTexture(np.array([[0.0, 1.0], [1.0, 0.0]], np.float32))
    assert abs(t.sample([0.5, 0.5], mode=BILINEAR)[0, 0] - 0.5) < 1e-6


def test_from_rgba_channel_extracts_requested_byte():
    from tpt.texture import Texture
    tex = Texture.from_rgba_channel(2, 1, bytes([10, 20, 30, 40, 110, 120, 130, 140]), 2)
    t = OracleTexture(tex.pixels)
    assert abs(t.sample([0.25, 0.5], mode=BILINEAR)[0, 0] - 30 / 255) < 1e-6
    assert abs(t.sample([0.75, 0.5], mode=BILINEAR)[0, 0] - 130 / 255) < 1e-6


def test_mip_pyramid_averages_to_top_level():
    t = tex_2x2()
    assert t.levels == 2
    assert np.allclose(t.level(1)[0, 0], 0.5, atol=1e-6)


def test_trilinear_sampling_uses_coarser_mip_for_wide_footprint():
    r = tex_2x2().sample([0.25, 0.25], [1.0, 0.0], [0.0, 1.0], mode=TRILINEAR)[0]
    assert np.allclose(r, 0.5, atol=1e-6)


def test_ewa_sampling_uses_coarser_mip_for_wide_footprint():
    r = tex_2x2().sample([0.25, 0.25], [1.0, 0.0], [0.0, 1.0], mode=FILTERED)[0]
    assert np.allclose(r, 0.5, atol=1e-6)


# ---- what the mathematics fixes -----------------------------------------------------------------
@pytest.mark.parametrize("shape,expect", [((5, 3), [(5, 3), (3, 2), (2, 1), (1, 1)]),
                                          ((8, 8), [(8, 8), (4, 4), (2, 2), (1, 1)]),
                                          ((1, 6), [(1, 6), (1, 3), (1, 2), (1, 1)])])
def test_level_sizes_halve_with_div_ceil(shape, expect):
    w, h = shape
    t = OracleTexture(np.zeros((h, w), np.float32))
    assert [(t.level(i).shape[1], t.level(i).shape[0]) for i in range(t.levels)] == expect


def test_pyramid_of_power_of_two_preserves_the_mean_and_block_averages():
    rng = np.random.default_rng(1)
    px = rng.random((16, 32, 3), dtype=np.float32)
    t = OracleTexture(px)
    prev = px.astype(np.float64)
    for i in range(1, t.levels):
        lv = t.level(i).astype(np.float64)
        h, w = prev.shape[:2]
        blocks = prev.reshape(max(h // 2, 1), min(h, 2), max(w // 2, 1), min(w, 2), 3)
        want = blocks.mean(axis=(1, 3)) if h > 1 and w > 1 else None
        if want is not None and want.shape == lv.shape:
            assert np.allclose(lv, want, atol=1e-6)
        assert np.allclose(lv.mean(axis=(0, 1)), px.astype(np.float64).mean(axis=(0, 1)), atol=1e-5)
        prev = lv


def test_pyramid_of_odd_size_wraps_the_last_column():
    px = np.arange(3 * 2, dtype=np.float32).reshape(2, 3)  # 3 wide, 2 high
    lv = OracleTexture(px).level(1)  # 2 x 1
    # texel (1, 0) averages columns 2 and 0 (wrapped) of rows 0 and 1
    assert lv.shape == (1, 2)
    assert lv[0, 0] == np.float32((((0 + 1) + 3) + 4) * 0.25)
    assert lv[0, 1] == np.float32((((2 + 0) + 5) + 3) * 0.25)


def test_bilinear_reproduces_an_affine_ramp_inside_the_image():
    h, w = 16, 32
    y, x = np.mgrid[0:h, 0:w].astype(np.float64)
    ramp = (0.25 * x + 0.5 * y + 1.0).astype(np.float32)  # exact in f32
    t = OracleTexture(ramp)
    rng = np.random.default_rng(2)
    uv = np.stack([rng.uniform(0.6 / w, 1 - 1.6 / w, 200), rng.uniform(0.6 / h, 1 - 1.6 / h, 200)], 1)
    uv = uv.astype(np.float32)
    got = t.sample(uv, mode=BILINEAR)[:, 0]
    sx = uv[:, 0].astype(np.float64) * w - 0.5
    sy = uv[:, 1].astype(np.float64) * h - 0.5
    assert np.allclose(got, 0.25 * sx + 0.5 * sy + 1.0, atol=2e-5)


def test_nearest_returns_the_texel_under_uv():
    rng = np.random.default_rng(3)
    px = rng.random((7, 5, 3), dtype=np.float32)
    t = OracleTexture(px)
    uv = rng.uniform(-2, 3, (300, 2)).astype(np.float32)
    got = t.sample(uv, mode=NEAREST)
    u, v = np.mod(uv[:, 0], 1.0), np.mod(uv[:, 1], 1.0)
    xi = np.floor(u * 5).astype(int) % 5
    yi = np.floor(v * 7).astype(int) % 7
    assert np.array_equal(got, px[yi, xi])


@pytest.mark.parametrize("mode", [FILTERED, TRILINEAR, MIP_BILINEAR])
def test_filters_of_a_constant_return_the_constant(mode):
    t = OracleTexture(np.full((13, 9, 3), 0.375, np.float32))
    rng = np.random.default_rng(4)
    n = 500
    uv = rng.uniform(-1, 2, (n, 2))
    d0 = rng.normal(size=(n, 2)) * 10.0 ** rng.uniform(-4, 0, (n, 1))
    d1 = rng.normal(size=(n, 2)) * 10.0 ** rng.uniform(-4, 0, (n, 1))
    got = t.sample(uv, d0, d1, mode=mode)
    assert np.allclose(got, 0.375, rtol=2e-6)


def test_ewa_centred_on_a_texel_of_a_linear_ramp_returns_the_ramp():
    h, w = 64, 64
    y, x = np.mgrid[0:h, 0:w].astype(np.float32)
    ramp = (0.01 * x + 0.02 * y).astype(np.float32)
    t = OracleTexture(ramp)
    # texel (20, 30) centre; axis-aligned footprints of 0.5..1.9 texels keep
    # level 0 (lod < 1) and stay away from the wrap seam
    xs, ys = 20, 30
    uv = np.array([[(xs + 0.5) / w, (ys + 0.5) / h]] * 4, np.float32)
    d0 = np.array([[0.6 / w, 0], [1.2 / w, 0], [0.9 / w, 0], [1.9 / w, 0]], np.float32)
    d1 = np.array([[0, 0.5 / h], [0, 1.0 / h], [0, 0.9 / h], [0, 0.3 / h]], np.float32)
    got = t.sample(uv, d0, d1, mode=FILTERED)[:, 0]
    # lod < 1 blends levels 0 and 1; level 1's value at that point is the ramp too
    assert np.allclose(got, 0.01 * xs + 0.02 * ys, atol=2e-4)


def test_ewa_is_a_convex_combination_of_texels():
    rng = np.random.default_rng(5)
    px = rng.random((32, 32), dtype=np.float32)
    t = OracleTexture(px)
    n = 400
    uv = rng.uniform(0, 1, (n, 2)).astype(np.float32)
    d0 = (rng.normal(size=(n, 2)) / 32 * rng.uniform(0.1, 8, (n, 1))).astype(np.float32)
    d1 = (rng.normal(size=(n, 2)) / 32 * rng.uniform(0.1, 8, (n, 1))).astype(np.float32)
    got = t.sample(uv, d0, d1, mode=FILTERED)[:, 0]
    assert np.all(got >= px.min() - 1e-6) and np.all(got <= px.max() + 1e-6)


def test_ewa_level_follows_log2_of_the_shorter_axis():
    n = 64
    y, x = np.mgrid[0:n, 0:n]
    checker = ((x + y) % 2).astype(np.float32)  # one-texel cells: grey 0.5 from level 1 on
    t = OracleTexture(checker)
    uv = np.array([[0.3, 0.7]], np.float32)
    # two texels wide: lod = 6 + log2(2 / 64) = 1 exactly -> level 1 only
    two = t.sample(uv, [[2 / n, 0]], [[0, 2 / n]], mode=FILTERED)[0, 0]
    assert abs(two - 0.5) < 1e-6
    # a quarter texel: lod clamps to 0, level 0 blended with level 1 at t = 0
    fine = t.sample(uv, [[0.25 / n, 0]], [[0, 0.25 / n]], mode=FILTERED)[0, 0]
    assert abs(fine - 0.5) > 0.01


def test_ewa_anisotropic_footprint_filters_along_its_long_axis_only():
    n = 64
    x = np.arange(n)
    stripes = np.tile((x % 2).astype(np.float32), (n, 1))  # vertical one-texel stripes
    t = OracleTexture(stripes)
    uv = np.array([[(10 + 0.5) / n, 0.5]], np.float32)  # centre of the 0 stripe at x = 10
    # long axis along v (constant direction), short axis below a texel in u
    along_v = t.sample(uv, [[0, 6 / n]], [[0.2 / n, 0]], mode=FILTERED)[0, 0]
    # long axis along u: averages the stripes
    along_u = t.sample(uv, [[6 / n, 0]], [[0, 0.2 / n]], mode=FILTERED)[0, 0]
    assert along_v < 0.3
    assert abs(along_u - 0.5) < 0.1


def test_non_finite_derivatives_fall_back_to_bilinear():
    rng = np.random.default_rng(6)
    t = OracleTexture(rng.random((8, 8, 3), dtype=np.float32))
    uv = np.array([[0.3, 0.6]], np.float32)
    want = t.sample(uv, mode=BILINEAR)
    for mode in (FILTERED, TRILINEAR, MIP_BILINEAR):
        assert np.array_equal(t.sample(uv, [[np.inf, 0]], [[0, 0.1]], mode=mode), want)


def test_max_value_is_the_largest_channel():
    px = np.zeros((3, 4, 3), np.float32)
    px[1, 2, 1] = 7.5
    px[2, 0, 0] = -9.0
    assert OracleTexture(px).max_value == 7.5
    assert OracleTexture(np.full((2, 2), -1.0, np.float32)).max_value == 0.0

