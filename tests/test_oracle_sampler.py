"""Pins the oracle's sampler and camera with the reference's own tests:
R/sampler/zsobol.rs:297-488 (OpenQMC examples, index uniqueness, Laine-Karras
and shuffle properties, (0,2)-sequence, Shirley remapping),
R/sampler/pcg.rs:35-55, R/scene/camera.rs:104-160."""
import ctypes

import numpy as np

import oracle

L = oracle.lib()
u32, u64, f32 = ctypes.c_uint32, ctypes.c_uint64, ctypes.c_float
for name, res, args in [
    ("orc_pcg_hash", u32, [u32]), ("orc_pcg_state_transition", u32, [u32]), ("orc_pcg_output", u32, [u32]),
    ("orc_laine_karras", u32, [u32, u32]), ("orc_reverse_and_shuffle", u32, [u32, u32]),
    ("orc_uint_to_float", f32, [u32]), ("orc_encode_morton_2", u64, [u32, u32]), ("orc_mix_bits", u64, [u64]),
    ("orc_shuffled_scrambled_sobol", None, [u32, u32, ctypes.c_int, ctypes.c_void_p]),
    ("orc_zsobol_sample_indices", None, [u32, u32, u32, u32, ctypes.c_void_p]),
    ("orc_aux_rng", None, [u32, ctypes.c_int, ctypes.c_void_p]),
    ("orc_path_sampler", None, [u32, u32, u32, u32, u32, u32, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]),
    ("orc_camera_rays", None, [ctypes.c_void_p, u32, u32, u32, u32, ctypes.c_void_p]),
]:
    getattr(L, name).restype = res
    getattr(L, name).argtypes = args

PRIMES = [2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37, 41, 43, 47, 53, 59, 61, 67, 71]
ONE_MINUS_EPS = np.frombuffer(np.uint32(0x3F7FFFFF).tobytes(), np.float32)[0]


def rev32(v):
    return int(f"{v:032b}"[::-1], 2)


def sobol(index, seed, depth):
    out = np.zeros(4, np.uint32)
    L.orc_shuffled_scrambled_sobol(index, seed, depth, out.ctypes.data)
    return out[:depth]


def test_uint_to_float_matches_openqmc_edge_cases():  # zsobol.rs:336-349
    f = L.orc_uint_to_float
    assert f(0) == 0.0 and 0.0 < f(1) < f(2)
    assert np.float32(f(0xFFFFFFFF)) == ONE_MINUS_EPS
    assert np.float32(f(0xFFFFFEFF)) < ONE_MINUS_EPS
    assert np.float32(f(0xFFFFFF00)) == ONE_MINUS_EPS
    assert f(0x7FFFFFFF) < 0.5 and f(0x80000000) == 0.5 and f(0x800000FF) == 0.5 and f(0x80000100) > 0.5


def test_encode_morton_2_interleaves():  # zsobol.rs:351-361
    m = L.orc_encode_morton_2
    assert [m(0, 0), m(1, 0), m(0, 1), m(1, 1), m(2, 0), m(0, 2), m(3, 3), m(4, 0)] == [0, 1, 2, 3, 4, 8, 15, 16]


def test_zsobol_sample_indices_are_unique_and_pixel_aligned():  # zsobol.rs:363-392
    rx, ry = 16, 9
    for log_s in range(0, 11):
        spp = 1 << log_s
        for dim in (0, 3, 6):
            out = np.zeros(rx * ry * spp, np.uint64)
            L.orc_zsobol_sample_indices(spp, rx, ry, dim, out.ctypes.data)
            assert len(np.unique(out)) == len(out)
            base = (out // spp).reshape(rx * ry, spp)
            assert np.all(base == base[:, :1])


def test_laine_karras_preserves_low_bits_under_high_bit_flip():  # zsobol.rs:394-414
    values = [0b01010101010101010011001100110011, 0b11111111000000001111000011110000,
              0b11111111111111110000000011111111, 0xFFFFFFFF, 0]
    for v in values:
        for p in PRIMES:
            a, b = L.orc_laine_karras(v, p), L.orc_laine_karras(v ^ 0x10000, p)
            assert a & 0xFFFF == b & 0xFFFF and a & ~0xFFFF != b & ~0xFFFF


def test_shuffle_is_a_permutation_over_low_bits():  # zsobol.rs:416-434
    for p in PRIMES:
        seen = set()
        for i in range(16):
            permuted = rev32(L.orc_reverse_and_shuffle(i, p))
            seen.add(permuted & 15)
        assert seen == set(range(16))


def test_sobol_is_a_0_2_sequence_for_256_samples():  # zsobol.rs:436-462
    seed = L.orc_pcg_hash(0)
    pts = np.array([sobol(i, seed, 2) for i in range(256)], np.uint64)
    for i in range(9):
        xr, yr = 1 << i, 1 << (8 - i)
        x = pts[:, 0] // (0xFFFFFFFF // xr)
        y = pts[:, 1] // (0xFFFFFFFF // yr)
        cells = x + y * xr
        assert len(np.unique(cells)) == 256


def test_sobol_shirley_remapping_property():  # zsobol.rs:464-488
    seed = L.orc_pcg_hash(0)
    w = 0xFFFFFFFF // 8
    pts = np.array([sobol(i, seed, 2) for i in range(64)], np.uint64)
    for i in range(8):
        ys = pts[pts[:, 0] // w == i][:, 1] // w
        assert sorted(ys.tolist()) == list(range(8))


def test_pcg_behaves_like_openqmc():  # pcg.rs:35-55
    st, out = L.orc_pcg_state_transition, L.orc_pcg_output
    assert st(0) != 0 and out(0) == 0
    for p in PRIMES:
        assert st(p) != p and out(p) != p and st(p) != out(p)
        state = (st(0) + p) & 0xFFFFFFFF  # init(seed)
        assert L.orc_pcg_hash(state) == out(st(state))  # hash == rng after init


def test_aux_rng_floats_are_24_bit_in_unit_interval():  # aux_rng.rs:26-28
    out = np.zeros(10000, np.float32)
    L.orc_aux_rng(12345, len(out), out.ctypes.data)
    assert np.all(out >= 0) and np.all(out < 1)
    assert np.all(out * 16777216.0 == np.floor(out * 16777216.0))
    assert abs(out.mean() - 0.5) < 0.01


def test_path_sampler_dimensions_are_stratified_per_pixel():
    # each path-vertex dimension over the spp samples of one pixel is a
    # scrambled Sobol (0,1)-sequence: one sample per 1/spp stratum
    spp, nd = 64, 3
    vals = np.zeros((spp, 2 + 14 * nd), np.float32)
    seeds = np.zeros((spp, 1 + nd), np.uint32)
    for s in range(spp):
        L.orc_path_sampler(32, 32, spp, 5, 7, s, nd, vals[s].ctypes.data, seeds[s].ctypes.data)
    for col in range(vals.shape[1]):
        strata = np.floor(vals[:, col] * spp).astype(int)
        assert len(np.unique(strata)) == spp, col
    assert len(np.unique(seeds)) == seeds.size


def _camera(eye, look_at, up, fov_deg, rx, ry):
    f32n = np.float32

    def norm(v):
        v = np.asarray(v, np.float32)
        return v * (f32n(1.0) / np.sqrt(np.dot(v, v, )).astype(np.float32))

    fwd = norm(np.asarray(look_at, np.float32) - np.asarray(eye, np.float32))
    right = norm(np.cross(fwd, np.asarray(up, np.float32)).astype(np.float32))
    iup = norm(np.cross(right, fwd).astype(np.float32))
    th = np.float32(np.tan(np.float32(0.5) * np.float32(np.radians(fov_deg))))
    aspect = np.float32(rx) / np.float32(ry)
    dx = np.float32(2) * aspect * th / np.float32(rx)
    dy = np.float32(2) * th / np.float32(ry)
    spread = np.float32(np.arctan(max(dx, dy)))
    return np.concatenate([np.asarray(eye, np.float32), fwd, right, iup, [th, spread]]).astype(np.float32)


def test_camera_rays():  # camera.rs:104-160
    cam = _camera((0, 0, 5), (0, 0, 0), (0, 1, 0), 60.0, 3, 3)
    out = np.zeros((9, 13), np.float32)
    L.orc_camera_rays(cam.ctypes.data, 3, 3, 1, 0, out.ctypes.data)
    centre = out[4]
    assert np.allclose(centre[0:3], (0, 0, 5))
    assert np.allclose(centre[3:6], (0, 0, -1), atol=0.2)  # jittered inside the centre pixel
    cam = _camera((0, 0, 5), (0, 0, 0), (0, 1, 0), 45.0, 64, 64)
    out = np.zeros((64 * 64, 13), np.float32)
    L.orc_camera_rays(cam.ctypes.data, 64, 64, 1, 0, out.ctypes.data)
    assert np.allclose(np.linalg.norm(out[:, 3:6], axis=1), 1.0, atol=1e-6)
    assert np.all(np.any(out[:, 6:9] != out[:, 3:6], axis=1)) and np.all(np.any(out[:, 9:12] != out[:, 3:6], axis=1))
    assert np.all(out[:, 12] > 0)
