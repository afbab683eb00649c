"""8-bit quantised nodes (no reference counterpart): every decoded child box,
evaluated in float32 exactly as the kernel does (origin + q * 2^e, one
rounding), contains the FP32 child box; the encoding is tight (one step
less would cut into the box); empty slots stay empty; children are kept."""
import numpy as np

import tpt
from tpt import scenes

LEAF_BIT = np.uint32(1 << 31)


def _decode(q):
    s = np.ldexp(np.float32(1.0), q["exp"].astype(np.int32)).astype(np.float32)  # (n,3)
    qq = q["q"].reshape(-1, 6, 4).astype(np.float32)  # lox, hix, loy, hiy, loz, hiz
    o = q["origin"].astype(np.float32)
    lo = np.stack([o[:, k, None] + qq[:, 2 * k] * s[:, k, None] for k in range(3)], 1)
    hi = np.stack([o[:, k, None] + qq[:, 2 * k + 1] * s[:, k, None] for k in range(3)], 1)
    return lo.astype(np.float32), hi.astype(np.float32), s


def _check(nodes):
    q = tpt.quantize_nodes(nodes)
    assert np.all(q["flags"] == 0)
    lo, hi, s = _decode(q)
    valid = nodes["children"] != LEAF_BIT
    flo = np.stack([nodes["min_x"], nodes["min_y"], nodes["min_z"]], 1)
    fhi = np.stack([nodes["max_x"], nodes["max_y"], nodes["max_z"]], 1)
    v3 = np.broadcast_to(valid[:, None, :], lo.shape)
    assert np.all(lo[v3] <= flo[v3]) and np.all(hi[v3] >= fhi[v3])
    # tightness: the next step inwards no longer contains the box
    qq = q["q"].reshape(-1, 6, 4).astype(np.int32)
    o = q["origin"]
    for k in range(3):
        ql, qh = qq[:, 2 * k], qq[:, 2 * k + 1]
        nl = (o[:, k, None] + np.minimum(ql + 1, 255).astype(np.float32) * s[:, k, None]).astype(np.float32)
        nh = (o[:, k, None] + np.maximum(qh - 1, 0).astype(np.float32) * s[:, k, None]).astype(np.float32)
        m = valid & (ql < 255)
        assert np.all(nl[m] > flo[:, k][m])
        m = valid & (qh > 0)
        assert np.all(nh[m] < fhi[:, k][m])
    assert np.array_equal(q["child"], nodes["children"] | np.where(valid, 0, LEAF_BIT).astype(np.uint32))
    return q


def test_sphere_field_nodes_contain_fp32_boxes():
    sf = scenes.sphere_field(n_spheres=30, seed=5)
    lo, hi = sf.mesh.triangle_bounds()
    _check(tpt.build_qbvh(lo, hi).nodes)


def test_extreme_coordinates_and_degenerate_boxes():
    rng = np.random.default_rng(3)
    c = rng.normal(size=(4000, 3)) * 10.0 ** rng.integers(-3, 6, (4000, 1))
    ext = rng.uniform(0, 1, (4000, 3)) * 10.0 ** rng.integers(-6, 2, (4000, 1))
    ext[:500] = 0.0  # zero-extent (flat / point) boxes
    q = tpt.build_qbvh(c - ext, c + ext)
    _check(q.nodes)


def test_layout_is_64_bytes_with_aligned_children():
    assert tpt.Q8_NODE_DTYPE.itemsize == 64
    assert tpt.Q8_NODE_DTYPE.fields["child"][1] == 48
