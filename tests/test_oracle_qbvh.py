"""Pins the oracle's QBVH build and traversal (R/qbvh.rs) with the reference's
own tests (R/qbvh.rs:686-1103): encoding, node invariants, root-leaf
special case, leaf coverage and brute-force closest box."""
import math

import numpy as np

import oracle

LEAF_BIT = 1 << 31
MAX_LEAF = 16


def decode(c):
    c = int(c)
    if c & LEAF_BIT:
        count = (c >> 24) & 0x7F
        return ("empty",) if count == 0 else ("leaf", c & ((1 << 24) - 1), count)
    return ("interior", c)


def boxes(rows):
    lo = np.array([r[0] for r in rows], np.float32)
    hi = np.array([r[1] for r in rows], np.float32)
    return lo, hi


def verify_invariants(q: oracle.OracleQbvh):
    """verify_invariants, R/qbvh.rs:1128-1157."""
    n = len(q.bmin)

    def subtree(ref, acc):
        kind = decode(ref)
        if kind[0] == "leaf":
            for k in range(kind[1], kind[1] + kind[2]):
                i = q.prim[k]
                acc[0] = np.minimum(acc[0], q.bmin[i])
                acc[1] = np.maximum(acc[1], q.bmax[i])
        elif kind[0] == "interior":
            for s in range(4):
                subtree(q.nodes[kind[1]]["children"][s], acc)

    total = 0
    for node in q.nodes:
        for s in range(4):
            ref = node["children"][s]
            kind = decode(ref)
            if kind[0] == "empty":
                continue
            acc = [np.full(3, np.inf, np.float32), np.full(3, -np.inf, np.float32)]
            subtree(ref, acc)
            cb_lo = np.array([node["min_x"][s], node["min_y"][s], node["min_z"][s]])
            cb_hi = np.array([node["max_x"][s], node["max_y"][s], node["max_z"][s]])
            assert np.all(cb_lo <= acc[0] + 1e-4) and np.all(cb_hi + 1e-4 >= acc[1])
            if kind[0] == "leaf":
                assert 1 <= kind[2] <= MAX_LEAF
                total += kind[2]
    if len(q.nodes) == 0:
        total = len(q.prim)
    assert total == n
    assert np.array_equal(np.sort(q.prim), np.arange(n))


def grid_bounds(n, f=(19, 11, 7), m=(41, 37, 31), size=0.4):
    rows = []
    for i in range(n):
        x, y, z = (i * f[0]) % m[0], (i * f[1]) % m[1], (i * f[2]) % m[2]
        rows.append(((x, y, z), (x + size, y + size, z + size)))
    return boxes(rows)


def slab(lo, hi, o, d):
    inv = np.array([np.inf if x == 0 else 1.0 / x for x in d], np.float64)
    with np.errstate(invalid="ignore"):
        t1, t2 = (lo - o) * inv, (hi - o) * inv
    tmin = max(np.fmax.reduce(np.fmin(t1, t2)), 0.0)
    tmax = np.fmin.reduce(np.fmax(t1, t2))
    return tmin if tmin <= tmax else None


def brute_closest(lo, hi, o, d):
    best = None
    for i in range(len(lo)):
        t = slab(lo[i].astype(np.float64), hi[i].astype(np.float64), np.asarray(o, np.float64), np.asarray(d))
        if t is not None and (best is None or t < best[1]):
            best = (i, t)
    return best


def test_empty_input_returns_none():  # qbvh.rs:692-695
    q = oracle.OracleQbvh(np.zeros((0, 3)), np.zeros((0, 3)))
    assert not q.h


def test_single_primitive_builds_root_leaf():  # qbvh.rs:697-703
    q = oracle.OracleQbvh(*boxes([((0, 0, 0), (1, 1, 1))]))
    assert len(q.nodes) == 0 and list(q.prim) == [0]


def test_build_separates_clusters():  # qbvh.rs:758-770
    rows = []
    for i in range(32):
        x, y, z = i % 4, (i // 4) % 4, i // 16
        rows.append(((x, y, z), (x + 0.5, y + 0.5, z + 0.5)))
    verify_invariants(oracle.OracleQbvh(*boxes(rows)))


def test_build_handles_coincident_primitives():  # qbvh.rs:772-779, 953-964
    for n in (20, 120):
        q = oracle.OracleQbvh(*boxes([((0, 0, 0), (1, 1, 1))] * n))
        verify_invariants(q)
        assert q.closest_box((0.5, 0.5, -5.0), (0, 0, 1)) is not None


def test_build_with_parallel_thresholds_preserves_invariants():  # qbvh.rs:781-794
    n = 8192 * 2 + 17
    rows = []
    for i in range(n):
        x, y, z = (i * 73) % 211, (i * 37) % 191, (i * 53) % 173
        rows.append(((x, y, z), (x + 0.5, y + 0.5, z + 0.5)))
    verify_invariants(oracle.OracleQbvh(*boxes(rows)))


def test_traverse_visits_every_box_under_the_ray():  # qbvh.rs:796-839
    lo, hi = grid_bounds(200)
    q = oracle.OracleQbvh(lo, hi)
    for i in range(32):
        o = (((i * 5) % 41) + 0.1, ((i * 9) % 37) + 0.1, 50.0)
        visited = set()
        for off, cnt in q.leaves(o, (0, 0, -1)):
            visited.update(int(q.prim[k]) for k in range(off, off + cnt))
        expected = {j for j in range(200) if lo[j][0] <= o[0] <= hi[j][0] and lo[j][1] <= o[1] <= hi[j][1]}
        assert expected <= visited


def test_closest_matches_brute_force_on_grid():  # qbvh.rs:902-951
    lo, hi = grid_bounds(200)
    q = oracle.OracleQbvh(lo, hi)
    center = 0.5 * (lo.min(0) + hi.max(0))
    dirs = [(0, 0, -1), (0, 0, 1), (-1, 0, 0), (1, 0, 0), (0, -1, 0), (0, 1, 0)]
    dirs += [tuple(np.array(v) / np.linalg.norm(v)) for v in ((1, 1, 1), (-1, 0.5, 0.3), (0.2, -0.7, 0.6))]
    hits = 0
    for d in dirs:
        d = np.array(d, np.float32)
        for i in range(50):
            jit = np.array([((i * 13) % 47) * 0.5 - 12, ((i * 17) % 43) * 0.5 - 11, ((i * 7) % 41) * 0.5 - 10])
            o = (center + jit - d * 80.0).astype(np.float32)
            bf = brute_closest(lo, hi, o, d)
            qb = q.closest_box(o, d)
            assert (bf is None) == (qb is None)
            if bf:
                tol = 1e-3 * max(abs(bf[1]), 1.0)
                assert abs(bf[1] - qb[1]) <= tol
                hits += 1
    assert hits > 0


def test_root_leaf_path_few_primitives():  # qbvh.rs:953-966
    for n in range(1, 5):
        q = oracle.OracleQbvh(*boxes([((i, i, i), (i + 0.5, i + 0.5, i + 0.5)) for i in range(n)]))
        assert len(q.nodes) == 0
        bf = brute_closest(q.bmin, q.bmax, (0.25, 0.25, -10.0), (0, 0, 1))
        qb = q.closest_box((0.25, 0.25, -10.0), (0, 0, 1))
        assert (bf and bf[0]) == (qb and qb[0])


def test_axis_aligned_rays():  # qbvh.rs:981-1007
    rows = []
    for i in range(32):
        x, y = i % 8, (i // 8) % 4
        rows.append(((x, y, 0), (x + 0.5, y + 0.5, 0.5)))
    q = oracle.OracleQbvh(*boxes(rows))
    cases = [((-10, 0.25, 0.25), (1, 0, 0)), ((10, 0.25, 0.25), (-1, 0, 0)), ((0.25, -10, 0.25), (0, 1, 0)),
             ((0.25, 10, 0.25), (0, -1, 0)), ((0.25, 0.25, -10), (0, 0, 1)), ((0.25, 0.25, 10), (0, 0, -1))]
    for o, d in cases:
        bf = brute_closest(q.bmin, q.bmax, o, d)
        qb = q.closest_box(o, d)
        assert (bf and bf[0]) == (qb and qb[0])


def test_ray_origin_inside_box():  # qbvh.rs:1009-1024
    q = oracle.OracleQbvh(*boxes([((0, 0, 0), (2, 2, 2)), ((5, 5, 5), (6, 6, 6)), ((-3, -3, -3), (-2, -2, -2)),
                                  ((10, 10, 10), (11, 11, 11)), ((20, 20, 20), (21, 21, 21))]))
    i, t = q.closest_box((1, 1, 1), (1, 0, 0))
    assert i == 0 and abs(t) < 1e-5


def test_skewed_distribution():  # qbvh.rs:1026-1064
    n = 500
    rows = []
    for i in range(n):
        t = i / n
        x, y, z = t * t * 100.0, math.sin(i) * 0.5, math.cos(i * 0.137) * 0.5
        rows.append(((x, y, z), (x + 0.2, y + 0.2, z + 0.2)))
    q = oracle.OracleQbvh(*boxes(rows))
    verify_invariants(q)
    dirs = [(1, 0, 0), (-1, 0, 0), tuple(np.array((1, 0.2, 0)) / np.linalg.norm((1, 0.2, 0))),
            tuple(np.array((-1, -0.1, 0.05)) / np.linalg.norm((-1, -0.1, 0.05)))]
    hits = 0
    for d in dirs:
        d = np.array(d, np.float32)
        for i in range(30):
            o = np.array([-50 + i * 0.3, math.sin(i) * 0.4, 0.0], np.float32)
            bf = brute_closest(q.bmin, q.bmax, o, d)
            qb = q.closest_box(o, d)
            assert (bf and bf[0]) == (qb and qb[0])
            hits += bf is not None
    assert hits > 0


def test_finds_correct_leaf_for_simple_grid():  # qbvh.rs:1066-1083
    q = oracle.OracleQbvh(*boxes([((2 * i, 0, 0), (2 * i + 1, 1, 1)) for i in range(16)]))
    visited = set()
    for off, cnt in q.leaves((20.5, 0.5, 5.0), (0, 0, -1)):
        visited.update(int(q.prim[k]) for k in range(off, off + cnt))
    assert 10 in visited


def test_leaf_children_visit_far_to_near():
    # R/qbvh.rs:650-671: within a node, leaves are handled in descending tmin.
    q = oracle.OracleQbvh(*boxes([((0, 0, z), (1, 1, z + 0.5)) for z in range(0, 80, 5)]))
    visits = q.leaves((0.5, 0.5, -10.0), (0, 0, 1))
    assert len(visits) > 1
