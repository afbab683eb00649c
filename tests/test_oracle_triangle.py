"""Pins the oracle's intersect_triangle (R/math/ray.rs:107-209).

1. The reference's own known answers (R/math/ray.rs:229-321), value for value.
2. difference_of_products / gamma against exact rational arithmetic.
3. A float64 brute force: on random triangles and rays the watertight test
   agrees with an exact (fractions) barycentric classification away from
   edges, and t agrees with the exact plane distance to a few ulp.  A
   dropped term, sign or permutation error in the restatement fails here.
"""
from fractions import Fraction

import numpy as np
import pytest

import oracle

UNIT = ((0, 0, 0), (1, 0, 0), (0, 1, 0))


def test_returns_hit():  # ray.rs:235-251
    h = oracle.intersect_triangle((0.25, 0.25, 1.0), (0, 0, -1), np.inf, *UNIT)
    assert h is not None
    t, b0, b1, b2 = h
    assert abs(t - 1.0) < 1e-6
    assert abs(b0 - 0.5) < 1e-6 and abs(b1 - 0.25) < 1e-6 and abs(b2 - 0.25) < 1e-6
    p = b0 * np.array(UNIT[0]) + b1 * np.array(UNIT[1]) + b2 * np.array(UNIT[2])
    assert np.allclose(p, (0.25, 0.25, 0.0), atol=1e-6)


def test_none_outside():  # ray.rs:253-259
    assert oracle.intersect_triangle((1.25, 1.25, 1.0), (0, 0, -1), np.inf, *UNIT) is None


def test_none_parallel():  # ray.rs:261-267
    assert oracle.intersect_triangle((0.25, 0.25, 1.0), (1, 0, 0), np.inf, *UNIT) is None


def test_none_degenerate():  # ray.rs:269-278
    assert oracle.intersect_triangle((0.25, 0, 1.0), (0, 0, -1), np.inf, (0, 0, 0), (1, 0, 0), (2, 0, 0)) is None


def test_hits_on_edge():  # ray.rs:280-290
    t, b0, b1, b2 = oracle.intersect_triangle((0.5, 0.0, 1.0), (0, 0, -1), np.inf, *UNIT)
    assert abs(b0 - 0.5) < 1e-6 and abs(b1 - 0.5) < 1e-6 and abs(b2) < 1e-6


def test_hits_on_vertex():  # ray.rs:292-301
    t, b0, b1, b2 = oracle.intersect_triangle((0.0, 0.0, 1.0), (0, 0, -1), np.inf, *UNIT)
    assert abs(b0 - 1.0) < 1e-6 and abs(b1) < 1e-6 and abs(b2) < 1e-6


def test_respects_t_max():  # ray.rs:303-310
    assert oracle.intersect_triangle((0.25, 0.25, 2.0), (0, 0, -1), 1.5, *UNIT) is None
    assert oracle.intersect_triangle((0.25, 0.25, 2.0), (0, 0, -1), 2.5, *UNIT) is not None


def test_rejects_hit_at_origin():  # ray.rs:312-318
    assert oracle.intersect_triangle((0.25, 0.25, 0.0), (0, 0, -1), np.inf, *UNIT) is None


def test_accepts_backface():  # ray.rs:320-321 (+ body)
    t, *_ = oracle.intersect_triangle((0.25, 0.25, -1.0), (0, 0, 1), np.inf, *UNIT)
    assert abs(t - 1.0) < 1e-6


def test_difference_of_products_is_exactly_rounded():
    rng = np.random.default_rng(0)
    for _ in range(2000):
        a, b, c, d = (np.float32(x) for x in rng.normal(size=4) * 10 ** rng.uniform(-3, 3, 4))
        got = oracle.lib().orc_difference_of_products(a, b, c, d)
        exact = Fraction(float(a)) * Fraction(float(b)) - Fraction(float(c)) * Fraction(float(d))
        # Kahan's algorithm: error <= 1.5 ulp of the result
        if exact == 0:
            assert got == 0.0
            continue
        ulp = np.spacing(np.float32(abs(float(exact))))
        assert abs(Fraction(float(got)) - exact) <= Fraction(1.5) * Fraction(float(ulp))


def test_gamma_matches_definition():
    eps = Fraction(2) ** -24  # f32::EPSILON * 0.5
    for n in (2, 3, 5, 7):
        exact = n * eps / (1 - n * eps)
        got = oracle.lib().orc_gamma(n)
        assert abs(Fraction(float(got)) - exact) <= exact * Fraction(2) ** -23


def _exact_classify(o, d, v0, v1, v2):
    """Exact barycentric edge functions of the ray/triangle (rational): the
    signs of the three scalar triple products decide the hit."""
    F = [Fraction(float(x)) for x in (*o, *d, *v0, *v1, *v2)]
    o, d = np.array(F[0:3], dtype=object), np.array(F[3:6], dtype=object)
    a, b, c = (np.array(F[6 + 3 * k:9 + 3 * k], dtype=object) - o for k in range(3))

    def triple(x, y, z):
        return (x[0] * (y[1] * z[2] - y[2] * z[1]) - x[1] * (y[0] * z[2] - y[2] * z[0])
                + x[2] * (y[0] * z[1] - y[1] * z[0]))

    e0, e1, e2 = triple(d, b, c), triple(d, c, a), triple(d, a, b)
    return e0, e1, e2


def test_fp64_brute_force_classification_and_t():
    rng = np.random.default_rng(42)
    agree = 0
    for _ in range(1500):
        v = rng.uniform(-1, 1, (3, 3)).astype(np.float32)
        o = rng.uniform(-3, 3, 3).astype(np.float32)
        target = (v.mean(0) + rng.normal(scale=0.6, size=3)).astype(np.float64)
        d = (target - o) / np.linalg.norm(target - o)
        d = d.astype(np.float32)
        h = oracle.intersect_triangle(o, d, np.inf, *v)
        e0, e1, e2 = _exact_classify(o, d, *v)
        inside = (e0 >= 0 and e1 >= 0 and e2 >= 0) or (e0 <= 0 and e1 <= 0 and e2 <= 0)
        s = e0 + e1 + e2
        if s == 0:
            continue
        # exact t along the ray of the plane intersection
        n = np.cross((v[1] - v[0]).astype(np.float64), (v[2] - v[0]).astype(np.float64))
        t_exact = float(np.dot(n, v[0].astype(np.float64) - o) / np.dot(n, d.astype(np.float64)))
        margin = min(abs(e0), abs(e1), abs(e2)) / abs(s)
        if margin < 1e-4:
            continue  # near an edge: the float test may round either way
        if inside and t_exact > 1e-3:
            assert h is not None, (o, d, v)
            assert abs(h[0] - t_exact) <= 8 * np.spacing(np.float32(t_exact)) + 1e-6 * abs(t_exact)
            assert abs(h[1] + h[2] + h[3] - 1.0) < 1e-5
            agree += 1
        elif not inside:
            assert h is None
            agree += 1
    assert agree > 1000


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_axis_permutation_is_consistent(axis):
    # the same hit seen along each major axis (kz = 0, 1, 2 branches)
    perm = np.roll(np.arange(3), axis)
    v = np.array(UNIT, np.float32)[:, perm]
    o = np.array([0.25, 0.25, 1.0], np.float32)[perm]
    d = np.array([0, 0, -1], np.float32)[perm]
    t, b0, b1, b2 = oracle.intersect_triangle(o, d, np.inf, *v)
    assert abs(t - 1.0) < 1e-6 and abs(b0 - 0.5) < 1e-6
