"""The conservative slab test of the any-hit walk and of 8-bit nodes accepts
every child the reference's slab test accepts (R/qbvh.rs:578-607).

The device test (slab4_cons in csrc/traverse.h) computes each near plane as
fma(p, ns, no) and each far plane as fma(p, fs, fo) from per-ray constants
with outward margins, and decodes 8-bit nodes ray-relative.  This file
restates that arithmetic in numpy float32 (a fused multiply-add is the
float64 product of two float32 values, exact, plus the addend, rounded once
to float32 by numpy after one float64 rounding: at most one extra half-ulp,
far inside the test's 16-ulp margins) and checks, on random and adversarial
rays and boxes, that no child the reference accepts is rejected.  The
reference side is the reference's own expression: fl(fl(p - o) * inv) with
compare-select min / max (x86 minps / maxps).
"""
import numpy as np

import tpt

F = np.float32
K_LO, K_HI = F(1.0 - 2.0 ** -20), F(1.0 + 2.0 ** -20)


def fma(a, b, c):
    return (a.astype(np.float64) * b.astype(np.float64) + c.astype(np.float64)).astype(F)


def safe_inv(d):
    with np.errstate(divide="ignore"):
        return np.where(d == 0, F(np.inf), F(1.0) / d).astype(F)


def sel_min(a, b):
    return np.where(a < b, a, b)


def sel_max(a, b):
    return np.where(a > b, a, b)


def reference_accept(lo, hi, o, inv, tmax):
    """lo, hi: (n, 3) child box; o, inv: (n, 3) ray; tmax: (n,)."""
    with np.errstate(invalid="ignore", over="ignore"):
        t1 = ((lo - o).astype(F) * inv).astype(F)
        t2 = ((hi - o).astype(F) * inv).astype(F)
    tmin = sel_min(t1, t2)
    tmx = sel_max(t1, t2)
    tn = sel_max(sel_max(sel_max(tmin[:, 0], tmin[:, 1]), tmin[:, 2]), F(0))
    tx = sel_min(sel_min(sel_min(tmx[:, 0], tmx[:, 1]), tmx[:, 2]), tmax)
    return tn <= tx


def cons_setup(o, inv):
    with np.errstate(over="ignore", invalid="ignore"):
        ai = np.abs(inv)
        ao = np.abs((o * inv).astype(F))
        ok = np.all((ai >= F(2.0 ** -60)) & (ai <= F(2.0 ** 60)) & (ao < F(2.0 ** 100)), axis=1)
        m = fma(ao, np.full_like(ao, 2.0 ** -21), np.full_like(ao, 2.0 ** -100))
        sn = (inv * K_LO).astype(F)
        sf = (inv * K_HI).astype(F)
        no = fma(-o, sn, -m)
        fo = fma(-o, sf, m)
    return ok, sn, no, sf, fo


def cons_accept_fp32(lo, hi, o, inv, tmax):
    ok, sn, no, sf, fo = cons_setup(o, inv)
    neg = inv < 0
    near = np.where(neg, hi, lo)
    far = np.where(neg, lo, hi)
    with np.errstate(over="ignore", invalid="ignore"):
        tn = fma(near, sn, no)
        tf = fma(far, sf, fo)
    l = np.maximum(np.maximum(tn[:, 0], tn[:, 1]), np.maximum(tn[:, 2], F(0)))
    h = np.minimum(np.minimum(tf[:, 0], tf[:, 1]), np.minimum(tf[:, 2], tmax))
    return ok, l <= h


def cons_accept_q8(origin, exp, qlo, qhi, o, inv, tmax):
    """origin (n,3) f32, exp (n,3) int, qlo/qhi (n,3) integer corners."""
    ok, sn_r, no, sf_r, fo = cons_setup(o, inv)
    neg = inv < 0
    with np.errstate(over="ignore", invalid="ignore"):
        sn = np.ldexp(sn_r, exp).astype(F)
        sf = np.ldexp(sf_r, exp).astype(F)
        bn = fma(origin, sn_r, no)
        bf = fma(origin, sf_r, fo)
        bn = fma(np.full_like(bn, -(2.0 ** -20)), fma(np.full_like(sn, 255.0), np.abs(sn), np.abs(bn)), bn)
        bf = fma(np.full_like(bf, 2.0 ** -20), fma(np.full_like(sf, 255.0), np.abs(sf), np.abs(bf)), bf)
        qn = np.where(neg, qhi, qlo).astype(F)
        qf = np.where(neg, qlo, qhi).astype(F)
        tn = fma(qn, sn, bn)
        tf = fma(qf, sf, bf)
    l = np.maximum(np.maximum(tn[:, 0], tn[:, 1]), np.maximum(tn[:, 2], F(0)))
    h = np.minimum(np.minimum(tf[:, 0], tf[:, 1]), np.minimum(tf[:, 2], tmax))
    return ok, l <= h


def random_cases(rng, n, scale):
    """Boxes near the ray, grazing and axis-touching rays, mixed magnitudes."""
    c = (rng.normal(size=(n, 3)) * scale).astype(F)
    half = (np.abs(rng.normal(size=(n, 3))) * scale * rng.choice([1e-6, 1e-3, 0.1, 1.0], size=(n, 1))).astype(F)
    lo = (c - half).astype(F)
    hi = (c + half).astype(F)
    o = (rng.normal(size=(n, 3)) * scale * rng.choice([0.0, 0.5, 2.0, 50.0], size=(n, 1))).astype(F)
    # aim at a point of the box (its corner, a face or inside), then perturb
    u = rng.choice([0.0, 1.0, 0.5, -1e-7, 1 + 1e-7], size=(n, 3)).astype(np.float64)
    u = np.where(rng.random((n, 3)) < 0.5, rng.random((n, 3)), u)
    target = lo + u * (hi - lo)
    d = (target - o).astype(F)
    d = np.where(rng.random((n, 3)) < 0.02, d * F(1e-7), d).astype(F)  # tiny components
    # origins exactly on a box plane
    on_plane = rng.random(n) < 0.1
    ax = rng.integers(0, 3, n)
    o[on_plane, ax[on_plane]] = np.where(rng.random(on_plane.sum()) < 0.5, lo[on_plane, ax[on_plane]],
                                         hi[on_plane, ax[on_plane]])
    tmax = np.where(rng.random(n) < 0.5, F(np.inf),
                    (rng.random(n) * 3).astype(F)).astype(F)
    return lo, hi, o, d, tmax


def test_conservative_fp32_slab_accepts_every_reference_child():
    rng = np.random.default_rng(7)
    total_ok = 0
    for scale in (1e-3, 1.0, 100.0, 1e4):
        lo, hi, o, d, tmax = random_cases(rng, 400_000, scale)
        inv = safe_inv(d)
        ref = reference_accept(lo, hi, o, inv, tmax)
        ok, got = cons_accept_fp32(lo, hi, o, inv, tmax)
        missed = ok & ref & ~got
        assert not missed.any(), (scale, np.flatnonzero(missed)[:5])
        total_ok += ok.sum()
    assert total_ok > 1_000_000  # most rays take the conservative test


def test_conservative_fp32_slab_margins_are_tight_on_generic_rays():
    """On rays and boxes of comparable scale the test accepts at most a few
    children per ten thousand beyond the reference's (efficiency, not
    correctness)."""
    rng = np.random.default_rng(3)
    n = 400_000
    for scale in (1.0, 100.0):
        c = (rng.normal(size=(n, 3)) * scale).astype(F)
        half = (np.abs(rng.normal(size=(n, 3))) * 0.1 * scale).astype(F)
        lo, hi = (c - half).astype(F), (c + half).astype(F)
        o = (rng.normal(size=(n, 3)) * scale * 3).astype(F)
        d = rng.normal(size=(n, 3)).astype(F)
        tmax = np.full(n, np.inf, F)
        inv = safe_inv(d)
        ref = reference_accept(lo, hi, o, inv, tmax)
        ok, got = cons_accept_fp32(lo, hi, o, inv, tmax)
        assert ok.all()
        extra = (got & ~ref).sum() / max(1, ref.sum())
        assert extra < 1e-3, (scale, extra)


def test_zero_direction_components_take_the_reference_test():
    o = np.array([[0.5, 0.5, -1.0]], F)
    d = np.array([[0.0, 0.0, 1.0]], F)
    ok, _, _, _, _ = cons_setup(o, safe_inv(d))
    assert not ok[0]


def test_conservative_q8_slab_accepts_every_reference_child():
    """8-bit nodes from the library's own quantiser (tpt_qbvh_quantize, host
    code): the ray-relative decode accepts every child the reference's test
    accepts on the FP32 box."""
    rng = np.random.default_rng(11)
    for scale, aligned in ((1e-2, False), (1.0, False), (300.0, False), (1.0, True), (300.0, True)):
        n_nodes = 50_000
        lo, hi, o, d, tmax = random_cases(rng, 4 * n_nodes, scale)
        if aligned:
            # boxes on the quantisation grid of their node: the decoded box is
            # the FP32 box, so only the decode's own margins separate the tests
            step = np.ldexp(F(1), np.floor(np.log2(scale * 4 / 255)).astype(int))
            org = (np.floor(lo.reshape(n_nodes, 4, 3).min(axis=1) / step) * step).astype(F)
            org4 = np.repeat(org, 4, axis=0)
            ql = np.clip(np.floor((lo - org4) / step), 0, 250).astype(np.int64)
            qh = np.clip(ql + 1 + rng.integers(0, 4, ql.shape), 0, 255)
            lo = (org4 + (ql * step).astype(F)).astype(F)
            hi = (org4 + (qh * step).astype(F)).astype(F)
            # top of each node exactly at origin + 255 step, so the quantiser keeps that step
            top = org4 + F(255) * F(step)
            first = np.zeros(4 * n_nodes, bool)
            first[::4] = True
            hi[first] = top[first]
            lo[first] = org4[first]
        nodes = np.zeros(n_nodes, tpt.NODE_DTYPE)
        for k, name in enumerate(("x", "y", "z")):
            nodes["min_" + name] = lo[:, k].reshape(n_nodes, 4)
            nodes["max_" + name] = hi[:, k].reshape(n_nodes, 4)
        nodes["children"] = np.arange(4 * n_nodes, dtype=np.uint32).reshape(n_nodes, 4) | np.uint32(0x80000000) \
            | np.uint32(1 << 24)
        nodes["valid_mask"] = 0xF
        q = tpt.quantize_nodes(nodes)
        assert not (q["flags"] & 1).any()
        qb = q["q"].reshape(n_nodes, 3, 2, 4)  # axis, (lo, hi), child
        origin = np.repeat(q["origin"], 4, axis=0)
        exp = np.repeat(q["exp"].astype(np.int32), 4, axis=0)
        qlo = qb[:, :, 0, :].transpose(0, 2, 1).reshape(-1, 3).astype(np.int32)
        qhi = qb[:, :, 1, :].transpose(0, 2, 1).reshape(-1, 3).astype(np.int32)
        # decoded boxes contain the FP32 boxes (the quantiser's guarantee)
        dec_lo = (origin + (qlo * np.ldexp(1.0, exp)).astype(F)).astype(F)
        dec_hi = (origin + (qhi * np.ldexp(1.0, exp)).astype(F)).astype(F)
        assert (dec_lo <= lo).all() and (dec_hi >= hi).all()
        inv = safe_inv(d)
        ref = reference_accept(lo, hi, o, inv, tmax)
        ok, got = cons_accept_q8(origin, exp, qlo, qhi, o, inv, tmax)
        missed = ok & ref & ~got
        assert not missed.any(), (scale, np.flatnonzero(missed)[:5])
