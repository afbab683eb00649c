"""Pins the oracle's two-level closest hit and shadow decision against the
reference's scene tests (R/scene/mod.rs:1028-1105) and a brute force over
every triangle of every instance (float64 classification)."""
import numpy as np

import oracle
import tpt
from tpt import scenes

UNIT_V = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
UNIT_I = np.array([[0, 1, 2]], np.uint32)


def _scene(instances):
    o = oracle.OracleScene()
    m = o.add_mesh(UNIT_V, UNIT_I)
    for mat in instances:
        o.add_instance(m, mat, np.linalg.inv(mat))
    o.build()
    return o


def _ray(o, d):
    return tpt.make_rays(np.array([o], np.float32), np.array([d], np.float32))


def test_closest_hit_returns_the_nearest_triangle():  # scene/mod.rs:1058-1086
    m = np.eye(4)
    m[2, 3] = -1.0
    h = _scene([np.eye(4), m]).closest_hit(_ray((0.25, 0.25, 2.0), (0, 0, -1)))
    assert h["instance"][0] == 0 and h["triangle"][0] == 0 and abs(h["t"][0] - 2.0) < 1e-6


def test_closest_hit_handles_scaled_instances():  # scene/mod.rs:1088-1110
    h = _scene([np.diag([2.0, 2.0, 2.0, 1.0])]).closest_hit(_ray((0.5, 0.5, 1.0), (0, 0, -1)))
    assert abs(h["t"][0] - 1.0) < 1e-6


def test_closest_hit_requires_bvh_build():  # scene/mod.rs:1112-1128
    o = oracle.OracleScene()
    m = o.add_mesh(UNIT_V, UNIT_I)
    o.add_instance(m, np.eye(4), np.eye(4))
    try:
        o.closest_hit(_ray((0.25, 0.25, 1.0), (0, 0, -1)))
    except RuntimeError as e:
        assert "QbvhNotBuilt" in str(e)
    else:
        raise AssertionError("expected QbvhNotBuilt")


def test_no_instances_misses():  # scene/mod.rs:346-348
    o = oracle.OracleScene()
    o.add_mesh(UNIT_V, UNIT_I)
    o.build()
    h = o.closest_hit(_ray((0.25, 0.25, 1.0), (0, 0, -1)))
    assert h["triangle"][0] == 0xFFFFFFFF


def _brute(scene: tpt.Scene, rays):
    """Closest hit over all triangles of all instances, float64 world space."""
    best_t = np.full(len(rays), np.inf)
    best = np.full((len(rays), 2), -1)
    o = rays["o"].astype(np.float64)
    d = rays["d"].astype(np.float64)
    for ii, inst in enumerate(scene.instances):
        m = scene.meshes[inst.mesh_index]
        v = m.positions[m.indices].astype(np.float64)
        l2w = inst.local_to_world.astype(np.float64)
        v = v @ l2w[:3, :3].T + l2w[:3, 3]
        e1 = v[:, 1] - v[:, 0]
        e2 = v[:, 2] - v[:, 0]
        for r in range(len(rays)):
            pvec = np.cross(d[r], e2)
            det = np.einsum("ij,ij->i", e1, pvec)
            with np.errstate(divide="ignore", invalid="ignore"):
                inv = 1.0 / det
                tvec = o[r] - v[:, 0]
                u = np.einsum("ij,ij->i", tvec, pvec) * inv
                qv = np.cross(tvec, e1)
                w = (qv @ d[r]) * inv
                t = np.einsum("ij,ij->i", e2, qv) * inv
            ok = (u >= 0) & (w >= 0) & (u + w <= 1) & (t > 1e-6) & np.isfinite(t)
            if ok.any():
                k = np.argmin(np.where(ok, t, np.inf))
                if t[k] < best_t[r]:
                    best_t[r] = t[k]
                    best[r] = (ii, k)
    return best, best_t


def test_cornell_matches_brute_force():
    cs = scenes.cornell()
    o = oracle.oracle_scene_from(cs.scene)
    rng = np.random.default_rng(3)
    org = rng.uniform(0.05, 0.95, (400, 3)).astype(np.float32)
    d = rng.normal(size=(400, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = tpt.make_rays(org, d.astype(np.float32))
    h = o.closest_hit(rays)
    best, bt = _brute(cs.scene, rays)
    # the box is open at z=0: rays leaving through it miss, all others hit
    hit = h["triangle"] != 0xFFFFFFFF
    assert np.array_equal(hit, best[:, 0] >= 0)
    assert hit.mean() > 0.5
    same = (h["instance"] == best[:, 0]) & (h["triangle"] == best[:, 1])
    # disagreements only where two triangles are hit at (almost) the same t
    assert np.all(np.abs(h["t"][hit & ~same] - bt[hit & ~same]) <= 1e-5 * np.maximum(1, bt[hit & ~same]))
    assert np.allclose(h["t"][hit], bt[hit], rtol=1e-5, atol=1e-6)


def test_unoccluded_decision_rule():  # integrator/mod.rs:239-265
    # blocker triangle at z=1 between origin (z=2) and a target at z=0
    o = oracle.OracleScene()
    m = o.add_mesh(UNIT_V, UNIT_I)
    o.add_instance(m, np.eye(4), np.eye(4))  # instance 0 at z=0 (target)
    up = np.eye(4)
    up[2, 3] = 1.0
    o.add_instance(m, up, np.linalg.inv(up))  # instance 1 at z=1 (blocker)
    o.build()
    r_block = _ray((0.25, 0.25, 2.0), (0, 0, -1))
    r_free = _ray((0.25, 0.25, 0.5), (0, 0, -1))
    r_miss = _ray((5.0, 5.0, 2.0), (0, 0, -1))
    tgt = np.array([[0, 0]], np.uint32)
    assert not o.unoccluded(r_block, tgt)[0]
    assert o.unoccluded(r_free, tgt)[0]
    assert o.unoccluded(r_miss, tgt)[0]  # a miss counts as unoccluded (reference K1)
    # finite distance: hit at t=1 vs distance 2 -> occluded; vs distance 1.0005 -> tolerance
    assert not o.unoccluded(r_block, None, np.array([2.0], np.float32))[0]
    assert o.unoccluded(r_block, None, np.array([1.0005], np.float32))[0]
    # infinite light: any hit occludes
    assert not o.unoccluded(r_block, None, np.array([np.inf], np.float32))[0]
    assert o.unoccluded(r_miss)[0]
