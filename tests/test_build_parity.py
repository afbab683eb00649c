"""The product's host QBVH builder (libtpt, threaded) produces exactly the
oracle's tree (the reference build restated): same nodes byte for byte and
same primitive order, for any thread count."""
import math

import numpy as np
import pytest

import oracle
import tpt
from tpt import scenes


def _cmp(lo, hi, threads=0):
    q = tpt.build_qbvh(lo, hi, threads)
    o = oracle.OracleQbvh(lo, hi)
    assert len(q.nodes) == len(o.nodes)
    a = q.nodes.view(np.uint8).reshape(len(q.nodes), 128)[:, :113]
    b = o.nodes.view(np.uint8).reshape(len(o.nodes), 128)[:, :113]
    assert np.array_equal(a, b)
    assert np.array_equal(q.primitive_indices, o.prim)
    return q


def _cases():
    rng = np.random.default_rng(1)
    c = rng.uniform(0, 50, (5000, 3))
    yield "random", (c - rng.uniform(0, 1, (5000, 3))), (c + rng.uniform(0, 1, (5000, 3)))
    yield "coincident", np.zeros((120, 3)), np.ones((120, 3))
    yield "single", np.zeros((1, 3)), np.ones((1, 3))
    yield "three", np.array([[0, 0, 0], [1, 1, 1], [2, 2, 2]]), np.array([[.5, .5, .5], [1.5, 1.5, 1.5], [3, 3, 3]])
    n = 500
    lo = np.array([[(i / n) ** 2 * 100, math.sin(i) * 0.5, math.cos(i * 0.137) * 0.5] for i in range(n)])
    yield "skewed", lo, lo + 0.2
    pts = rng.uniform(0, 1, (3000, 3))
    pts[:1500, 0] = 0.5  # half of the centroids coincide on one axis
    yield "flat", pts, pts + 1e-3
    sf = scenes.sphere_field(n_spheres=16, seed=9)
    lo, hi = sf.mesh.triangle_bounds()
    yield "spheres", lo, hi


@pytest.mark.parametrize("name,lo,hi", list(_cases()), ids=lambda x: x if isinstance(x, str) else "")
def test_builder_matches_oracle(name, lo, hi):
    _cmp(np.asarray(lo, np.float32), np.asarray(hi, np.float32))


def test_thread_count_does_not_change_the_tree():
    sf = scenes.sphere_field(n_spheres=30, seed=2)
    lo, hi = sf.mesh.triangle_bounds()
    a = tpt.build_qbvh(lo, hi, threads=1)
    b = tpt.build_qbvh(lo, hi, threads=8)
    assert np.array_equal(a.nodes.view(np.uint8), b.nodes.view(np.uint8))
    assert np.array_equal(a.primitive_indices, b.primitive_indices)


def test_scene_build_matches_oracle_for_every_mesh_and_tlas():
    cs = scenes.cornell()
    o = oracle.oracle_scene_from(cs.scene)
    for i, m in enumerate(cs.scene.meshes):
        nodes, prim = o.mesh_qbvh(i)
        assert np.array_equal(nodes.view(np.uint8), m.qbvh.nodes.view(np.uint8))
        assert np.array_equal(prim, m.qbvh.primitive_indices)
    nodes, prim = o.tlas()
    assert np.array_equal(nodes.view(np.uint8), cs.scene.qbvh.nodes.view(np.uint8))
    assert np.array_equal(prim, cs.scene.qbvh.primitive_indices)


def test_not_built_scene_refuses_queries():  # ClosestHitError::QbvhNotBuilt, R/scene/mod.rs:100-113
    cs = scenes.cornell(build=False)
    with pytest.raises(tpt.ClosestHitError) as e:
        cs.scene.upload(None)
    assert e.value.kind is tpt.ClosestHitError.Kind.QbvhNotBuilt
