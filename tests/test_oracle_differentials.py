"""Ray differentials and ray cones of the oracle, pinned by what the geometry
fixes (the reference has no tests for them):

* compute_shading_differentials (R/scene/mod.rs:748-819): the offset rays'
  intersections with the tangent plane, and dudx... solving dp = dpdu du +
  dpdv dv, against float64 closed forms on planes with orthogonal and skewed
  partials;
* its ray-cone fallback (:821-859): width = width0 + spread t projected by
  1 / cos(theta), over |dpdu| and |dpdv|;
* spawn_scattered_ray (R/integrator/mod.rs:58-75): the cone's width at the
  hit and its spread plus the lobe's, so a mirror (spread 0) keeps it;
* the specular ray differential (:88-157): on a flat mirror the reflected
  offset directions are the mirror images of the incoming ones and the offset
  origins the offset rays' hit points; on a sphere it agrees with the exact
  reflection of the offset rays to first order (the error shrinks with the
  square of the offset);
* the pinhole camera's differentials (R/scene/camera.rs:48-96): the offset
  rays' footprint on a plane at distance D is the pixel pitch there times
  max(1/sqrt(spp), 0.125), and the camera cone's spread is the angle of one
  pixel.
"""
import ctypes

import numpy as np

import oracle

F = np.float32


def _lib():
    L = oracle.lib()
    vp = ctypes.c_void_p
    L.orc_shading_differentials.argtypes = [vp] * 6
    L.orc_spawn_scattered.argtypes = [vp, ctypes.c_float, vp, ctypes.c_int, ctypes.c_float, vp, vp]
    L.orc_camera_rays.argtypes = [vp] + [ctypes.c_uint32] * 4 + [vp]
    return L


def ray_rec(o, d, rx=None, ry=None, cone=(0.0, 0.0)):
    r = np.zeros(21, F)
    r[0:3], r[3:6] = o, d
    if rx is not None:
        r[6] = 1.0
        r[7:10], r[13:16] = rx[0], rx[1]
        r[10:13], r[16:19] = ry[0], ry[1]
    r[19:21] = cone
    return r


def shading_diff(p, n, dpdu, dpdv, ray):
    out = np.zeros(10, F)
    a = [np.ascontiguousarray(x, F) for x in (p, n, dpdu, dpdv, ray)]
    _lib().orc_shading_differentials(*[x.ctypes.data for x in a], out.ctypes.data)
    return out


def spawn(ray, hit_t, vtx, delta, cone_spread, wi):
    out = np.zeros(21, F)
    r, v, w = (np.ascontiguousarray(x, F) for x in (ray, vtx, wi))
    _lib().orc_spawn_scattered(r.ctypes.data, hit_t, v.ctypes.data, int(delta), cone_spread, w.ctypes.data,
                               out.ctypes.data)
    return out


def unit(v):
    v = np.asarray(v, np.float64)
    return v / np.linalg.norm(v)


def plane_hit(o, d, p, n):
    t = np.dot(n, p - o) / np.dot(n, d)
    return o + t * d


def solve_uv(dp, dpdu, dpdv):
    a = np.array([[dpdu @ dpdu, dpdu @ dpdv], [dpdu @ dpdv, dpdv @ dpdv]])
    return np.linalg.solve(a, np.array([dpdu @ dp, dpdv @ dp]))


def test_differentials_on_planes_match_the_geometry():
    rng = np.random.default_rng(1)
    for case in range(200):
        n = unit(rng.normal(size=3))
        t1 = unit(np.cross(n, rng.normal(size=3)))
        t2 = np.cross(n, t1)
        # orthogonal or skewed partials of different lengths in the plane
        dpdu = rng.uniform(0.5, 3.0) * t1
        dpdv = rng.uniform(0.5, 3.0) * (t2 if case % 2 == 0 else unit(t2 + rng.uniform(-0.8, 0.8) * t1))
        p = rng.normal(size=3)
        o = p + unit(n + 0.7 * rng.normal(size=3)) * rng.uniform(1, 5)
        if np.dot(n, o - p) < 0.3:
            continue
        d = unit(p - o)
        rxd = unit(d + 1e-3 * rng.normal(size=3))
        ryd = unit(d + 1e-3 * rng.normal(size=3))
        ray = ray_rec(o, d, (o, rxd), (o, ryd))
        got = shading_diff(p, n, dpdu, dpdv, ray)
        px, py = plane_hit(o, rxd, p, n), plane_hit(o, ryd, p, n)
        du_x, dv_x = solve_uv(px - p, dpdu, dpdv)
        du_y, dv_y = solve_uv(py - p, dpdu, dpdv)
        want = np.array([du_x, dv_x, du_y, dv_y])
        scale = np.abs(want).max()
        assert np.allclose(got[:4], want, rtol=0, atol=2e-3 * scale), (case, got[:4], want)
        assert np.allclose(got[4:7], px - p, atol=2e-3 * np.linalg.norm(px - p))
        assert np.allclose(got[7:10], py - p, atol=2e-3 * np.linalg.norm(py - p))


def test_cone_footprint_is_width_over_cosine():
    p = np.array([0.2, -0.1, 0.0])
    n = np.array([0.0, 0.0, 1.0])
    dpdu, dpdv = np.array([2.0, 0.0, 0.0]), np.array([0.0, 0.5, 0.0])
    for theta_deg in (0.0, 30.0, 60.0, 89.99):
        th = np.radians(theta_deg)
        d = np.array([np.sin(th), 0.0, -np.cos(th)])
        dist = 3.0
        o = p - dist * d
        w0, spread = 0.01, 0.002
        got = shading_diff(p, n, dpdu, dpdv, ray_rec(o, d, cone=(w0, spread)))
        width = w0 + spread * dist
        proj = width / max(np.cos(th), 1e-3)  # cos_theta clamped at 1e-3 (R/scene/mod.rs:835)
        assert np.isclose(got[0], proj / 2.0, rtol=1e-5), (theta_deg, got[0], proj / 2.0)
        assert np.isclose(got[3], proj / 0.5, rtol=1e-5)
        assert got[1] == 0.0 and got[2] == 0.0
        assert np.allclose(got[4:7], got[0] * dpdu, rtol=1e-6) and np.allclose(got[7:10], got[3] * dpdv, rtol=1e-6)
    # no cone, no differential: no footprint
    got = shading_diff(p, n, dpdu, dpdv, ray_rec(p + n, -n))
    assert np.all(got == 0.0)


def _vertex(p, ng, ns, wo, dpdx=(0, 0, 0), dpdy=(0, 0, 0), dndu=(0, 0, 0), dndv=(0, 0, 0), uvd=(0, 0, 0, 0)):
    return np.concatenate([np.asarray(x, np.float64) for x in (p, ng, ns, wo, dpdx, dpdy, dndu, dndv, uvd)])


def test_spawned_cone_grows_by_the_lobe_spread_and_a_mirror_keeps_it():
    o, d = np.zeros(3), unit([0.0, 0.0, -1.0])
    p = np.array([0.0, 0.0, -2.0])
    n = np.array([0.0, 0.0, 1.0])
    v = _vertex(p, n, n, -d)
    wi = unit([0.3, 0.0, 1.0])
    for w0, spread, lobe in ((0.01, 0.002, 0.5), (0.0, 0.01, 0.0), (0.05, 0.0, 0.6)):
        for delta in (False, True):
            out = spawn(ray_rec(o, d, cone=(w0, spread)), 2.0, v, delta, lobe, wi)
            assert np.isclose(out[19], w0 + spread * 2.0, rtol=1e-6)
            assert np.isclose(out[20], spread + lobe, rtol=1e-6)
            assert np.allclose(out[3:6], wi, atol=1e-7)
    # a mirror bounce (cone spread 0) keeps the incoming spread
    out = spawn(ray_rec(o, d, cone=(0.0, 0.0042)), 2.0, v, True, 0.0, wi)
    assert np.isclose(out[20], 0.0042, rtol=1e-6)


def reflect(w, n):
    return w - 2.0 * np.dot(w, n) * n


def test_flat_mirror_reflects_the_offset_rays():
    rng = np.random.default_rng(4)
    n = np.array([0.0, 0.0, 1.0])
    for _ in range(50):
        o = np.array([rng.uniform(-1, 1), rng.uniform(-1, 1), rng.uniform(1, 3)])
        p = np.array([rng.uniform(-0.5, 0.5), rng.uniform(-0.5, 0.5), 0.0])
        d = unit(p - o)
        rxd, ryd = unit(d + 2e-3 * rng.normal(size=3)), unit(d + 2e-3 * rng.normal(size=3))
        px, py = plane_hit(o, rxd, p, n), plane_hit(o, ryd, p, n)
        wi = reflect(d, n)
        v = _vertex(p, n, n, -d, dpdx=px - p, dpdy=py - p)
        out = spawn(ray_rec(o, d, (o, rxd), (o, ryd)), float(np.linalg.norm(p - o)), v, True, 0.0, wi)
        assert out[6] == 1.0
        # directions: the mirror images of the incoming offset directions
        assert np.allclose(out[13:16], reflect(rxd, n), atol=2e-6)
        assert np.allclose(out[16:19], reflect(ryd, n), atol=2e-6)
        # origins: the offset rays' hit points, moved off the surface by the
        # spawn offset (1e-4 along the normal on the side wi leaves)
        assert np.allclose(out[7:10], px + 1e-4 * n, atol=2e-6)
        assert np.allclose(out[10:13], py + 1e-4 * n, atol=2e-6)


def test_curved_mirror_differential_is_first_order_exact():
    """A sphere of radius R: at the hit, dn/du = dp/du / R (the outward
    normal of a sphere moves with the point), so the specular differential
    predicts the reflected offset ray; the exact reflection of the offset ray
    at its own hit point differs from the prediction by O(offset^2)."""
    R = 2.0
    c = np.zeros(3)
    errs = []
    for eps in (4e-3, 2e-3, 1e-3):
        o = np.array([0.3, 0.2, 5.0])
        p = unit([0.1, 0.05, 1.0]) * R
        n = unit(p - c)
        d = unit(p - o)
        t1 = unit(np.cross(n, [0.0, 1.0, 0.0]))
        t2 = np.cross(n, t1)
        dpdu, dpdv = t1, t2  # unit-speed parameterisation at p
        dndu, dndv = dpdu / R, dpdv / R
        rxd = unit(d + eps * t1)

        def sphere_hit(oo, dd):
            b = np.dot(oo - c, dd)
            cc = np.dot(oo - c, oo - c) - R * R
            return oo + (-b - np.sqrt(b * b - cc)) * dd

        px = sphere_hit(o, rxd)
        exact_dir = reflect(rxd, unit(px - c))
        # the vertex's screen-space derivatives from the tangent plane
        pxp = plane_hit(o, rxd, p, n)
        du, dv = solve_uv(pxp - p, dpdu, dpdv)
        v = _vertex(p, n, n, -d, dpdx=pxp - p, dpdy=pxp - p, dndu=dndu, dndv=dndv, uvd=(du, dv, du, dv))
        out = spawn(ray_rec(o, d, (o, rxd), (o, rxd)), float(np.linalg.norm(p - o)), v, True, 0.0, reflect(d, n))
        pred = unit(out[13:16].astype(np.float64))
        errs.append(np.linalg.norm(pred - exact_dir))
        # first order: far better than ignoring the curvature
        flat = unit(reflect(rxd, n))
        assert errs[-1] < 0.2 * np.linalg.norm(flat - exact_dir), (eps, errs[-1])
    # halving the offset quarters the error (O(eps^2)), within float noise
    assert errs[1] < 0.35 * errs[0] and errs[2] < 0.35 * errs[1], errs


def test_pinhole_differential_footprint_on_a_plane():
    from tpt import camera
    W = H = 64
    for spp in (1, 16, 256):
        cam = camera.pinhole(eye=(0.0, 0.0, 0.0), look_at=(0.0, 0.0, -1.0), up=(0.0, 1.0, 0.0), fov_y_deg=40.0,
                             width=W, height=H)
        c = camera.camera_array(cam)
        out = np.zeros((H * W, 13), F)
        _lib().orc_camera_rays(c.ctypes.data, W, H, spp, 0, out.ctypes.data)
        k = (H // 2) * W + W // 2  # a pixel at the centre
        d, rxd, ryd = (out[k, a:a + 3].astype(np.float64) for a in (3, 6, 9))
        D = 10.0
        n = np.array([0.0, 0.0, 1.0])
        p0 = plane_hit(np.zeros(3), d, np.array([0.0, 0.0, -D]), n)
        px = plane_hit(np.zeros(3), rxd, np.array([0.0, 0.0, -D]), n)
        py = plane_hit(np.zeros(3), ryd, np.array([0.0, 0.0, -D]), n)
        pitch = 2.0 * D * np.tan(np.radians(20.0)) / H
        scale = max(1.0 / np.sqrt(spp), 0.125)
        assert np.isclose(np.linalg.norm(px - p0), pitch * scale, rtol=2e-3), (spp, np.linalg.norm(px - p0))
        assert np.isclose(np.linalg.norm(py - p0), pitch * scale, rtol=2e-3)
        assert abs(px[1] - p0[1]) < 1e-3 * pitch and abs(py[0] - p0[0]) < 1e-3 * pitch
        # the camera cone: one pixel's angle (pixel_spread_angle)
        assert np.isclose(cam.pixel_spread, np.arctan(2.0 * np.tan(np.radians(20.0)) / H), rtol=1e-5)
