"""Pins the oracle's OpenPBR transmission (oracle/bsdf.h): the specular BTDF,
rough and smooth, its multiple-scattering compensation, and thin walls.

1. The reference's own tests of these lobes (R/bsdf/open_pbr.rs:2431-2642):
   the multiple-scattering transmission ratio's reciprocity constraint, the
   thin-walled smooth BTDF passing straight through, and the smooth BTDF
   refracting by the transmission IOR.
2. What the mathematics fixes, in float64: Snell's law and total internal
   reflection for smooth refraction; the smooth sample's weight carries
   (1 - Fresnel) / eta_rel^2 with the textbook unpolarised Fresnel term;
   the thin-wall slab's reflectance and transmittance are the textbook
   geometric series (R + T = 1 without absorption, the absorbing slab's
   closed form with it); rough samples' density is the pdf (histogram
   against quadrature of pdf); sample() weights are f |cos| / pdf; a rough
   interface never gains energy (R + eta_rel^2 T <= 1, radiance scaling
   undone); transmission samples leave below the surface and carry the
   relative IOR.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle as orc
from tpt.material import OpenPbr


def fresnel(cos_i, eta):
    """Unpolarised dielectric Fresnel reflectance from the eta_i = 1 side."""
    cos_i = np.clip(np.asarray(cos_i, np.float64), 0.0, 1.0)
    sin_t2 = (1.0 - cos_i**2) / eta**2
    cos_t = np.sqrt(np.maximum(1.0 - sin_t2, 0.0))
    rs = (cos_i - eta * cos_t) / (cos_i + eta * cos_t)
    rp = (eta * cos_i - cos_t) / (eta * cos_i + cos_t)
    return np.where(sin_t2 >= 1.0, 1.0, 0.5 * (rs**2 + rp**2))


def unit(v):
    v = np.asarray(v, np.float64)
    return (v / np.linalg.norm(v)).astype(np.float32)


def sphere(n):
    """Midpoint grid over the whole sphere: directions and the cell solid angle."""
    z = -1 + (np.arange(2 * n) + 0.5) / n
    phi = (np.arange(2 * n) + 0.5) * (np.pi / n)
    zz, pp = np.meshgrid(z, phi, indexing="ij")
    r = np.sqrt(np.maximum(1 - zz * zz, 0))
    wi = np.stack([r * np.cos(pp), r * np.sin(pp), zz], -1).reshape(-1, 3)
    return wi.astype(np.float32), (1.0 / n) * (np.pi / n)


GLASS = OpenPbr(base_weight=0.0, transmission_weight=1.0, specular_roughness=0.0)


# ---- the reference's tests -----------------------------------------------------------------
@pytest.mark.parametrize("eta", [1.1, 1.5, 2.4])
@pytest.mark.parametrize("roughness", [0.35, 0.65, 0.88, 1.0])
def test_ms_transmission_ratio_reciprocity(eta, roughness):
    """dielectric_transmission_ms_ratio_satisfies_reciprocity_constraint
    (R/bsdf/open_pbr.rs:2431-2449): (1 - r_enter) / (1 - E_avg(eta)) =
    (1 - r_exit) eta^2 / (1 - E_avg(1 / eta))."""
    m = OpenPbr.new((0.8, 0.6, 0.4))
    alpha = roughness * roughness
    _, r_in, ome_in = orc.openpbr_probe(m, 4, a=alpha, b=eta)
    _, r_out, ome_out = orc.openpbr_probe(m, 4, a=alpha, b=1.0 / eta)
    enter = (1.0 - r_in) / ome_in
    leave = (1.0 - r_out) * eta * eta / ome_out
    assert abs(enter - leave) <= 1e-4, (enter, leave)


def test_thin_walled_smooth_btdf_passes_straight_through():
    """thin_walled_specular_btdf_yields_pass_through (R/bsdf/open_pbr.rs:2560-2584)."""
    m = OpenPbr(base_weight=0.0, specular_weight=0.0, transmission_weight=1.0, specular_roughness=0.0,
                thin_walled=True)
    wo = unit([0.2, -0.3, 0.9327379])
    u = np.random.default_rng(0).random((32, 3), dtype=np.float32)
    s = orc.openpbr_sample(m, wo, u)
    t = s["ok"] & s["transmission"]
    assert t.any()
    assert np.allclose(s["wi"][t], -wo, atol=1e-5)
    assert np.all(s["delta"][t]) and np.all(s["eta"][t] == 1.0)


def test_smooth_transmission_refracts_by_the_transmission_ior():
    """smooth_transmission_eta_controls_refraction_direction
    (R/bsdf/open_pbr.rs:2618-2642), with Snell's law in float64."""
    wo = unit([0.35, 0.0, 0.93675])
    u = np.random.default_rng(1).random((64, 3), dtype=np.float32)
    s = orc.openpbr_sample(GLASS, wo, u)
    t = s["ok"] & s["transmission"]
    assert t.any()
    eta = 1.0 / 1.5
    w = wo.astype(np.float64)
    expected = np.array([-eta * w[0], -eta * w[1], -np.sqrt(1 - eta**2 * (1 - w[2] ** 2))])
    assert np.allclose(s["wi"][t], expected, atol=1e-5)
    assert np.allclose(s["eta"][t], eta, rtol=1e-6)


# ---- float64 pins ----------------------------------------------------------------------------
@pytest.mark.parametrize("ior", [1.33, 1.5, 2.0])
@pytest.mark.parametrize("front", [True, False])
def test_smooth_refraction_snell_and_total_internal_reflection(ior, front):
    """Front face: eta_rel = 1 / ior; back face (inside the medium): ior,
    with no refraction past the critical angle."""
    m = OpenPbr(base_weight=0.0, transmission_weight=1.0, specular_roughness=0.0, specular_ior=ior)
    eta = 1.0 / ior if front else ior
    rng = np.random.default_rng(int(ior * 100) + front)
    for theta in np.linspace(0.05, 1.5, 12):
        wo = unit([np.sin(theta) * np.cos(1.0), np.sin(theta) * np.sin(1.0), np.cos(theta)])
        s = orc.openpbr_sample(m, wo, rng.random((200, 3), dtype=np.float32), front_face=front)
        t = s["ok"] & s["transmission"]
        sin_t = eta * np.sin(theta)
        if sin_t >= 1.0:
            assert not t.any(), theta
            continue
        assert t.any(), theta
        w = wo.astype(np.float64)
        expected = np.array([-eta * w[0], -eta * w[1], -np.sqrt(1 - sin_t**2)])
        assert np.allclose(s["wi"][t], expected, atol=2e-6)


@pytest.mark.parametrize("ior", [1.33, 1.5, 2.0])
def test_smooth_transmission_weight_carries_fresnel_and_radiance_scaling(ior):
    """A smooth glass sample's weight: refraction (1 - F) / eta_rel^2 and
    reflection F over their lobe probabilities, which split as the layer
    weights 1 and 1 - E_spec (E_spec, the smooth specular albedo, is the
    Fresnel term up to the table's interpolation): the transmitted weight is
    (1 - F)(2 - E) / eta_rel^2 and the reflected one F (2 - E)."""
    m = OpenPbr(base_weight=0.0, transmission_weight=1.0, specular_roughness=0.0, specular_ior=ior)
    rng = np.random.default_rng(7)
    for theta in (0.1, 0.6, 1.0, 1.3):
        wo = unit([np.sin(theta), 0.0, np.cos(theta)])
        s = orc.openpbr_sample(m, wo, rng.random((400, 3), dtype=np.float32))
        f = float(fresnel(np.cos(theta), ior))
        t = s["ok"] & s["transmission"]
        r = s["ok"] & ~s["transmission"]
        assert t.any() and r.any()
        wt = s["weight"][t][:, 0].astype(np.float64)
        wr = s["weight"][r][:, 0].astype(np.float64)
        assert np.ptp(wt) < 1e-6 * wt.max() and np.ptp(wr) < 1e-6 * wr.max()
        two_minus_e_r = wr[0] / f  # (2 - E) from the reflected weight
        assert 1.0 < two_minus_e_r <= 2.0
        expected_t = (1.0 - f) * two_minus_e_r * ior**2
        assert abs(wt[0] - expected_t) <= 1e-4 * expected_t, (theta, wt[0], expected_t)
        # E_spec is the Fresnel term up to the table's interpolation
        assert abs((2.0 - two_minus_e_r) - f) < 0.02, (theta, 2.0 - two_minus_e_r, f)


@pytest.mark.parametrize("ior", [1.2, 1.5, 2.2])
def test_thin_wall_slab_closed_forms(ior):
    """thin_wall_coefficients against the textbook slab: with interface
    reflectance F (the same from both sides) and one-pass transmittance a,
    T = (1 - F)^2 a / (1 - F^2 a^2), R = F + (1 - F)^2 F a^2 / (1 - F^2 a^2);
    R + T = 1 without absorption.  At grazing angles the internal Fresnel
    term is recomputed through 1 - eta^2 sin^2(theta_i), which cancels to a
    few 1e-4 in FP32, so the tolerance widens below cos 0.1."""
    clear = OpenPbr(transmission_weight=1.0, specular_ior=ior, thin_walled=True)
    tint = (0.9, 0.5, 0.2)
    tinted = clear.with_(transmission_color=tint)
    for c in np.linspace(0.02, 1.0, 25):
        f = float(fresnel(c, ior))
        r = orc.openpbr_probe(clear, 5, a=float(c)).astype(np.float64)
        t = orc.openpbr_probe(clear, 6, a=float(c)).astype(np.float64)
        assert np.allclose(r + t, 1.0, atol=2e-6), (c, r, t)
        assert np.allclose(t, (1 - f) / (1 + f), atol=2e-6)
        cos_i = np.sqrt(1.0 - (1.0 - c * c) / ior**2)
        a = np.asarray(tint, np.float64) ** (1.0 / cos_i)
        rt = orc.openpbr_probe(tinted, 5, a=float(c)).astype(np.float64)
        tt = orc.openpbr_probe(tinted, 6, a=float(c)).astype(np.float64)
        den = 1.0 - f * f * a * a
        tol = 2e-5 if c >= 0.1 else 1e-3
        assert np.allclose(tt, (1 - f) ** 2 * a / den, rtol=tol, atol=1e-7), c
        assert np.allclose(rt, f + (1 - f) ** 2 * f * a * a / den, rtol=tol, atol=1e-7), c


def test_thin_wall_alphas():
    """thin_wall_transmission_alpha_xy: alpha scaled by 3.7 (eta-1)(eta-1/2)^2 / eta^3."""
    for ior in (1.1, 1.5, 2.0):
        for rough in (0.2, 0.6):
            m = OpenPbr(transmission_weight=1.0, specular_ior=ior, specular_roughness=rough, thin_walled=True)
            ax, ay, _ = orc.openpbr_probe(m, 7)
            alpha = rough * rough
            k = 3.7 * (ior - 1) * (ior - 0.5) ** 2 / ior**3
            assert abs(ax - np.clip(alpha * k, 1e-4, 1.0)) < 1e-6 and ax == ay


ROUGH = [
    OpenPbr(base_weight=0.0, transmission_weight=1.0, specular_roughness=0.3),
    OpenPbr(base_weight=0.0, transmission_weight=1.0, specular_roughness=0.5, specular_ior=1.33),
    OpenPbr(base_color=(0.5, 0.4, 0.3), transmission_weight=0.6, transmission_color=(0.8, 0.6, 0.9),
            specular_roughness=0.4, coat_weight=0.3, coat_roughness=0.2),
    OpenPbr(base_color=(0.7, 0.7, 0.2), transmission_weight=0.8, specular_roughness=0.45, thin_walled=True),
]


@pytest.mark.parametrize("mi", range(len(ROUGH)))
def test_rough_samples_follow_the_pdf(mi):
    """The directions sample() produces have the density pdf() states: a
    histogram of 400k samples over (cos theta, phi) cells of the sphere
    against the float64 quadrature of pdf over each cell.  Front faces: the
    reference's BTDF pdf also assigns density to directions whose
    generalised half vector faces away from them (no refraction reaches
    those); their share is small at these roughnesses and is excluded here
    by comparing only cells whose expected count is large."""
    m = ROUGH[mi]
    wo = unit([0.45, 0.2, 0.87])
    n = 400_000
    s = orc.openpbr_sample(m, wo, np.random.default_rng(11 + mi).random((n, 3), dtype=np.float32))
    ok = s["ok"] & ~s["delta"]
    wi = s["wi"][ok].astype(np.float64)
    nz, nphi = 24, 16
    iz = np.clip(((wi[:, 2] + 1) * 0.5 * nz).astype(int), 0, nz - 1)
    iphi = np.clip(((np.arctan2(wi[:, 1], wi[:, 0]) + np.pi) / (2 * np.pi) * nphi).astype(int), 0, nphi - 1)
    counts = np.bincount(iz * nphi + iphi, minlength=nz * nphi).astype(np.float64)
    grid, dw = sphere(480)
    _, pdf = orc.openpbr_eval(m, wo, grid)
    g = grid.astype(np.float64)
    gz = np.clip(((g[:, 2] + 1) * 0.5 * nz).astype(int), 0, nz - 1)
    gphi = np.clip(((np.arctan2(g[:, 1], g[:, 0]) + np.pi) / (2 * np.pi) * nphi).astype(int), 0, nphi - 1)
    expected = np.bincount(gz * nphi + gphi, weights=pdf.astype(np.float64) * dw, minlength=nz * nphi) * n
    big = expected > 400
    assert big.sum() > 20
    dev = np.abs(counts[big] - expected[big]) / np.sqrt(expected[big])
    rel = np.abs(counts[big] - expected[big]) / expected[big]
    assert np.all((dev < 6.0) | (rel < 0.03)), (dev.max(), rel.max())


@pytest.mark.parametrize("mi", range(len(ROUGH)))
@pytest.mark.parametrize("front", [True, False])
def test_rough_sample_weights_are_f_cos_over_pdf(mi, front):
    """finalize_rough_sample: weight = eval |cos| / pdf and the returned pdf is
    pdf(wo, wi) (the thin-walled rough BTDF returns its own lobe's value and
    density instead, sample_spec_btdf :1700-1727, and is left out);
    transmission samples go below the surface, reflections above."""
    m = ROUGH[mi]
    wo = unit([0.3, -0.4, 0.86])
    s = orc.openpbr_sample(m, wo, np.random.default_rng(3 + mi).random((3000, 3), dtype=np.float32),
                           front_face=front)
    ok = s["ok"] & ~s["delta"]
    assert ok.any()
    assert np.all(s["wi"][ok & s["transmission"]][:, 2] < 0)
    assert np.all(s["wi"][ok & ~s["transmission"]][:, 2] > 0)
    sel = ok & ~(s["transmission"] & m.thin_walled)
    f, pdf = orc.openpbr_eval(m, wo, s["wi"][sel], front_face=front)
    assert np.allclose(s["pdf"][sel], pdf, rtol=1e-6)
    expected = f.astype(np.float64) * np.abs(s["wi"][sel][:, 2:3]) / pdf[:, None]
    assert np.allclose(s["weight"][sel], expected, rtol=2e-5, atol=1e-8)
    eta = 1.0 / 1.5 if front else 1.5
    if not m.thin_walled and m.specular_ior == 1.5:
        assert np.allclose(s["eta"][sel & s["transmission"]], eta, rtol=1e-6)


@pytest.mark.parametrize("rough", [0.3, 0.5, 0.8])
@pytest.mark.parametrize("front", [True, False])
def test_rough_interface_does_not_gain_energy(rough, front):
    """R + eta_rel^2 T <= 1 over the sphere (the radiance scaling 1 / eta_rel^2
    of the BTDF undone), and the interface transmits more than nothing."""
    m = OpenPbr(base_weight=0.0, transmission_weight=1.0, specular_roughness=rough)
    eta = 1.0 / 1.5 if front else 1.5
    grid, dw = sphere(300)
    up = grid[:, 2] > 0
    for theta in (0.2, 0.8, 1.3):
        wo = unit([np.sin(theta), 0.0, np.cos(theta)])
        f, _ = orc.openpbr_eval(m, wo, grid, front_face=front)
        c = np.abs(grid[:, 2]).astype(np.float64)
        r = (f[up, 0] * c[up]).sum() * dw
        t = (f[~up, 0] * c[~up]).sum() * dw
        assert r + eta * eta * t <= 1.01, (theta, r, t)
        assert t > 0.0
