"""Pins the oracle's OpenPBR BSDF and GGX tables (oracle/bsdf.h).

1. The reference's own OpenPBR BSDF tests (R/bsdf/open_pbr.rs:2298-2560),
   restated on the oracle: white-furnace bounds for rough dielectric and
   metal specular, diffuse + specular, low and zero specular weight, the
   multiple-scattering transmission reciprocity constraint, F82 Fresnel at
   normal incidence and at cos = 1/7, and its hemispherical average.
2. The GGX tables (R/bsdf/directional_albedo.rs) against what they estimate,
   computed here independently in float64: Fresnel in the smooth cells,
   hemisphere / sphere quadrature of the GGX reflection BRDF and of the
   Walter et al. rough-dielectric BTDF in the rough cells, and the
   cosine-weighted averages.
3. BSDF invariants: the pdf integrates to one, sample() returns eval/pdf
   consistent values, the mean sample weight is the directional albedo.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle as orc
from tpt.material import OpenPbr

WO = np.array([0.25, -0.15, 0.956556], np.float32)
WO = (WO / np.linalg.norm(WO)).astype(np.float32)


def hemisphere(n=128):
    """integrate_upper_hemisphere_vec3's midpoint grid (R/bsdf/mod.rs:45-63)."""
    z = (np.arange(n) + 0.5) / n
    phi = (np.arange(n) + 0.5) * (2 * np.pi / n)
    zz, pp = np.meshgrid(z, phi, indexing="ij")
    r = np.sqrt(np.maximum(1 - zz * zz, 0))
    wi = np.stack([r * np.cos(pp), r * np.sin(pp), zz], -1).reshape(-1, 3)
    return wi.astype(np.float32), (1.0 / n) * (2 * np.pi / n)


WI, DW = hemisphere()


def energy(mat, wo=WO):
    f, _ = orc.openpbr_eval(mat, wo, WI)
    return (f.astype(np.float64) * WI[:, 2:3]).sum(0) * DW


def pdf_mass(mat, wo=WO):
    _, pdf = orc.openpbr_eval(mat, wo, WI)
    return float(pdf.astype(np.float64).sum() * DW)


def dielectric(r, weight=1.0, eta=1.5, base=0.0, base_color=(0.0, 0.0, 0.0), diffuse_roughness=0.0):
    return OpenPbr(base_weight=base, base_color=base_color, specular_weight=weight, specular_roughness=r,
                   specular_ior=eta, base_diffuse_roughness=diffuse_roughness)


# ---- the reference's tests -------------------------------------------------------------------
def test_default_diffuse_dominant_evaluates_finite():
    m = OpenPbr(base_weight=0.8, base_color=(0.8, 0.6, 0.4), specular_roughness=0.2)
    wi = np.array([0.2, 0.3, 0.9327379], np.float32)
    f, pdf = orc.openpbr_eval(m, (0, 0, 1), wi / np.linalg.norm(wi))
    assert np.all(np.isfinite(f)) and np.isfinite(pdf).all()


@pytest.mark.parametrize("r", [0.35, 0.65, 0.88, 1.0])
def test_rough_dielectric_specular_within_white_furnace(r):
    e = energy(dielectric(r))
    assert np.all(np.isfinite(e)) and e.max() <= 1.0 + 1.0e-2


@pytest.mark.parametrize("r", [0.35, 0.65, 0.88, 1.0])
def test_rough_diffuse_plus_specular_does_not_gain_energy(r):
    e = energy(dielectric(r, base=1.0, base_color=(1.0, 1.0, 1.0), diffuse_roughness=1.0))
    assert np.all(np.isfinite(e)) and e.max() <= 1.0 + 1.0e-2


def test_nonmetal_rough_specular_stays_near_dielectric_reflectance():
    spec = energy(dielectric(0.88, diffuse_roughness=1.0))
    assert 0.005 <= spec[0] <= 0.06 and np.allclose(spec, spec[0], atol=1e-5)
    total = energy(dielectric(0.88, base=1.0, base_color=(0.28, 0.04, 0.1), diffuse_roughness=1.0))
    assert total.max() <= 1.0 + 1.0e-2


def test_low_weight_specular_remains_low_energy():
    assert energy(dielectric(0.88, weight=0.01, diffuse_roughness=1.0)).max() <= 0.004


def test_zero_weight_specular_vanishes():
    assert energy(dielectric(0.88, weight=0.0)).max() <= 1.0e-4


@pytest.mark.parametrize("eta", [1.1, 1.5, 2.4])
@pytest.mark.parametrize("r", [0.35, 0.65, 0.88, 1.0])
def test_dielectric_ms_transmission_reciprocity(eta, r):
    m = OpenPbr()
    a = r * r
    ent = orc.openpbr_probe(m, 4, a=a, b=eta)
    ext = orc.openpbr_probe(m, 4, a=a, b=1.0 / eta)
    enter = (1 - ent[1]) / ent[2]
    exit_ = (1 - ext[1]) * eta * eta / ext[2]
    assert abs(enter - exit_) <= 1.0e-4


@pytest.mark.parametrize("r", [0.35, 0.65, 0.88, 1.0])
def test_rough_metal_within_white_furnace(r):
    m = OpenPbr(base_weight=1.0, base_color=(1.0, 1.0, 1.0), base_metalness=1.0, specular_roughness=r)
    e = energy(m)
    assert np.all(np.isfinite(e)) and e.max() <= 1.0 + 1.5e-2


def test_pure_metal_samples_reflection():
    m = OpenPbr(base_metalness=1.0, specular_roughness=float(np.sqrt(0.2)))
    u = np.random.default_rng(0).random((64, 3), dtype=np.float32)
    s = orc.openpbr_sample(m, WO, u)
    # a VNDF-sampled microfacet can reflect below the horizon (no sample)
    assert s["ok"].mean() > 0.9 and np.all(s["wi"][s["ok"], 2] > 0)


F82 = OpenPbr(base_weight=0.8, base_color=(0.9, 0.45, 0.2), specular_weight=0.7, specular_color=(0.4, 1.0, 0.25))


def schlick(f0, c):
    return f0 + (1 - f0) * (1 - c) ** 5


def test_metal_fresnel_uses_f82_tint():
    bc, sc = np.array(F82.base_color), np.array(F82.specular_color)
    f0 = 0.8 * bc
    assert np.allclose(orc.openpbr_probe(F82, 0, a=1.0), 0.7 * f0, atol=1e-6)
    assert np.allclose(orc.openpbr_probe(F82, 0, a=1.0 / 7.0), 0.7 * sc * schlick(f0, 1 / 7), atol=1e-5)


def test_metal_fresnel_average_is_the_hemispherical_average():
    avg = orc.openpbr_probe(F82, 1)
    f0, sc = 0.8 * np.array(F82.base_color), np.array(F82.specular_color)
    b = (schlick(f0, 1 / 7) - sc * schlick(f0, 1 / 7)) / ((1 / 7) * (6 / 7) ** 6)
    assert np.allclose(avg, np.clip(0.7 * (f0 + (1 - f0) / 21 - b / 126), 0, 1), atol=1e-6)
    # independent: 2 * integral F(mu) mu dmu by the midpoint rule in float64
    mu = (np.arange(4096) + 0.5) / 4096
    quad = sum(2 * m * orc.openpbr_probe(F82, 0, a=float(m)).astype(np.float64) for m in mu[::16]) / len(mu[::16])
    assert np.allclose(avg, quad, atol=2e-3)


# ---- GGX tables against what they estimate -----------------------------------------------------
def fresnel64(c, eta):
    c = np.clip(c, 0, 1)
    s = np.sqrt(np.maximum(1 - c * c, 0)) / eta
    ct = np.sqrt(np.maximum(1 - s * s, 0))
    rp = (eta * c - ct) / (eta * c + ct)
    rs = (c - eta * ct) / (c + eta * ct)
    return np.where(s >= 1, 1.0, 0.5 * (rp * rp + rs * rs))


def ggx64(h, ax, ay):
    t = h[..., 0] ** 2 / ax ** 2 + h[..., 1] ** 2 / ay ** 2 + h[..., 2] ** 2
    return np.where(h[..., 2] > 0, 1 / (np.pi * ax * ay * t * t), 0.0)


def lambda64(w, ax, ay):
    c = np.abs(w[..., 2])
    return 0.5 * (-1 + np.sqrt(1 + (ax ** 2 * w[..., 0] ** 2 + ay ** 2 * w[..., 1] ** 2) / (c * c)))


def sphere_grid(n, lower=False):
    z = (np.arange(n) + 0.5) / n
    phi = (np.arange(2 * n) + 0.5) * (np.pi / n)
    zz, pp = np.meshgrid(z, phi, indexing="ij")
    r = np.sqrt(1 - zz * zz)
    w = np.stack([r * np.cos(pp), r * np.sin(pp), -zz if lower else zz], -1)
    return w.reshape(-1, 3), (1.0 / n) * (np.pi / n)


def reflect_albedo64(wo, ax, ay, eta, n=1024):
    wi, dw = sphere_grid(n)
    h = wo + wi
    h /= np.linalg.norm(h, axis=-1, keepdims=True)
    g2 = 1 / (1 + lambda64(wo[None], ax, ay) + lambda64(wi, ax, ay))
    f = ggx64(h, ax, ay) * g2 * fresnel64(np.abs(h @ wo), eta) / (4 * wo[2] * wi[:, 2])
    return float((f * wi[:, 2]).sum() * dw)


def lut_cell_wo(ci, pi, ncos=64, nphi=16):
    sc = (ci + 0.5) / ncos
    c = sc * sc
    s = np.sqrt(max(1 - c * c, 0))
    phi = (pi + 0.5) / nphi * (np.pi / 2)
    return np.array([s * np.cos(phi), s * np.sin(phi), c])


@pytest.fixture(scope="module")
def albedo15():
    return orc.lut_albedo(1.5)


def test_albedo_table_smooth_cells_are_fresnel(albedo15):
    for ri in (0, 1):  # alpha below 1e-3 in both stored anisotropy slices
        for ci in (0, 17, 40, 63):
            c = lut_cell_wo(ci, 5)[2]
            assert np.allclose(albedo15[ci, 5, ri], fresnel64(c, 1.5), atol=2e-6)


def vdc(i):
    return int("{:032b}".format(i)[::-1], 2) * 2.3283064e-10


def albedo_estimator64(wo, ax, ay, eta):
    """The table's 64-direction QMC estimator (:292-341) in float64."""
    i = np.arange(64)
    z = (i + 0.5) / 64
    phi = 2 * np.pi * np.array([vdc(k) for k in i])
    r = np.sqrt(1 - z * z)
    wi = np.stack([r * np.cos(phi), r * np.sin(phi), z], -1)
    h = wo + wi
    h /= np.linalg.norm(h, axis=-1, keepdims=True)
    g2 = 1 / (1 + lambda64(wo[None], ax, ay) + lambda64(wi, ax, ay))
    f = ggx64(h, ax, ay) * g2 * fresnel64(np.abs(h @ wo), eta) / (4 * wo[2] * wi[:, 2])
    return float(np.clip(np.mean(f * z * 2 * np.pi), 0, 1))


def cell_alphas(r, an):
    a = 2 * (an + 0.5) / 16 - 1
    alpha, asp = r * r, np.sqrt(1 - 0.9 * abs(a))
    ax, ay = (alpha / asp, alpha * asp) if a >= 0 else (alpha * asp, alpha / asp)
    return float(np.clip(ax, 1e-4, 1.0)), float(np.clip(ay, 1e-4, 1.0))


@pytest.mark.parametrize("ci,pi,ri", [(8, 2, 20), (30, 9, 40), (63, 15, 63), (50, 0, 30), (3, 12, 5)])
def test_albedo_table_is_its_qmc_estimator(albedo15, ci, pi, ri):
    wo = lut_cell_wo(ci, pi)
    for slot, an in ((0, 7), (1, 8)):
        ax, ay = cell_alphas((ri + 0.5) / 64, an)
        ref = albedo_estimator64(wo, ax, ay, 1.5)
        assert abs(albedo15[ci, pi, ri, slot] - ref) <= 2e-5 + 2e-4 * ref


@pytest.mark.parametrize("ci,pi,ri", [(40, 4, 45), (30, 9, 40), (63, 15, 63), (50, 0, 30)])
def test_albedo_table_rough_cells_match_quadrature(albedo15, ci, pi, ri):
    wo = lut_cell_wo(ci, pi)
    r = (ri + 0.5) / 64
    for slot, an in ((0, 7), (1, 8)):
        ax, ay = cell_alphas(r, an)
        ref = reflect_albedo64(wo, ax, ay, 1.5)
        assert abs(albedo15[ci, pi, ri, slot] - ref) < 0.02 + 0.1 * ref


def test_albedo_lookup_interpolates_cell_centres(albedo15):
    # at a cell centre of sqrt(cos), phi and roughness the lookup returns the
    # mean of the two anisotropy slices
    wo = lut_cell_wo(30, 9)
    r = (40 + 0.5) / 64
    got = orc.lut_lookup(0, 1.5, wo.astype(np.float32), r)
    assert abs(got - 0.5 * (albedo15[30, 9, 40, 0] + albedo15[30, 9, 40, 1])) < 2e-5


@pytest.fixture(scope="module")
def conductor():
    return orc.lut_conductor()


def test_conductor_table_smooth_rows_are_one(conductor):
    e, ea = conductor
    assert np.all(e[:2] == 1.0) and np.allclose(ea[:2], 1.0, atol=2e-4)


@pytest.mark.parametrize("ci,ri", [(10, 20), (40, 40), (63, 63), (5, 60)])
def test_conductor_table_matches_white_furnace_quadrature(conductor, ci, ri):
    e, _ = conductor
    sc = (ci + 0.5) / 64
    c = sc * sc
    wo = np.array([np.sqrt(max(1 - c * c, 0)), 0.0, c])
    a = max(((ri + 0.5) / 64) ** 2, 1e-4)
    wi, dw = sphere_grid(1024)
    h = wo + wi
    h /= np.linalg.norm(h, axis=-1, keepdims=True)
    g2 = 1 / (1 + lambda64(wo[None], a, a) + lambda64(wi, a, a))
    ref = float((ggx64(h, a, a) * g2 / (4 * wo[2])).sum() * dw)
    assert abs(e[ri, ci] - ref) < 0.01


def test_conductor_average_is_cosine_weighted_mean(conductor):
    _, ea = conductor
    for ri in (10, 33, 63):
        r = (ri + 0.5) / 64
        mu = (np.arange(2048) + 0.5) / 2048
        vals = np.array([orc.lut_lookup(1, a=float(m), b=r) for m in mu])
        assert abs(ea[ri] - 2 * np.mean(vals * mu)) < 3e-3


def btdf_albedo64(wo, a, eta, n=1024):
    """Full-sphere albedo of a rough dielectric: GGX reflection plus the
    Walter et al. (2007) BTDF without the radiance 1/eta^2 scaling."""
    refl = reflect_albedo64(wo, a, a, eta, n)
    wi, dw = sphere_grid(n, lower=True)
    h = wo + eta * wi
    h /= np.linalg.norm(h, axis=-1, keepdims=True)
    h *= np.sign(h[:, 2:3])
    ho, hi = h @ wo, np.sum(h * wi, -1)
    g2 = 1 / (1 + lambda64(wo[None], a, a) + lambda64(wi, a, a))
    ok = (ho > 0) & (hi < 0)
    ft = (np.abs(ho * hi) / (wo[2] * np.abs(wi[:, 2])) * eta * eta * (1 - fresnel64(ho, eta)) * g2 * ggx64(h, a, a)
          / (ho + eta * hi) ** 2)
    return refl + float((np.where(ok, ft, 0) * np.abs(wi[:, 2])).sum() * dw)


@pytest.fixture(scope="module")
def dielectric_table():
    return orc.lut_dielectric()


def test_dielectric_table_smooth_rows_are_one(dielectric_table):
    e, _ = dielectric_table
    assert np.all(e[:, 0, :] == 1.0)  # roughness 1/64: alpha below 1e-3


@pytest.mark.parametrize("ei,ri,ci", [(25, 15, 20), (28, 25, 28), (20, 31, 10), (31, 20, 31)])
def test_dielectric_table_matches_sphere_quadrature(dielectric_table, ei, ri, ci):
    e, _ = dielectric_table
    t = (ei + 0.5) / 32
    eta = 1 + (t - 0.5) * 4 if t >= 0.5 else 1 / (1 + (0.5 - t) * 4)
    sc = (ci + 0.5) / 32
    c = sc * sc
    wo = np.array([np.sqrt(max(1 - c * c, 0)), 0.0, c])
    a = ((ri + 0.5) / 32) ** 2
    ref = btdf_albedo64(wo, a, eta)
    assert abs(e[ei, ri, ci] - ref) < 0.015


# ---- BSDF invariants -------------------------------------------------------------------------
LAYERED = OpenPbr(base_weight=0.9, base_color=(0.7, 0.3, 0.2), base_diffuse_roughness=0.4, base_metalness=0.3,
                  specular_roughness=0.45, specular_color=(0.9, 0.8, 1.0), coat_weight=0.6, coat_roughness=0.3,
                  coat_color=(0.9, 0.95, 0.7), fuzz_weight=0.4, fuzz_roughness=0.6, fuzz_color=(0.5, 0.9, 0.6))
MATERIALS = {
    "layered": LAYERED,
    "metal": OpenPbr(base_metalness=1.0, base_color=(0.95, 0.6, 0.3), specular_roughness=0.35),
    "coat": OpenPbr(base_color=(0.2, 0.4, 0.8), coat_weight=1.0, coat_roughness=0.25, specular_roughness=0.5),
    "fuzz": OpenPbr(base_weight=0.0, specular_color=(0.0, 0.0, 0.0), fuzz_weight=1.0, fuzz_roughness=0.5),
    "plastic": dielectric(0.4, base=1.0, base_color=(0.5, 0.6, 0.2)),
}


@pytest.mark.parametrize("name", list(MATERIALS))
def test_pdf_integrates_to_one(name):
    assert abs(pdf_mass(MATERIALS[name]) - 1.0) < 0.03


@pytest.mark.parametrize("name", list(MATERIALS))
def test_energy_bounded(name):
    e = energy(MATERIALS[name])
    assert np.all(np.isfinite(e)) and np.all(e >= 0) and e.max() <= 1.0 + 1.5e-2


@pytest.mark.parametrize("name", list(MATERIALS))
def test_sample_matches_eval_and_pdf(name):
    m = MATERIALS[name]
    u = np.random.default_rng(7).random((3000, 3), dtype=np.float32)
    s = orc.openpbr_sample(m, WO, u)
    ok = s["ok"] & ~s["delta"]
    assert ok.sum() > 1000
    f, pdf = orc.openpbr_eval(m, WO, s["wi"][ok])
    assert np.allclose(s["pdf"][ok], pdf, rtol=1e-5, atol=0)
    w = f * (np.abs(s["wi"][ok, 2]) / pdf)[:, None]
    assert np.allclose(s["weight"][ok], w, rtol=1e-5, atol=1e-7)
    # the mean sample weight estimates the directional albedo
    est = np.where(s["ok"][:, None], s["weight"], 0).astype(np.float64)
    ref = energy(m)
    sigma = est.std(0) / np.sqrt(len(est))
    assert np.all(np.abs(est.mean(0) - ref) < 4 * sigma + 0.01)


def test_zero_specular_weight_keeps_its_lobe_probability():
    # R/bsdf/open_pbr.rs:392-422 gives the specular lobe probability
    # luminance(below_metal * specular_color) even at specular weight 0,
    # while pdf() and eval() skip the lobe (spec_brdf_is_smooth): the pdf then
    # integrates to the other lobes' share, and picking the specular lobe
    # gives its smooth (delta) mirror sample, weighted by the eta-1 Fresnel
    # term (zero up to rounding).
    m = OpenPbr(base_weight=0.0, specular_weight=0.0, fuzz_weight=1.0, fuzz_roughness=0.5)
    mass = pdf_mass(m)
    assert 0.3 < mass < 0.7
    u = np.random.default_rng(3).random((4000, 3), dtype=np.float32)
    s = orc.openpbr_sample(m, WO, u)
    rough = s["ok"] & ~s["delta"]
    assert abs(rough.mean() - mass) < 0.05
    assert np.all(np.abs(s["weight"][s["delta"]]) < 1e-6)


def test_darkening_and_exit_tint_are_one_without_coat_and_fuzz():
    m = OpenPbr(base_color=(0.3, 0.6, 0.9))
    assert np.array_equal(orc.openpbr_probe(m, 3, wo=WO), np.ones(3, np.float32))
    assert np.array_equal(orc.openpbr_probe(m, 2, wo=WO), np.ones(3, np.float32))


def test_coat_exit_tint_follows_the_refracted_path():
    # white coat: no absorption; tinted coat: colour^(1/mu_t) blended by weight
    m = OpenPbr(coat_weight=0.5, coat_ior=1.5, coat_color=(1.0, 1.0, 1.0))
    assert np.allclose(orc.openpbr_probe(m, 2, wo=WO), 1.0, atol=1e-6)
    m = OpenPbr(coat_weight=0.5, coat_ior=1.5, coat_color=(0.5, 0.7, 0.9))
    c = float(WO[2])
    mu = np.sqrt(1 - (1 - c * c) / 1.5 ** 2)
    expect = 1 + (np.array([0.5, 0.7, 0.9]) ** (1 / mu) - 1) * 0.5
    assert np.allclose(orc.openpbr_probe(m, 2, wo=WO), expect, rtol=2e-5)
