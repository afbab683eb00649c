"""The oracle's thin film (R/bsdf/thin_film.rs, restated in oracle/bsdf.h) and
the artist-friendly complex IOR (R/bsdf/conductor_complex.rs:132-147).

Pinned by the reference's own unit tests (thin_film.rs:246-296: finite and
non-negative at zero thickness, the response changes with thickness, a
conductor film stays in [0, 1]) and by what the physics fixes independently
of the restatement (textbook Fresnel equations with numpy complex
arithmetic):
- a film of the base's own index (r23 = 0) reflects like the bare
  outside/film interface, whatever the thickness;
- a film far thicker than the coherence length (the interference terms'
  Gaussian envelopes vanish) reflects the reference's incoherent
  two-interface sum R12 + T R23 / (1 - R12 R23) per polarisation, with the
  reference's single factor T = 1 - R12 (thin_film.rs:67-91; the physical
  sum has T^2), for dielectric and conductor bases;
- total internal reflection into the film gives 1;
- the complex IOR reproduces the face colour at normal incidence and the
  edge colour's limits (g = 1: n = (1 - r) / (1 + r)).
The interference orders themselves (the sensitivity fit's constants) are
pinned only through the reference's qualitative tests above."""
import numpy as np
import pytest

import oracle


def fresnel_s_p(n1, n2, cos1):
    """Amplitude reflectances |r_s|^2, |r_p|^2 from medium n1 into n2 (complex)."""
    n1 = complex(n1)
    n2 = complex(n2)
    sin1 = np.sqrt(1 - cos1 ** 2)
    sin2 = n1 * sin1 / n2
    cos2 = np.sqrt(1 - sin2 ** 2 + 0j)
    if cos2.imag < 0:
        cos2 = -cos2
    rs = (n1 * cos1 - n2 * cos2) / (n1 * cos1 + n2 * cos2)
    rp = (n2 * cos1 - n1 * cos2) / (n2 * cos1 + n1 * cos2)
    return abs(rs) ** 2, abs(rp) ** 2


def incoherent(cos_i, eta_out, eta_film, n_base, k_base):
    sin_i = np.sqrt(1 - cos_i ** 2)
    if (eta_out / eta_film * sin_i) ** 2 >= 1:  # total internal reflection into the film
        return np.ones(3)
    cos_f = np.sqrt(1 - (eta_out / eta_film * sin_i) ** 2)
    out = []
    for n, k in zip(n_base, k_base):
        r12 = fresnel_s_p(eta_out, eta_film, cos_i)
        r23 = fresnel_s_p(eta_film, complex(n, k), cos_f)
        pol = [r12[j] + (1 - r12[j]) * r23[j] / (1 - r12[j] * r23[j]) for j in range(2)]
        out.append(0.5 * (pol[0] + pol[1]))
    return np.clip(np.array(out), 0, 1)


def test_reference_unit_tests():
    f = oracle.thin_film(0.7, 1.0, 1.5, [1.5] * 3, [0.0] * 3, 0.0)
    assert np.all(np.isfinite(f)) and np.all(f >= 0)
    f0, f250, f500 = (oracle.thin_film(0.85, 1.0, 2.4, [1.5] * 3, [0.0] * 3, t) for t in (0.0, 250.0, 500.0))
    assert max(np.abs(f0 - f250).max(), np.abs(f250 - f500).max()) > 1e-3
    c = oracle.thin_film(0.7, 1.0, 2.0, [0.18, 0.42, 1.37], [3.42, 2.35, 1.77], 300.0)
    assert np.all(np.isfinite(c)) and np.all(c >= 0) and np.all(c <= 1)


@pytest.mark.parametrize("cos_i", [1.0, 0.8, 0.5, 0.2])
@pytest.mark.parametrize("thickness", [0.0, 120.0, 700.0])
def test_film_of_the_base_index_is_the_bare_interface(cos_i, thickness):
    for eta_out, eta in ((1.0, 1.5), (1.5, 2.2), (1.0, 1.33)):
        f = oracle.thin_film(cos_i, eta_out, eta, [eta] * 3, [0.0] * 3, thickness)
        rs, rp = fresnel_s_p(eta_out, eta, cos_i)
        np.testing.assert_allclose(f, 0.5 * (rs + rp), rtol=2e-5, atol=2e-7)


@pytest.mark.parametrize("cos_i", [1.0, 0.9, 0.6, 0.3])
def test_thick_film_is_the_incoherent_sum(cos_i):
    cases = [(1.0, 1.4, [1.5] * 3, [0.0] * 3), (1.0, 2.0, [1.45] * 3, [0.0] * 3),
             (1.5, 1.3, [1.8] * 3, [0.0] * 3),  # under a coat
             (1.0, 1.6, [0.18, 0.42, 1.37], [3.42, 2.35, 1.77]),  # gold-like base
             (1.0, 2.0, [1.2, 0.9, 0.6], [2.0, 2.5, 3.0])]
    for eta_out, eta_film, n, k in cases:
        f = oracle.thin_film(cos_i, eta_out, eta_film, n, k, 1.0e6)
        np.testing.assert_allclose(f, incoherent(cos_i, eta_out, eta_film, n, k), rtol=3e-5, atol=1e-6)


def test_total_internal_reflection_into_the_film():
    # from a dense outside medium into a thin film at grazing angles
    f = oracle.thin_film(0.2, 2.0, 1.2, [1.5] * 3, [0.0] * 3, 300.0)
    np.testing.assert_array_equal(f, np.ones(3, np.float32))


@pytest.mark.parametrize("face", [(0.9, 0.6, 0.3), (0.05, 0.5, 0.95), (0.5, 0.5, 0.5)])
def test_artist_friendly_complex_ior(face):
    for edge in ((1.0, 1.0, 1.0), (0.8, 0.9, 0.6), (0.0, 0.2, 0.4)):
        n, k = oracle.artist_ior(face, edge)
        r0 = ((n - 1) ** 2 + k ** 2) / ((n + 1) ** 2 + k ** 2)
        np.testing.assert_allclose(r0, face, rtol=2e-4, atol=1e-6)
    n, k = oracle.artist_ior(face, (1.0, 1.0, 1.0))
    r = np.asarray(face)
    np.testing.assert_allclose(n, (1 - r) / (1 + r), rtol=1e-6)
