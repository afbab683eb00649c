"""OpenPBR thin film in the GPU path tracers (R/bsdf/thin_film.rs): the film
Cornell box (rough metal under a film, a smooth coated dielectric with the
film under the coat, a thin-walled pane whose slab sees the film from
inside; and the same box with a textured floor) against the oracle's
restatement of the reference renderer.

* Pixel by pixel against the oracle at equal samples: most pixels
  bit-identical (a last-place difference of a transcendental can send a
  path down another branch; such a pixel differs by that path).
* Against the CPU oracle's 16,384-spp image of the film box the GPU
  image's relMSE at 64 spp is no worse than the CPU image's.
* The wavefront batch equals the megakernel bit for bit.
* Scenes without a film launch kernels built without the film code (the
  host checks the material table); a film-free scene renders the same bits
  through either build: the same scene with one extra, unused filmed
  material (which selects the film build) gives the identical image.
* The parameter adjoint (base colour, roughness, metalness) of filmed
  materials against central differences of the oracle's detached render
  (the paths of the unperturbed parameters, the BSDF values of the stepped
  ones).
"""
import time

import numpy as np
import pytest

import golden_images
import oracle
from tpt import camera, scenes
from tpt.material import OpenPbr

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", params=[False, True], ids=["plain", "textured"])
def film(device, request):
    cs = scenes.cornell_film(textured=request.param)
    ds = cs.scene.upload(device)
    yield cs, ds
    ds.close()


def test_film_render_matches_oracle_pixel_by_pixel(film):
    cs, ds = film
    w = h = 40
    cam = camera.pinhole(**cs.camera, width=w, height=h)
    gpu = ds.render(cam, w, h, 8, 8)
    cpu = oracle.OracleRenderer(cs.scene, cs.materials).render(cam, w, h, 8, 8, threads=16)
    assert np.all(np.isfinite(gpu))
    same = np.all(gpu.view(np.uint32) == cpu.view(np.uint32), axis=2)
    close = np.all(np.abs(gpu - cpu) <= 1e-5 * np.maximum(np.abs(cpu), 1.0), axis=2)
    print(f"bit-identical pixels: {same.mean():.4f}, within 1e-5: {close.mean():.4f}")
    assert same.mean() > 0.9
    assert close.mean() > 0.95


def test_film_relmse_no_worse_than_cpu(device):
    """Against the CPU oracle's 16,384-spp image of the film box
    (tests/golden/reference_images.npz) the GPU image's relMSE at 64 spp is
    no worse than the CPU image's at 64 spp."""
    cs = scenes.cornell_film()
    ds = cs.scene.upload(device)
    e_gpu, e_cpu = golden_images.gpu_and_cpu_relmse("cornell_film", cs, ds, 64)
    ds.close()
    print(f"cornell_film relMSE at 64 spp against the CPU reference: gpu {e_gpu:.5f} cpu {e_cpu:.5f}")
    assert e_gpu <= 1.1 * e_cpu + 1e-6


def test_film_wavefront_equals_megakernel(film):
    cs, ds = film
    w = h = 64
    cam = camera.pinhole(**cs.camera, width=w, height=h)
    mega = ds.render(cam, w, h, 16, 8)
    wave = ds.render_batch([(cam, w, h, 16, 8)])[0]
    assert np.array_equal(mega.view(np.uint32), wave.view(np.uint32))


def test_film_free_scene_same_bits_through_the_film_build(device):
    """cornell_glass has no film: tpt_render and tpt_render_batch launch the
    film-free kernels.  Appending an unused material with a film makes the
    host pick the film build; the image must not change by a bit."""
    w = h = 48
    cs = scenes.cornell_glass(textured=True)
    cam = camera.pinhole(**cs.camera, width=w, height=h)
    ds = cs.scene.upload(device)
    plain_mega = ds.render(cam, w, h, 8, 8)
    plain_wave = ds.render_batch([(cam, w, h, 8, 8)])[0]
    ds.close()
    cs2 = scenes.cornell_glass(textured=True)
    cs2.scene.add_material(OpenPbr.new((0.5, 0.5, 0.5)).with_(thin_film_weight=1.0, thin_film_thickness=0.5))
    ds2 = cs2.scene.upload(device)
    film_mega = ds2.render(cam, w, h, 8, 8)
    film_wave = ds2.render_batch([(cam, w, h, 8, 8)])[0]
    ds2.close()
    assert np.array_equal(plain_mega.view(np.uint32), film_mega.view(np.uint32))
    assert np.array_equal(plain_wave.view(np.uint32), film_wave.view(np.uint32))


W = H = 64
# (material, parameter): 0-2 base colour, 3 roughness, 4 metalness
PARAMS = [(4, 0), (4, 2), (4, 3), (4, 4), (3, 1), (3, 3), (5, 0), (5, 3), (0, 3)]


def test_params_adjoint_through_film_matches_detached_differences(device):
    """The parameter adjoint of a weighted image sum against central
    differences (h = 1e-3, one-sided where a step would leave the range the
    BSDF clamps roughness and metalness to, as the adjoint's own
    differences are) of the oracle's detached render: the oracle
    traces the unperturbed parameters' paths (sampling decisions,
    densities, MIS weights, roulette: the same samples as the GPU) and
    evaluates the BSDF values at the stepped parameters, which is the
    derivative the path-replay adjoint accumulates sample by sample.
    Central differences of plain renders with common random numbers
    estimate the same gradient through moved samples and are far noisier
    on the roughness rows (measured on MI355X on the film Cornell box at 256 x 256, 1,024 spp:
    30-80 % apart from the adjoint there, with or without the films, and
    10-50 % apart between steps 5e-3 and 2e-2)."""
    cs = scenes.cornell_film(light_size=0.9, light_radiance=2.0)
    ds = cs.scene.upload(device)
    cam = camera.pinhole(**cs.camera, width=W, height=H)
    spp, depth, h = 64, 3, 1e-3
    wts = np.random.default_rng(5).uniform(0.5, 1.5, (H, W, 3)).astype(np.float32)
    t0 = time.perf_counter()
    g = ds.render_adjoint_params(cam, W, H, spp, depth, wts)[0].astype(np.float64)
    print(f"adjoint {time.perf_counter() - t0:.1f} s", flush=True)
    ds.close()
    rs = oracle.OracleRenderer(cs.scene, cs.materials)
    fd = []
    for i, c in PARAMS:
        side = []
        x = (list(cs.materials[i].base_color) + [cs.materials[i].specular_roughness,
                                                 cs.materials[i].base_metalness])[c]
        # one-sided at the bounds the BSDF clamps roughness and metalness to
        hi, lo = (x + h if c < 3 else min(x + h, 1.0)), max(x - h, 0.0)
        for v in (hi, lo):
            mats = list(cs.materials)
            row = list(mats[i].base_color) + [mats[i].specular_roughness, mats[i].base_metalness]
            row[c] = v
            mats[i] = mats[i].with_(base_color=tuple(float(x) for x in row[:3]), specular_roughness=float(row[3]),
                                    base_metalness=float(row[4]))
            ev = oracle.OracleRenderer(cs.scene, mats)
            side.append(rs.render_rect_detached_f64(ev, cam, W, H, spp, depth, 0, W, 0, H, threads=16))
        fd.append(float((wts * (side[0] - side[1])).sum() / (hi - lo)))
        print(f"({i}, {c}) adjoint {g[i, c]:.6g} detached difference {fd[-1]:.6g}", flush=True)
    fd = np.array(fd)
    ga = np.array([g[i, c] for i, c in PARAMS])
    rough = np.array([c == 3 for _, c in PARAMS])
    rel = float(np.linalg.norm(ga[~rough] - fd[~rough]) / np.linalg.norm(fd[~rough]))
    rel_r = float(np.linalg.norm(ga[rough] - fd[rough]) / np.linalg.norm(fd[rough]))
    print(f"rel L2 base colour + metalness {rel:.4f}, roughness {rel_r:.4f}")
    assert rel < 0.01
    assert rel_r < 0.01
