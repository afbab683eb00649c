"""The CPU reference images of tests/golden/reference_images.npz (the
oracle's renderer at 16,384 samples per pixel; made by
tests/golden/make_reference_images.py) and the relMSE comparison the GPU
image-parity tests run against them."""
from pathlib import Path

import numpy as np

import oracle
from tpt import camera
from tpt.imageio import relmse

PATH = Path(__file__).resolve().parent / "golden" / "reference_images.npz"


def reference(name, scene):
    import sys
    sys.path.insert(0, str(PATH.parent))
    from make_reference_images import CASES, mesh_digest
    z = np.load(PATH)
    assert str(z[name + "_digest"]) == mesh_digest(scene), f"{name}: scene changed; regenerate {PATH.name}"
    _, size, depth = CASES[name]
    return z[name], size, depth, int(z["spp"])


def gpu_and_cpu_relmse(name, cs, ds, spp):
    """relMSE of the GPU image and of the oracle's image, both at spp samples
    per pixel, against the CPU reference image."""
    ref, size, depth, _ = reference(name, cs.scene)
    cam = camera.pinhole(**cs.camera, width=size, height=size)
    gpu = ds.render(cam, size, size, spp, depth)
    cpu = oracle.OracleRenderer(cs.scene, cs.materials).render(cam, size, size, spp, depth, threads=16)
    return relmse(gpu, ref), relmse(cpu, ref)
