#!/usr/bin/env python3
"""Writes tests/golden/reference_images.npz: high-sample-count reference
images rendered by the oracle (the CPU restatement of the reference's render
loop, R/main.rs:98-126 with IntegratorKind::Mis), 16,384 samples per pixel.

The GPU image-parity tests compare the GPU's and the oracle's images at a
low sample count against these: the GPU image's relMSE must be no worse than
the CPU image's at equal samples per pixel.  A reference rendered by the GPU
itself would carry any bias of the GPU path into both sides of that
comparison; this one is independent of the device code.

Each entry records the scene generator's geometry digest, so a changed
generator makes the CPU suite fail until the images are regenerated.

Run from the repository root after building the oracle (about 5 minutes on
8 host threads for the first six cases, 10 more for the two glass boxes,
about 7 for the thin-film box):
    python tests/golden/make_reference_images.py
Entries whose scene digest still matches are kept; --all renders every case.
"""
from __future__ import annotations

import sys
import time
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent
sys.path.insert(0, str(ROOT / "tests"))
sys.path.insert(0, str(ROOT / "toy-path-tracing_amd"))

import oracle  # noqa: E402
from tpt import camera, scenes  # noqa: E402

REF_SPP = 16384

# name -> (scene factory, width = height, max depth)
CASES = {
    "cornell_diffuse": (lambda: scenes.cornell(specular_weight=0.0), 64, 8),
    "cornell_glossy": (lambda: scenes.cornell(specular_weight=1.0), 64, 8),
    "cornell_lambert_nee": (lambda: scenes.cornell_lambert(), 64, 8),
    "cornell_textured": (lambda: scenes.cornell_textured(), 32, 8),
    "outdoor_openpbr": (lambda: scenes.outdoor("openpbr"), 24, 6),
    "outdoor_lambert": (lambda: scenes.outdoor("lambert"), 24, 6),
    "cornell_glass": (lambda: scenes.cornell_glass(), 48, 8),
    "cornell_glass_textured": (lambda: scenes.cornell_glass(textured=True), 32, 8),
    "cornell_film": (lambda: scenes.cornell_film(), 48, 8),
}


def mesh_digest(scene) -> str:
    import hashlib
    h = hashlib.sha256()
    for m in scene.meshes:
        h.update(np.ascontiguousarray(m.positions, np.float32).tobytes())
        h.update(np.ascontiguousarray(m.indices, np.uint32).tobytes())
    for inst in scene.instances:
        h.update(np.ascontiguousarray(inst.local_to_world, np.float32).tobytes())
    return h.hexdigest()


def render_reference(name: str, threads: int = 0):
    import os
    make, size, depth = CASES[name]
    cs = make()
    cam = camera.pinhole(**cs.camera, width=size, height=size)
    r = oracle.OracleRenderer(cs.scene, cs.materials)
    img = r.render(cam, size, size, REF_SPP, depth, threads=threads or (os.cpu_count() or 8))
    return img, mesh_digest(cs.scene)


def main():
    out = {}
    path = HERE / "reference_images.npz"
    if path.exists() and "--all" not in sys.argv:
        old = np.load(path)
        if int(old["spp"]) == REF_SPP:
            for name, (make, _, _) in CASES.items():
                if name in old and str(old[name + "_digest"]) == mesh_digest(make().scene):
                    out[name] = old[name]
                    out[name + "_digest"] = old[name + "_digest"]
    for name in CASES:
        if name in out:
            continue
        t = time.time()
        img, dig = render_reference(name)
        out[name] = img
        out[name + "_digest"] = np.array(dig)
        print(f"{name}: {img.shape[1]}x{img.shape[0]} at {REF_SPP} spp in {time.time() - t:.0f}s, "
              f"mean {img.mean():.5f}", flush=True)
    out["spp"] = np.array(REF_SPP)
    np.savez_compressed(HERE / "reference_images.npz", **out)


if __name__ == "__main__":
    main()
