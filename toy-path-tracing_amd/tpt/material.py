"""OpenPBR material parameters (R/material/open_pbr.rs:23-143).

``OpenPbr.new(base_color)`` gives the reference's defaults
(``OpenPbrMaterial::new``); ``with_*`` mirror its builder methods.  The
record crosses the ABI as ``TptOpenPbr`` (include/tpt.h).
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field, replace


class OpenPbrRecord(ctypes.Structure):
    _fields_ = [
        ("base_weight", ctypes.c_float), ("base_color", ctypes.c_float * 3),
        ("base_diffuse_roughness", ctypes.c_float), ("base_metalness", ctypes.c_float),
        ("specular_weight", ctypes.c_float), ("specular_color", ctypes.c_float * 3),
        ("specular_roughness", ctypes.c_float), ("specular_ior", ctypes.c_float),
        ("specular_roughness_anisotropy", ctypes.c_float),
        ("transmission_weight", ctypes.c_float), ("coat_weight", ctypes.c_float),
        ("fuzz_weight", ctypes.c_float), ("thin_film_weight", ctypes.c_float),
        ("emission_luminance", ctypes.c_float), ("emission_color", ctypes.c_float * 3),
        ("thin_walled", ctypes.c_uint32),
    ]


assert ctypes.sizeof(OpenPbrRecord) == 88


@dataclass(frozen=True)
class OpenPbr:
    base_weight: float = 1.0
    base_color: tuple = (0.8, 0.8, 0.8)
    base_diffuse_roughness: float = 0.0
    base_metalness: float = 0.0
    specular_weight: float = 1.0
    specular_color: tuple = (1.0, 1.0, 1.0)
    specular_roughness: float = 0.3
    specular_ior: float = 1.5
    specular_roughness_anisotropy: float = 0.0
    transmission_weight: float = 0.0
    coat_weight: float = 0.0
    fuzz_weight: float = 0.0
    thin_film_weight: float = 0.0
    emission_luminance: float = 0.0
    emission_color: tuple = field(default=(1.0, 1.0, 1.0))
    thin_walled: bool = False

    @staticmethod
    def new(base_color) -> "OpenPbr":
        """OpenPbrMaterial::new (R/material/open_pbr.rs:85-143)."""
        return OpenPbr(base_color=tuple(float(c) for c in base_color))

    def with_(self, **kw) -> "OpenPbr":
        """The reference's with_* builders, by field name."""
        return replace(self, **kw)

    def record(self) -> OpenPbrRecord:
        r = OpenPbrRecord()
        for name, _ in OpenPbrRecord._fields_:
            v = getattr(self, name)
            if isinstance(v, tuple):
                getattr(r, name)[:] = [float(x) for x in v]
            else:
                setattr(r, name, int(v) if name == "thin_walled" else float(v))
        return r


def records(materials) -> ctypes.Array:
    arr = (OpenPbrRecord * max(1, len(materials)))()
    for i, m in enumerate(materials):
        arr[i] = m.record()
    return arr
