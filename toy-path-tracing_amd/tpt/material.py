"""Materials of the backend (R/material/mod.rs:100-115 variants carried).

``OpenPbr.new(base_color)`` gives the reference's defaults
(``OpenPbrMaterial::new``, R/material/open_pbr.rs:23-143); ``with_*`` mirror
its builder methods.  ``NormalizedLambert``, ``Emissive`` and ``Mirror``
are the reference's small materials (R/material/normalized_lambert.rs,
emissive.rs, mirror.rs).  Every material crosses the ABI as a
``TptMaterial`` record (kind + parameters, include/tpt.h).
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
        ("coat_color", ctypes.c_float * 3), ("coat_roughness", ctypes.c_float),
        ("coat_roughness_anisotropy", ctypes.c_float), ("coat_ior", ctypes.c_float),
        ("coat_darkening", ctypes.c_float),
        ("fuzz_color", ctypes.c_float * 3), ("fuzz_roughness", ctypes.c_float),
        ("base_color_texture", ctypes.c_uint32), ("base_metalness_texture", ctypes.c_uint32),
        ("specular_roughness_texture", ctypes.c_uint32), ("fuzz_weight_texture", ctypes.c_uint32),
        ("fuzz_roughness_texture", ctypes.c_uint32), ("emission_color_texture", ctypes.c_uint32),
        ("transmission_color", ctypes.c_float * 3), ("transmission_dispersion_scale", ctypes.c_float),
        ("transmission_dispersion_abbe_number", ctypes.c_float),
        ("thin_film_thickness", ctypes.c_float), ("thin_film_ior", ctypes.c_float),
        ("pad", ctypes.c_uint32 * 2),
    ]


assert ctypes.sizeof(OpenPbrRecord) == 192
NO_TEXTURE = 0xFFFFFFFF
TEXTURE_FIELDS = ("base_color_texture", "base_metalness_texture", "specular_roughness_texture",
                  "fuzz_weight_texture", "fuzz_roughness_texture", "emission_color_texture")

KIND_OPEN_PBR, KIND_NORMALIZED_LAMBERT, KIND_EMISSIVE, KIND_MIRROR = 0, 1, 2, 3


class MaterialRecord(ctypes.Structure):
    """TptMaterial: kind, pad, then the kind's parameters (OpenPBR: the
    192-byte TptOpenPbr; Lambert: rho[3]; Emissive: color[3], strength;
    Mirror: color[3])."""
    _fields_ = [("kind", ctypes.c_uint32), ("pad", ctypes.c_uint32), ("params", ctypes.c_float * 48)]


assert ctypes.sizeof(MaterialRecord) == 200


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
    coat_color: tuple = (1.0, 1.0, 1.0)
    coat_roughness: float = 0.0
    coat_roughness_anisotropy: float = 0.0
    coat_ior: float = 1.6
    coat_darkening: float = 1.0
    fuzz_color: tuple = (1.0, 1.0, 1.0)
    fuzz_roughness: float = 0.5
    transmission_color: tuple = (1.0, 1.0, 1.0)
    transmission_dispersion_scale: float = 0.0
    transmission_dispersion_abbe_number: float = 20.0
    thin_film_thickness: float = 0.5
    thin_film_ior: float = 1.4
    # indices into the scene's texture table (Scene.add_texture), None: no
    # texture (R/material/open_pbr.rs:66-72)
    base_color_texture: int | None = None
    base_metalness_texture: int | None = None
    specular_roughness_texture: int | None = None
    fuzz_weight_texture: int | None = None
    fuzz_roughness_texture: int | None = None
    emission_color_texture: int | None = None

    @staticmethod
    def new(base_color) -> "OpenPbr":
        """OpenPbrMaterial::new (R/material/open_pbr.rs:85-143)."""
        return OpenPbr(base_color=tuple(float(c) for c in base_color))

    def with_(self, **kw) -> "OpenPbr":
        """The reference's with_* builders, by field name."""
        return replace(self, **kw)

    def open_pbr_record(self) -> OpenPbrRecord:
        r = OpenPbrRecord()
        for name, _ in OpenPbrRecord._fields_:
            if name == "pad":
                continue
            v = getattr(self, name)
            if name in TEXTURE_FIELDS:
                setattr(r, name, NO_TEXTURE if v is None else int(v))
            elif isinstance(v, tuple):
                getattr(r, name)[:] = [float(x) for x in v]
            else:
                setattr(r, name, int(v) if name == "thin_walled" else float(v))
        return r

    def record(self) -> MaterialRecord:
        r = MaterialRecord()
        r.kind = KIND_OPEN_PBR
        rec = self.open_pbr_record()  # kept alive across the copy
        ctypes.memmove(ctypes.addressof(r.params), ctypes.addressof(rec), ctypes.sizeof(OpenPbrRecord))
        return r


@dataclass(frozen=True)
class NormalizedLambert:
    """NormalizedLambertMaterial::new(rho) (R/material/normalized_lambert.rs:27-37)."""
    rho: tuple = (0.8, 0.8, 0.8)

    def record(self) -> MaterialRecord:
        r = MaterialRecord()
        r.kind = KIND_NORMALIZED_LAMBERT
        r.params[0:3] = [float(x) for x in self.rho]
        return r


@dataclass(frozen=True)
class Emissive:
    """EmissiveMaterial::new(color, strength) (R/material/emissive.rs:17-26)."""
    color: tuple = (1.0, 1.0, 1.0)
    strength: float = 1.0

    def record(self) -> MaterialRecord:
        r = MaterialRecord()
        r.kind = KIND_EMISSIVE
        r.params[0:4] = [float(x) for x in self.color] + [float(self.strength)]
        return r


@dataclass(frozen=True)
class Mirror:
    """MirrorMaterial::new(color) (R/material/mirror.rs)."""
    color: tuple = (1.0, 1.0, 1.0)

    def record(self) -> MaterialRecord:
        r = MaterialRecord()
        r.kind = KIND_MIRROR
        r.params[0:3] = [float(x) for x in self.color]
        return r


def records(materials) -> ctypes.Array:
    arr = (MaterialRecord * max(1, len(materials)))()
    for i, m in enumerate(materials):
        arr[i] = m.record()
    return arr
