"""Textures of the backend: the reference's ``Texture<Vec3>`` / ``ScalarTexture``
(R/material/texture.rs:53-126) as host pixel data that crosses the ABI in a
``TptTextureDesc``.  The mip pyramid is built on the device by
``tpt_scene_create`` and every lookup (``DeviceScene.texture_sample``) runs
there; nothing is sampled on the host.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass
class Texture:
    """Level-0 pixels, (height, width, 3) for colour or (height, width) for a
    scalar texture, float32, row 0 first (texel (x, y) covers uv
    [x/w, (x+1)/w) x [y/h, (y+1)/h))."""

    pixels: np.ndarray

    @staticmethod
    def from_pixels(width: int, height: int, pixels) -> "Texture":
        """Texture::from_pixels (R/material/texture.rs:90-115): `pixels` holds
        width * height RGB triples or scalars in row-major order."""
        if width <= 0 or height <= 0:
            raise ValueError("texture must be non-empty")
        a = np.asarray(pixels, np.float32)
        if a.size == width * height * 3:
            a = a.reshape(height, width, 3)
        elif a.size == width * height:
            a = a.reshape(height, width)
        else:
            raise ValueError("pixel buffer length does not match width * height")
        return Texture(np.ascontiguousarray(a))

    @staticmethod
    def from_rgba_channel(width: int, height: int, rgba, channel: int) -> "Texture":
        """ScalarTexture::from_rgba_channel (R/material/texture.rs:390-405):
        one byte channel of a packed RGBA8 buffer, / 255."""
        if not 0 <= channel < 4:
            raise ValueError("channel index must be in 0..4")
        b = np.frombuffer(bytes(rgba), np.uint8) if not isinstance(rgba, np.ndarray) else rgba.astype(np.uint8)
        if b.size != width * height * 4:
            raise ValueError("rgba buffer length does not match width * height * 4")
        px = b.reshape(-1, 4)[:, channel].astype(np.float32) / np.float32(255.0)
        return Texture(np.ascontiguousarray(px.reshape(height, width)))

    @property
    def width(self) -> int:
        return int(self.pixels.shape[1])

    @property
    def height(self) -> int:
        return int(self.pixels.shape[0])

    @property
    def channels(self) -> int:
        return 3 if self.pixels.ndim == 3 else 1

    def max_value(self) -> float:
        """Texture::max_value: the largest channel over level 0, at least 0."""
        return float(max(0.0, float(np.nanmax(self.pixels)) if self.pixels.size else 0.0))


def checkerboard(width: int, height: int, cells: int, a=(0.9, 0.9, 0.9), b=(0.1, 0.1, 0.1)) -> Texture:
    """An RGB checkerboard of cells x cells squares (a test and bench texture)."""
    y, x = np.mgrid[0:height, 0:width]
    odd = ((x * cells // width) + (y * cells // height)) % 2 == 1
    px = np.where(odd[..., None], np.asarray(b, np.float32), np.asarray(a, np.float32))
    return Texture(np.ascontiguousarray(px.astype(np.float32)))
