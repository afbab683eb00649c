"""Pinhole camera set-up (R/scene/camera.rs:21-42, 89-101) on the host.

The basis is normalised with glam's rule (v * (1 / sqrt(dot(v, v))), f32);
tan(fov/2) and the pixel spread angle use the host libm, as the renderer's
own PinholeCamera::new does, so the device only does +, -, *, /, sqrt.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from ._lib import RAY_DTYPE, check, lib, ptr

F = np.float32


def _normalize(v: np.ndarray) -> np.ndarray:
    v = np.asarray(v, F)
    d = v[0] * v[0] + v[1] * v[1] + v[2] * v[2]
    r = F(1.0) / np.sqrt(F(d))
    return (v * r).astype(F)


def _cross(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    return np.array([a[1] * b[2] - b[1] * a[2], a[2] * b[0] - b[2] * a[0], a[0] * b[1] - b[0] * a[1]], F)


def pinhole(eye, look_at, up, fov_y_deg: float, width: int, height: int, exposure: float = 1.0) -> _lib.Camera:
    """PinholeCamera::new(eye, look_at, up, fov_y.to_radians(), exposure) + pixel_spread_angle(resolution)."""
    eye = np.asarray(eye, F)
    fwd = _normalize(np.asarray(look_at, F) - eye)
    right = _normalize(_cross(fwd, np.asarray(up, F)))
    iup = _normalize(_cross(right, fwd))
    fov = F(np.deg2rad(F(fov_y_deg)))
    th = F(np.tan(F(0.5) * fov))
    aspect = F(width) / F(height)
    dx = F(2.0) * aspect * th / F(width)
    dy = F(2.0) * th / F(height)
    spread = F(np.arctan(max(dx, dy)))
    c = _lib.Camera()
    c.eye[:] = eye.tolist()
    c.forward[:] = fwd.tolist()
    c.right[:] = right.tolist()
    c.image_up[:] = iup.tolist()
    c.tan_half_fov_y = float(th)
    c.pixel_spread = float(spread)
    c.exposure = float(exposure)
    return c


def camera_rays(device, cam: _lib.Camera, width: int, height: int, spp: int, sample_index: int,
                differentials: bool = False):
    """Primary rays (and optionally differentials) of one sample index on the GPU."""
    rays = np.zeros(width * height, RAY_DTYPE)
    diff = np.zeros((width * height, 7), np.float32) if differentials else None
    check(lib().tpt_camera_rays(device.handle, ctypes.byref(cam), width, height, spp, sample_index, ptr(rays),
                                ptr(diff), 0))
    return (rays, diff) if differentials else rays


def path_randoms(device, width: int, height: int, spp: int, sample_index: int, depth_count: int):
    """Per-pixel integrator random numbers and aux seeds of one sample index."""
    vals = np.zeros((width * height, 2 + 14 * depth_count), np.float32)
    seeds = np.zeros((width * height, 1 + depth_count), np.uint32)
    check(lib().tpt_path_randoms(device.handle, width, height, spp, sample_index, depth_count, ptr(vals),
                                 ptr(seeds), 0))
    return vals, seeds


def camera_array(cam: _lib.Camera) -> np.ndarray:
    """The oracle's camera layout: eye, forward, right, image_up, tan_half, spread."""
    return np.array(list(cam.eye) + list(cam.forward) + list(cam.right) + list(cam.image_up)
                    + [cam.tan_half_fov_y, cam.pixel_spread], np.float32)
