"""tpt: MI355X backend for the reference renderer's ray queries.

Host-side package over libtpt.so (include/tpt.h).  ``tpt.scene`` mirrors the
reference Scene geometry API; ``tpt.scenes`` generates the seeded workloads.
"""
from . import _lib
from .scene import (
    HIT_DTYPE, INVALID_INDEX, NODE_DTYPE, Q8_NODE_DTYPE, RAY_DTYPE, ClosestHitError, Device, DeviceScene,
    Environment, Instance, Mesh, Qbvh, Scene, build_qbvh, glam_cols, make_rays, quantize_nodes, transform_bounds,
    transform_point3,
)

__all__ = [
    "Environment",
    "_lib", "HIT_DTYPE", "INVALID_INDEX", "NODE_DTYPE", "Q8_NODE_DTYPE", "RAY_DTYPE", "ClosestHitError", "Device",
    "DeviceScene", "Instance", "Mesh", "Qbvh", "Scene", "build_qbvh", "glam_cols", "make_rays", "quantize_nodes",
    "transform_bounds", "transform_point3",
]
