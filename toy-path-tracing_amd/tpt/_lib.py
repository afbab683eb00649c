"""ctypes binding of libtpt.so (the C ABI declared in include/tpt.h).

The shared library is built in-tree (``toy-path-tracing_amd/libtpt.so``) by
``__graft_entry__.build()``.  Loading fails loudly when it is missing: there
is no CPU fallback for any query.
"""
from __future__ import annotations

import ctypes
import os
from pathlib import Path

import numpy as np

PKG_DIR = Path(__file__).resolve().parent.parent
LIB_PATH = PKG_DIR / "libtpt.so"

TPT_OK = 0
TPT_ERROR = 1
TPT_ERROR_NOT_BUILT = 2
TPT_ERROR_BAD_ARGUMENT = 3
TPT_ERROR_DEVICE = 4
TPT_ERROR_OVERFLOW = 5

TPT_FLAG_DEVICE_POINTERS = 0x1
TPT_FLAG_QUANTIZED = 0x2
TPT_FLAG_GRAD_FP64 = 0x8
TPT_FLAG_ANY_EXACT = 0x4

INVALID_INDEX = 0xFFFFFFFF

# numpy mirrors of the ABI records --------------------------------------------
RAY_DTYPE = np.dtype([("o", "<f4", 3), ("tmax", "<f4"), ("d", "<f4", 3), ("pad", "<f4")])
HIT_DTYPE = np.dtype([("instance", "<u4"), ("triangle", "<u4"), ("t", "<f4"), ("b", "<f4", 3)])
NODE_DTYPE = np.dtype(
    [
        ("min_x", "<f4", 4), ("min_y", "<f4", 4), ("min_z", "<f4", 4),
        ("max_x", "<f4", 4), ("max_y", "<f4", 4), ("max_z", "<f4", 4),
        ("children", "<u4", 4), ("valid_mask", "u1"), ("pad", "u1", 15),
    ]
)
Q8_NODE_DTYPE = np.dtype(
    [
        ("origin", "<f4", 3), ("exp", "i1", 3), ("flags", "u1"), ("q", "u1", 24),
        ("pad", "<u4", 2), ("child", "<u4", 4),
    ]
)
assert RAY_DTYPE.itemsize == 32 and HIT_DTYPE.itemsize == 24
assert NODE_DTYPE.itemsize == 128 and Q8_NODE_DTYPE.itemsize == 64


class MeshDesc(ctypes.Structure):
    _fields_ = [
        ("positions", ctypes.c_void_p),
        ("vertex_count", ctypes.c_uint32),
        ("triangle_count", ctypes.c_uint32),
        ("indices", ctypes.c_void_p),
        ("nodes", ctypes.c_void_p),
        ("node_count", ctypes.c_uint32),
        ("pad", ctypes.c_uint32),
        ("primitive_indices", ctypes.c_void_p),
        ("normals", ctypes.c_void_p),
        ("uvs", ctypes.c_void_p),
    ]


class InstanceDesc(ctypes.Structure):
    _fields_ = [
        ("mesh", ctypes.c_uint32),
        ("material", ctypes.c_uint32),
        ("local_to_world", ctypes.c_float * 16),
        ("world_to_local", ctypes.c_float * 16),
    ]


# TptLightNode (80 bytes)
LIGHT_NODE_DTYPE = np.dtype([("flux", "<f4"), ("mu", "<f4", 3), ("sigma_s2", "<f4"), ("nu", "<f4", 3),
                             ("lambda", "<f4"), ("radius", "<f4"), ("area", "<f4"), ("left", "<u4"),
                             ("right", "<u4"), ("parent", "<u4"), ("instance", "<u4"), ("triangle", "<u4"),
                             ("leaf", "<u4"), ("pad", "<u4", 3)])
assert LIGHT_NODE_DTYPE.itemsize == 80


class SceneDesc(ctypes.Structure):
    _fields_ = [
        ("meshes", ctypes.c_void_p),
        ("mesh_count", ctypes.c_uint32),
        ("instance_count", ctypes.c_uint32),
        ("instances", ctypes.c_void_p),
        ("tlas_nodes", ctypes.c_void_p),
        ("tlas_node_count", ctypes.c_uint32),
        ("material_count", ctypes.c_uint32),
        ("tlas_primitive_indices", ctypes.c_void_p),
        ("materials", ctypes.c_void_p),
        ("env_pixels", ctypes.c_void_p),
        ("env_width", ctypes.c_uint32),
        ("env_height", ctypes.c_uint32),
        ("env_scale", ctypes.c_float),
        ("env_rotate_y", ctypes.c_float),
        ("textures", ctypes.c_void_p),
        ("texture_count", ctypes.c_uint32),
        ("directional_count", ctypes.c_uint32),
        ("directional_lights", ctypes.c_void_p),
    ]


# TptDirectionalLight (32 bytes)
DIRECTIONAL_DTYPE = np.dtype([("direction", "<f4", 3), ("intensity", "<f4"), ("color", "<f4", 3), ("pad", "<f4")])
assert DIRECTIONAL_DTYPE.itemsize == 32


class TextureDesc(ctypes.Structure):
    _fields_ = [
        ("pixels", ctypes.c_void_p),
        ("width", ctypes.c_uint32),
        ("height", ctypes.c_uint32),
        ("channels", ctypes.c_uint32),
        ("pad", ctypes.c_uint32),
    ]


# TptTextureQuery (24 bytes) and the lookup modes
TEXTURE_QUERY_DTYPE = np.dtype([("uv", "<f4", 2), ("dstdx", "<f4", 2), ("dstdy", "<f4", 2)])
assert TEXTURE_QUERY_DTYPE.itemsize == 24
TEXTURE_FILTERED, TEXTURE_BILINEAR, TEXTURE_NEAREST, TEXTURE_TRILINEAR, TEXTURE_MIP_BILINEAR = 0, 1, 2, 3, 4

assert ctypes.sizeof(MeshDesc) == 64 and ctypes.sizeof(SceneDesc) == 104 and ctypes.sizeof(TextureDesc) == 24
assert ctypes.sizeof(InstanceDesc) == 136


class Camera(ctypes.Structure):
    """TptCamera: PinholeCamera after ::new plus the pixel spread angle."""

    _fields_ = [
        ("eye", ctypes.c_float * 3), ("tan_half_fov_y", ctypes.c_float),
        ("forward", ctypes.c_float * 3), ("pixel_spread", ctypes.c_float),
        ("right", ctypes.c_float * 3), ("exposure", ctypes.c_float),
        ("image_up", ctypes.c_float * 3), ("pad", ctypes.c_float),
    ]


assert ctypes.sizeof(Camera) == 64


class RenderDesc(ctypes.Structure):
    _fields_ = [("camera", Camera), ("width", ctypes.c_uint32), ("height", ctypes.c_uint32),
                ("samples_per_pixel", ctypes.c_uint32), ("max_depth", ctypes.c_uint32)]


assert ctypes.sizeof(RenderDesc) == 80


BSDF_QUERY_DTYPE = np.dtype([("wo", "<f4", 3), ("wi", "<f4", 3), ("u", "<f4", 3), ("material", "<u4"),
                             ("front_face", "<u4"), ("pad", "<u4")])
BSDF_RESULT_DTYPE = np.dtype([("f", "<f4", 3), ("pdf", "<f4"), ("weight", "<f4", 3), ("sample_pdf", "<f4"),
                              ("wi", "<f4", 3), ("flags", "<u4")])
assert BSDF_QUERY_DTYPE.itemsize == 48 and BSDF_RESULT_DTYPE.itemsize == 48


class TuneResult(ctypes.Structure):
    _fields_ = [("wf_paths", ctypes.c_uint32), ("blocks_per_cu", ctypes.c_uint32), ("refill_below", ctypes.c_uint32),
                ("trials", ctypes.c_uint32), ("wf_sort", ctypes.c_uint32), ("ms_default", ctypes.c_float),
                ("ms_best", ctypes.c_float), ("lds_depth", ctypes.c_uint32), ("wf_tile", ctypes.c_uint32)]


class FrameDesc(ctypes.Structure):
    """TptFrameDesc: one frame of tpt_render_batch."""
    _fields_ = [("render", RenderDesc), ("rgb", ctypes.c_void_p)]


assert ctypes.sizeof(FrameDesc) == 88


class AdjointFrameDesc(ctypes.Structure):
    """TptAdjointFrameDesc: one view of tpt_render_adjoint_batch."""
    _fields_ = [("render", RenderDesc), ("dloss_dimage", ctypes.c_void_p)]


assert ctypes.sizeof(AdjointFrameDesc) == 88


class TptError(RuntimeError):
    """A non-zero status from libtpt, carrying tpt_last_error()."""

    def __init__(self, status: int, message: str):
        super().__init__(f"libtpt status {status}: {message}")
        self.status = status


_lib = None

# (name, restype, argtypes) for every symbol of include/tpt.h
_SIGNATURES = [
    ("tpt_last_error", ctypes.c_char_p, []),
    ("tpt_device_create", ctypes.c_void_p, [ctypes.c_int32]),
    ("tpt_device_release", None, [ctypes.c_void_p]),
    ("tpt_device_set_stream", ctypes.c_int32, [ctypes.c_void_p, ctypes.c_void_p]),
    ("tpt_device_synchronize", ctypes.c_int32, [ctypes.c_void_p]),
    ("tpt_device_set_param", ctypes.c_int32, [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int64]),
    ("tpt_qbvh_build", ctypes.c_void_p, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int32]),
    ("tpt_qbvh_release", None, [ctypes.c_void_p]),
    ("tpt_qbvh_node_count", ctypes.c_size_t, [ctypes.c_void_p]),
    ("tpt_qbvh_primitive_count", ctypes.c_size_t, [ctypes.c_void_p]),
    ("tpt_qbvh_export", ctypes.c_int32, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    ("tpt_qbvh_quantize", ctypes.c_int32, [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p]),
    ("tpt_scene_create", ctypes.c_void_p, [ctypes.c_void_p, ctypes.c_void_p]),
    ("tpt_qbvh_build_device", ctypes.c_void_p,
     [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_uint32]),
    ("tpt_scene_q8_nodes", ctypes.c_int32,
     [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p, ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t)]),
    ("tpt_scene_desc_save", ctypes.c_int32, [ctypes.c_void_p, ctypes.c_char_p]),
    ("tpt_scene_desc_load", ctypes.c_void_p, [ctypes.c_char_p]),
    ("tpt_scene_desc_free", None, [ctypes.c_void_p]),
    ("tpt_scene_release", None, [ctypes.c_void_p]),
    ("tpt_scene_device_bytes", ctypes.c_size_t, [ctypes.c_void_p]),
    ("tpt_scene_light_tree", ctypes.c_int32, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p]),
    ("tpt_light_tree_build", ctypes.c_int32, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p]),
    ("tpt_trace_closest", ctypes.c_int32,
     [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p, ctypes.c_uint32]),
    ("tpt_trace_any", ctypes.c_int32,
     [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
      ctypes.c_uint32]),
    ("tpt_camera_rays", ctypes.c_int32,
     [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
      ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32]),
    ("tpt_path_randoms", ctypes.c_int32,
     [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
      ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32]),
    ("tpt_render", ctypes.c_int32, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32]),
    ("tpt_render_batch", ctypes.c_int32, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32]),
    ("tpt_render_adjoint", ctypes.c_int32,
     [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32]),
    ("tpt_render_adjoint_texels", ctypes.c_int32,
     [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32]),
    ("tpt_render_adjoint_params", ctypes.c_int32,
     [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32]),
    ("tpt_render_adjoint_batch", ctypes.c_int32,
     [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32]),
    ("tpt_scene_set_materials", ctypes.c_int32, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32]),
    ("tpt_scene_set_nee_policy", ctypes.c_int32, [ctypes.c_void_p, ctypes.c_uint32]),
    ("tpt_scene_set_texture", ctypes.c_int32, [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p, ctypes.c_uint32]),
    ("tpt_material_bsdf", ctypes.c_int32,
     [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p, ctypes.c_uint32]),
    ("tpt_texture_sample", ctypes.c_int32,
     [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p,
      ctypes.c_uint32]),
    ("tpt_texture_info", ctypes.c_int32,
     [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    ("tpt_texture_level", ctypes.c_int32,
     [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p,
      ctypes.c_void_p]),
    ("tpt_autotune_render", ctypes.c_int32,
     [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_void_p]),
    ("tpt_trace_noalloc", ctypes.c_int32,
     [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p, ctypes.c_void_p,
      ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32]),
    ("tpt_trace_counts", ctypes.c_int32,
     [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p, ctypes.c_void_p,
      ctypes.c_void_p, ctypes.c_uint32]),
]


def lib() -> ctypes.CDLL:
    """Load libtpt.so once; raise if it has not been built."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise ImportError(
                f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'` "
                "(there is no CPU fallback)"
            )
        handle = ctypes.CDLL(os.fspath(LIB_PATH))
        for name, res, args in _SIGNATURES:
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _lib = handle
    return _lib


def last_error() -> str:
    return lib().tpt_last_error().decode("utf-8", "replace")


def check(status: int) -> None:
    if status != TPT_OK:
        raise TptError(status, last_error())


def ptr(a) -> int | None:
    """Address of a numpy array or torch tensor (None for None)."""
    if a is None:
        return None
    if isinstance(a, np.ndarray):
        if not a.flags["C_CONTIGUOUS"]:
            raise ValueError("array must be C-contiguous")
        return a.ctypes.data
    return a.data_ptr()  # torch tensor
