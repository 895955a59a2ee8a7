"""ctypes binding of ``include/vmapping_capi.h`` (libvmapping_amd.so).

The library is loaded from the package directory; there is no fallback: if
it is missing, :func:`lib` raises with the build command to run.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libvmapping_amd.so")

VM_OK = 0
VOXEL_I16_U8 = 0
VOXEL_I16_U8_RGB = 1
VOXEL_F32_F32_RGB8 = 2
DEPTH_F32 = 0
DEPTH_U16 = 1
STATUS_HEAP_FULL = 1
STATUS_EXCESS_FULL = 2
STATUS_CHAIN_BOUND = 4
STATUS_MESH_CLAMPED = 8

ARRAY_HASH_TABLE = 0
ARRAY_HEAP = 1
ARRAY_VOXELS = 2
ARRAY_COUNTERS = 3
ARRAY_VISIBLE_TABLE = 4
ARRAY_BLOCK_POS = 5


class MapDesc(C.Structure):
    _fields_ = [
        ("n_entry", C.c_int32),
        ("n_bucket", C.c_int32),
        ("n_block", C.c_int32),
        ("voxel_size", C.c_float),
        ("trunc_dist", C.c_float),
        ("depth_min", C.c_float),
        ("depth_max", C.c_float),
        ("voxel_format", C.c_int32),
        ("max_weight", C.c_float),
    ]


class MapArrays(C.Structure):
    _fields_ = [
        ("hash_table", C.c_void_p),
        ("visible_table", C.c_void_p),
        ("heap", C.c_void_p),
        ("heap_ptr", C.c_void_p),
        ("excess_ptr", C.c_void_p),
        ("bucket_mutex", C.c_void_p),
        ("visible_block_num", C.c_void_p),
        ("voxel_block", C.c_void_p),
        ("block_pos", C.c_void_p),
    ]


class MapStats(C.Structure):
    _fields_ = [
        ("allocated_blocks", C.c_int32),
        ("heap_ptr", C.c_int32),
        ("excess_used", C.c_int32),
        ("visible_blocks", C.c_int32),
        ("status", C.c_int32),
        ("frames_fused", C.c_int32),
        ("sum_visible", C.c_int64),
    ]


class BatchScene(C.Structure):
    _fields_ = [
        ("map", C.c_void_p),
        ("width", C.c_int32),
        ("height", C.c_int32),
        ("K", C.c_float * 4),
        ("depth_format", C.c_int32),
        ("depth_scale", C.c_float),
        ("depth_frames", C.c_void_p),
        ("depth_pitch", C.c_int64),
        ("depth_frame_stride", C.c_int64),
        ("rgb_frames", C.c_void_p),
        ("rgb_pitch", C.c_int64),
        ("rgb_frame_stride", C.c_int64),
        ("poses", C.c_void_p),
        ("n_frames", C.c_int32),
        ("vmap", C.c_void_p),
        ("vmap_pitch", C.c_int64),
        ("nmap", C.c_void_p),
        ("nmap_pitch", C.c_int64),
    ]


class BatchTuning(C.Structure):
    _fields_ = [
        ("ray_tile_width", C.c_int32),
        ("xcd_bands", C.c_int32),
        ("visible_wg_per_cu", C.c_int32),
        ("integrate_wg_per_cu", C.c_int32),
        ("scene_groups", C.c_int32),
        ("persistent", C.c_int32),
        ("persist_wg_per_cu", C.c_int32),
    ]


class BackwardStats(C.Structure):
    _fields_ = [("hit_pixels", C.c_int32), ("overflow_pixels", C.c_int32)]


_P = C.c_void_p
_F4 = C.POINTER(C.c_float)
_U32P = C.POINTER(C.c_uint32)

# name -> (restype, argtypes); mirrors include/vmapping_capi.h
SIGNATURES = {
    "vm_last_error": (C.c_char_p, []),
    "vm_version": (C.c_char_p, []),
    "vm_map_create": (C.c_int, [C.POINTER(MapDesc), C.POINTER(_P)]),
    "vm_map_destroy": (C.c_int, [_P]),
    "vm_map_reset": (C.c_int, [_P, _P]),
    "vm_map_get_desc": (C.c_int, [_P, C.POINTER(MapDesc)]),
    "vm_map_get_arrays": (C.c_int, [_P, C.POINTER(MapArrays)]),
    "vm_map_get_stats": (C.c_int, [_P, _P, C.POINTER(MapStats)]),
    "vm_fuse_depth": (C.c_int, [_P, _P, C.c_int64, C.c_int32, C.c_int32, C.c_int32, C.c_float, _P, C.c_int64,
                                _F4, _F4, _P, _U32P]),
    "vm_allocate_blocks": (C.c_int, [_P, _P, C.c_int64, C.c_int32, C.c_int32, C.c_int32, C.c_float, _F4, _F4,
                                     _P]),
    "vm_integrate": (C.c_int, [_P, _P, C.c_int64, C.c_int32, C.c_int32, C.c_int32, C.c_float, _P, C.c_int64,
                               _F4, _F4, _P, _U32P]),
    "vm_check_visible_blocks": (C.c_int, [_P, C.c_int32, C.c_int32, _F4, _F4, _P, _U32P]),
    "vm_raycast": (C.c_int, [_P, _P, C.c_int64, C.c_int32, C.c_int32, _F4, _F4, _P]),
    "vm_render_scene": (C.c_int, [_P, _P, C.c_int64, C.c_int32, C.c_int32, _F4, _F4, _P]),
    "vm_project_rendering_blocks": (C.c_int, [_P, _P, _P, C.c_int64, C.c_int32, C.c_int32, _F4, _F4, _P,
                                              _U32P]),
    "vm_compute_normals": (C.c_int, [_P, C.c_int64, _P, C.c_int64, C.c_int32, C.c_int32, _P]),
    "vm_render_shading": (C.c_int, [_P, C.c_int64, _P, C.c_int64, _P, C.c_int64, C.c_int32, C.c_int32, _P]),
    "vm_polygonize": (C.c_int, [_P, _P, _P, C.c_int64, _P, _U32P, _U32P]),
    "vm_polygonize_noalloc": (C.c_int, [_P, _P, _P, C.c_int64, _P, _U32P, _U32P]),
    "vm_surface_points": (C.c_int, [_P, _P, C.c_int64, _P]),
    "vm_raycast_backward": (C.c_int, [_P, _P, C.c_int64, _P, C.c_int64, _P, C.c_int64, C.c_int32, C.c_int32, _F4,
                                      _F4, _P, _P, _P, C.POINTER(BackwardStats)]),
    "vm_map_download": (C.c_int, [_P, C.c_int32, _P, C.c_int64]),
    "vm_map_upload": (C.c_int, [_P, C.c_int32, _P, C.c_int64]),
    "vm_refine_poses": (C.c_int, [_P, _P, C.c_int64, C.c_int32, C.c_int32, C.c_int32, C.c_float, _F4,
                                  C.POINTER(C.c_float), C.c_int32, C.c_int32, C.c_int32, C.c_float, C.c_float,
                                  C.POINTER(C.c_float), C.POINTER(C.c_float), C.POINTER(C.c_double), _P]),
    "vm_map_download_blocks": (C.c_int, [_P, C.c_int32, C.c_int32, _P]),
    "vm_map_upload_blocks": (C.c_int, [_P, C.c_int32, C.c_int32, _P]),
    "vm_batch_create": (C.c_int, [C.POINTER(BatchScene), C.c_int32, C.POINTER(_P)]),
    "vm_batch_destroy": (C.c_int, [_P]),
    "vm_batch_set_frame": (C.c_int, [_P, C.c_int32, _P]),
    "vm_batch_get_frame": (C.c_int, [_P, _P, C.POINTER(C.c_int32)]),
    "vm_batch_step": (C.c_int, [_P, _P]),
    "vm_batch_capture": (C.c_int, [_P, _P]),
    "vm_batch_get_tuning": (C.c_int, [_P, C.POINTER(BatchTuning)]),
    "vm_batch_set_tuning": (C.c_int, [_P, C.POINTER(BatchTuning)]),
    "vm_batch_tune": (C.c_int, [_P, _P, C.c_int32, C.POINTER(BatchTuning)]),
}

_lib = None


def lib() -> C.CDLL:
    """Load libvmapping_amd.so (once). Raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `python -m voxelmapping_amd.build` "
                "(there is no CPU fallback for the HIP pipeline)")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


class VMError(RuntimeError):
    pass


def check(rc: int) -> None:
    if rc != VM_OK:
        msg = lib().vm_last_error().decode(errors="replace")
        raise VMError(f"vmapping error {rc}: {msg}")


def floats(values, n: int):
    arr = (C.c_float * n)(*[float(v) for v in values])
    return arr
