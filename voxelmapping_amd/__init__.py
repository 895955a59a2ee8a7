"""MI355X (gfx950) voxel-hashing TSDF fusion pipeline.

Hash-table block allocation, visibility, TSDF / weight / colour integration,
raycasting, vertex-map normals and marching cubes as hand-written HIP
kernels behind the C ABI of ``include/vmapping_capi.h``; this package is the
Python side of that ABI (see :mod:`voxelmapping_amd.mapping`).
"""
from . import _capi
from ._capi import VMError
from .mapping import DeviceMap, VoxelMapping, computeNormal, renderScene, surface_points

__all__ = ["DeviceMap", "VoxelMapping", "computeNormal", "renderScene", "surface_points", "VMError", "_capi"]
