"""Python mirror of the reference's fusion API on top of the C ABI.

* :class:`DeviceMap` mirrors ``MapStruct<TVoxel>`` (include/map_struct.h:44-68)
  with the stage functions of src/device_functions.h:11-66 as methods.
* :class:`VoxelMapping` mirrors ``vmap::VoxelMapping`` (include/vmapping.h:10-27,
  src/vmapping.cpp:150-190), same method names.
* :func:`computeNormal` / :func:`renderScene` mirror include/image_proc.h:10-11.

Images are torch CUDA tensors (device memory); poses are 4x4 camera-to-world
matrices (numpy or torch, row-major as printed) or column-major float[16]
arrays. Every call goes to the gfx950 kernels; there is no CPU path.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _capi
from ._capi import check, lib

try:
    import torch
except ImportError:  # pragma: no cover - torch is part of the image
    torch = None


def _stream(stream):
    if stream is None:
        if torch is not None and torch.cuda.is_available():
            return C.c_void_p(torch.cuda.current_stream().cuda_stream)
        return None
    if hasattr(stream, "cuda_stream"):
        return C.c_void_p(stream.cuda_stream)
    return C.c_void_p(int(stream))


def pose_colmajor(T) -> "C.Array":
    """4x4 matrix (row-major indexing) or 16 column-major floats -> float[16]."""
    a = np.asarray(T.cpu() if hasattr(T, "cpu") else T, dtype=np.float32)
    flat = a.T.reshape(-1) if a.shape == (4, 4) else a.reshape(-1)
    if flat.size != 16:
        raise ValueError("pose must be 4x4 or 16 column-major floats")
    return _capi.floats(flat, 16)


def intrinsics(K) -> "C.Array":
    """3x3 K or (fx, fy, cx, cy) -> float[4]."""
    a = np.asarray(K.cpu() if hasattr(K, "cpu") else K, dtype=np.float32)
    if a.shape == (3, 3):
        a = np.array([a[0, 0], a[1, 1], a[0, 2], a[1, 2]], np.float32)
    if a.size != 4:
        raise ValueError("K must be 3x3 or (fx, fy, cx, cy)")
    return _capi.floats(a.reshape(-1), 4)


def _image(t, channels_bytes: int, name: str):
    """(ptr, pitch, width, height) of a CUDA tensor laid out HxW[xC] with unit inner stride."""
    if torch is None or not isinstance(t, torch.Tensor):
        raise TypeError(f"{name} must be a torch tensor in device memory")
    if not t.is_cuda:
        raise TypeError(f"{name} must be a CUDA (HIP) tensor")
    if t.dim() < 2:
        raise ValueError(f"{name} must be at least 2-D")
    H, W = t.shape[0], t.shape[1]
    es = t.element_size()
    row_elems = 1
    for d in range(t.dim() - 1, 0, -1):
        if t.stride(d) != row_elems:
            raise ValueError(f"{name} rows must be contiguous")
        row_elems *= t.shape[d]
    if row_elems * es != W * channels_bytes:
        raise ValueError(f"{name}: expected {channels_bytes} bytes per pixel")
    return C.c_void_p(t.data_ptr()), t.stride(0) * es, W, H


class DeviceMap:
    """MapStruct<TVoxel>: hash table + excess list + block heap + voxels on the GPU."""

    def __init__(self, n_entry: int, n_bucket: int, n_block: int, voxel_size: float, trunc_dist: float,
                 voxel_format: int = _capi.VOXEL_I16_U8, depth_min: float = 0.1, depth_max: float = 5.0,
                 max_weight: float = 255.0):
        d = _capi.MapDesc(n_entry, n_bucket, n_block, voxel_size, trunc_dist, depth_min, depth_max,
                          voxel_format, max_weight)
        h = C.c_void_p()
        self._lib = lib()  # kept for release() during interpreter shutdown
        check(self._lib.vm_map_create(C.byref(d), C.byref(h)))
        self._h = h
        self.desc = d

    # MapStruct::create / release / reset / empty
    @classmethod
    def create(cls, nEntry, nBucket, nVBlock, voxelSize, truncationDist, **kw):
        return cls(nEntry, nBucket, nVBlock, voxelSize, truncationDist, **kw)

    def release(self):
        if getattr(self, "_h", None):
            self._lib.vm_map_destroy(self._h)
            self._h = None

    __del__ = release

    def empty(self) -> bool:
        return not self._h

    def reset(self, stream=None):
        check(lib().vm_map_reset(self._h, _stream(stream)))

    @property
    def handle(self):
        return self._h

    @property
    def voxel_bytes(self) -> int:
        return {0: 4, 1: 6, 2: 12}[self.desc.voxel_format]

    def stats(self, stream=None) -> dict:
        s = _capi.MapStats()
        check(lib().vm_map_get_stats(self._h, _stream(stream), C.byref(s)))
        return {k: getattr(s, k) for k, _ in _capi.MapStats._fields_}

    def getVisibleBlockCount(self) -> int:
        return self.stats()["visible_blocks"]

    def arrays(self) -> dict:
        a = _capi.MapArrays()
        check(lib().vm_map_get_arrays(self._h, C.byref(a)))
        return {k: getattr(a, k) for k, _ in _capi.MapArrays._fields_}

    # --- stages (device_functions.h) ---------------------------------------
    def _depth(self, depth, depth_scale):
        if depth.dtype == torch.float32:
            p, pitch, W, H = _image(depth, 4, "depth")
            return p, pitch, W, H, _capi.DEPTH_F32, 1.0
        if depth.dtype in (torch.uint16, torch.int16):
            p, pitch, W, H = _image(depth, 2, "depth")
            return p, pitch, W, H, _capi.DEPTH_U16, float(depth_scale)
        raise TypeError("depth must be float32 metres or uint16 raw units")

    def fuse(self, depth, K, cam_to_world, rgb=None, depth_scale: float = 1e-3, stream=None,
             return_count: bool = False):
        """FuseImage / FuseDepthAndImage. Returns the visible-block count if asked (synchronises)."""
        p, pitch, W, H, fmt, sc = self._depth(depth, depth_scale)
        rp, rpitch = None, 0
        if rgb is not None:
            rp, rpitch, rw, rh = _image(rgb, 3, "rgb")
            if (rw, rh) != (W, H):
                raise ValueError("rgb and depth sizes differ")
        cnt = C.c_uint32(0)
        check(lib().vm_fuse_depth(self._h, p, pitch, W, H, fmt, sc, rp, rpitch, intrinsics(K),
                                  pose_colmajor(cam_to_world), _stream(stream),
                                  C.byref(cnt) if return_count else None))
        return int(cnt.value) if return_count else None

    def allocate(self, depth, K, cam_to_world, depth_scale: float = 1e-3, stream=None):
        p, pitch, W, H, fmt, sc = self._depth(depth, depth_scale)
        check(lib().vm_allocate_blocks(self._h, p, pitch, W, H, fmt, sc, intrinsics(K),
                                       pose_colmajor(cam_to_world), _stream(stream)))

    def integrate(self, depth, K, cam_to_world, rgb=None, depth_scale: float = 1e-3, stream=None,
                  return_count: bool = False):
        """Visibility + integration after :meth:`allocate` (the rest of FuseImage)."""
        p, pitch, W, H, fmt, sc = self._depth(depth, depth_scale)
        rp, rpitch = None, 0
        if rgb is not None:
            rp, rpitch, rw, rh = _image(rgb, 3, "rgb")
            if (rw, rh) != (W, H):
                raise ValueError("rgb and depth sizes differ")
        cnt = C.c_uint32(0)
        check(lib().vm_integrate(self._h, p, pitch, W, H, fmt, sc, rp, rpitch, intrinsics(K),
                                 pose_colmajor(cam_to_world), _stream(stream),
                                 C.byref(cnt) if return_count else None))
        return int(cnt.value) if return_count else None

    def check_visible(self, width, height, K, cam_to_world, stream=None) -> int:
        cnt = C.c_uint32(0)
        check(lib().vm_check_visible_blocks(self._h, width, height, intrinsics(K), pose_colmajor(cam_to_world),
                                            _stream(stream), C.byref(cnt)))
        return int(cnt.value)

    def raycast(self, vmap, K, cam_to_world, stream=None):
        p, pitch, W, H = _image(vmap, 16, "vmap")
        check(lib().vm_raycast(self._h, p, pitch, W, H, intrinsics(K), pose_colmajor(cam_to_world),
                               _stream(stream)))
        return vmap

    def render_scene(self, vmap, K, cam_to_world, stream=None):
        p, pitch, W, H = _image(vmap, 16, "vmap")
        check(lib().vm_render_scene(self._h, p, pitch, W, H, intrinsics(K), pose_colmajor(cam_to_world),
                                    _stream(stream)))
        return vmap

    def raycast_backward(self, vmap, K, cam_to_world, grad_tsdf, grad_pose, grad_vertex=None, grad_normal=None,
                         stream=None, stats=True):
        """Adjoint of raycast + normals (float voxel format): accumulates
        d loss / d TSDF into grad_tsdf (float32, n_block * 512) and d loss / d xi
        into grad_pose (float64, 6) for camToWorld = T exp(xi). With stats,
        waits for the result and returns {hit_pixels, overflow_pixels};
        without, returns None and leaves the work queued on the stream."""
        p, pitch, W, H = _image(vmap, 16, "vmap")
        gv, gvp = (None, 0) if grad_vertex is None else _image(grad_vertex, 16, "grad_vertex")[:2]
        gn, gnp = (None, 0) if grad_normal is None else _image(grad_normal, 16, "grad_normal")[:2]
        if grad_tsdf.numel() != self.desc.n_block * 512 or grad_tsdf.dtype != torch.float32:
            raise ValueError("grad_tsdf must be float32 with n_block * 512 elements")
        if grad_pose.numel() != 6 or grad_pose.dtype != torch.float64:
            raise ValueError("grad_pose must be 6 float64")
        st = _capi.BackwardStats()
        check(lib().vm_raycast_backward(self._h, p, pitch, gv, gvp, gn, gnp, W, H, intrinsics(K),
                                        pose_colmajor(cam_to_world), C.c_void_p(grad_tsdf.data_ptr()),
                                        C.c_void_p(grad_pose.data_ptr()), _stream(stream),
                                        C.byref(st) if stats else None))
        return {"hit_pixels": st.hit_pixels, "overflow_pixels": st.overflow_pixels} if stats else None

    def project_rendering_blocks(self, zrange_x, zrange_y, K, world_to_cam, stream=None) -> int:
        px, pitch, W8, H8 = _image(zrange_x, 4, "zrange_x")
        py, pitch2, _, _ = _image(zrange_y, 4, "zrange_y")
        if pitch2 != pitch:
            raise ValueError("zrange maps must share a pitch")
        cnt = C.c_uint32(0)
        check(lib().vm_project_rendering_blocks(self._h, px, py, pitch, W8, H8, intrinsics(K),
                                                pose_colmajor(world_to_cam), _stream(stream), C.byref(cnt)))
        return int(cnt.value)

    def polygonize(self, verts, norms, stream=None):
        """Marching cubes into device buffers of shape (cap, 9). Returns (blocks, triangles)."""
        cap = verts.shape[0]
        if norms.shape[0] != cap or verts.numel() != cap * 9 or norms.numel() != cap * 9:
            raise ValueError("verts and norms must be (cap, 9) float32")
        b, t = C.c_uint32(0), C.c_uint32(0)
        check(lib().vm_polygonize(self._h, C.c_void_p(verts.data_ptr()), C.c_void_p(norms.data_ptr()), cap,
                                  _stream(stream), C.byref(b), C.byref(t)))
        return int(b.value), int(t.value)

    def polygonize_noalloc(self, verts, norms, stream=None):
        """polygonize through vm_polygonize_noalloc: no device allocation, the
        library's fixed mesh work array (at most VM_NOALLOC_MAX_BLOCKS
        allocated blocks)."""
        cap = verts.shape[0]
        if norms.shape[0] != cap or verts.numel() != cap * 9 or norms.numel() != cap * 9:
            raise ValueError("verts and norms must be (cap, 9) float32")
        b, t = C.c_uint32(0), C.c_uint32(0)
        check(lib().vm_polygonize_noalloc(self._h, C.c_void_p(verts.data_ptr()), C.c_void_p(norms.data_ptr()), cap,
                                          _stream(stream), C.byref(b), C.byref(t)))
        return int(b.value), int(t.value)

    # --- state transfer -------------------------------------------------------
    def _nbytes(self, which):
        d = self.desc
        return {0: 20 * d.n_entry, 1: 4 * d.n_block, 2: self.voxel_bytes * 512 * d.n_block, 3: 64,
                4: 20 * d.n_entry, 5: 16 * d.n_block}[which]

    def refine_poses(self, depth, K, init_poses, iters: int = 8, stride: int = 4, max_residual: float = 0.05,
                     depth_scale: float = 1e-3, stream=None, return_system: bool = False, max_depth: float = 0.0):
        """Gauss-Newton refinement of camera-to-world hypotheses against one
        depth image (vm_refine_poses). init_poses: (n, 4, 4) matrices or
        (n, 16) column-major. Pixels deeper than max_depth metres (if > 0)
        are left out. Returns (poses (n, 4, 4) float32, cost
        (n, iters + 1, 2) = {mean squared residual, count} per evaluation),
        plus the normal equations (n, iters + 1, 27) if return_system."""
        p, pitch, W, H, fmt, sc = self._depth(depth, depth_scale)
        P = np.asarray(init_poses, np.float32)
        if P.ndim == 3 and P.shape[1:] == (4, 4):
            P = np.ascontiguousarray(P.transpose(0, 2, 1).reshape(-1, 16))
        elif P.ndim != 2 or P.shape[1] != 16:
            raise ValueError("init_poses must be (n, 4, 4) or (n, 16)")
        n = P.shape[0]
        P = np.ascontiguousarray(P)
        out = np.zeros((n, 16), np.float32)
        cost = np.zeros((n, iters + 1, 2), np.float32)
        system = np.zeros((n, iters + 1, 27), np.float64)
        fp = lambda a: a.ctypes.data_as(C.POINTER(C.c_float))
        check(lib().vm_refine_poses(self._h, p, pitch, W, H, fmt, sc, intrinsics(K), fp(P), n, iters, stride,
                                    max_residual, max_depth, fp(out), fp(cost),
                                    system.ctypes.data_as(C.POINTER(C.c_double)),
                                    _stream(stream)))
        poses = out.reshape(n, 4, 4).transpose(0, 2, 1).copy()
        return (poses, cost, system) if return_system else (poses, cost)

    def download(self, which: int) -> np.ndarray:
        n = self._nbytes(which)
        out = np.empty(n, np.uint8)
        check(lib().vm_map_download(self._h, which, out.ctypes.data_as(C.c_void_p), n))
        return out

    def upload(self, which: int, data: np.ndarray):
        data = np.ascontiguousarray(data).view(np.uint8).reshape(-1)
        check(lib().vm_map_upload(self._h, which, data.ctypes.data_as(C.c_void_p), data.size))

    def hash_table(self) -> np.ndarray:
        return self.download(0).view(np.int32).reshape(-1, 5)

    def counters(self) -> np.ndarray:
        return self.download(3).view(np.int32)

    def download_blocks(self, first: int, count: int) -> np.ndarray:
        out = np.empty(count * 512 * self.voxel_bytes, np.uint8)
        check(lib().vm_map_download_blocks(self._h, first, count, out.ctypes.data_as(C.c_void_p)))
        return out

    def upload_blocks(self, first: int, data: np.ndarray):
        data = np.ascontiguousarray(data).view(np.uint8).reshape(-1)
        per = 512 * self.voxel_bytes
        if data.size % per:
            raise ValueError(f"voxel data must be a whole number of {per}-byte blocks")
        check(lib().vm_map_upload_blocks(self._h, first, data.size // per, data.ctypes.data_as(C.c_void_p)))

    # map file: numpy .npz (no pickled objects) holding the descriptor, the
    # hash table, the heap, the counters and the voxels of the allocated
    # blocks [0, allocated); everything else of a map is its reset state
    MAP_FILE_VERSION = 1

    def save(self, path: str) -> None:
        d = self.desc
        counters = self.counters()
        alloc = self.stats()["allocated_blocks"]
        np.savez(path, version=np.int32(self.MAP_FILE_VERSION),
                 desc_int=np.array([d.n_entry, d.n_bucket, d.n_block, d.voxel_format], np.int64),
                 desc_float=np.array([d.voxel_size, d.trunc_dist, d.depth_min, d.depth_max, d.max_weight],
                                     np.float32),
                 hash_table=self.hash_table(), heap=self.download(1).view(np.int32), counters=counters,
                 voxels=self.download_blocks(0, alloc))

    @classmethod
    def load(cls, path: str) -> "DeviceMap":
        with np.load(path, allow_pickle=False) as z:
            if int(z["version"]) != cls.MAP_FILE_VERSION:
                raise ValueError(f"{path}: map file version {int(z['version'])} is not supported")
            ne, nb, nv, fmt = (int(v) for v in z["desc_int"])
            vs, tr, dmin, dmax, mw = (float(v) for v in z["desc_float"])
            m = cls(ne, nb, nv, vs, tr, fmt, dmin, dmax, mw)
            m.upload(0, z["hash_table"])
            m.upload(1, z["heap"])
            m.upload(3, z["counters"])
            m.upload_blocks(0, z["voxels"])
        return m


def computeNormal(vmap, nmap=None, stream=None):
    """vmap::computeNormal (image_proc.cu:95-104)."""
    if nmap is None:
        nmap = torch.empty_like(vmap)
    p, pitch, W, H = _image(vmap, 16, "vmap")
    q, qpitch, W2, H2 = _image(nmap, 16, "nmap")
    if (W, H) != (W2, H2):
        raise ValueError("vmap and nmap sizes differ")
    check(lib().vm_compute_normals(p, pitch, q, qpitch, W, H, _stream(stream)))
    return nmap


def renderScene(vmap, nmap, image=None, stream=None):
    """vmap::renderScene (image_proc.cu:63-72): Phong shading to RGBA8."""
    H, W = vmap.shape[0], vmap.shape[1]
    if image is None:
        image = torch.empty((H, W, 4), dtype=torch.uint8, device=vmap.device)
    p, pitch, _, _ = _image(vmap, 16, "vmap")
    q, qpitch, _, _ = _image(nmap, 16, "nmap")
    r, rpitch, _, _ = _image(image, 4, "image")
    check(lib().vm_render_shading(p, pitch, q, qpitch, r, rpitch, W, H, _stream(stream)))
    return image


def surface_points(verts, n_triangles: int, points=None, stream=None):
    """GetSurfacePoints (marching_cubes.cu:330-344): triangle centroids."""
    if points is None:
        points = torch.empty((max(n_triangles, 1), 3), dtype=torch.float32, device=verts.device)
    check(lib().vm_surface_points(C.c_void_p(verts.data_ptr()), C.c_void_p(points.data_ptr()), n_triangles,
                                  _stream(stream)))
    return points[:n_triangles]


class Batch:
    """Many scenes fused and raycast in lockstep (vm_batch_* of the C ABI).

    scenes: list of dicts with keys map (DeviceMap), K, depth (CUDA tensor
    (F, H, W) of all frames, float32 metres or uint16), depth_scale, rgb
    (CUDA uint8 (F, H, W, 3) or None), poses ((F, 4, 4) or (F, 16)
    column-major camera-to-world). Output vertex/normal maps are allocated
    here: ``vmaps[i]`` / ``nmaps[i]`` (H, W, 4) float32.
    """

    def __init__(self, scenes):
        import torch
        self._keep = []
        descs = (_capi.BatchScene * len(scenes))()
        self.vmaps, self.nmaps = [], []
        for i, sc in enumerate(scenes):
            dep = sc["depth"]
            if dep.dim() != 3 or not dep.is_cuda:
                raise ValueError("depth must be a CUDA tensor (frames, H, W)")
            F, H, W = dep.shape
            u16 = dep.dtype in (torch.uint16, torch.int16)
            es = 2 if u16 else 4
            vm = torch.zeros((H, W, 4), dtype=torch.float32, device=dep.device)
            nm = torch.zeros_like(vm)
            self.vmaps.append(vm)
            self.nmaps.append(nm)
            pa = np.asarray(sc["poses"], np.float32)
            P = pa.transpose(0, 2, 1).reshape(F, 16) if pa.ndim == 3 else pa.reshape(F, 16)
            P = np.ascontiguousarray(P)
            self._keep.append(P)
            d = descs[i]
            d.map = sc["map"].handle
            d.width, d.height = W, H
            d.K[:] = list(intrinsics(sc["K"]))
            d.depth_format = _capi.DEPTH_U16 if u16 else _capi.DEPTH_F32
            d.depth_scale = float(sc.get("depth_scale", 1e-3))
            d.depth_frames = dep.data_ptr()
            d.depth_pitch = dep.stride(1) * es
            d.depth_frame_stride = dep.stride(0) * es
            rgb = sc.get("rgb")
            if rgb is not None:
                d.rgb_frames = rgb.data_ptr()
                d.rgb_pitch = rgb.stride(1)
                d.rgb_frame_stride = rgb.stride(0)
            d.poses = P.ctypes.data
            d.n_frames = F
            d.vmap, d.vmap_pitch = vm.data_ptr(), W * 16
            d.nmap, d.nmap_pitch = nm.data_ptr(), W * 16
        self._scenes = scenes
        h = C.c_void_p()
        self._lib = lib()  # kept for close() during interpreter shutdown
        check(self._lib.vm_batch_create(descs, len(scenes), C.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            self._lib.vm_batch_destroy(self._h)
            self._h = None

    __del__ = close

    def set_frame(self, f: int, stream=None):
        check(lib().vm_batch_set_frame(self._h, int(f), _stream(stream)))

    def frame(self, stream=None) -> int:
        v = C.c_int32(0)
        check(lib().vm_batch_get_frame(self._h, _stream(stream), C.byref(v)))
        return int(v.value)

    def step(self, stream=None):
        check(lib().vm_batch_step(self._h, _stream(stream)))

    def capture(self, stream=None):
        check(lib().vm_batch_capture(self._h, _stream(stream)))

    def tuning(self) -> dict:
        t = _capi.BatchTuning()
        check(lib().vm_batch_get_tuning(self._h, C.byref(t)))
        return {k: getattr(t, k) for k, _ in _capi.BatchTuning._fields_}

    def set_tuning(self, **kw):
        t = _capi.BatchTuning(**{**self.tuning(), **kw})
        check(lib().vm_batch_set_tuning(self._h, C.byref(t)))

    def tune(self, steps: int = 2, stream=None) -> dict:
        t = _capi.BatchTuning()
        check(lib().vm_batch_tune(self._h, _stream(stream), steps, C.byref(t)))
        return {k: getattr(t, k) for k, _ in _capi.BatchTuning._fields_}


MAX_VERTS_BUFFER = 20_000_000  # triangles, vmapping.cpp:6


class VoxelMapping:
    """vmap::VoxelMapping (include/vmapping.h:10-27). Same methods and defaults:
    CreateMap uses nBucket = int(0.9 * numVoxels), truncation = 4 * voxelSize and
    the reference Voxel format (vmapping.cpp:57-75)."""

    def __init__(self, w: int, h: int, K, mesh_capacity: int = MAX_VERTS_BUFFER):
        self.width, self.height = int(w), int(h)
        self.K = intrinsics(K)
        self.map: DeviceMap | None = None
        self.numVisBlock = 0
        self.mesh_capacity = int(mesh_capacity)
        self._verts = self._norms = None

    def reset(self):
        self._need_map().reset()

    def CreateMap(self, numEntries: int, numVoxels: int, voxelSize: float = 0.01):
        nb = int(numVoxels * 0.9)
        self.map = DeviceMap(numEntries, nb, numVoxels, voxelSize, voxelSize * 4)

    def _need_map(self) -> DeviceMap:
        if self.map is None:
            raise RuntimeError("CreateMap must be called first")
        return self.map

    def FuseDepth(self, depth, camToWorld):
        self.numVisBlock = self._need_map().fuse(depth, self.K, camToWorld, return_count=True)

    def FuseDepthAndImage(self, rgb, depth, camToWorld):
        """The reference leaves this a stub (vmapping.cpp:167-170); here it fuses colour
        when the map's format carries colour (FuseDepthAndImage, depth_fusion.cu:383-488)."""
        self.numVisBlock = self._need_map().fuse(depth, self.K, camToWorld, rgb=rgb, return_count=True)

    def RenderScene(self, vmap, camToWorld):
        m = self._need_map()
        if vmap is None:
            vmap = torch.empty((self.height, self.width, 4), dtype=torch.float32, device="cuda")
        m.render_scene(vmap, self.K, camToWorld)
        return vmap

    def _mesh(self):
        m = self._need_map()
        if self._verts is None:
            self._verts = torch.empty((self.mesh_capacity, 9), dtype=torch.float32, device="cuda")
            self._norms = torch.empty((self.mesh_capacity, 9), dtype=torch.float32, device="cuda")
        blocks, tris = m.polygonize(self._verts, self._norms)
        self.numVisBlock = blocks
        return tris

    def Polygonize(self):
        """Returns (triangle_count, verts[T*9], norms[T*9]) as host arrays."""
        t = self._mesh()
        return t, self._verts[:t].cpu().numpy().reshape(-1), self._norms[:t].cpu().numpy().reshape(-1)

    def GetSurfacePoints(self):
        t = self._mesh()
        if t == 0:
            return np.zeros((0, 3), np.float32)
        return surface_points(self._verts, t).cpu().numpy()
