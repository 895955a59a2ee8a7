"""ctypes wrapper of the CPU oracle (liboracle_vm.so). TEST INFRASTRUCTURE.

Only tests, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` use this module, as the checker of the HIP library.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# VM_ORACLE_LIB points the tests' mutant check at a copy of the oracle built
# with a planted mistake (tests/test_oracle_mutants.py)
LIB_PATH = os.environ.get("VM_ORACLE_LIB") or os.path.join(HERE, "liboracle_vm.so")
_lib = None

_P = C.c_void_p


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} missing: run `make -C oracle`")
        L = C.CDLL(LIB_PATH)
        L.or_map_create.restype = _P
        L.or_map_create.argtypes = [C.c_int, C.c_int, C.c_int, C.c_float, C.c_float, C.c_float, C.c_float,
                                    C.c_int, C.c_float]
        L.or_map_destroy.argtypes = [_P]
        L.or_map_reset.argtypes = [_P]
        L.or_set_threads.argtypes = [C.c_int]
        L.or_fuse.restype = C.c_int
        L.or_fuse.argtypes = [_P, _P, _P, C.c_int, C.c_int, _P, _P]
        L.or_allocate.argtypes = [_P, _P, C.c_int, C.c_int, _P, _P]
        L.or_check_visible.restype = C.c_int
        L.or_check_visible.argtypes = [_P, C.c_int, C.c_int, _P, _P]
        L.or_raycast.argtypes = [_P, _P, C.c_int, C.c_int, _P, _P]
        L.or_render_scene.restype = C.c_int
        L.or_render_scene.argtypes = [_P, _P, C.c_int, C.c_int, _P, _P]
        L.or_compute_normals.argtypes = [_P, _P, C.c_int, C.c_int]
        L.or_render_shading.argtypes = [_P, _P, _P, C.c_int, C.c_int]
        L.or_polygonize.restype = C.c_int64
        L.or_polygonize.argtypes = [_P, _P, _P, C.c_int64, _P]
        L.or_surface_points.argtypes = [_P, _P, C.c_int64]
        L.or_project_rendering_blocks.restype = C.c_int
        L.or_project_rendering_blocks.argtypes = [_P, _P, _P, C.c_int, C.c_int, _P, _P]
        L.or_raycast_pixels_f32.argtypes = [_P, C.c_int, _P, _P, _P, _P, _P, _P, _P]
        L.or_raycast_pixels_f64.argtypes = [_P, C.c_int, _P, _P, _P, _P, _P, _P, C.c_int, _P, _P, _P]
        L.or_surface_depth_f64.argtypes = [_P, C.c_int, _P, _P, _P, _P, _P, _P]
        L.or_insert_block.restype = C.c_int
        L.or_insert_block.argtypes = [_P, C.c_int, C.c_int, C.c_int]
        L.or_map_counters.restype = C.c_int
        L.or_map_counters.argtypes = [_P, _P]
        for name in ("or_map_hash_table", "or_map_heap", "or_map_visible_table"):
            getattr(L, name).restype = C.POINTER(C.c_int32)
            getattr(L, name).argtypes = [_P]
        L.or_map_voxels.restype = C.POINTER(C.c_uint8)
        L.or_map_voxels.argtypes = [_P]
        L.or_voxel_bytes.restype = C.c_int
        L.or_voxel_bytes.argtypes = [C.c_int]
        L.or_hash.restype = C.c_int
        L.or_hash.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int]
        L.or_voxel_to_block.argtypes = [C.c_int, C.c_int, C.c_int, _P]
        L.or_voxel_to_local.argtypes = [C.c_int, C.c_int, C.c_int, _P]
        L.or_float2int_rd.restype = C.c_int
        L.or_float2int_rd.argtypes = [C.c_float]
        L.or_pose_inverse.argtypes = [_P, _P]
        L.or_mc_edge_table.restype = C.c_int
        L.or_mc_edge_table.argtypes = [C.c_int]
        L.or_mc_tri_table.restype = C.c_int
        L.or_mc_tri_table.argtypes = [C.c_int, C.c_int]
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(_P)


def f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def set_threads(n: int):
    lib().or_set_threads(int(n))


class OracleMap:
    """Serial restatement of MapStruct<TVoxel> + the fusion stages, in host memory."""

    def __init__(self, n_entry, n_bucket, n_block, voxel_size, trunc, voxel_format=0, depth_min=0.1,
                 depth_max=5.0, max_weight=255.0):
        self.h = lib().or_map_create(n_entry, n_bucket, n_block, voxel_size, trunc, depth_min, depth_max,
                                     voxel_format, max_weight)
        if not self.h:
            raise ValueError("invalid oracle map geometry")
        self.n_entry, self.n_bucket, self.n_block = n_entry, n_bucket, n_block
        self.voxel_size, self.trunc, self.fmt = voxel_size, trunc, voxel_format
        self.vbytes = lib().or_voxel_bytes(voxel_format)

    def __del__(self, _destroy=None):
        if getattr(self, "h", None) and _lib is not None:
            _lib.or_map_destroy(self.h)
            self.h = None

    def insert_block(self, x, y, z) -> int:
        return lib().or_insert_block(self.h, int(x), int(y), int(z))

    def voxels_view(self) -> np.ndarray:
        """Writable view of the oracle's voxel bytes."""
        p = lib().or_map_voxels(self.h)
        return np.ctypeslib.as_array(p, shape=(self.n_block * 512 * self.vbytes,))

    def reset(self):
        lib().or_map_reset(self.h)

    def fuse(self, depth_m, K, T, rgb=None) -> int:
        d = f32(depth_m)
        H, W = d.shape
        c = None if rgb is None else np.ascontiguousarray(rgb, np.uint8)
        return lib().or_fuse(self.h, _p(d), _p(c), W, H, _p(f32(K)), _p(f32(T)))

    def allocate(self, depth_m, K, T):
        d = f32(depth_m)
        H, W = d.shape
        lib().or_allocate(self.h, _p(d), W, H, _p(f32(K)), _p(f32(T)))

    def check_visible(self, W, H, K, T) -> int:
        return lib().or_check_visible(self.h, W, H, _p(f32(K)), _p(f32(T)))

    def raycast(self, W, H, K, T, init=0.0) -> np.ndarray:
        v = np.full((H, W, 4), init, np.float32)
        lib().or_raycast(self.h, _p(v), W, H, _p(f32(K)), _p(f32(T)))
        return v

    def render_scene(self, vmap, K, T) -> int:
        H, W = vmap.shape[:2]
        return lib().or_render_scene(self.h, _p(vmap), W, H, _p(f32(K)), _p(f32(T)))

    def polygonize(self, cap):
        v = np.zeros((cap, 9), np.float32)
        n = np.zeros((cap, 9), np.float32)
        bc = np.zeros(1, np.int32)
        t = lib().or_polygonize(self.h, _p(v), _p(n), cap, _p(bc))
        return int(bc[0]), int(t), v[:t], n[:t]

    def raycast_pixels_f32(self, px, py, K, T):
        """(vertices n x 4, decision signatures, hit flags) of the float march."""
        px = np.ascontiguousarray(px, np.int32)
        py = np.ascontiguousarray(py, np.int32)
        n = len(px)
        v = np.zeros((n, 4), np.float32)
        sig = np.zeros(n, np.uint64)
        hit = np.zeros(n, np.int32)
        lib().or_raycast_pixels_f32(self.h, n, _p(px), _p(py), _p(f32(K)), _p(f32(T)), _p(v), _p(sig), _p(hit))
        return v, sig, hit.astype(bool)

    def raycast_pixels_f64(self, px, py, K, T, pert_idx=None, pert_val=None):
        """FP64 march with TSDF slot pert_idx[k] += pert_val[k]; T: 16 column-major doubles."""
        px = np.ascontiguousarray(px, np.int32)
        py = np.ascontiguousarray(py, np.int32)
        n = len(px)
        v = np.zeros((n, 4), np.float64)
        sig = np.zeros(n, np.uint64)
        hit = np.zeros(n, np.int32)
        pi = np.ascontiguousarray(np.zeros(0) if pert_idx is None else pert_idx, np.int64)
        pv = np.ascontiguousarray(np.zeros(0) if pert_val is None else pert_val, np.float64)
        lib().or_raycast_pixels_f64(self.h, n, _p(px), _p(py), _p(np.ascontiguousarray(K, np.float64)),
                                    _p(np.ascontiguousarray(T, np.float64).reshape(-1)), _p(pi), _p(pv), len(pi),
                                    _p(v), _p(sig), _p(hit))
        return v, sig, hit.astype(bool)

    def surface_depth_f64(self, px, py, K, T):
        """(z, ok) per pixel: the pose refinement's depth in FP64 (march,
        corrections, three Newton steps onto the trilinear zero level);
        T: 4x4 camera-to-world (row-major) or 16 column-major doubles."""
        px = np.ascontiguousarray(px, np.int32)
        py = np.ascontiguousarray(py, np.int32)
        T = np.asarray(T, np.float64)
        Tc = np.ascontiguousarray(T.T.reshape(-1) if T.shape == (4, 4) else T.reshape(-1), np.float64)
        z = np.zeros(len(px), np.float64)
        ok = np.zeros(len(px), np.int32)
        lib().or_surface_depth_f64(self.h, len(px), _p(px), _p(py), _p(np.ascontiguousarray(K, np.float64)), _p(Tc),
                                   _p(z), _p(ok))
        return z, ok.astype(bool)

    def project_rendering_blocks(self, cols8, rows8, K, Twc):
        zx = np.zeros((rows8, cols8), np.float32)
        zy = np.zeros((rows8, cols8), np.float32)
        n = lib().or_project_rendering_blocks(self.h, _p(zx), _p(zy), cols8, rows8, _p(f32(K)), _p(f32(Twc)))
        return n, zx, zy

    def counters(self) -> dict:
        out = np.zeros(8, np.int32)
        lib().or_map_counters(self.h, _p(out))
        keys = ["heap_ptr", "excess_ptr", "visible", "alias_events", "alloc_failures", "chain_bound_hits"]
        return dict(zip(keys, out[:6].tolist()))

    def hash_table(self) -> np.ndarray:
        p = lib().or_map_hash_table(self.h)
        return np.ctypeslib.as_array(p, shape=(self.n_entry * 5,)).reshape(-1, 5).copy()

    def heap(self) -> np.ndarray:
        return np.ctypeslib.as_array(lib().or_map_heap(self.h), shape=(self.n_block,)).copy()

    def voxels(self) -> np.ndarray:
        p = lib().or_map_voxels(self.h)
        return np.ctypeslib.as_array(p, shape=(self.n_block * 512 * self.vbytes,)).copy()

    def visible_table(self) -> np.ndarray:
        p = lib().or_map_visible_table(self.h)
        n = self.counters()["visible"]
        return np.ctypeslib.as_array(p, shape=(self.n_entry * 5,)).reshape(-1, 5)[:n].copy()


def compute_normals(vmap):
    H, W = vmap.shape[:2]
    n = np.zeros_like(vmap)
    lib().or_compute_normals(_p(np.ascontiguousarray(vmap, np.float32)), _p(n), W, H)
    return n


def render_shading(vmap, nmap):
    H, W = vmap.shape[:2]
    out = np.zeros((H, W, 4), np.uint8)
    lib().or_render_shading(_p(f32(vmap)), _p(f32(nmap)), _p(out), W, H)
    return out


def surface_points(verts):
    v = f32(verts).reshape(-1, 9)
    p = np.zeros((len(v), 3), np.float32)
    lib().or_surface_points(_p(v), _p(p), len(v))
    return p


def hash_block(x, y, z, n):
    return lib().or_hash(int(x), int(y), int(z), int(n))


def voxel_to_block(v):
    o = np.zeros(3, np.int32)
    lib().or_voxel_to_block(int(v[0]), int(v[1]), int(v[2]), _p(o))
    return o


def voxel_to_local(v):
    o = np.zeros(3, np.int32)
    lib().or_voxel_to_local(int(v[0]), int(v[1]), int(v[2]), _p(o))
    return o


def pose_inverse(T):
    o = np.zeros(16, np.float32)
    lib().or_pose_inverse(_p(f32(T).reshape(-1)), _p(o))
    return o


def mc_tables():
    L = lib()
    edge = np.array([L.or_mc_edge_table(c) for c in range(256)], np.int32)
    tri = np.array([[L.or_mc_tri_table(c, i) for i in range(16)] for c in range(256)], np.int32)
    return edge, tri


# --- decoding helpers shared by tests -------------------------------------------

def decode_voxels(raw: np.ndarray, fmt: int):
    """raw bytes -> (sdf float32, weight float32, rgb uint8 Nx3 or None)."""
    if fmt == 0:
        r = raw.reshape(-1, 4)
        sdf = r[:, 0:2].copy().view(np.int16).reshape(-1).astype(np.float32) / np.float32(32767.0)
        return sdf, r[:, 2].astype(np.float32), None
    if fmt == 1:
        r = raw.reshape(-1, 6)
        sdf = r[:, 0:2].copy().view(np.int16).reshape(-1).astype(np.float32) / np.float32(32767.0)
        return sdf, r[:, 2].astype(np.float32), r[:, 3:6].copy()
    r = raw.reshape(-1, 12)
    sdf = r[:, 0:4].copy().view(np.float32).reshape(-1)
    w = r[:, 4:8].copy().view(np.float32).reshape(-1)
    return sdf, w, r[:, 8:11].copy()


def allocated_blocks(table: np.ndarray):
    """{(x, y, z): ptr} of all entries with ptr >= 0 (table: N x 5 ints)."""
    occ = table[table[:, 0] >= 0]
    return {(int(e[2]), int(e[3]), int(e[4])): int(e[0]) for e in occ}
