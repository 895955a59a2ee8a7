"""Writes the golden fixtures of tests/golden from the CPU oracle.

The reference holds no test vectors or fixtures for its fusion pipeline, so
these are the oracle's outputs on two tiny seeded scenes (80x60 depth, six
frames of the 4x4x3 m room, 1 cm voxels), one per voxel format family:

* ``tiny_ref.npz``    -- reference ``Voxel`` (int16 TSDF, uint8 weight);
* ``tiny_f32rgb.npz`` -- float TSDF, float weight, RGB8, colour fused.

Each holds its inputs (the six depth frames in metres, colour for the colour
format, camera-to-world
poses, intrinsics and map geometry) and the oracle's outputs on them: the
sorted allocated block keys, the per-frame visible counts, the
full voxel state of eight sampled blocks, a SHA-256 of the whole voxel state
in key order, the raycast vertex map and normal map at the last pose, and the
marching-cubes triangle count with a SHA-256 of the canonically sorted soup
and its first 32 triangles.

Run from the repository root after the build:

    python tests/golden/make_golden.py

The oracle is deterministic (serial allocation; per-voxel and per-pixel
arithmetic independent of the thread count; the mesh is compared sorted), so
rerunning reproduces the files byte for byte.
"""
from __future__ import annotations

import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
for p in (ROOT, os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)

CASES = {"tiny_ref": 0, "tiny_f32rgb": 2}
N_FRAMES = 6
MESH_CAP = 400_000


def workload(fmt: int):
    from voxelmapping_amd import synth
    return synth.room_small(seed=7, width=80, height=60, K=(65.625, 65.625, 39.5, 29.5), n_frames=N_FRAMES,
                            voxel_format=fmt, n_bucket=1 << 14, n_excess=1 << 11, n_block=1 << 13)


def sorted_keys(table: np.ndarray) -> np.ndarray:
    occ = table[table[:, 0] >= 0]
    keys = occ[:, 2:5].astype(np.int32)
    order = np.lexsort(keys.T[::-1])
    return keys[order], occ[order, 0].astype(np.int64)


def voxel_state(raw: np.ndarray, fmt: int, ptrs: np.ndarray):
    import oracle as O
    sdf, w, rgb = O.decode_voxels(raw, fmt)
    idx = (ptrs[:, None] + np.arange(512)[None, :]).reshape(-1)
    return sdf[idx].reshape(-1, 512), w[idx].reshape(-1, 512), None if rgb is None else rgb[idx].reshape(-1, 512, 3)


def state_sha(sdf, w, rgb) -> str:
    h = hashlib.sha256()
    h.update(np.ascontiguousarray(sdf, np.float32).tobytes())
    h.update(np.ascontiguousarray(w, np.float32).tobytes())
    if rgb is not None:
        h.update(np.ascontiguousarray(rgb, np.uint8).tobytes())
    return h.hexdigest()


def canonical_mesh(verts, norms):
    tri = np.concatenate([np.asarray(verts, np.float32).reshape(-1, 9), np.asarray(norms, np.float32).reshape(-1, 9)],
                         axis=1)
    order = np.lexsort(tri.T[::-1])
    return tri[order]


def sample_rows(n: int) -> np.ndarray:
    return np.sort(np.random.default_rng(12345).choice(n, size=min(8, n), replace=False))


def case_inputs(fmt: int) -> dict:
    """The scene's frames, poses, intrinsics and map geometry, from the generator."""
    w = workload(fmt)
    depth = np.zeros((N_FRAMES, w.height, w.width), np.float32)
    rgb = np.zeros((N_FRAMES, w.height, w.width, 3), np.uint8)
    for i in range(N_FRAMES):
        d, c = w.frame(i, want_rgb=True)
        depth[i] = w.depth_metres(d)
        rgb[i] = c
    geom = np.array([w.n_bucket + w.n_excess, w.n_bucket, w.n_block, fmt], np.int64)
    out = dict(depth=depth, poses=np.asarray(w.poses(N_FRAMES), np.float32), K=np.asarray(w.K, np.float32),
               geometry=geom, voxel=np.array([w.voxel_size, w.trunc], np.float32))
    if fmt != 0:
        out["rgb"] = rgb
    return out


def oracle_outputs(inp: dict) -> dict:
    """The oracle's map state, raycast, normals and mesh for a case's inputs."""
    import oracle as O
    n_entry, n_bucket, n_block, fmt = (int(v) for v in inp["geometry"])
    vs, trunc = (float(v) for v in inp["voxel"])
    depth, rgb, poses, K = inp["depth"], inp.get("rgb"), inp["poses"], inp["K"]
    H, W = depth.shape[1:]
    o = O.OracleMap(n_entry, n_bucket, n_block, vs, trunc, fmt, 0.1, 5.0, 255.0)
    counts = [o.fuse(depth[i], K, poses[i], rgb=None if rgb is None else rgb[i]) for i in range(len(depth))]
    keys, ptrs = sorted_keys(o.hash_table())
    sdf, wt, col = voxel_state(o.voxels(), fmt, ptrs)
    rows = sample_rows(len(keys))
    vmap = o.raycast(W, H, K, poses[-1], init=0.0)
    nmap = O.compute_normals(vmap)
    _, t, v, n = o.polygonize(MESH_CAP)
    mesh = canonical_mesh(v, n)
    out = dict(keys=keys, counts=np.asarray(counts, np.int32), sample_rows=rows.astype(np.int32),
               sample_sdf=sdf[rows], sample_weight=wt[rows], state_sha=np.array(state_sha(sdf, wt, col)),
               vmap=vmap[..., :3].copy(), nmap=nmap[..., :3].copy(), n_triangles=np.int64(t),
               mesh_sha=np.array(hashlib.sha256(mesh.tobytes()).hexdigest()), mesh_head=mesh[:32].copy())
    if col is not None:
        out["sample_rgb"] = col[rows]
    return out


def main() -> None:
    import oracle as O
    O.set_threads(min(8, os.cpu_count() or 1))
    for name, fmt in CASES.items():
        data = case_inputs(fmt)
        data.update(oracle_outputs(data))
        path = os.path.join(HERE, name + ".npz")
        np.savez_compressed(path, **data)
        print(f"wrote {path}: {len(data['keys'])} blocks, {int(data['n_triangles'])} triangles, "
              f"{os.path.getsize(path)} bytes")


if __name__ == "__main__":
    main()
