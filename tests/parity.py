"""Shared helpers of the GPU parity tests: drive the HIP library and the CPU
oracle with identical inputs and compare their states."""
from __future__ import annotations

import numpy as np

import oracle as O


def colmajor(T):
    return np.asarray(T, np.float32).T.reshape(-1)


def make_pair(w, fmt=None, n_block=None, n_bucket=None, n_excess=None, max_weight=255.0):
    """GPU DeviceMap and OracleMap with the workload's geometry."""
    from voxelmapping_amd import DeviceMap
    fmt = w.voxel_format if fmt is None else fmt
    nb = w.n_block if n_block is None else n_block
    nbk = w.n_bucket if n_bucket is None else n_bucket
    nex = w.n_excess if n_excess is None else n_excess
    g = DeviceMap(nbk + nex, nbk, nb, w.voxel_size, w.trunc, fmt, 0.1, 5.0, max_weight)
    o = O.OracleMap(nbk + nex, nbk, nb, w.voxel_size, w.trunc, fmt, 0.1, 5.0, max_weight)
    return g, o


def gpu_frame(w, i, device="cuda", rgb=True):
    import torch
    d, c = w.frame(i, want_rgb=rgb)
    dt = torch.from_numpy(d.view(np.int16) if d.dtype == np.uint16 else d).to(device)
    if d.dtype == np.uint16:
        dt = dt.view(torch.uint16)
    ct = torch.from_numpy(c).to(device) if c is not None else None
    return d, c, dt, ct


def fuse_both(w, g, o, frames, rgb=True, check_counts=True):
    """Fuse frames into both maps; returns per-frame visible counts (gpu, oracle)."""
    counts = []
    for i in frames:
        d, c, dt, ct = gpu_frame(w, i, rgb=rgb)
        T = w.poses(i + 1)[i]
        ng = g.fuse(dt, w.K, T, rgb=ct if rgb else None, depth_scale=w.depth_scale, return_count=True)
        no = o.fuse(w.depth_metres(d), w.K, T, rgb=c if rgb else None)
        counts.append((ng, no))
        if check_counts:
            assert ng == no, f"frame {i}: visible blocks gpu {ng} vs oracle {no}"
    return counts


def block_map(table):
    """{key: ptr} of allocated entries; asserts there are no duplicate keys."""
    occ = table[table[:, 0] >= 0]
    keys = [tuple(int(v) for v in e[2:]) for e in occ]
    assert len(set(keys)) == len(keys), "duplicate keys in the hash table"
    return dict(zip(keys, (int(p) for p in occ[:, 0])))


def chain_lookup_all(table, n_bucket, keys):
    """Every key must be reachable from its bucket by following offsets."""
    index = {}
    for k in keys:
        h = O.hash_block(k[0], k[1], k[2], n_bucket)
        slot, found, guard = h, False, len(table)
        while guard > 0:
            e = table[slot]
            if e[0] >= 0 and tuple(int(v) for v in e[2:]) == k:
                found = True
                break
            if e[1] < 0:
                break
            slot = n_bucket + int(e[1]) - 1
            guard -= 1
        assert found, f"key {k} not reachable from bucket {h}"
        index[k] = slot
    return index


def compare_maps(g, o, fmt, sdf_tol=1e-5):
    """Allocated key sets bit-exact; per-voxel TSDF within sdf_tol, weight and colour exact."""
    gt = g.hash_table()
    ot = o.hash_table()
    gb, ob = block_map(gt), block_map(ot)
    assert set(gb) == set(ob), (f"allocated sets differ: gpu-only {len(set(gb) - set(ob))}, "
                                f"oracle-only {len(set(ob) - set(gb))}")
    chain_lookup_all(gt, g.desc.n_bucket, list(gb)[:20000])
    if not gb:
        return gb, ob
    keys = sorted(gb)
    gi = np.array([gb[k] for k in keys], np.int64)[:, None] + np.arange(512)[None, :]
    oi = np.array([ob[k] for k in keys], np.int64)[:, None] + np.arange(512)[None, :]
    gs, gw, gc = O.decode_voxels(g.download(2), fmt)
    os_, ow, oc = O.decode_voxels(o.voxels(), fmt)
    assert np.array_equal(gw[gi], ow[oi]), "weights differ"
    diff = np.abs(gs[gi] - os_[oi])
    assert diff.max() <= sdf_tol, f"TSDF differs by {diff.max()}"
    if gc is not None:
        assert np.array_equal(gc[gi.reshape(-1)], oc[oi.reshape(-1)]), "colours differ"
    return gb, ob


def canonical_triangles(v):
    """Sort triangles (rows of 9 floats) lexicographically."""
    v = np.asarray(v, np.float32).reshape(-1, 9)
    order = np.lexsort(v.T[::-1])
    return v[order], order
