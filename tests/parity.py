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


def compare_map_blocks(g, o, fmt, sdf_tol=1e-5, check_chains=20000):
    """compare_maps reading only what is allocated: the GPU's blocks [0, A)
    (vm_map_download_blocks) and the oracle's pool in place, so a 2^20-block
    pool is not copied whole. Same bar: key sets bit-exact, TSDF within
    sdf_tol, weights and colours exact. Returns the number of blocks."""
    gt, ot = g.hash_table(), o.hash_table()
    gb, ob = block_map(gt), block_map(ot)
    assert set(gb) == set(ob), (f"allocated sets differ: gpu-only {len(set(gb) - set(ob))}, "
                                f"oracle-only {len(set(ob) - set(gb))}")
    if check_chains:
        chain_lookup_all(gt, g.desc.n_bucket, list(gb)[:check_chains])
    if not gb:
        return 0
    keys = sorted(gb)
    A = g.stats()["allocated_blocks"]
    vb = {0: 4, 1: 6, 2: 12}[fmt]
    graw = g.download_blocks(0, A).reshape(A, 512 * vb)
    oraw = o.voxels_view().reshape(-1, 512 * vb)
    gsel = graw[np.array([gb[k] // 512 for k in keys])].reshape(-1)
    osel = oraw[np.array([ob[k] // 512 for k in keys])].reshape(-1)
    gs, gw, gc = O.decode_voxels(gsel, fmt)
    os_, ow, oc = O.decode_voxels(osel, fmt)
    assert np.array_equal(gw, ow), f"weights differ at {np.sum(gw != ow)} voxels"
    diff = np.abs(gs - os_)
    assert diff.max() <= sdf_tol, f"TSDF differs by {diff.max()}"
    if gc is not None:
        assert np.array_equal(gc, oc), "colours differ"
    return len(keys)


def _soup_hash(v, n):
    """64-bit hash of each triangle's 18 floats (bit patterns), wrapping."""
    rows = np.concatenate([np.asarray(v, np.float32).reshape(-1, 9), np.asarray(n, np.float32).reshape(-1, 9)],
                          axis=1).view(np.uint32).astype(np.uint64)
    mult = (np.arange(1, 19, dtype=np.uint64) * np.uint64(0x9E3779B97F4A7C15)) | np.uint64(1)
    h = np.zeros(len(rows), np.uint64)
    for j in range(18):
        h = (h ^ (rows[:, j] * mult[j])) * np.uint64(0xBF58476D1CE4E5B9)
    return np.sort(h)


def compare_soups(gv, gn, ov, on, tol=1e-5):
    """Two triangle soups (vertices and face normals, 9 floats per triangle)
    hold the same triangles in any order: first as multisets of exact bit
    patterns (sorted row hashes); where that fails, after canonical sorting,
    vertices and normals within tol."""
    assert len(gv) == len(ov), f"triangle counts {len(gv)} vs {len(ov)}"
    if len(gv) == 0 or np.array_equal(_soup_hash(gv, gn), _soup_hash(ov, on)):
        return "bit-exact"
    gs, gi = canonical_triangles(gv)
    os_, oi = canonical_triangles(ov)
    assert np.max(np.abs(gs - os_)) <= tol, f"vertices differ by {np.max(np.abs(gs - os_))}"
    gnn = np.asarray(gn, np.float32).reshape(-1, 9)[gi]
    onn = np.asarray(on, np.float32).reshape(-1, 9)[oi]
    assert np.max(np.abs(gnn - onn)) <= 1e-4, f"normals differ by {np.max(np.abs(gnn - onn))}"
    return "within tolerance"


def compare_vmaps(gv, ov, tol=1e-4, min_hits=0.0):
    """Hit masks equal, hit vertices within tol metres; returns the hit rate."""
    gh, oh = ~np.isnan(gv[..., 0]), ~np.isnan(ov[..., 0])
    assert np.array_equal(gh, oh), f"hit masks differ at {np.sum(gh != oh)} pixels"
    d = np.abs(gv[gh][:, :3] - ov[oh][:, :3])
    assert d.size == 0 or d.max() <= tol, f"vertex maps differ by {d.max()}"
    assert gh.mean() >= min_hits, f"hit rate {gh.mean()}"
    return gh.mean()


def compare_normals(gn, on, tol=1e-4):
    """Normal maps: NaN in the same places, finite values within tol (the
    last row and column, never written, are left out)."""
    gn, on = gn[:-1, :-1], on[:-1, :-1]
    assert np.array_equal(np.isnan(gn), np.isnan(on))
    fin = ~np.isnan(on)
    assert not fin.any() or np.max(np.abs(gn[fin] - on[fin])) <= tol
