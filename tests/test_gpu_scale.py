"""The large-room workload at full size: 1280x720 uint16 depth with colour,
2 mm voxels, 1 cm truncation, 2^23 buckets + 2^20 excess slots, 2^22
float-TSDF blocks (24 GiB of voxels). The first two frames are checked against
the oracle (visible counts, allocated key set, TSDF/weight/colour of sampled
blocks, raycast vertices of sampled pixels, the whole mesh); forty frames are
then checked through size-independent properties (no capacity status, every
sampled key reachable from its bucket, weights and TSDF in range,
deterministic rerun of the last frame's raycast; the final mesh: the same
buffers from two calls and a closed surface inside sampled boxes). Pose
refinement at full size on the same geometry without depth noise."""
import numpy as np
import pytest

import oracle as O
from parity import chain_lookup_all, colmajor, compare_soups, gpu_frame

pytestmark = pytest.mark.gpu

ORACLE_BLOCKS = 1 << 19


@pytest.fixture(scope="module")
def large(oracle_threads):
    from voxelmapping_amd import DeviceMap, synth
    w = synth.room_large()
    g = DeviceMap(w.n_entry, w.n_bucket, w.n_block, w.voxel_size, w.trunc, w.voxel_format)
    o = O.OracleMap(w.n_entry, w.n_bucket, ORACLE_BLOCKS, w.voxel_size, w.trunc, w.voxel_format)
    counts = []
    for i in range(2):
        d, c, dt, ct = gpu_frame(w, i)
        T = w.poses(i + 1)[i]
        ng = g.fuse(dt, w.K, T, rgb=ct, depth_scale=w.depth_scale, return_count=True)
        no = o.fuse(w.depth_metres(d), w.K, T, rgb=c)
        counts.append((ng, no))
    yield w, g, o, counts
    g.release()


def test_first_frames_match_oracle(large):
    w, g, o, counts = large
    for ng, no in counts:
        assert ng == no > 100000
    gt, ot = g.hash_table(), o.hash_table()
    gk = gt[gt[:, 0] >= 0]
    ok = ot[ot[:, 0] >= 0]
    gkeys = {tuple(map(int, e[2:])): int(e[0]) for e in gk}
    okeys = {tuple(map(int, e[2:])): int(e[0]) for e in ok}
    assert len(gkeys) == len(gk) and set(gkeys) == set(okeys)
    assert o.counters()["alias_events"] == 0
    # voxels of 20000 sampled blocks
    rng = np.random.default_rng(1)
    keys = list(gkeys)
    pick = [keys[i] for i in rng.choice(len(keys), 20000, replace=False)]
    A = g.stats()["allocated_blocks"]
    gv = g.download_blocks(0, A)
    gs, gw, gc = O.decode_voxels(gv, 2)
    os_, ow, oc = O.decode_voxels(o.voxels(), 2)
    gi = np.array([gkeys[k] for k in pick], np.int64)[:, None] + np.arange(512)[None, :]
    oi = np.array([okeys[k] for k in pick], np.int64)[:, None] + np.arange(512)[None, :]
    assert np.array_equal(gw[gi], ow[oi])
    assert np.abs(gs[gi] - os_[oi]).max() <= 1e-5
    assert np.array_equal(gc[gi.reshape(-1)], oc[oi.reshape(-1)])


def test_raycast_pixels_match_oracle(large):
    import torch
    w, g, o, _ = large
    T = w.poses(2)[1]
    vm = torch.zeros((w.height, w.width, 4), dtype=torch.float32, device="cuda")
    g.render_scene(vm, w.K, T)
    v = vm.cpu().numpy()
    rng = np.random.default_rng(2)
    px = rng.integers(0, w.width, 3000)
    py = rng.integers(0, w.height, 3000)
    ov, _, hit = o.raycast_pixels_f32(px, py, np.asarray(w.K, np.float32), colmajor(T.reshape(4, 4).T))
    gvp = v[py, px]
    assert hit.mean() > 0.5
    assert np.array_equal(np.isnan(gvp[:, 0]), ~hit)
    assert np.array_equal(gvp[hit], ov[hit])


def test_two_frame_mesh_matches_oracle(large):
    """Marching cubes over the 2-frame map (every allocated block): the same
    block and triangle counts as the oracle and the same triangles."""
    import torch
    w, g, o, _ = large
    cap = 40_000_000
    verts = torch.empty((cap, 9), dtype=torch.float32, device="cuda")
    norms = torch.empty_like(verts)
    gb, gt = g.polygonize(verts, norms)
    ob, ot, ov, on = o.polygonize(cap)
    print(f"2-frame mesh: {gb} blocks, {gt} triangles", flush=True)
    assert (gb, gt) == (ob, ot) and gt > 1_000_000
    how = compare_soups(verts[:gt].cpu().numpy(), norms[:gt].cpu().numpy(), ov, on)
    print("triangles", how)
    del verts, norms
    torch.cuda.empty_cache()


def _vertex_grid(t, vs):
    """The grid edge each marching-cubes vertex lies on: its lower voxel
    corner and axis (the coordinate off the integer grid; 3 for a vertex at
    a corner). Neighbouring cubes compute a shared edge's vertex from
    opposite ends, so its floats may differ in the last bit, never its edge."""
    u = t.astype(np.float64) / vs
    r = np.round(u)
    fr = np.abs(u - r)
    axis = np.argmax(fr, axis=-1)
    corner = fr.max(axis=-1) < 1e-3
    lo = r.copy()
    ax_u = np.take_along_axis(u, axis[..., None], -1)[..., 0]
    np.put_along_axis(lo, axis[..., None], np.floor(ax_u)[..., None], -1)
    return lo.astype(np.int64), np.where(corner, 3, axis)


def _open_edges_by_plane(v, vs, lo_b, hi_b):
    """Triangles with all vertices in the box [lo_b, hi_b) (metres): of their
    edges lying in a grid plane at a block boundary (voxel coordinate 0 or 7
    mod 8) and of those in a plane inside a block (3 or 4 mod 8), the
    fractions not shared by exactly two triangles. A block whose halo or
    neighbour layer is missing opens the surface along block boundaries
    only; noise and unobserved voxels open it anywhere."""
    t = v.reshape(-1, 3, 3)
    near = np.all((t >= lo_b - 2 * vs) & (t < hi_b + 2 * vs), axis=(1, 2))
    t = t[near]
    inside = np.all((t >= lo_b) & (t < hi_b), axis=(1, 2))
    lo, ax = _vertex_grid(t, vs)
    key = ((lo[..., 0] * 73856093) ^ (lo[..., 1] * 19349669) ^ (lo[..., 2] * 83492791)) * 4 + ax
    e, plane = [], []
    for a, b in ((0, 1), (1, 2), (2, 0)):
        ka, kb = key[:, a], key[:, b]
        e.append(np.where(ka < kb, ka * 1000003 + kb, kb * 1000003 + ka))
        f = np.zeros(len(ka), np.int8)  # 1: block-boundary plane, 2: inner plane
        for k in range(3):
            on = (ax[:, a] != k) & (ax[:, b] != k) & (lo[:, a, k] == lo[:, b, k])
            md = lo[:, a, k] % 8
            f = np.where(on & ((md == 0) | (md == 7)), 1, np.where(on & ((md == 3) | (md == 4)) & (f == 0), 2, f))
        plane.append(f)
    e, plane = np.stack(e, axis=1), np.stack(plane, axis=1)
    uniq, cnt = np.unique(e.reshape(-1), return_counts=True)
    mine, pm = e[inside].reshape(-1), plane[inside].reshape(-1)
    opened = cnt[np.searchsorted(uniq, mine)] != 2
    return float(opened[pm == 1].mean()), float(opened[pm == 2].mean()), int(inside.sum())


def test_forty_frames_invariants(large):
    import torch
    w, g, _, _ = large
    for i in range(2, 40):
        _, _, dt, ct = gpu_frame(w, i)
        g.fuse(dt, w.K, w.poses(i + 1)[i], rgb=ct, depth_scale=w.depth_scale)
    st = g.stats()
    assert st["status"] == 0 and st["frames_fused"] == 40
    assert 0 < st["allocated_blocks"] < w.n_block
    gt = g.hash_table()
    occ = gt[gt[:, 0] >= 0]
    assert len(occ) == st["allocated_blocks"]
    keys = [tuple(map(int, e[2:])) for e in occ]
    assert len(set(keys)) == len(keys)
    rng = np.random.default_rng(3)
    chain_lookup_all(gt, w.n_bucket, [keys[i] for i in rng.choice(len(keys), 20000, replace=False)])
    s, wt, _ = O.decode_voxels(g.download_blocks(0, st["allocated_blocks"]), 2)
    assert wt.max() <= 255.0 and wt.min() >= 0.0
    seen = wt > 0
    assert np.all(np.abs(s[seen]) <= 1.0)
    T = w.poses(40)[39]
    a = torch.zeros((w.height, w.width, 4), dtype=torch.float32, device="cuda")
    b = torch.zeros_like(a)
    g.render_scene(a, w.K, T)
    g.render_scene(b, w.K, T)
    torch.cuda.synchronize()
    assert torch.equal(torch.nan_to_num(a, 3.0), torch.nan_to_num(b, 3.0))
    assert (~torch.isnan(a[..., 0])).float().mean().item() > 0.5


def test_forty_frame_mesh(large):
    """The full mesh of the 40-frame map (runs after the forty frames are
    fused): two calls give identical buffers and counts, and inside sampled
    boxes on the walls the edges left open (shared by other than two
    triangles) are no more frequent on block-boundary planes than inside
    blocks. This workload's depth noise at 5 m (3 cm) is three times its
    truncation, so its far walls have holes everywhere; a missing halo layer
    or neighbour block would open them along block faces."""
    import torch
    w, g, _, _ = large
    st = g.stats()
    assert st["frames_fused"] == 40
    cap = 120_000_000
    verts = torch.empty((cap, 9), dtype=torch.float32, device="cuda")
    norms = torch.empty_like(verts)
    gb, gt = g.polygonize(verts, norms)
    print(f"40-frame mesh: {gb} blocks, {gt} triangles", flush=True)
    assert gb == st["allocated_blocks"] and 1_000_000 < gt < cap
    assert g.stats()["status"] == 0
    first_v, first_n = verts[:gt].clone(), norms[:gt].clone()
    gb2, gt2 = g.polygonize(verts, norms)
    assert (gb2, gt2) == (gb, gt)
    assert torch.equal(first_v, verts[:gt]) and torch.equal(first_n, norms[:gt])
    del first_v, first_n
    v = verts[:gt].cpu().numpy()
    del verts, norms
    torch.cuda.empty_cache()
    assert np.all(np.isfinite(v))
    # box centres independent of the triangle order (which follows the
    # blocks' allocation order): the first vertices of the wall triangles
    # (normal within ~45 degrees of horizontal) with the smallest hashes of
    # their voxel cells. Floors seen at grazing angles get sparse blocks
    # along the truncation band, so their surfaces open at block faces in
    # the reference's scheme too; walls do not, unless a halo layer is lost.
    t = v.reshape(-1, 9)
    q = np.floor(t[:, 0:3] / w.voxel_size).astype(np.int64)
    h = ((q[:, 0] * 73856093) ^ (q[:, 1] * 19349669) ^ (q[:, 2] * 83492791)) & 0x7FFFFFFF
    cand = np.argpartition(h, 4000)[:4000]
    tc = t[cand].astype(np.float64)
    tc = tc[np.lexsort((*tc.T[::-1], h[cand]))]  # ties broken by the vertices
    n = np.cross(tc[:, 3:6] - tc[:, 0:3], tc[:, 6:9] - tc[:, 0:3])
    # wall triangles whose box stays clear of the floor and the ceiling
    wall = (np.abs(n[:, 2]) < 0.7 * np.linalg.norm(n, axis=1)) & (tc[:, 2] > 0.3) & (tc[:, 2] < 2.5)
    centres = []
    for c in tc[wall][:, 0:3]:
        if all(np.abs(c - d).max() > 0.35 for d in centres):  # boxes that do not overlap
            centres.append(c)
    centres = np.asarray(centres[:6], np.float32)
    assert len(centres) == 6
    del t, q, h
    excess = []
    for c in centres:
        fb, fi, n = _open_edges_by_plane(v, w.voxel_size, c - 0.15, c + 0.15)
        excess.append(fb - 2 * fi)
        print(f"box at {c}: {n} triangles, open edges {fb:.4f} on block-boundary planes, {fi:.4f} inside blocks")
    # the oracle's meshes of the smaller rooms give -0.01..0.002 here; a
    # missing layer along block boundaries gives 0.05..0.55
    assert np.median(excess) <= 0.01, excess


def test_noisy_tracking_at_scale(large):
    """Tracking on the workload itself -- its depth noise (0.0012 z^2 m, 3 cm
    at 5 m, against a 1 cm truncation) and dropout -- into the 40-frame map
    of the tests above: 16 hypotheses up to 2 cm / 1.5 degrees around the
    pose of frame 40, 8 Gauss-Newton iterations over every 4th pixel of its
    1280x720 depth, with the residuals limited to depth below 2 m (where the
    noise stays below half the truncation; beyond, one-sided averaging
    biases the fused surface by a range-dependent amount). The hypothesis
    of lowest cost ends closer to the true pose than it started, within
    1 cm, and the median hypothesis improves."""
    from voxelmapping_amd.autograd import se3_exp
    w, g, _, _ = large
    assert g.stats()["frames_fused"] == 40
    _, _, dt, _ = gpu_frame(w, 40, rgb=False)
    T0 = w.poses(41)[40].reshape(4, 4).T.astype(np.float64)
    rng = np.random.default_rng(12)
    hyps = np.stack([T0 @ se3_exp(np.r_[rng.uniform(-1, 1, 3) * 0.02, np.radians(rng.uniform(-1, 1, 3) * 1.5)])
                     for _ in range(16)]).astype(np.float32)
    poses, cost = g.refine_poses(dt, w.K, hyps, iters=8, stride=4, depth_scale=w.depth_scale, max_depth=2.0)
    err0 = np.array([np.linalg.norm((np.linalg.inv(T0) @ h.astype(np.float64))[:3, 3]) for h in hyps])
    err = np.array([np.linalg.norm((np.linalg.inv(T0) @ p.astype(np.float64))[:3, 3]) for p in poses])
    best = int(np.argmin(cost[:, -1, 0] + (cost[:, -1, 0] < 0) * 1e9))
    print(f"start {np.round(err0 * 1000, 1)} mm\nend   {np.round(err * 1000, 1)} mm; best hypothesis {best}")
    assert err[best] < err0[best] and err[best] < 0.01
    assert np.median(err) < np.median(err0)


def test_pose_refinement_at_scale():
    """Tracking at full size on the large room's geometry seen without depth
    noise (frames 0-23 fused into a fresh 2^22-block map): 16 hypotheses up
    to 2 cm / 1.5 degrees around the pose of frame 24 are refined by 8
    Gauss-Newton iterations (one launch per iteration for all hypotheses)
    against frame 24's 1280x720 depth; every hypothesis ends within 1 cm /
    0.3 degrees of the true pose and the one of lowest cost within 1 mm.
    (With the workload's depth noise, 0.0012 z^2 m against a 1 cm
    truncation, the fused surface is biased at range and the refinement
    settles centimetres away; the bench reports that case.)"""
    import torch
    from voxelmapping_amd import DeviceMap, synth
    from voxelmapping_amd.autograd import se3_exp
    w = synth.room_large(noise=0.0, dropout=0.0)
    g = DeviceMap(w.n_entry, w.n_bucket, w.n_block, w.voxel_size, w.trunc, w.voxel_format)
    try:
        poses_all = w.poses(25)
        for i in range(24):
            d, _ = w.frame(i, want_rgb=False)
            g.fuse(torch.from_numpy(d.view(np.int16)).cuda().view(torch.uint16), w.K, poses_all[i],
                   depth_scale=w.depth_scale)
        d, _ = w.frame(24, want_rgb=False)
        dt = torch.from_numpy(d.view(np.int16)).cuda().view(torch.uint16)
        T0 = poses_all[24].reshape(4, 4).T.astype(np.float64)
        rng = np.random.default_rng(4)
        hyps = np.stack([T0 @ se3_exp(np.r_[rng.uniform(-1, 1, 3) * 0.02, np.radians(rng.uniform(-1, 1, 3) * 1.5)])
                         for _ in range(16)]).astype(np.float32)
        poses, cost = g.refine_poses(dt, w.K, hyps, iters=8, stride=4, depth_scale=w.depth_scale)
    finally:
        g.release()
    assert np.all(cost[:, 0, 1] > 1000)
    err = []
    for h in range(16):
        dT = np.linalg.inv(T0) @ poses[h].astype(np.float64)
        t_err = np.linalg.norm(dT[:3, 3])
        r_err = np.degrees(np.arccos(np.clip((np.trace(dT[:3, :3]) - 1) / 2, -1, 1)))
        print(f"hyp {h}: {t_err * 1000:.2f} mm, {r_err:.3f} deg; cost {cost[h, 0, 0]:.2e} -> {cost[h, -1, 0]:.2e}")
        err.append((t_err, r_err))
    err = np.array(err)
    assert np.all(err[:, 0] < 0.01) and np.all(err[:, 1] < 0.3)
    best = int(np.argmin(cost[:, -1, 0]))
    assert err[best, 0] < 0.001
