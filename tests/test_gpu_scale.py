"""The large-room workload at full size: 1280x720 uint16 depth with colour,
2 mm voxels, 1 cm truncation, 2^23 buckets + 2^20 excess slots, 2^22
float-TSDF blocks (24 GiB of voxels). The first two frames are checked against
the oracle (visible counts, allocated key set, TSDF/weight/colour of sampled
blocks, raycast vertices of sampled pixels); forty frames are then checked
through size-independent properties (no capacity status, every sampled key
reachable from its bucket, weights and TSDF in range, deterministic rerun
of the last frame's raycast). Pose refinement at full size on the same
geometry without depth noise."""
import numpy as np
import pytest

import oracle as O
from parity import chain_lookup_all, colmajor, gpu_frame

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
