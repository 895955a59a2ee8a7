"""The HIP path against the committed golden fixtures (tests/golden/*.npz):
the stored depth frames are fused on the GPU and the results compared to the
oracle's stored outputs -- block keys and visible counts exact, TSDF within
1e-5 and weights / colours exact on the sampled blocks, vertex map within
1e-4 m with an equal hit mask, normals within 1e-4 where finite, the same
triangle count and the first canonical triangles within 1e-5 m."""
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
import make_golden as G  # noqa: E402

pytestmark = pytest.mark.gpu


def load(name):
    with np.load(os.path.join(HERE, "golden", name + ".npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


@pytest.mark.parametrize("name", sorted(G.CASES))
def test_gpu_matches_golden(name):
    import torch
    from voxelmapping_amd import DeviceMap, computeNormal
    gold = load(name)
    n_entry, n_bucket, n_block, fmt = (int(v) for v in gold["geometry"])
    vs, trunc = (float(v) for v in gold["voxel"])
    depth, poses, K = gold["depth"], gold["poses"], gold["K"]
    rgb = gold.get("rgb")
    H, W = depth.shape[1:]
    print(f"golden case {name}: {W}x{H}, {len(depth)} frames, format {fmt}", flush=True)
    g = DeviceMap(n_entry, n_bucket, n_block, vs, trunc, fmt, 0.1, 5.0, 255.0)
    counts = []
    for i in range(len(depth)):
        d = torch.from_numpy(depth[i]).cuda()
        c = None if rgb is None else torch.from_numpy(rgb[i]).cuda()
        counts.append(g.fuse(d, K, poses[i], rgb=c, return_count=True))
    assert counts == gold["counts"].tolist()

    keys, ptrs = G.sorted_keys(g.hash_table())
    assert np.array_equal(keys, gold["keys"])
    rows = gold["sample_rows"]
    sdf, wt, col = G.voxel_state(g.download(2), fmt, ptrs[rows])
    assert np.max(np.abs(sdf - gold["sample_sdf"])) <= 1e-5
    assert np.array_equal(wt, gold["sample_weight"])
    if col is not None:
        assert np.array_equal(col, gold["sample_rgb"])

    vm = torch.zeros((H, W, 4), dtype=torch.float32, device="cuda")
    g.raycast(vm, K, poses[-1])
    nm = computeNormal(vm, torch.zeros_like(vm))
    torch.cuda.synchronize()
    gv, gn = vm.cpu().numpy()[..., :3], nm.cpu().numpy()[..., :3]
    ov, on = gold["vmap"], gold["nmap"]
    gh, oh = ~np.isnan(gv[..., 0]), ~np.isnan(ov[..., 0])
    assert np.array_equal(gh, oh) and gh.mean() > 0.9
    assert np.max(np.abs(gv[gh] - ov[oh])) <= 1e-4
    assert np.array_equal(np.isnan(gn), np.isnan(on))
    fin = ~np.isnan(on)
    assert np.max(np.abs(gn[fin] - on[fin])) <= 1e-4

    cap = G.MESH_CAP
    verts = torch.zeros((cap, 9), dtype=torch.float32, device="cuda")
    norms = torch.zeros_like(verts)
    _, t = g.polygonize(verts, norms)
    assert t == int(gold["n_triangles"])
    mesh = G.canonical_mesh(verts[:t].cpu().numpy(), norms[:t].cpu().numpy())
    head = gold["mesh_head"]
    assert np.max(np.abs(mesh[:len(head), :9] - head[:, :9])) <= 1e-5
    assert np.max(np.abs(mesh[:len(head), 9:] - head[:, 9:])) <= 1e-4
