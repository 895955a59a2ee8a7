"""The Python mirror of vmap::VoxelMapping (voxelmapping_amd/mapping.py),
driven the way the reference's facade is (include/vmapping.h:10-27:
CreateMap, FuseDepth, RenderScene, Polygonize, GetSurfacePoints, reset),
against the oracle at the geometry CreateMap picks (vmapping.cpp:57-75:
nBucket = int(0.9 numVoxels), truncation 4 voxels, the reference Voxel
format)."""
import numpy as np
import pytest

import oracle as O
from parity import compare_soups, compare_vmaps

pytestmark = pytest.mark.gpu


def test_voxelmapping_mirror_end_to_end(oracle_threads):
    import torch
    from voxelmapping_amd import VoxelMapping, synth
    w = synth.room_small(seed=2, voxel_format=0, n_frames=8)
    fx, fy, cx, cy = (float(v) for v in w.K)
    K3 = np.array([[fx, 0, cx], [0, fy, cy], [0, 0, 1]], np.float32)
    vm = VoxelMapping(w.width, w.height, K3, mesh_capacity=2_000_000)
    T0 = w.poses(8)[0]
    with pytest.raises(RuntimeError):
        vm.RenderScene(None, T0)  # before CreateMap
    numEntries, numVoxels = 90_000, 1 << 16
    vm.CreateMap(numEntries, numVoxels, w.voxel_size)
    assert vm.map.desc.n_bucket == int(0.9 * numVoxels) and vm.map.desc.voxel_format == 0
    o = O.OracleMap(numEntries, int(0.9 * numVoxels), numVoxels, w.voxel_size, 4 * w.voxel_size, 0)
    for i in range(8):
        d, _ = w.frame(i, want_rgb=False)
        metres = np.ascontiguousarray(w.depth_metres(d), dtype=np.float32)
        T = w.poses(8)[i]
        vm.FuseDepth(torch.from_numpy(metres).cuda(), T)
        assert vm.numVisBlock == o.fuse(metres, w.K, T) > 0, f"frame {i}"
    assert o.counters()["alias_events"] == 0
    T = w.poses(8)[7]
    v = vm.RenderScene(None, T)
    torch.cuda.synchronize()
    ov = np.zeros((w.height, w.width, 4), np.float32)
    o.render_scene(ov, w.K, T)
    compare_vmaps(v.cpu().numpy(), ov, min_hits=0.5)
    t, verts, norms = vm.Polygonize()
    ob, ot, over, onorm = o.polygonize(2_000_000)
    assert t == ot > 1000 and vm.numVisBlock == ob
    compare_soups(verts.reshape(-1, 9), norms.reshape(-1, 9), over, onorm)
    pts = vm.GetSurfacePoints()
    assert pts.shape == (t, 3)
    assert np.allclose(pts, verts.reshape(-1, 3, 3).mean(1), atol=1e-6)
    # after reset nothing is visible: RenderScene leaves the caller's buffer
    # untouched, as the reference's gate does (vmapping.cpp:87-88)
    vm.reset()
    v2 = torch.full((w.height, w.width, 4), 7.0, dtype=torch.float32, device="cuda")
    assert vm.RenderScene(v2, T) is v2
    torch.cuda.synchronize()
    assert (v2 == 7.0).all()
