"""Map file round trip (save -> load gives an identical map that renders
identically and keeps fusing identically) and a TUM-format sequence fused
from disk through DatasetLoader, checked against the oracle."""
import numpy as np
import pytest
from PIL import Image

from parity import colmajor, compare_maps, fuse_both, make_pair

pytestmark = pytest.mark.gpu


def _mesh(m, cap=600_000):
    import torch
    verts = torch.zeros((cap, 9), dtype=torch.float32, device="cuda")
    norms = torch.zeros_like(verts)
    _, t = m.polygonize(verts, norms)
    return verts[:t].cpu().numpy(), norms[:t].cpu().numpy()


@pytest.mark.parametrize("fmt", [1, 2])
def test_map_save_load_round_trip(tmp_path, oracle_threads, fmt):
    import torch
    from voxelmapping_amd import DeviceMap, synth
    w = synth.room_small(seed=2, voxel_format=fmt, n_frames=6)
    g, o = make_pair(w)
    fuse_both(w, g, o, range(4))
    path = str(tmp_path / "room.npz")
    g.save(path)
    h = DeviceMap.load(path)
    assert h.stats()["allocated_blocks"] == g.stats()["allocated_blocks"] > 0
    assert np.array_equal(h.hash_table(), g.hash_table())
    assert np.array_equal(h.download(1), g.download(1))
    assert np.array_equal(h.download(2), g.download(2))
    assert np.array_equal(h.download(5), g.download(5))
    T = w.poses(6)[3]
    va = torch.zeros((w.height, w.width, 4), dtype=torch.float32, device="cuda")
    vb = torch.zeros_like(va)
    g.raycast(va, w.K, T)
    h.raycast(vb, w.K, T)
    torch.cuda.synchronize()
    assert torch.equal(torch.nan_to_num(va, 7.0), torch.nan_to_num(vb, 7.0))
    # the same mesh from the loaded map
    mg, mh = _mesh(g), _mesh(h)
    assert len(mg[0]) > 1000 and np.array_equal(mg[0], mh[0]) and np.array_equal(mg[1], mh[1])
    # voxels written from the host (a range of blocks loses its weights):
    # the mesh follows them, and equals that of the saved and reloaded map
    k = h.stats()["allocated_blocks"] // 3
    vox = h.download_blocks(0, k).reshape(k * 512, -1).copy()
    if fmt == 2:
        vox[:, 4:8] = 0  # float weight
    else:
        vox[:, 2] = 0  # weight byte
    h.upload_blocks(0, vox)
    path2 = str(tmp_path / "room2.npz")
    h.save(path2)
    h2 = DeviceMap.load(path2)
    m1, m2 = _mesh(h), _mesh(h2)
    assert 0 < len(m1[0]) < len(mg[0]) and np.array_equal(m1[0], m2[0]) and np.array_equal(m1[1], m2[1])
    h2.release()
    # the loaded map keeps fusing exactly as the oracle does
    h = DeviceMap.load(path)
    fuse_both(w, h, o, range(4, 6))
    compare_maps(h, o, fmt)


def test_tum_sequence_from_disk(tmp_path, oracle_threads):
    import torch
    from voxelmapping_amd import synth
    from voxelmapping_amd.dataset import TUM_DEPTH_SCALE, DatasetLoader
    w = synth.room_small(seed=5, voxel_format=1, n_frames=3)
    (tmp_path / "rgb").mkdir()
    (tmp_path / "depth").mkdir()
    assoc, gt = [], []
    from scipy.spatial.transform import Rotation
    for i in range(3):
        d, c = w.frame(i)
        metres = w.depth_metres(d)
        units = np.where(np.isfinite(metres) & (metres > 0), np.round(metres / TUM_DEPTH_SCALE), 0)
        Image.fromarray(np.clip(units, 0, 65535).astype(np.uint16)).save(tmp_path / "depth" / f"{i}.png")
        Image.fromarray(c).save(tmp_path / "rgb" / f"{i}.png")
        t = 1000.0 + i / 30.0
        assoc.append(f"{t:.6f} rgb/{i}.png {t:.6f} depth/{i}.png")
        T = w.poses(3)[i].reshape(4, 4).T.astype(np.float64)
        q = Rotation.from_matrix(T[:3, :3]).as_quat()
        gt.append(f"{t:.6f} " + " ".join(f"{v:.12f}" for v in (*T[:3, 3], *q)))
    (tmp_path / "associated.txt").write_text("\n".join(assoc) + "\n")
    (tmp_path / "groundtruth.txt").write_text("\n".join(gt) + "\n")
    K = w.K if np.ndim(w.K) == 1 else np.array([w.K[0, 0], w.K[1, 1], w.K[0, 2], w.K[1, 2]])
    (tmp_path / "calibration.txt").write_text(" ".join(repr(float(v)) for v in K) + "\n")

    L = DatasetLoader(str(tmp_path))
    L.loadImages(True)
    assert L.loadGroundTruth()
    Kc = L.loadCalibration()
    g, o = make_pair(w)
    while (item := L.GetNext()) is not None:
        d, c, t, T = item
        dt = torch.from_numpy(d.view(np.int16)).cuda().view(torch.uint16)
        ng = g.fuse(dt, Kc, T.astype(np.float32), rgb=torch.from_numpy(c).cuda(), depth_scale=TUM_DEPTH_SCALE,
                    return_count=True)
        dm = d.astype(np.float32) * np.float32(TUM_DEPTH_SCALE)
        K4 = np.array([Kc[0, 0], Kc[1, 1], Kc[0, 2], Kc[1, 2]], np.float32)
        no = o.fuse(dm, K4, colmajor(T), rgb=c)
        assert ng == no > 0
    compare_maps(g, o, 1)
