"""GPU parity of the raycaster, normals, shading, rendering-block z-ranges,
marching cubes and surface points against the CPU oracle (raycast depth
within 1e-4 m with equal hit masks; same triangle count, vertices within
1e-5 m after canonical sorting)."""
import numpy as np
import pytest

import oracle as O
from parity import canonical_triangles, compare_maps, fuse_both, make_pair

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def fused(oracle_threads):
    from voxelmapping_amd import synth
    w = synth.room_small()
    g, o = make_pair(w, fmt=2)
    fuse_both(w, g, o, range(10), rgb=True)
    return w, g, o


def _gpu_raycast(g, w, T, facade=False):
    import torch
    vm = torch.zeros((w.height, w.width, 4), dtype=torch.float32, device="cuda")
    if facade:
        g.render_scene(vm, w.K, T)
    else:
        g.raycast(vm, w.K, T)
    torch.cuda.synchronize()
    return vm


def _compare_vmaps(gv, ov, tol=1e-4):
    gh, oh = ~np.isnan(gv[..., 0]), ~np.isnan(ov[..., 0])
    assert np.array_equal(gh, oh), f"hit masks differ at {np.sum(gh != oh)} pixels"
    d = np.abs(gv[gh][:, :3] - ov[oh][:, :3])
    assert d.size == 0 or d.max() <= tol, f"vertex maps differ by {d.max()}"
    assert np.array_equal(gv[~gh][:, 1:], ov[~oh][:, 1:])
    return gh.mean()


@pytest.mark.parametrize("which", ["final", "perturbed"])
def test_raycast_parity(fused, which):
    w, g, o = fused
    T = w.poses(10)[9].reshape(4, 4).T.copy()
    if which == "perturbed":
        T[:3, 3] += np.array([0.013, -0.007, 0.004], np.float32)
    Tc = T.T.reshape(-1)
    gv = _gpu_raycast(g, w, Tc).cpu().numpy()
    ov = o.raycast(w.width, w.height, w.K, Tc, init=0.0)
    assert _compare_vmaps(gv, ov) > 0.9


def _look_at(eye, target, up=(0.0, 0.0, 1.0)):
    """camToWorld (4x4, row-major) of a camera at eye looking at target
    (camera z forward, x right, y down)."""
    eye, target, up = (np.asarray(a, np.float64) for a in (eye, target, up))
    z = target - eye
    z /= np.linalg.norm(z)
    x = np.cross(z, up)
    x /= np.linalg.norm(x)
    y = np.cross(z, x)
    T = np.eye(4)
    T[:3, 0], T[:3, 1], T[:3, 2], T[:3, 3] = x, y, z, eye
    return T.astype(np.float32)


@pytest.mark.parametrize("which", ["backward", "surface_3cm", "surface_12cm", "oblique", "outside", "corner"])
def test_raycast_parity_unusual_poses(fused, which):
    """Poses that stress the per-tile ray-distance bounds: blocks reaching
    behind the camera plane (surface_3cm, surface_12cm: the camera moved
    along its axis to that distance from the surface of the centre pixel),
    rays leaving the map through the far side (backward, outside), an
    oblique view from farther away, the camera outside the room looking in
    through a fused wall (outside), a long diagonal (corner)."""
    w, g, o = fused
    T = w.poses(10)[9].reshape(4, 4).T.copy()
    if which == "backward":
        T[:3, 0] *= -1.0
        T[:3, 2] *= -1.0
    elif which.startswith("surface_"):
        ov = o.raycast(w.width, w.height, w.K, T.T.reshape(-1), init=0.0)
        c = ov[w.height // 2, w.width // 2]
        assert not np.isnan(c[0])
        T[:3, 3] += T[:3, 2] * (float(c[2]) - float(which[8:-2]) / 100.0)
    elif which == "oblique":
        a = np.radians(25.0)
        R = np.array([[np.cos(a), 0, np.sin(a)], [0, 1, 0], [-np.sin(a), 0, np.cos(a)]], np.float32)
        T[:3, :3] = T[:3, :3] @ R
        T[:3, 3] -= T[:3, 2] * 0.8
    elif which == "outside":
        T = _look_at((2.0, -0.6, 1.4), (2.0, 2.0, 1.2))
    else:
        T = _look_at((0.3, 0.3, 0.3), (3.8, 3.8, 2.8))
    Tc = T.T.reshape(-1).copy()
    gv = _gpu_raycast(g, w, Tc).cpu().numpy()
    ov = o.raycast(w.width, w.height, w.K, Tc, init=0.0)
    _compare_vmaps(gv, ov)
    gv2 = _gpu_raycast(g, w, Tc, facade=True).cpu().numpy()
    ov2 = np.zeros_like(gv2)
    o.render_scene(ov2, w.K, Tc)
    _compare_vmaps(gv2, ov2)


def test_raycast_bounds_across_image_sizes(fused):
    """One map raycast at several image sizes in turn: the per-tile bounds
    keep their layout by capacity, and 4096x1040 (66560 tiles, more than the
    map's initial 2^16) reallocates and refills them; every size matches the
    oracle."""
    import torch
    w, g, o = fused
    T = w.poses(10)[9]
    for W, H in [(320, 240), (4096, 1040), (160, 120), (640, 480)]:
        s = W / 320.0
        K = (262.5 * s, 262.5 * s, (W - 1) / 2.0, (H - 1) / 2.0)
        vm = torch.zeros((H, W, 4), dtype=torch.float32, device="cuda")
        g.raycast(vm, K, T)
        torch.cuda.synchronize()
        gv = vm.cpu().numpy()
        ov = o.raycast(W, H, K, T, init=0.0)
        assert _compare_vmaps(gv, ov) > 0.5, (W, H)


def test_render_scene_facade_and_gate(fused):
    w, g, o = fused
    Tc = w.poses(10)[9]
    gv = _gpu_raycast(g, w, Tc, facade=True).cpu().numpy()
    ov = np.zeros_like(gv)
    assert o.render_scene(ov, w.K, Tc) > 0
    _compare_vmaps(gv, ov)
    # nothing visible (camera looking away from every block): vmap untouched
    import torch
    from voxelmapping_amd import DeviceMap
    e = DeviceMap(1 << 12, 1 << 11, 64, 0.01, 0.04, 0)
    vm = torch.full((w.height, w.width, 4), 7.0, device="cuda")
    e.render_scene(vm, w.K, Tc)
    torch.cuda.synchronize()
    assert torch.all(vm == 7.0)


def test_normals_parity(fused):
    import torch
    from voxelmapping_amd import computeNormal
    w, g, o = fused
    vm = _gpu_raycast(g, w, w.poses(10)[9])
    nm = torch.zeros_like(vm)
    computeNormal(vm, nm)
    torch.cuda.synchronize()
    gn = nm.cpu().numpy()
    on = O.compute_normals(vm.cpu().numpy())
    both = np.isnan(gn) & np.isnan(on)
    assert np.array_equal(np.isnan(gn), np.isnan(on))
    assert np.max(np.abs(np.where(both, 0, gn - on))) <= 1e-6


def test_shading_parity(fused):
    import torch
    from voxelmapping_amd import computeNormal, renderScene
    w, g, o = fused
    vm = _gpu_raycast(g, w, w.poses(10)[9])
    nm = computeNormal(vm, torch.zeros_like(vm))
    img = renderScene(vm, nm)
    torch.cuda.synchronize()
    ref = O.render_shading(vm.cpu().numpy(), nm.cpu().numpy())
    assert np.max(np.abs(img.cpu().numpy().astype(int) - ref.astype(int))) <= 1


def test_rendering_blocks_parity(fused):
    import torch
    w, g, o = fused
    T = w.poses(10)[9]
    # both sides compute the visible table at this pose first
    assert g.check_visible(w.width, w.height, w.K, T) == o.check_visible(w.width, w.height, w.K, T)
    Twc = O.pose_inverse(T)
    c8, r8 = w.width // 8, w.height // 8
    zx = torch.zeros((r8, c8), dtype=torch.float32, device="cuda")
    zy = torch.zeros_like(zx)
    n = g.project_rendering_blocks(zx, zy, w.K, Twc.reshape(-1))
    no, ozx, ozy = o.project_rendering_blocks(c8, r8, w.K, Twc)
    assert n == no and n > 0
    assert np.array_equal(zx.cpu().numpy(), ozx) and np.array_equal(zy.cpu().numpy(), ozy)


def test_polygonize_parity(fused):
    import torch
    from voxelmapping_amd import surface_points
    w, g, o = fused
    cap = 3_000_000
    verts = torch.zeros((cap, 9), dtype=torch.float32, device="cuda")
    norms = torch.zeros_like(verts)
    gb, gt = g.polygonize(verts, norms)
    ob, ot, ov, on = o.polygonize(cap)
    assert (gb, gt) == (ob, ot) and gt > 1000
    gv = verts[:gt].cpu().numpy()
    gn = norms[:gt].cpu().numpy()
    sg, og_ = canonical_triangles(gv)
    so, oo = canonical_triangles(ov)
    assert np.max(np.abs(sg - so)) <= 1e-5
    assert np.max(np.abs(gn[og_] - on[oo])) <= 1e-5
    pts = surface_points(verts, gt).cpu().numpy()
    assert np.max(np.abs(pts - O.surface_points(gv))) <= 1e-6
    # the triangle order is a function of the map: a second call is identical
    verts2 = torch.zeros_like(verts)
    norms2 = torch.zeros_like(verts)
    assert g.polygonize(verts2, norms2) == (gb, gt)
    assert torch.equal(verts2[:gt], verts[:gt]) and torch.equal(norms2[:gt], norms[:gt])


def test_polygonize_cap_clamps(fused):
    import torch
    w, g, o = fused
    verts = torch.zeros((1000, 9), dtype=torch.float32, device="cuda")
    norms = torch.zeros_like(verts)
    _, t = g.polygonize(verts, norms)
    assert t == 1000
    assert g.stats()["status"] & 8


def test_polygonize_empty_map_and_zero_cap(fused):
    """An empty map meshes to nothing; cap_triangles = 0 accepts NULL
    buffers and still counts the allocated blocks."""
    import ctypes as C
    import torch
    from voxelmapping_amd import _capi
    w, g, o = fused
    e, _ = make_pair(w, fmt=2)
    verts = torch.zeros((16, 9), dtype=torch.float32, device="cuda")
    norms = torch.zeros_like(verts)
    assert e.polygonize(verts, norms) == (0, 0)
    b, t = C.c_uint32(7), C.c_uint32(7)
    _capi.check(_capi.lib().vm_polygonize(g.handle, None, None, 0, None, C.byref(b), C.byref(t)))
    assert (b.value, t.value) == (g.stats()["allocated_blocks"], 0)


def test_uploaded_oracle_state_raycast_and_mesh(oracle_threads):
    """Load an oracle-built map into the GPU map and compare stage outputs."""
    import torch
    from voxelmapping_amd import DeviceMap
    from voxelmapping_amd import synth
    w = synth.room_small(seed=5)
    o = O.OracleMap(w.n_entry, w.n_bucket, w.n_block, w.voxel_size, w.trunc, 0)
    for i in range(4):
        d, _ = w.frame(i, want_rgb=False)
        o.fuse(d, w.K, w.poses(i + 1)[i])
    g = DeviceMap(w.n_entry, w.n_bucket, w.n_block, w.voxel_size, w.trunc, 0)
    g.upload(2, o.voxels())
    g.upload(1, o.heap())
    c = np.zeros(16, np.int32)
    oc = o.counters()
    c[0], c[1] = oc["heap_ptr"], oc["excess_ptr"]
    g.upload(3, c)
    g.upload(0, o.hash_table())
    compare_maps(g, o, 0)
    T = w.poses(4)[3]
    gv = _gpu_raycast(g, w, T).cpu().numpy()
    _compare_vmaps(gv, o.raycast(w.width, w.height, w.K, T, init=0.0))
    verts = torch.zeros((2_000_000, 9), dtype=torch.float32, device="cuda")
    norms = torch.zeros_like(verts)
    gb, gt = g.polygonize(verts, norms)
    ob, ot, ov, _ = o.polygonize(2_000_000)
    assert (gb, gt) == (ob, ot)
    sg, _ = canonical_triangles(verts[:gt].cpu().numpy())
    so, _ = canonical_triangles(ov)
    assert np.max(np.abs(sg - so)) <= 1e-5


@pytest.mark.parametrize("scale", [1.0 + 3e-5, 1.02])
def test_raycast_non_rigid_pose(fused, scale):
    """A camToWorld whose rotation is scaled (not orthonormal): the per-tile
    ray-distance bounds assume a rigid rotation, so the library marches
    these rays in full; the result still equals the oracle's march (which
    has no bounds) -- raycast and the facade's render_scene alike."""
    w, g, o = fused
    T = w.poses(10)[9].reshape(4, 4).T.copy()
    T[:3, :3] *= np.float32(scale)
    Tc = T.T.reshape(-1)
    ov = o.raycast(w.width, w.height, w.K, Tc, init=0.0)
    assert _compare_vmaps(_gpu_raycast(g, w, Tc).cpu().numpy(), ov) > 0.5
    assert _compare_vmaps(_gpu_raycast(g, w, Tc, facade=True).cpu().numpy(), ov) > 0.5


def test_render_reuses_fusion_bounds_only_at_its_view(oracle_threads):
    """A fusion computes the raycast's per-tile ray bounds at its own pose in
    its visibility launch; a render at that pose (render_scene and raycast)
    reuses them, a render at another pose or after an allocation computes
    its own. Every render equals the oracle's."""
    import torch
    from voxelmapping_amd import synth
    w = synth.room_small(n_frames=8)
    g, o = make_pair(w, fmt=2)
    fuse_both(w, g, o, range(6), rgb=True)

    def both(T, facade=True):
        vm = _gpu_raycast(g, w, T, facade=facade).cpu().numpy()
        ov = np.zeros_like(vm)
        o.render_scene(ov, w.K, T)
        return _compare_vmaps(vm, ov)

    T5 = w.poses(6)[5]
    assert both(T5) > 0.9  # the fusion's bounds
    Tm = T5.reshape(4, 4).T.copy()
    Tm[:3, 3] += np.array([0.001, 0.0, -0.001], np.float32)
    assert both(Tm.T.reshape(-1)) > 0.9  # another pose: bounds of its own
    assert both(T5, facade=False) > 0.9  # back at the fusion's pose
    # an allocation alone adds empty blocks, which render as absent ones:
    # the render after it equals the oracle's map without them
    from parity import gpu_frame
    d, c, dt, ct = gpu_frame(w, 7)
    T7 = w.poses(8)[7]
    g.allocate(dt, w.K, T7, depth_scale=w.depth_scale)
    assert both(T5) > 0.9
    # the integration of that frame computes the bounds at its pose, which
    # the renders at that pose reuse
    g.integrate(dt, w.K, T7, rgb=ct, depth_scale=w.depth_scale)
    o.fuse(w.depth_metres(d), w.K, T7, rgb=c)
    torch.cuda.synchronize()
    assert both(T7) > 0.9
    assert both(T7, facade=False) > 0.9
    compare_maps(g, o, 2)
