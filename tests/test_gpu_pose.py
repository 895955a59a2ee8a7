"""Batched Gauss-Newton pose refinement (vm_refine_poses).

* The gradient J^T r of the normal equations equals central finite
  differences of the cost over the six twist components: of the kernel's
  own FP32 cost, and of the oracle's FP64 restatement of the depth
  (or_surface_depth_f64: march, corrections, Newton steps onto the zero
  level) on a noisy room_vga map, with the residual set and the cost
  compared too.
* Hypotheses perturbed by up to 2 cm / 1.5 degrees around the true pose of a
  well-conditioned view converge to it, with the cost decreasing.
* Arguments are validated."""
import numpy as np
import pytest

from parity import fuse_both, make_pair

pytestmark = pytest.mark.gpu

STRIDE, MAXRES = 2, 0.05


@pytest.fixture(scope="module")
def fused(oracle_threads):
    from voxelmapping_amd import synth
    w = synth.room_small(seed=3, voxel_format=2, n_frames=12)
    g, o = make_pair(w)
    fuse_both(w, g, o, range(12), rgb=False)
    return w, g


def true_pose(w, i):
    return w.poses(i + 1)[i].reshape(4, 4).T.astype(np.float64)


def perturb(T, rng, t_sigma, r_deg):
    from voxelmapping_amd.autograd import se3_exp
    xi = np.r_[rng.uniform(-1, 1, 3) * t_sigma, np.radians(rng.uniform(-1, 1, 3) * r_deg)]
    return T @ se3_exp(xi)


def fused_map(w, n):
    import torch
    from voxelmapping_amd import DeviceMap
    g = DeviceMap(w.n_entry, w.n_bucket, w.n_block, w.voxel_size, w.trunc, w.voxel_format)
    for i in range(n):
        d, _ = w.frame(i, want_rgb=False)
        g.fuse(torch.from_numpy(d).cuda(), w.K, w.poses(n)[i])
    return g


def test_gradient_matches_finite_differences():
    """J^T r of the normal equations against central differences of the cost
    sum r^2 over the six twist components, in an empty room seen without
    noise, with the observed depth set to the map's own raycast depth plus a
    smooth offset of a few millimetres (residuals continuous in the pose)."""
    import torch
    from voxelmapping_amd import synth
    from voxelmapping_amd.autograd import se3_exp
    w = synth.room_small(seed=6, voxel_format=2, n_frames=30, n_spheres=0, n_boxes=0, noise=0.0, dropout=0.0)
    g = fused_map(w, 30)
    rng = np.random.default_rng(2)
    T = perturb(true_pose(w, 20), rng, 0.01, 0.5)
    vm = torch.zeros((w.height, w.width, 4), dtype=torch.float32, device="cuda")
    g.raycast(vm, w.K, T.astype(np.float32))
    yy, xx = torch.meshgrid(torch.arange(w.height, device="cuda"), torch.arange(w.width, device="cuda"),
                            indexing="ij")
    D = vm[..., 2] + 0.003 * torch.sin(xx * 0.05) * torch.cos(yy * 0.07)
    D = torch.where(torch.isnan(vm[..., 0]), torch.full_like(D, float("nan")), D).contiguous()

    def evaluate(Tk):
        _, cost, system = g.refine_poses(D, w.K, Tk.astype(np.float32)[None], iters=0, stride=STRIDE,
                                         max_residual=MAXRES, return_system=True)
        return float(cost[0, 0, 0]) * float(cost[0, 0, 1]), int(cost[0, 0, 1]), system[0, 0]

    E0, n0, system = evaluate(T)
    assert n0 > 10000
    grad = 2.0 * system[21:27]  # d/dxi sum r^2 = 2 J^T r
    h = 1e-4
    fd = np.zeros(6)
    for k in range(6):
        e = np.zeros(6)
        e[k] = h
        Ep, np_, _ = evaluate(T @ se3_exp(e))
        Em, nm, _ = evaluate(T @ se3_exp(-e))
        # a ray at the edge of the fused region may switch between hit and miss
        assert abs(np_ - n0) <= 1 and abs(nm - n0) <= 1, (k, np_, nm, n0)
        fd[k] = (Ep - Em) / (2 * h)
    rel = np.abs(fd - grad) / np.abs(grad).max()
    print(f"gradient {grad}\nfinite differences {fd}\nrel {rel}")
    assert rel.max() < 5e-2


def test_gradient_matches_fp64_host_evaluation(oracle_threads):
    """room_vga (640x480 uint16 depth with its noise and dropout), 12 frames
    fused into the GPU map and the oracle; frame 12's depth against a pose
    1 cm / 0.5 degrees off, every 8th pixel: the residual count and the cost
    match the FP64 host evaluation, and J^T r matches FP64 central
    differences of 0.5 sum r^2 over the six twist components (h = 1e-6)."""
    from voxelmapping_amd import synth
    from voxelmapping_amd.autograd import se3_exp
    from parity import gpu_frame
    w = synth.room_vga()
    g, o = make_pair(w)
    fuse_both(w, g, o, range(12), rgb=False)
    d, _, dt, _ = gpu_frame(w, 12, rgb=False)
    rng = np.random.default_rng(5)
    T32 = perturb(true_pose(w, 12), rng, 0.01, 0.5).astype(np.float32)
    _, cost, system = g.refine_poses(dt, w.K, T32[None], iters=0, stride=8, max_residual=0.05,
                                     depth_scale=w.depth_scale, return_system=True)
    n_gpu, mse_gpu, jtr_gpu = int(cost[0, 0, 1]), float(cost[0, 0, 0]), system[0, 0, 21:27]
    ys, xs = np.mgrid[0:w.height:8, 0:w.width:8]
    dep = w.depth_metres(d)[ys, xs].astype(np.float64)
    keep = (dep >= 0.1) & (dep <= 5.0)
    xs, ys, dep = xs[keep], ys[keep], dep[keep]
    K = np.asarray(w.K, np.float64)
    T = T32.astype(np.float64)
    z, ok = o.surface_depth_f64(xs, ys, K, T)
    r = z - dep
    inset = ok & (np.abs(r) < 0.05)
    n_host = int(inset.sum())
    print(f"residuals gpu {n_gpu} host {n_host}; mean r^2 gpu {mse_gpu:.6e} host {np.mean(r[inset] ** 2):.6e}")
    assert n_gpu > 2000 and abs(n_gpu - n_host) <= 0.002 * n_host + 2
    assert abs(mse_gpu - np.mean(r[inset] ** 2)) <= 0.01 * mse_gpu
    h = 1e-6
    fd = np.zeros(6)
    for k in range(6):
        e = np.zeros(6)
        e[k] = h
        zp, okp = o.surface_depth_f64(xs, ys, K, T @ se3_exp(e))
        zm, okm = o.surface_depth_f64(xs, ys, K, T @ se3_exp(-e))
        both = inset & okp & okm
        assert both.sum() >= inset.sum() - 2
        fd[k] = np.sum(r[both] * (zp[both] - zm[both]) / (2 * h))
    rel = np.abs(fd - jtr_gpu) / np.abs(fd).max()
    print(f"J^T r gpu {jtr_gpu}\nfp64 central differences {fd}\nrelative error {rel}")
    assert rel.max() < 1e-2


def test_hypotheses_converge():
    """Noise-free views of a room with 20 objects; the frame whose J^T J at
    the true pose is best conditioned is refined from 8 hypotheses up to 2 cm
    and 1.5 degrees away."""
    import torch
    from voxelmapping_amd import synth
    w = synth.room_small(seed=3, voxel_format=2, n_frames=30, n_spheres=12, n_boxes=8, noise=0.0, dropout=0.0)
    g = fused_map(w, 30)
    best, cond = None, np.inf
    for i in range(15, 30):
        d, _ = w.frame(i, want_rgb=False)
        _, _, sysm = g.refine_poses(torch.from_numpy(d).cuda(), w.K, true_pose(w, i).astype(np.float32)[None],
                                    iters=0, stride=STRIDE, max_residual=MAXRES, return_system=True)
        H = np.zeros((6, 6))
        H[np.triu_indices(6)] = sysm[0, 0, :21]
        H = H + np.triu(H, 1).T
        # translation in metres, rotation in radians scaled by a 2 m lever
        S = np.diag([1, 1, 1, 0.5, 0.5, 0.5])
        ev = np.linalg.eigvalsh(S @ H @ S)
        if ev[-1] / max(ev[0], 1e-30) < cond:
            best, cond = i, ev[-1] / max(ev[0], 1e-30)
    print(f"frame {best}: condition number {cond:.1f}")
    rng = np.random.default_rng(9)
    T0 = true_pose(w, best)
    hyps = np.stack([perturb(T0, rng, 0.02, 1.5) for _ in range(8)]).astype(np.float32)
    d, _ = w.frame(best, want_rgb=False)
    dt = torch.from_numpy(d).cuda()
    _, c_true = g.refine_poses(dt, w.K, T0.astype(np.float32)[None], iters=0, stride=STRIDE, max_residual=MAXRES)
    poses, cost = g.refine_poses(dt, w.K, hyps, iters=8, stride=STRIDE, max_residual=MAXRES)
    print(f"cost at the true pose {c_true[0, 0, 0]:.2e}")
    close = []
    for h in range(len(hyps)):
        dT = np.linalg.inv(T0) @ poses[h].astype(np.float64)
        t_err = np.linalg.norm(dT[:3, 3])
        r_err = np.degrees(np.arccos(np.clip((np.trace(dT[:3, :3]) - 1) / 2, -1, 1)))
        print(f"hyp {h}: {t_err * 1000:.2f} mm, {r_err:.3f} deg; cost {cost[h, 0, 0]:.2e} -> {cost[h, -1, 0]:.2e}")
        # every hypothesis is optimised down to (at least) the true pose's cost
        assert cost[h, -1, 0] < 1.5 * c_true[0, 0, 0]
        close.append(t_err < 0.006 and r_err < 0.4)
    # Depth alone leaves shallow valleys (the fused 1 cm TSDF's zero level is
    # itself millimetres off the geometry): most hypotheses, and the one of
    # lowest cost, end within 6 mm / 0.4 degrees of the true pose.
    assert sum(close) >= len(hyps) // 2
    assert close[int(np.argmin(cost[:, -1, 0]))]


def test_arguments_are_checked(fused):
    import torch
    from voxelmapping_amd import VMError
    w, g = fused
    d = torch.zeros((w.height, w.width), dtype=torch.float32, device="cuda")
    T = true_pose(w, 0).astype(np.float32)[None]
    with pytest.raises(VMError):
        g.refine_poses(d, w.K, T, iters=-1)
    with pytest.raises(VMError):
        g.refine_poses(d, w.K, T, stride=0)
    with pytest.raises(VMError):
        g.refine_poses(d, w.K, T, max_residual=0.0)
    bad = T.copy()
    bad[0, 0, 0] = np.nan
    with pytest.raises(VMError):
        g.refine_poses(d, w.K, bad)
    # hypotheses are rigid: a rotation scaled by 1.001 is refused
    scaled = T.copy()
    scaled[0, :3, :3] *= np.float32(1.001)
    with pytest.raises(VMError):
        g.refine_poses(d, w.K, scaled)
    # an empty depth image leaves every pose as it was and reports no residuals
    poses, cost = g.refine_poses(d, w.K, T, iters=2)
    assert np.array_equal(poses, T) and np.all(cost[..., 1] == 0) and np.all(cost[..., 0] == -1)
