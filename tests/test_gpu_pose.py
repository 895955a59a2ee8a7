"""Batched Gauss-Newton pose refinement (vm_refine_poses).

* The gradient J^T r of the normal equations equals central finite
  differences of the cost over the six twist components.
* Hypotheses perturbed by up to 2 cm / 1.5 degrees around the true pose of a
  frame converge to it, with the mean squared residual decreasing.
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


def test_gradient_matches_finite_differences(oracle_threads):
    """J^T r of the normal equations against central differences of the
    cost sum r^2 over pose twists, in an empty room seen without noise (no
    depth discontinuities, so the residuals are continuous in the pose)."""
    import torch
    from voxelmapping_amd import synth
    from voxelmapping_amd.autograd import se3_exp
    w = synth.room_small(seed=6, voxel_format=2, n_frames=30, n_spheres=0, n_boxes=0, noise=0.0, dropout=0.0)
    g, o = make_pair(w)
    fuse_both(w, g, o, range(30), rgb=False, check_counts=False)
    rng = np.random.default_rng(2)
    T = perturb(true_pose(w, 20), rng, 0.01, 0.5)
    d, _ = w.frame(20, want_rgb=False)
    dt = torch.from_numpy(d).cuda()

    def evaluate(Tk):
        _, cost, system = g.refine_poses(dt, w.K, Tk.astype(np.float32)[None], iters=0, stride=STRIDE,
                                         max_residual=0.1, return_system=True)
        return float(cost[0, 0, 0]) * float(cost[0, 0, 1]), int(cost[0, 0, 1]), system[0, 0]

    E0, n0, system = evaluate(T)
    assert n0 > 5000
    grad = 2.0 * system[21:27]  # d/dxi sum r^2 = 2 J^T r
    h = 1e-4
    fd = np.zeros(6)
    for k in range(6):
        e = np.zeros(6)
        e[k] = h
        Ep, np_, _ = evaluate(T @ se3_exp(e))
        Em, nm, _ = evaluate(T @ se3_exp(-e))
        assert np_ == nm == n0, (k, np_, nm, n0)
        fd[k] = (Ep - Em) / (2 * h)
    rel = np.abs(fd - grad) / np.abs(grad).max()
    print(f"gradient {grad}\nfinite differences {fd}\nrel {rel}")
    assert rel.max() < 2e-2


def test_hypotheses_converge(fused):
    import torch
    w, g = fused
    rng = np.random.default_rng(9)
    T0 = true_pose(w, 11)
    hyps = np.stack([perturb(T0, rng, 0.02, 1.5) for _ in range(8)]).astype(np.float32)
    d, _ = w.frame(11, want_rgb=False)
    dt = torch.from_numpy(d).cuda()
    _, c0 = g.refine_poses(dt, w.K, T0.astype(np.float32)[None], iters=0, stride=STRIDE, max_residual=MAXRES)
    _, c1 = g.refine_poses(dt, w.K, T0.astype(np.float32)[None], iters=8, stride=STRIDE, max_residual=MAXRES)
    print(f"cost at the true pose {c0[0, 0, 0]:.3e} (n {c0[0, 0, 1]:.0f}); refined from it {c1[0, :, 0]}")
    poses, cost = g.refine_poses(dt, w.K, hyps, iters=8, stride=STRIDE, max_residual=MAXRES)
    for h in range(len(hyps)):
        print("  costs", cost[h, :, 0])
        dT = np.linalg.inv(T0) @ poses[h].astype(np.float64)
        t_err = np.linalg.norm(dT[:3, 3])
        r_err = np.degrees(np.arccos(np.clip((np.trace(dT[:3, :3]) - 1) / 2, -1, 1)))
        t0 = np.linalg.norm((np.linalg.inv(T0) @ hyps[h].astype(np.float64))[:3, 3])
        print(f"hyp {h}: start {t0 * 1000:.1f} mm -> {t_err * 1000:.2f} mm, {r_err:.3f} deg; "
              f"cost {cost[h, 0, 0]:.2e} -> {cost[h, -1, 0]:.2e} (n {cost[h, -1, 1]:.0f})")
    for h in range(len(hyps)):
        dT = np.linalg.inv(T0) @ poses[h].astype(np.float64)
        t_err = np.linalg.norm(dT[:3, 3])
        r_err = np.degrees(np.arccos(np.clip((np.trace(dT[:3, :3]) - 1) / 2, -1, 1)))
        assert cost[h, -1, 0] < cost[h, 0, 0]
        assert t_err < 0.003 and r_err < 0.15


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
    # an empty depth image leaves every pose as it was and reports no residuals
    poses, cost = g.refine_poses(d, w.K, T, iters=2)
    assert np.array_equal(poses, T) and np.all(cost[..., 1] == 0) and np.all(cost[..., 0] == -1)
