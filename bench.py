#!/usr/bin/env python3
"""Fuse + raycast throughput on one MI355X.

One step = one frame of the plain drop-in pipeline at one pose: block
allocation, visibility + TSDF/weight/colour integration, raycast and vertex
normals, on the room_vga workload (640x480 uint16-millimetre depth with
colour, 5 mm voxels, 2 cm truncation, 2^20 buckets + 2^17 excess slots,
2^20-block heap, float TSDF + float weight + uint8 colour, seeded random-walk
trajectory). All frames are uploaded to HBM before the timed region.

Prints one JSON line (see the "roofline" and "cpu_baseline" objects).
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "frames/s fuse+raycast 640x480 @5mm, batched vs plain; % of HBM peak"

# algorithmic bytes (per pixel / per visible block) of one frame step
B_DEPTH, B_COLOR, B_VMAP, B_NMAP = 2, 3, 16, 16
S_VOXEL, S_RAY, S_ENTRY = 11, 8, 16
S_VOXEL_NO_COLOR = 8  # float TSDF + float weight when no colour is fused
# timed steps whose every stage is bracketed by HIP events (each event is a
# command on the stream; the raycast stage is bracketed on every step)
STAGE_EVERY = 8
# a spin kernel of this many GPU clock cycles (~10 ms) holds the stream while
# the host queues a timed batch of short calls behind it
HOLD_CYCLES = 20_000_000


def stage_bytes(P, V):
    """Algorithmic bytes per launch of each kernel for P pixels, V visible blocks."""
    return {
        "allocate": P * B_DEPTH,
        "integrate": P * (B_DEPTH + B_COLOR) + V * (512 * 2 * S_VOXEL + S_ENTRY),
        "raycast": V * 512 * S_RAY + P * B_VMAP,
        "normals": P * (B_VMAP + B_NMAP),
    }


def profiled_traffic(kernel: str, workload: str):
    """HBM bytes per launch of `kernel` on `workload` measured by the two
    rocprofv3 --pmc passes (FETCH_SIZE, WRITE_SIZE, scaled by a copy of known
    size) and the kernel's L2 hit rate (TCC_HIT_sum, TCC_MISS_sum), recorded
    in profiles/<kernel>_<workload>_traffic.md; None where absent."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_{workload}_traffic.md")
    traffic, l2 = None, None
    try:
        with open(path) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    traffic = int(line.split(":", 1)[1].strip())
                elif line.startswith("l2_hit_rate:"):
                    l2 = float(line.split(":", 1)[1].strip())
    except OSError:
        return None, None, None
    return traffic, l2, os.path.relpath(path, ROOT)


def frame_bytes(P, V):
    return P * (B_DEPTH + B_COLOR + B_VMAP + B_NMAP) + V * (512 * (2 * S_VOXEL + S_RAY) + S_ENTRY)


def _cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(w, seconds: float, max_frames: int = 64) -> dict:
    """The oracle (OpenMP over blocks and pixels, serial allocation) on the
    first frames of the same sequence -- frame i fused at pose i, then the
    raycast and normals at pose i, exactly the GPU step -- until `seconds` of
    work or max_frames frames are done. Threads: the host's CPU share of one
    GPU, at most 16 (what a one-GPU job may use on a shared machine), below
    that the cores this process may run on."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    avail = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    threads = max(1, min(16, avail))
    O.set_threads(threads)
    o = O.OracleMap(w.n_entry, w.n_bucket, w.n_block, w.voxel_size, w.trunc, w.voxel_format)
    poses = w.poses(max_frames)
    vm = np.zeros((w.height, w.width, 4), np.float32)
    el, n = 0.0, 0
    while n < max_frames and el < seconds:
        d, c = w.frame(n)  # input generation is outside the timed work
        t0 = time.perf_counter()
        o.fuse(w.depth_metres(d), w.K, poses[n], rgb=c)
        o.render_scene(vm, w.K, poses[n])
        O.compute_normals(vm)
        el += time.perf_counter() - t0
        n += 1
    return {"value": round(n / el, 4), "unit": "frames/s", "cores": threads, "kind": "port",
            "nproc": os.cpu_count(), "affinity_cpus": avail, "cpu_model": _cpu_model(),
            "sample": f"oracle fuse+raycast+normals of frames 0-{n - 1} of {w.name}, frame i at pose i "
                      f"({el:.1f} s of work; {threads} OpenMP threads for integration, visibility and "
                      f"raycast, serial allocation; g++ -O2)"}


def batched_bench(warmup: int, steps: int) -> dict:
    """Plain per-scene loop vs one batched step (autotuned, hipGraph) over the
    same frames of the 16-scene workload; aggregate frames/s of both."""
    import torch
    from voxelmapping_amd import _capi, synth
    from voxelmapping_amd.mapping import Batch, DeviceMap, intrinsics, pose_colmajor
    L = _capi.lib()
    ws = synth.batched16()
    nfr = warmup + steps
    dev = torch.device("cuda:0")
    maps, specs, plain_args = [], [], []
    for w in ws:
        d = np.stack([w.frame(i, want_rgb=False)[0] for i in range(nfr)])
        dt = torch.from_numpy(d.view(np.int16)).to(dev).view(torch.uint16)
        m = DeviceMap(w.n_entry, w.n_bucket, w.n_block, w.voxel_size, w.trunc, w.voxel_format)
        maps.append(m)
        specs.append(dict(map=m, K=w.K, depth=dt, depth_scale=w.depth_scale, poses=w.poses(nfr)))
        plain_args.append((w, dt, [pose_colmajor(w.poses(nfr)[i]) for i in range(nfr)], intrinsics(w.K)))
    b = Batch(specs)
    stream = torch.cuda.Stream()
    sp = C.c_void_p(stream.cuda_stream)

    def plain_step(f):
        for (w, dt, poses, K), m, vm, nm in zip(plain_args, maps, b.vmaps, b.nmaps):
            W, H = w.width, w.height
            _capi.check(L.vm_fuse_depth(m.handle, C.c_void_p(dt[f].data_ptr()), W * 2, W, H, 1,
                                        C.c_float(w.depth_scale), None, 0, K, poses[f], sp, None))
            _capi.check(L.vm_render_scene(m.handle, C.c_void_p(vm.data_ptr()), W * 16, W, H, K, poses[f], sp))
            _capi.check(L.vm_compute_normals(C.c_void_p(vm.data_ptr()), W * 16, C.c_void_p(nm.data_ptr()), W * 16,
                                             W, H, sp))

    def sums():
        return sum(m.stats(stream)["sum_visible"] for m in maps)

    # plain loop
    with torch.cuda.stream(stream):
        for f in range(warmup):
            plain_step(f)
        stream.synchronize()
        v0 = sums()
        t0 = time.perf_counter()
        for f in range(warmup, nfr):
            plain_step(f)
        stream.synchronize()
        t_plain = time.perf_counter() - t0
        v_plain = sums() - v0
        # autotune over the first frames (18 launch shapes, 3 scene-group
        # counts and the one-kernel step, 3 steps each), then the batched
        # steps over the same frames as the plain loop, from empty maps
        tune_steps = 3
        if 22 * tune_steps > nfr:
            raise SystemExit("the autotuner needs 66 frames")
        for m in maps:
            m.reset(stream)
        b.set_frame(0, stream)
        best = b.tune(steps=tune_steps, stream=stream)
        for m in maps:
            m.reset(stream)
        b.set_frame(0, stream)
        for _ in range(warmup):
            b.step(stream)
        b.capture(stream)
        stream.synchronize()
        v0 = sums()
        t0 = time.perf_counter()
        for _ in range(steps):
            b.step(stream)
        stream.synchronize()
        t_batch = time.perf_counter() - t0
        v_batch = sums() - v0
        # the same frames again with the step as one kernel (grid barriers
        # between the stages), graph-replayed: reported beside the tuned step
        for m in maps:
            m.reset(stream)
        b.set_frame(0, stream)
        b.set_tuning(**{**best, "scene_groups": 1, "persistent": 0})
        for _ in range(warmup):
            b.step(stream)
        b.set_tuning(**{**best, "scene_groups": 1, "persistent": 1, "persist_wg_per_cu": 4})
        b.capture(stream)
        stream.synchronize()
        v0 = sums()
        t0 = time.perf_counter()
        for _ in range(steps):
            b.step(stream)
        stream.synchronize()
        t_one = time.perf_counter() - t0
        v_one = sums() - v0
        b.frame(stream)  # raises if a grid-barrier wait ran out
    n = len(ws)
    P_total = sum(w.width * w.height for w in ws)
    bytes_frame_pixels = P_total * (B_DEPTH + B_VMAP + B_NMAP)  # no colour in this workload
    per_vis = 512 * (2 * S_VOXEL_NO_COLOR + S_RAY) + S_ENTRY
    batch_bytes = steps * bytes_frame_pixels + v_batch * per_vis
    return {
        "workload": "16 scenes, 160x120..640x480, 4/6/10 mm, 2^16..2^20 buckets, uint16 depth, float TSDF",
        "scenes": [w.name for w in ws],
        "frames": [warmup, nfr],
        "value": round(n * steps / t_batch, 2), "unit": "frames/s (all scenes)",
        "ms_per_step": round(1000 * t_batch / steps, 4),
        "plain_loop_value": round(n * steps / t_plain, 2),
        "plain_loop_ms_per_step": round(1000 * t_plain / steps, 4),
        "speedup_vs_plain_loop": round(t_plain / t_batch, 3),
        "hbm_frac": round(batch_bytes / t_batch / 1e9 / HBM_PEAK_GBS, 4),
        "visible_blocks_equal": bool(v_plain == v_batch),
        "tuning": best, "graph": True,
        "one_kernel_step": {"ms_per_step": round(1000 * t_one / steps, 4),
                            "over_tuned_step": round(t_one / t_batch, 3),
                            "visible_blocks_equal": bool(v_one == v_batch),
                            "persist_wg_per_cu": 4},
    }


def pose_bench(m, w, depth, T, n_hyp: int = 32, iters: int = 8, stride: int = 4, reps: int = 3) -> dict:
    """Batched pose refinement on the fused room_vga map: n_hyp hypotheses
    around the frame's pose (U[-2, 2] cm, U[-1.5, 1.5] deg), iters
    Gauss-Newton iterations over every stride-th pixel of the 640x480 depth."""
    from voxelmapping_amd.autograd import se3_exp
    rng = np.random.default_rng(7)
    T0 = np.asarray(T, np.float64)
    hyps = np.stack([T0 @ se3_exp(np.r_[rng.uniform(-1, 1, 3) * 0.02, np.radians(rng.uniform(-1, 1, 3) * 1.5)])
                     for _ in range(n_hyp)]).astype(np.float32)
    m.refine_poses(depth, w.K, hyps, iters=iters, stride=stride, depth_scale=w.depth_scale)
    t0 = time.perf_counter()
    for _ in range(reps):
        poses, cost = m.refine_poses(depth, w.K, hyps, iters=iters, stride=stride, depth_scale=w.depth_scale)
    dt = (time.perf_counter() - t0) / reps
    rays = n_hyp * (iters + 1) * ((w.width + stride - 1) // stride) * ((w.height + stride - 1) // stride)
    best = int(np.argmin(cost[:, -1, 0] + (cost[:, -1, 0] < 0) * 1e9))
    dT = np.linalg.inv(T0) @ poses[best].astype(np.float64)
    return {"workload": f"{n_hyp} hypotheses x {iters} Gauss-Newton iterations, 640x480 depth every {stride}th pixel, "
                        "room_vga map", "ms_per_refinement": round(1000 * dt, 3),
            "rays_per_s": round(rays / dt, 1), "launches_per_iteration": 3,
            "best_hypothesis_error_mm": round(1000 * float(np.linalg.norm(dT[:3, 3])), 2)}


def scale_bench(frames: int = 40, warmup: int = 4, track_every: int = 10, n_hyp: int = 32, iters: int = 8,
                stride: int = 4, track_max_depth: float = 2.0, **synth_kw) -> dict:
    """The large-room workload at full size: 1280x720 uint16 depth fused at
    2 mm into 2^23 buckets / 2^22 float-TSDF blocks (24 GiB of voxels), fuse +
    raycast + normals per frame. Every track_every-th frame, before it is
    fused, n_hyp pose hypotheses around its true pose (U[-2, 2] cm,
    U[-1.5, 1.5] deg) are refined against its depth by iters Gauss-Newton
    iterations over every stride-th pixel (one Gauss-Newton launch per
    iteration for all hypotheses); refinement time is reported apart from the fusion step.
    Tracking uses the depth below track_max_depth metres (0: all): the
    workload's noise, 0.0012 z^2 m, stays below half its 1 cm truncation up to
    2 m, and beyond that the fused surface is biased by one-sided averaging.
    synth_kw overrides workload fields (e.g. noise=0 for the geometry
    without depth noise)."""
    import torch
    from voxelmapping_amd import computeNormal, synth
    from voxelmapping_amd.autograd import se3_exp
    from voxelmapping_amd.mapping import DeviceMap
    w = synth.room_large(**synth_kw)
    m = DeviceMap(w.n_entry, w.n_bucket, w.n_block, w.voxel_size, w.trunc, w.voxel_format, 0.1, 5.0)
    n = warmup + frames
    deps = [torch.from_numpy(w.frame(i, want_rgb=False)[0].view(np.int16)).cuda().view(torch.uint16)
            for i in range(n)]
    poses = w.poses(n)
    vm = torch.zeros((w.height, w.width, 4), dtype=torch.float32, device="cuda")
    nm = torch.zeros_like(vm)
    rng = np.random.default_rng(11)
    refine_s, errs, errs0 = [], [], []

    def step(i):
        m.fuse(deps[i], w.K, poses[i], depth_scale=w.depth_scale)
        m.render_scene(vm, w.K, poses[i])
        computeNormal(vm, nm)

    def track(i):
        T0 = poses[i].reshape(4, 4).T.astype(np.float64)
        hyps = np.stack([T0 @ se3_exp(np.r_[rng.uniform(-1, 1, 3) * 0.02, np.radians(rng.uniform(-1, 1, 3) * 1.5)])
                         for _ in range(n_hyp)]).astype(np.float32)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        out, cost = m.refine_poses(deps[i], w.K, hyps, iters=iters, stride=stride, depth_scale=w.depth_scale,
                                   max_depth=track_max_depth)
        refine_s.append(time.perf_counter() - t1)
        best = int(np.argmin(cost[:, -1, 0] + (cost[:, -1, 0] < 0) * 1e9))
        errs.append(float(np.linalg.norm((np.linalg.inv(T0) @ out[best].astype(np.float64))[:3, 3])))
        errs0.append(float(np.linalg.norm((np.linalg.inv(T0) @ hyps[best].astype(np.float64))[:3, 3])))

    for i in range(warmup):
        step(i)
    torch.cuda.synchronize()
    v0 = m.stats()["sum_visible"]
    t0 = time.perf_counter()
    for i in range(warmup, n):
        if i % track_every == 0:
            track(i)
        step(i)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0 - sum(refine_s)
    st = m.stats()
    m.release()
    # algorithmic bytes of the fused frames (no colour): depth, vertex and
    # normal maps per pixel, TSDF + weight read and written and raycast reads
    # per voxel of every visible block
    vis = st["sum_visible"] - v0
    frame_bytes_total = frames * w.width * w.height * (B_DEPTH + B_VMAP + B_NMAP) + vis * (
        512 * (2 * S_VOXEL_NO_COLOR + S_RAY) + S_ENTRY)
    return {"workload": w.name, "frames": [warmup, n], "value": round(frames / dt, 2), "unit": "frames/s",
            "ms_per_step": round(1000 * dt / frames, 4), "mean_visible_blocks": round(vis / frames, 1),
            "frame_hbm_frac": round(frame_bytes_total / dt / 1e9 / HBM_PEAK_GBS, 4),
            "allocated_blocks": st["allocated_blocks"],
            "status": st["status"], "voxel_pool_gib": round(w.n_block * 512 * 12 / 2**30, 1),
            "tracking": {"every": track_every, "hypotheses": n_hyp, "iterations": iters,
                         "pixels": ((w.width + stride - 1) // stride) * ((w.height + stride - 1) // stride),
                         "refinements": len(refine_s),
                         "ms_per_refinement": round(1000 * float(np.mean(refine_s)), 3) if refine_s else None,
                         "max_depth_m": track_max_depth, "frames_tracked": [i for i in range(warmup, n)
                                                                            if i % track_every == 0],
                         "best_hypothesis_error_mm": [round(1000 * e, 2) for e in errs],
                         "best_hypothesis_start_error_mm": [round(1000 * e, 2) for e in errs0]}}


def mesh_bench(m, reps: int = 5, cap: int = 20_000_000) -> dict:
    """Marching cubes over every allocated block of the fused room_vga map
    (vm_polygonize): mean time of reps calls, triangle count, and the
    algorithmic bytes (8 B of TSDF + weight per voxel of each allocated
    block read once, 72 B of vertices + face normals per triangle written)
    over that time."""
    import torch
    verts = torch.empty((cap, 9), dtype=torch.float32, device="cuda")
    norms = torch.empty_like(verts)
    m.polygonize(verts, norms)
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    torch.cuda.synchronize()
    torch.cuda._sleep(HOLD_CYCLES)  # the calls are queued before the GPU reaches them
    ev[0].record()
    for _ in range(reps):
        _, tris = m.polygonize(verts, norms)
    ev[1].record()
    torch.cuda.synchronize()
    ms = ev[0].elapsed_time(ev[1]) / reps
    A = m.stats()["allocated_blocks"]
    byts = A * 512 * 8 + tris * 72
    return {"workload": "room_vga map after the bench frames", "allocated_blocks": A, "triangles": tris,
            "ms": round(ms, 3), "achieved_gbs": round(byts / (ms * 1e-3) / 1e9, 1),
            "hbm_frac": round(byts / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}


def adjoint_bench(reps: int = 20) -> dict:
    """Raycast forward vs its adjoint: 640x480 rays (fx = fy = 525) from a
    pose perturbed by (1 cm, 0.5 deg) into the fused room_small volume (50
    frames, 1 cm voxels, float TSDF), upstream gradients U[-1, 1] on every
    pixel's depth and normal."""
    import torch
    from voxelmapping_amd import computeNormal, synth
    from voxelmapping_amd.autograd import se3_exp
    from voxelmapping_amd.mapping import DeviceMap
    w = synth.room_small(voxel_format=2)
    m = DeviceMap(w.n_entry, w.n_bucket, w.n_block, w.voxel_size, w.trunc, 2)
    for i in range(w.n_frames):
        d, _ = w.frame(i, want_rgb=False)
        m.fuse(torch.from_numpy(d).cuda(), w.K, w.poses(i + 1)[i])
    rng = np.random.default_rng(3)
    T = (w.poses()[-1].reshape(4, 4).T.astype(np.float64) @
         se3_exp(np.r_[rng.normal(0, 0.01, 3), rng.normal(0, np.radians(0.5), 3)])).astype(np.float32)
    K = np.array([525.0, 525.0, 319.5, 239.5], np.float32)
    H, W = 480, 640
    vm = torch.zeros((H, W, 4), dtype=torch.float32, device="cuda")
    nm = torch.zeros_like(vm)
    gv = torch.zeros_like(vm)
    gv[..., 2] = torch.rand((H, W), device="cuda") * 2 - 1
    gn = torch.rand((H, W, 4), device="cuda") * 2 - 1
    gt = torch.zeros(w.n_block * 512, dtype=torch.float32, device="cuda")
    gp = torch.zeros(6, dtype=torch.float64, device="cuda")
    m.raycast(vm, K, T)
    st = m.raycast_backward(vm, K, T, gt, gp, grad_vertex=gv, grad_normal=gn)
    # each pass queued reps times back to back behind a spin kernel that
    # holds the GPU until all of them are queued, so the clock sees GPU time
    # rather than the launches' host latency
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
    fwd = bwd = 0.0
    for _ in range(3):
        torch.cuda.synchronize()
        torch.cuda._sleep(HOLD_CYCLES)
        ev[0].record()
        for _ in range(reps):
            m.raycast(vm, K, T)
            computeNormal(vm, nm)
        ev[1].record()
        for _ in range(reps):
            m.raycast_backward(vm, K, T, gt, gp, grad_vertex=gv, grad_normal=gn, stats=False)
        ev[2].record()
        torch.cuda.synchronize()
        fwd += ev[0].elapsed_time(ev[1])
        bwd += ev[1].elapsed_time(ev[2])
    reps *= 3
    return {"workload": "640x480 rays into room_small (1 cm, 50 frames), perturbed pose",
            "forward_us": round(1000 * fwd / reps, 2), "backward_us": round(1000 * bwd / reps, 2),
            "backward_over_forward": round(bwd / fwd, 3), "hit_pixels": st["hit_pixels"],
            "overflow_pixels": st["overflow_pixels"]}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=480)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-batched", action="store_true", help="skip the 16-scene batched comparison")
    ap.add_argument("--batched-warmup", type=int, default=24)
    ap.add_argument("--batched-steps", type=int, default=180)
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a 256 MiB device-to-device copy 5 times (scales rocprofv3 byte counters)")
    args = ap.parse_args()
    if args.calibrate_copy:
        import torch
        n = 64 << 20  # floats
        x = torch.ones(n, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        for _ in range(5):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"calibration": "copy", "bytes_read": 4 * n, "bytes_written": 4 * n, "launches": 5}))
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              "(scenes shard per GPU with no exchange; see the design notes)", file=sys.stderr)
        sys.exit(2)

    import torch
    from voxelmapping_amd import _capi, synth
    from voxelmapping_amd.mapping import DeviceMap, intrinsics, pose_colmajor

    L = _capi.lib()
    dev = torch.device("cuda:0")
    nframes = args.warmup + args.steps
    w = synth.room_vga(n_frames=max(500, nframes))
    W, H = w.width, w.height
    P = W * H

    # inputs: generated on the host, resident in HBM before timing
    depth_d, rgb_d = [], []
    for i in range(nframes):
        d, c = w.frame(i)
        depth_d.append(torch.from_numpy(d.view(np.int16)).to(dev))
        rgb_d.append(torch.from_numpy(c).to(dev))
    poses = [pose_colmajor(w.poses(nframes)[i]) for i in range(nframes)]
    K = intrinsics(w.K)
    vmap = torch.zeros((H, W, 4), dtype=torch.float32, device=dev)
    nmap = torch.zeros_like(vmap)
    m = DeviceMap(w.n_entry, w.n_bucket, w.n_block, w.voxel_size, w.trunc, w.voxel_format)
    h = m.handle
    stream = torch.cuda.current_stream()
    sp = C.c_void_p(stream.cuda_stream)
    dpitch, rpitch, vpitch = W * 2, W * 3, W * 16
    scale = C.c_float(w.depth_scale)
    vptr, nptr = C.c_void_p(vmap.data_ptr()), C.c_void_p(nmap.data_ptr())

    def step(i, ev=None, all_stages=False):
        # ev: HIP events on the launch stream; the raycast stage (the
        # dominant one, whose roofline is reported) is bracketed on every
        # timed step, the other stages on every STAGE_EVERY-th step only
        dp, cp = C.c_void_p(depth_d[i].data_ptr()), C.c_void_p(rgb_d[i].data_ptr())
        T = poses[i]
        if ev and all_stages: ev[0].record(stream)
        _capi.check(L.vm_allocate_blocks(h, dp, dpitch, W, H, 1, scale, K, T, sp))
        if ev and all_stages: ev[1].record(stream)
        _capi.check(L.vm_integrate(h, dp, dpitch, W, H, 1, scale, cp, rpitch, K, T, sp, None))
        if ev: ev[2].record(stream)
        _capi.check(L.vm_render_scene(h, vptr, vpitch, W, H, K, T, sp))
        if ev: ev[3].record(stream)
        _capi.check(L.vm_compute_normals(vptr, vpitch, nptr, vpitch, W, H, sp))
        if ev and all_stages: ev[4].record(stream)

    for i in range(args.warmup):
        step(i)
    torch.cuda.synchronize()
    s0 = m.stats()

    names = ["allocate", "integrate", "raycast", "normals"]
    events = [[torch.cuda.Event(enable_timing=True) for _ in range(5)] for _ in range(args.steps)]
    sampled = [k for k in range(args.steps) if k % STAGE_EVERY == 0]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        step(args.warmup + k, events[k], k % STAGE_EVERY == 0)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    s1 = m.stats()

    stage_avg_us = {}
    for j, n in enumerate(names):
        ks = range(args.steps) if n == "raycast" else sampled
        stage_avg_us[n] = 1000.0 * sum(events[k][j].elapsed_time(events[k][j + 1]) for k in ks) / len(ks)
    sum_vis = s1["sum_visible"] - s0["sum_visible"]
    mean_vis = sum_vis / args.steps
    total_bytes = args.steps * P * (B_DEPTH + B_COLOR + B_VMAP + B_NMAP) + sum_vis * (
        512 * (2 * S_VOXEL + S_RAY) + S_ENTRY)
    fps = args.steps / wall
    per_launch = stage_bytes(P, mean_vis)
    dom = max(names, key=lambda n: stage_avg_us[n])
    ach = per_launch[dom] / (stage_avg_us[dom] * 1e-6) / 1e9
    traffic, l2_hit, traffic_src = profiled_traffic(dom, w.name)
    roofline = {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": traffic, "kernel": dom,
                "bytes_per_launch": int(per_launch[dom]), "avg_launch_us": round(stage_avg_us[dom], 2)}
    if traffic is not None:
        roofline["traffic_source"] = traffic_src
    if l2_hit is not None:
        roofline["l2_hit_rate"] = l2_hit
    if dom == "raycast":
        # the raycast stage is one launch, the march: its per-tile
        # ray-distance bounds were computed by the fusion's visibility launch
        # at the same pose (k_visible_bounds, in the integrate stage)
        roofline["stage_kernels"] = ["k_raycast<2>"]
    out = {
        "metric": METRIC, "value": round(fps, 2), "unit": "frames/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(1000.0 * wall / args.steps, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": w.name, "form": "plain", "width": W, "height": H, "voxel_size_m": w.voxel_size,
                   "trunc_m": w.trunc, "n_bucket": w.n_bucket, "n_excess": w.n_excess, "n_block": w.n_block,
                   "voxel_format": "f32_tsdf+f32_weight+rgb8", "depth": "uint16 mm", "colour": True,
                   "frames": [args.warmup, args.warmup + args.steps]},
        "roofline": roofline,
        "frame_hbm_frac": round(total_bytes / wall / 1e9 / HBM_PEAK_GBS, 4),
        "mean_visible_blocks": round(mean_vis, 1),
        "allocated_blocks": s1["allocated_blocks"],
        "stage_avg_us": {n: round(v, 2) for n, v in stage_avg_us.items()},
        "stage_events": f"raycast on every timed step, the other stages on every {STAGE_EVERY}th",
        "status": s1["status"],
        "mesh": mesh_bench(m),
    }
    if not args.no_batched:
        out["batched"] = batched_bench(args.batched_warmup, args.batched_steps)
        out["adjoint"] = adjoint_bench()
        last = args.warmup + args.steps - 1
        out["pose_refine"] = pose_bench(m, w, depth_d[last].view(torch.uint16), w.poses(nframes)[last].reshape(4, 4).T)
        out["scale"] = scale_bench()
        out["scale"]["tracking_without_depth_noise"] = scale_bench(noise=0.0, dropout=0.0,
                                                                   track_max_depth=0.0)["tracking"]
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(synth.room_vga(), args.cpu_seconds)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
