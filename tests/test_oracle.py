"""Pins of the CPU oracle (oracle/vm_oracle.cpp) against facts that do not
come from the oracle itself: hand-computed known answers, brute force over
integer ranges, closed forms of simple scenes, numpy's matrix inverse, the
marching-cubes table invariants and the reference's table text.

Pinned here, per oracle function:
  hash3                 hand-computed KATs + range invariant
  voxel_to_block/local  brute force vs Python floor division / modulo
  f2i_rd                special values
  pose_inverse          numpy.linalg.inv (float64) and rigid closed form
  MC tables             820 triangles, edge masks, reference text (if present)
  allocate_pixels +
  create_block          hand-derived block sets (1-pixel images), capacity
                        exhaustion, and the excess-slot alias of the reference
  check_visible +
  integrate             fronto-parallel plane closed form, n-frame weights,
                        int16 packing
  raycast               plane closed form (hit depth <= 1e-4 m)
  compute_normals       plane closed form (normal (0, 0, 1))
  polygonize            all 256 single-cube cases vs the table, sphere mesh
                        (vertex radius, closed surface, consistent winding)
  surface_points        centroid closed form
  project_rendering_blocks  two blocks whose corners project to exact
                        integers (power-of-two voxel size and focal length):
                        rectangles, clamping at the border, min / max merge
  render_shading        special geometries of the Phong terms (light and eye
                        along the normal, normal perpendicular to or facing
                        away from the light) and the background colour
"""
import math
import os
import re

import numpy as np
import pytest

import oracle as O

REF_TABLE = "/root/reference/src/triangle_table.h"


def colmajor(T):
    return np.asarray(T, np.float32).T.reshape(-1)


def c_int32(v):
    v &= 0xFFFFFFFF
    return v - (1 << 32) if v >= (1 << 31) else v


# --- hash ----------------------------------------------------------------------

def test_hash_known_answers():
    n = 1 << 20
    assert O.hash_block(1, 0, 0, n) == 455773        # 73856093 mod 2^20
    assert O.hash_block(0, 1, 0, n) == 475301        # 19349669 mod 2^20
    assert O.hash_block(0, 0, 1, n) == 655287        # 83492791 mod 2^20
    assert O.hash_block(-1, 0, 0, n) == 592803       # 2^20 - 455773
    assert O.hash_block(100, 0, 0, n) == 488532      # 7385609300 wraps to -1204325292
    assert O.hash_block(0, 0, 0, 943718) == 0
    # non power of two: 73856093 = 78 * 943718 + 246089
    assert O.hash_block(1, 0, 0, 943718) == 246089


def test_hash_range_and_wrap():
    rng = np.random.default_rng(0)
    for n in (1, 7, 1 << 17, 943718, (1 << 20) + 3):
        for _ in range(200):
            x, y, z = (int(v) for v in rng.integers(-(1 << 28), 1 << 28, 3))
            h = O.hash_block(x, y, z, n)
            assert 0 <= h < n
            r = c_int32(x * 73856093) ^ c_int32(y * 19349669) ^ c_int32(z * 83492791)
            assert h == r - n * int(r / n) + (n if r - n * int(r / n) < 0 else 0)


# --- coordinates -----------------------------------------------------------------

def test_voxel_block_local_bruteforce():
    for v in range(-64, 65):
        b = O.voxel_to_block((v, -v, v // 3))
        assert list(b) == [v // 8, (-v) // 8, (v // 3) // 8]
        l = O.voxel_to_local((v, -v, v // 3))
        assert list(l) == [v % 8, (-v) % 8, (v // 3) % 8]


def test_float2int_rd_special_values():
    f = O.lib().or_float2int_rd
    assert f(-0.5) == -1 and f(2.999) == 2 and f(-3.0) == -3 and f(0.0) == 0
    assert f(float("nan")) == 0
    assert f(1e10) == 2**31 - 1 and f(-1e10) == -(2**31)


# --- pose inverse ----------------------------------------------------------------

def random_pose(rng, scale=3.0):
    a = rng.normal(size=3)
    th = rng.uniform(0, math.pi)
    k = a / np.linalg.norm(a)
    Kx = np.array([[0, -k[2], k[1]], [k[2], 0, -k[0]], [-k[1], k[0], 0]])
    R = np.eye(3) + math.sin(th) * Kx + (1 - math.cos(th)) * Kx @ Kx
    T = np.eye(4)
    T[:3, :3] = R
    T[:3, 3] = rng.uniform(-scale, scale, 3)
    return T.astype(np.float32)


def test_pose_inverse_matches_numpy():
    rng = np.random.default_rng(1)
    for _ in range(50):
        T = random_pose(rng)
        inv = O.pose_inverse(colmajor(T)).reshape(4, 4).T
        ref = np.linalg.inv(T.astype(np.float64))
        assert np.allclose(inv, ref, atol=2e-7, rtol=1e-6)
        # rigid closed form [R^T | -R^T t]
        R, t = T[:3, :3].astype(np.float64), T[:3, 3].astype(np.float64)
        assert np.allclose(inv[:3, :3], R.T, atol=1e-6)
        assert np.allclose(inv[:3, 3], -R.T @ t, atol=1e-5)


# --- marching-cubes tables -------------------------------------------------------

CORNERS = [(0, 1), (1, 2), (2, 3), (3, 0), (4, 5), (5, 6), (6, 7), (7, 4), (0, 4), (1, 5), (2, 6), (3, 7)]


def test_mc_table_invariants():
    edge, tri = O.mc_tables()
    counts = [(row != -1).sum() for row in tri]
    assert sum(counts) == 820 * 3
    assert max(counts) == 15 and all(c % 3 == 0 for c in counts)
    for c in range(256):
        assert edge[c] == edge[255 - c]
        used = 0
        for e in tri[c][tri[c] != -1]:
            used |= 1 << int(e)
        assert used == edge[c], c
        cut = sum(1 << k for k, (a, b) in enumerate(CORNERS) if ((c >> a) & 1) != ((c >> b) & 1))
        assert cut == edge[c]


@pytest.mark.skipif(not os.path.exists(REF_TABLE), reason="reference tree not mounted")
def test_mc_table_equals_reference_text():
    edge, tri = O.mc_tables()
    text = open(REF_TABLE).read()
    etext = text[text.index("edgeTable"):text.index("triTable")]
    ref_edge = [int(v, 16) for v in re.findall(r"0x[0-9a-fA-F]+", etext)]
    assert ref_edge == edge.tolist()
    rows = re.findall(r"\{([-0-9, ]+)\}", text[text.index("triTable"):])
    ref_tri = [[int(v) for v in r.split(",")] for r in rows]
    assert ref_tri == tri.tolist()


# --- allocation: hand-derived block sets -------------------------------------------

K1 = np.array([1.0, 1.0, 0.0, 0.0], np.float32)  # 1x1 image, pixel (0,0) on the optical axis


def one_pixel_map(n_bucket=64, n_excess=16, n_block=64, fmt=0):
    return O.OracleMap(n_bucket + n_excess, n_bucket, n_block, 0.01, 0.04, fmt)


def allocated(m):
    return set(O.allocated_blocks(m.hash_table()).keys())


def test_allocation_single_block():
    m = one_pixel_map()
    # band [0.983, 1.023] m -> voxels z 98..102 -> block 12 (voxels 96..103)
    m.allocate(np.array([[1.003]], np.float32), K1, colmajor(np.eye(4)))
    assert allocated(m) == {(0, 0, 12)}


def test_allocation_crosses_block_boundary():
    m = one_pixel_map()
    # band [1.021, 1.061] -> voxels 102..106 -> blocks 12 and 13
    m.allocate(np.array([[1.041]], np.float32), K1, colmajor(np.eye(4)))
    assert allocated(m) == {(0, 0, 12), (0, 0, 13)}
    c = m.counters()
    assert c["heap_ptr"] == 64 - 1 - 2 and c["alias_events"] == 0
    # first pop is block 0 (heap[i] = n - 1 - i), and the nearer block is traced first
    tab = m.hash_table()
    ptr = {tuple(e[2:]): e[0] for e in tab if e[0] >= 0}
    assert ptr[(0, 0, 12)] == 0 and ptr[(0, 0, 13)] == 512


def test_allocation_half_truncation_band():
    """The traced segment is depth -/+ HALF the truncation (depth_fusion.cu:
    43-44, truncation * 0.5 at :226): at 0.9851 m the band [0.9651, 1.0051]
    starts at voxel 96 (block 12); the full truncation would start it at
    voxel 94 and add block 11."""
    m = one_pixel_map()
    m.allocate(np.array([[0.9851]], np.float32), K1, colmajor(np.eye(4)))
    assert allocated(m) == {(0, 0, 12)}


def test_allocation_negative_coordinates():
    m = one_pixel_map()
    flip = np.diag([1.0, -1.0, -1.0, 1.0])  # camera looks down world -z
    m.allocate(np.array([[1.041]], np.float32), K1, colmajor(flip))
    # voxels -103 .. -107 (samples truncate toward zero) -> blocks -13, -14
    assert allocated(m) == {(0, 0, -13), (0, 0, -14)}


def test_allocation_invalid_depths_allocate_nothing():
    m = one_pixel_map()
    for d in (0.0, float("nan"), 0.05, 5.5):
        m.allocate(np.array([[d]], np.float32), K1, colmajor(np.eye(4)))
    assert allocated(m) == set()


def test_allocation_heap_exhaustion():
    m = one_pixel_map(n_block=1)
    m.allocate(np.array([[1.041]], np.float32), K1, colmajor(np.eye(4)))
    assert allocated(m) == {(0, 0, 12)}
    c = m.counters()
    assert c["heap_ptr"] == -1 and c["alloc_failures"] >= 1


def test_excess_alias_quirk_of_reference():
    # one bucket: the second key's excess offset 0 maps to slot nBucket-1 = 0,
    # i.e. onto the occupied head (map_struct.h:344); the reference overwrites it
    m = O.OracleMap(3, 1, 8, 0.01, 0.04, 0)
    m.allocate(np.array([[1.041]], np.float32), K1, colmajor(np.eye(4)))
    c = m.counters()
    assert c["alias_events"] == 1
    tab = m.hash_table()
    assert tuple(tab[0][2:]) == (0, 0, 13) and tab[0][1] == 0  # self-linked head
    assert allocated(m) == {(0, 0, 13)}


# --- integration: plane closed form ----------------------------------------------------

W, H = 32, 24
KP = np.array([30.0, 30.0, 15.5, 11.5], np.float32)
D = 1.0037
VS, TR = 0.01, 0.04


def plane_map(fmt=2, frames=1):
    m = O.OracleMap(4096 + 1024, 4096, 2048, VS, TR, fmt)
    depth = np.full((H, W), D, np.float32)
    for _ in range(frames):
        m.fuse(depth, KP, colmajor(np.eye(4)))
    return m


def voxel_coords(table, n_block):
    """(vx, vy, vz, voxel index) of every voxel of every allocated block."""
    out = []
    lz, ly, lx = np.meshgrid(np.arange(8), np.arange(8), np.arange(8), indexing="ij")
    local = (lz * 64 + ly * 8 + lx).reshape(-1)
    for e in table[table[:, 0] >= 0]:
        ptr, bx, by, bz = e[0], e[2], e[3], e[4]
        out.append(np.stack([bx * 8 + lx.reshape(-1), by * 8 + ly.reshape(-1), bz * 8 + lz.reshape(-1),
                             ptr + local], 1))
    return np.concatenate(out)


def expected_plane(vc):
    x, y, z = (vc[:, i].astype(np.float64) * VS for i in range(3))
    u = KP[0] * x / z + KP[2] + 0.5
    v = KP[1] * y / z + KP[3] + 0.5
    inside = (u >= 0) & (v >= 0) & (u < W) & (v < H)
    border = (np.abs(u - np.round(u)) < 1e-3) & ((np.round(u) == 0) | (np.round(u) == W))
    border |= (np.abs(v - np.round(v)) < 1e-3) & ((np.round(v) == 0) | (np.round(v) == H))
    sd = D - z
    upd = inside & (sd >= -TR)
    near_edge = border | (np.abs(sd + TR) < 1e-6)
    return upd, np.minimum(1.0, sd / TR), near_edge


def test_integration_plane_first_frame():
    m = plane_map(fmt=2)
    tab = m.hash_table()
    vis = {tuple(e[2:]) for e in m.visible_table()}
    assert vis == set(O.allocated_blocks(tab).keys())  # every allocated block is in view
    vc = voxel_coords(tab, m.n_block)
    sdf, w, _ = O.decode_voxels(m.voxels(), 2)
    upd, exp, edge = expected_plane(vc)
    s, ww = sdf[vc[:, 3]], w[vc[:, 3]]
    ok = ~edge
    assert np.array_equal(ww[ok] == 1.0, upd[ok])
    assert np.all(ww[ok & ~upd] == 0)
    assert np.max(np.abs(s[ok & upd] - exp[ok & upd])) < 2e-6
    assert upd.sum() > 1000


def test_integration_plane_weights_accumulate():
    m = plane_map(fmt=2, frames=7)
    tab = m.hash_table()
    vc = voxel_coords(tab, m.n_block)
    sdf, w, _ = O.decode_voxels(m.voxels(), 2)
    upd, exp, edge = expected_plane(vc)
    ok = ~edge & upd
    assert np.all(w[vc[ok, 3]] == 7.0)
    assert np.max(np.abs(sdf[vc[ok, 3]] - exp[ok])) < 4e-6


def test_integration_plane_int16_packing():
    m = plane_map(fmt=0, frames=3)
    tab = m.hash_table()
    vc = voxel_coords(tab, m.n_block)
    sdf, w, _ = O.decode_voxels(m.voxels(), 0)
    upd, exp, edge = expected_plane(vc)
    ok = ~edge & upd
    assert np.all(w[vc[ok, 3]] == 3)
    # packFloat truncates: stored = trunc(v * 32767) / 32767, repeated 3 times
    assert np.max(np.abs(sdf[vc[ok, 3]] - exp[ok])) <= 3.0 / 32767 + 1e-6


def test_weight_saturates_at_255():
    m = O.OracleMap(4096 + 1024, 4096, 2048, VS, TR, 0)
    depth = np.full((4, 4), D, np.float32)
    k = np.array([3.0, 3.0, 1.5, 1.5], np.float32)
    for _ in range(260):
        m.fuse(depth, k, colmajor(np.eye(4)))
    _, w, _ = O.decode_voxels(m.voxels(), 0)
    assert w.max() == 255


def test_colour_running_average():
    m = O.OracleMap(4096 + 1024, 4096, 2048, VS, TR, 1)
    depth = np.full((H, W), D, np.float32)
    c1 = np.full((H, W, 3), 200, np.uint8)
    c2 = np.full((H, W, 3), 100, np.uint8)
    m.fuse(depth, KP, colmajor(np.eye(4)), rgb=c1)
    m.fuse(depth, KP, colmajor(np.eye(4)), rgb=c2)
    _, w, rgb = O.decode_voxels(m.voxels(), 1)
    sel = w == 2
    assert sel.sum() > 0
    # first observation copies, then u8(0.7*old + 0.3*new) = u8(170.0)
    exp = int(np.float32(0.7) * np.float32(200) + np.float32(0.3) * np.float32(100))
    assert np.all(rgb[sel] == exp)


# --- raycast / normals: plane closed form ----------------------------------------------

def test_raycast_plane_depth():
    m = plane_map(fmt=2)
    v = m.raycast(W, H, KP, colmajor(np.eye(4)))
    ys, xs = np.mgrid[0:H, 0:W]
    dz = 1.0 / np.sqrt(((xs - KP[2]) / KP[0]) ** 2 + ((ys - KP[3]) / KP[1]) ** 2 + 1.0)
    sel = (dz > 0.985) & (xs > 0) & (ys > 0) & (xs < W - 1) & (ys < H - 1)
    assert sel.sum() > 50
    hit = ~np.isnan(v[..., 0])
    assert hit[sel].all()
    assert np.max(np.abs(v[..., 2][sel] - D)) < 1e-4
    assert np.all(v[..., 3][hit] == 1.0)


def test_normals_plane():
    ys, xs = np.mgrid[0:H, 0:W].astype(np.float32)
    z = np.float32(D)
    vmap = np.stack([(xs - KP[2]) / KP[0] * z, (ys - KP[3]) / KP[1] * z, np.full_like(xs, z),
                     np.ones_like(xs)], -1).astype(np.float32)
    n = O.compute_normals(vmap)
    inner = n[:H - 1, :W - 1]
    assert np.allclose(inner[..., :3], [0, 0, 1], atol=1e-6)
    assert np.all(inner[..., 3] == 1)
    assert np.all(n[H - 1] == 0) and np.all(n[:, W - 1] == 0)


def test_surface_points_centroid():
    v = np.arange(18, dtype=np.float32).reshape(2, 9)
    p = O.surface_points(v)
    assert np.allclose(p, v.reshape(2, 3, 3).mean(1))


# --- marching cubes ----------------------------------------------------------------

OFFS = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]
EDGE_MID = {k: tuple((np.array(OFFS[a]) + np.array(OFFS[b])) / 2.0) for k, (a, b) in enumerate(CORNERS)}


def set_voxel(raw, fmt, idx, sdf, w):
    if fmt == 2:
        rec = raw.reshape(-1, 12)
        rec[idx, 0:4] = np.frombuffer(np.float32(sdf).tobytes(), np.uint8)
        rec[idx, 4:8] = np.frombuffer(np.float32(w).tobytes(), np.uint8)
    else:
        rec = raw.reshape(-1, 4)
        rec[idx, 0:2] = np.frombuffer(np.int16(int(sdf * 32767)).tobytes(), np.uint8)
        rec[idx, 2] = w


def test_polygonize_single_cube_cases():
    _, tri = O.mc_tables()
    for case in range(256):
        m = O.OracleMap(64 + 16, 64, 8, 1.0, 4.0, 2)
        ptr = m.insert_block(0, 0, 0)
        raw = m.voxels_view()
        for i, (dx, dy, dz) in enumerate(OFFS):
            set_voxel(raw, 2, ptr + dz * 64 + dy * 8 + dx, -0.5 if (case >> i) & 1 else 0.5, 1.0)
        _, n, v, _ = m.polygonize(16)
        row = [e for e in tri[case] if e != -1]
        assert n == len(row) // 3, case
        for t in range(n):
            a, b, c = row[3 * t:3 * t + 3]
            got = v[t].reshape(3, 3)
            assert np.allclose(got, [EDGE_MID[c], EDGE_MID[b], EDGE_MID[a]]), case


def sphere_map(r=0.1, c=(0.3037, 0.2961, 0.3013), vs=0.01, tr=0.04):
    m = O.OracleMap(4096 + 1024, 4096, 512, vs, tr, 2)
    raw = m.voxels_view()
    lo = int(math.floor((c[0] - r - 2 * tr) / vs)) // 8
    hi = int(math.floor((c[0] + r + 2 * tr) / vs)) // 8
    lz, ly, lx = np.meshgrid(np.arange(8), np.arange(8), np.arange(8), indexing="ij")
    rec = raw.reshape(-1, 12)
    for bx in range(lo, hi + 1):
        for by in range(lo, hi + 1):
            for bz in range(lo, hi + 1):
                ptr = m.insert_block(bx, by, bz)
                p = np.stack([(bx * 8 + lx) * vs, (by * 8 + ly) * vs, (bz * 8 + lz) * vs], -1).reshape(-1, 3)
                d = np.sqrt(((p - np.array(c)) ** 2).sum(1)) - r
                s = np.clip(d / tr, -1, 1).astype(np.float32)
                idx = ptr + (lz * 64 + ly * 8 + lx).reshape(-1)
                rec[idx, 0:4] = s.view(np.uint8).reshape(-1, 4)
                rec[idx, 4:8] = np.frombuffer(np.float32(1).tobytes(), np.uint8)
    return m


def merged_vertex_ids(pts, tol=1e-6):
    """Vertex ids after merging points closer than tol (the same edge crossing
    computed from two neighbouring cubes can differ in the last bit)."""
    from scipy.spatial import cKDTree
    parent = np.arange(len(pts))

    def find(i):
        while parent[i] != i:
            parent[i] = parent[parent[i]]
            i = parent[i]
        return i

    for a, b in cKDTree(pts).query_pairs(tol):
        ra, rb = find(a), find(b)
        if ra != rb:
            parent[max(ra, rb)] = min(ra, rb)
    return np.array([find(i) for i in range(len(pts))])


def test_polygonize_sphere():
    c, r = np.array([0.3037, 0.2961, 0.3013]), 0.1
    m = sphere_map()
    blocks, n, v, nrm = m.polygonize(200000)
    assert n > 1000 and blocks == 6 ** 3  # blocks 1..6 per axis cover [0.12, 0.48] m
    pts = v.reshape(-1, 3)
    assert np.max(np.abs(np.linalg.norm(pts - c, axis=1) - r)) < 5e-4
    # closed surface: every undirected edge is shared by an even number of
    # triangles (2, or 4 where a corner value is ~0 and vertices coincide)
    q = merged_vertex_ids(pts).reshape(-1, 3)
    edges = {}
    for t in q:
        for i in range(3):
            a, b = int(t[i]), int(t[(i + 1) % 3])
            if a == b:
                continue
            key = (a, b) if a < b else (b, a)
            edges[key] = edges.get(key, 0) + 1
    counts = np.array(list(edges.values()))
    assert np.all(counts % 2 == 0)
    assert np.mean(counts == 2) > 0.99
    # consistent winding: face normals all point the same way relative to the centre
    cen = v.reshape(-1, 3, 3).mean(1)
    s = np.sign(((cen - c) * nrm[:, :3]).sum(1))
    assert abs(s.mean()) > 0.99
    # surface points are the centroids
    assert np.allclose(O.surface_points(v), cen, atol=1e-7)


# --- FP32 / FP64 pixel marches used by the adjoint's finite-difference checks ----------

def test_pixel_march_f32_equals_raycast():
    m = plane_map(fmt=2, frames=2)
    T = colmajor(np.eye(4))
    full = m.raycast(W, H, KP, T, init=0.0)
    ys, xs = np.mgrid[0:H, 0:W]
    v, sig, hit = m.raycast_pixels_f32(xs.ravel(), ys.ravel(), KP, T)
    assert np.array_equal(hit, ~np.isnan(full[..., 0]).ravel())
    assert np.array_equal(v[hit], full.reshape(-1, 4)[hit])


def test_pixel_march_f64_plane_and_pose_derivative():
    # plane z = D in front of an identity camera: the centre ray's depth is D,
    # and moving the camera along its own +z by h shortens it by h
    from voxelmapping_amd.autograd import se3_exp
    m = plane_map(fmt=2, frames=1)
    px, py = np.array([15, 16]), np.array([11, 12])
    v0, s0, h0 = m.raycast_pixels_f64(px, py, KP, np.eye(4).T.reshape(-1))
    assert h0.all() and np.all(np.abs(v0[:, 2] - D) < 1e-4)
    h = 1e-5
    Tp = (np.eye(4) @ se3_exp([0, 0, h, 0, 0, 0])).T.reshape(-1)
    Tm = (np.eye(4) @ se3_exp([0, 0, -h, 0, 0, 0])).T.reshape(-1)
    vp, sp, _ = m.raycast_pixels_f64(px, py, KP, Tp)
    vm, sm, _ = m.raycast_pixels_f64(px, py, KP, Tm)
    assert np.array_equal(sp, s0) and np.array_equal(sm, s0)
    dz = (vp[:, 2] - vm[:, 2]) / (2 * h)
    assert np.allclose(dz, -1.0, atol=1e-3)


def test_rendering_blocks_closed_form():
    """Voxel 1/64 m (block edge 0.125 m), camera at the origin looking down
    +z, fx = fy = 48, c = (0, 0), an 8x8 z-range map (1/8 resolution). Block
    (0, 0, 2) spans z 0.25..0.375 and projects to x, y in [0, 3] (corner
    projections 0, 2, 3); block (-1, 0, 3) spans z 0.375..0.5 and projects to
    x in [-2, 0] (clamped to [0, 0]) and y in [0, 2]; block (1, 1, 5) spans z
    0.625..0.75 and its corners project to 1.0, 1.2, 2.0, 2.4 on each axis,
    so floor / ceil give [1, 3]. Cells covered by several blocks take the
    smallest near z and the largest far z; the rest keep 100 / 0."""
    m = O.OracleMap(4096 + 512, 4096, 64, 1.0 / 64, 4.0 / 64, 0)
    m.insert_block(0, 0, 2)
    m.insert_block(-1, 0, 3)
    m.insert_block(1, 1, 5)
    m.insert_block(0, 0, -4)  # behind the camera: not visible
    K = np.array([48.0, 48.0, 0.0, 0.0], np.float32)
    eye = np.eye(4, dtype=np.float32).reshape(-1)
    assert m.check_visible(64, 64, K, eye) == 3
    n, zx, zy = m.project_rendering_blocks(8, 8, K, eye)
    assert n == 3  # one 16x16 tile per block
    ex = np.full((8, 8), 100.0, np.float32)
    ey = np.zeros((8, 8), np.float32)
    ex[0:4, 0:4] = 0.25
    ey[0:4, 0:4] = 0.375
    ex[0:3, 0] = 0.25   # min(0.25, 0.375)
    ey[0:3, 0] = 0.5    # max(0.375, 0.5)
    ey[1:4, 1:4] = 0.75  # block (1, 1, 5): max(0.375, 0.75); near z stays 0.25
    assert np.array_equal(zx, ex), zx
    assert np.array_equal(zy, ey), zy


def test_render_shading_closed_form():
    """Phong shading with ambient 0.3, diffuse 0.5, specular 0.2 (power 20),
    white material, light at (5, 5, 5), eye at the origin. For a point on the
    ray through (1, 1, 1) behind the origin, the light and eye directions are
    both u = (1, 1, 1)/sqrt(3): normal u gives 0.3 + 0.5 + 0.2 = 1 (255);
    a normal perpendicular to u gives ambient only (the mirror direction is
    -u), 0.3 -> 76; normal -u gives ambient + specular, 0.5 -> 127. A NaN
    vertex gets the background (4, 2, 2). One unit of rounding is allowed."""
    u = np.ones(3) / np.sqrt(3.0)
    p = np.array([-1.0, -1.0, -1.0])
    perp = np.array([1.0, -1.0, 0.0]) / np.sqrt(2.0)
    vmap = np.zeros((1, 4, 4), np.float32)
    nmap = np.zeros((1, 4, 4), np.float32)
    vmap[0, 0, :3] = np.nan
    for i, n in enumerate((u, perp, -u)):
        vmap[0, i + 1, :3] = p
        vmap[0, i + 1, 3] = 1.0
        nmap[0, i + 1, :3] = n
        nmap[0, i + 1, 3] = 1.0
    img = O.render_shading(vmap, nmap).astype(int)
    assert img[0, 0].tolist() == [4, 2, 2, 255]
    for i, want in enumerate((255, 76, 127)):
        px = img[0, i + 1]
        assert px[3] == 255 and px[0] == px[1] == px[2], px
        assert abs(int(px[0]) - want) <= 1, (i, px, want)


# --- raycast: sphere closed form ----------------------------------------------------
#
# A sphere of radius 0.25 m about 1 m in front of the camera, its exact
# signed distance written into every voxel of the blocks around it
# (clip(d / trunc, -1, 1), weight 1) at 5 mm voxels and 2 cm truncation.
# Trilinear interpolation of that field moves the zero level by at most
# h^2 / (8 r) = 1.25e-5 m. The march's two corrections after the first
# sample with sdf <= 0 (ray_tracing.cu:381-393) each step by the normal
# distance, which leaves (1 - cos t)^2 of the overshoot on a ray meeting the
# surface at incidence t: below 1e-4 m for cos t >= 0.9, where one
# correction alone leaves up to 10 %.

SPH_C = np.array([0.0311, -0.0187, 1.0093])
SPH_R, SPH_VS, SPH_TR = 0.25, 0.005, 0.02
SPH_K = np.array([200.0, 200.0, 79.5, 63.5], np.float32)
SPH_W, SPH_H = 160, 128


def sphere_raycast_map():
    m = O.OracleMap(16384 + 8192, 16384, 8192, SPH_VS, SPH_TR, 2)
    rec = m.voxels_view().reshape(-1, 12)
    lo = np.floor((SPH_C - SPH_R - 2 * SPH_TR) / SPH_VS).astype(int) // 8
    hi = np.floor((SPH_C + SPH_R + 2 * SPH_TR) / SPH_VS).astype(int) // 8
    lz, ly, lx = np.meshgrid(np.arange(8), np.arange(8), np.arange(8), indexing="ij")
    loc = np.stack([lx, ly, lz], -1).reshape(-1, 3)
    one = np.frombuffer(np.float32(1).tobytes(), np.uint8)
    for bx in range(lo[0], hi[0] + 1):
        for by in range(lo[1], hi[1] + 1):
            for bz in range(lo[2], hi[2] + 1):
                ptr = m.insert_block(bx, by, bz)
                assert ptr >= 0
                p = (np.array([bx, by, bz]) * 8 + loc) * SPH_VS
                s = np.clip((np.linalg.norm(p - SPH_C, axis=1) - SPH_R) / SPH_TR, -1, 1).astype(np.float32)
                idx = ptr + (loc[:, 2] * 64 + loc[:, 1] * 8 + loc[:, 0])
                rec[idx, 0:4] = s.view(np.uint8).reshape(-1, 4)
                rec[idx, 4:8] = one
    return m


def sphere_rays():
    """Per pixel: unit ray direction, analytic hit point (NaN on a miss),
    cosine of the incidence angle, and the closest approach of the ray to the
    centre."""
    ys, xs = np.mgrid[0:SPH_H, 0:SPH_W]
    d = np.stack([(xs - SPH_K[2]) / SPH_K[0], (ys - SPH_K[3]) / SPH_K[1], np.ones_like(xs, float)], -1)
    d /= np.linalg.norm(d, axis=-1, keepdims=True)
    b = d @ SPH_C
    disc = b * b - (SPH_C @ SPH_C - SPH_R ** 2)
    t = b - np.sqrt(np.maximum(disc, 0))
    P = t[..., None] * d
    nrm = (P - SPH_C) / SPH_R
    cos = np.abs((nrm * d).sum(-1))
    P[disc < 0] = np.nan
    approach = np.sqrt(np.maximum(SPH_C @ SPH_C - b * b, 0))
    return P, cos, approach


def test_raycast_sphere_depth():
    m = sphere_raycast_map()
    v = m.raycast(SPH_W, SPH_H, SPH_K, colmajor(np.eye(4)))
    P, cos, approach = sphere_rays()
    hit = ~np.isnan(v[..., 0])
    sel = cos >= 0.9
    assert sel.sum() > 1000 and (sel & (cos < 0.97)).sum() > 100
    assert hit[sel].all()
    err = np.abs(v[..., :3][sel] - P[sel]).max()
    assert err <= 1e-4, f"hit points off the sphere by {err} m"
    # rays passing at least 2 voxels outside the sphere miss
    away = approach > SPH_R + 2 * SPH_VS
    assert away.sum() > 1000 and not hit[away].any()
    # the per-pixel march (used by the adjoint's finite differences) agrees
    ys, xs = np.mgrid[0:SPH_H, 0:SPH_W]
    pv, _, ph = m.raycast_pixels_f32(xs.ravel(), ys.ravel(), SPH_K, colmajor(np.eye(4)))
    assert np.array_equal(ph, hit.ravel()) and np.array_equal(pv[ph], v.reshape(-1, 4)[ph])


# --- visibility at the image border and the depth limits ------------------------------
#
# checkBlockVisible (map_struct.h:204-244) keeps a block when any corner
# projects to u in [0, cols - 1], v in [0, rows - 1] with z in [0.1, 5],
# borders included. With an identity pose, power-of-two voxel sizes and
# focal length 8, every corner projects exactly; each block below has only
# one group of corners on the tested border line and every other corner
# outside, so moving the line half a pixel (or using < for <=) flips it.

def one_block_visible(block, vs, K, W, H):
    m = O.OracleMap(64 + 16, 64, 16, vs, 4 * vs, 0)
    assert m.insert_block(*block) >= 0
    n = m.check_visible(W, H, np.asarray(K, np.float32), colmajor(np.eye(4)))
    keys = {tuple(int(v) for v in e[2:]) for e in m.visible_table()}
    assert n == len(keys) and keys <= {tuple(block)}
    return n == 1


VIS_EDGE_CASES = [
    # block, voxel size, K, W, H, visible
    ((16, 0, 7), 1 / 64, (8, 8, 15, 11.5), 32, 24, True),      # min u of the block = 31 = cols - 1
    ((16, 0, 7), 1 / 64, (8, 8, 15.5, 11.5), 32, 24, False),   # min u = 31.5
    ((-17, 0, 7), 1 / 64, (8, 8, 16, 11.5), 32, 24, True),     # max u = 0
    ((-17, 0, 7), 1 / 64, (8, 8, 15.5, 11.5), 32, 24, False),  # max u = -0.5
    ((0, 12, 7), 1 / 64, (8, 8, 15, 11), 32, 24, True),        # min v = 23 = rows - 1
    ((0, 12, 7), 1 / 64, (8, 8, 15, 11.25), 32, 24, False),    # min v = 23.25
    ((0, -13, 7), 1 / 64, (8, 8, 15, 12), 32, 24, True),       # max v = 0
    ((0, 0, 0), 0.0125, (8, 8, 15, 11.5), 32, 24, True),       # nearest in-view corners at z = 0.1 exactly
    ((0, 0, 8), 0.078125, (8, 8, 15, 11.5), 32, 24, True),     # nearest corners at z = 5.0 exactly
    ((0, 0, 9), 0.078125, (8, 8, 15, 11.5), 32, 24, False),    # z in [5.625, 6.25]
]


@pytest.mark.parametrize("case", range(len(VIS_EDGE_CASES)))
def test_visibility_border_and_depth_limits(case):
    block, vs, K, W, H, expect = VIS_EDGE_CASES[case]
    assert one_block_visible(block, vs, K, W, H) == expect


def test_surface_depth_f64_sphere():
    """The pose refinement's depth restated in FP64 (the march, its
    corrections, three Newton steps onto the trilinear zero level): on the
    analytic sphere its camera z lies on the sphere within the trilinear
    interpolation's own error, h^2 / 8 * (2 / r) = 2.5e-5 m along the normal,
    divided by the incidence cosine along the ray, at every incidence up to
    70 degrees -- where the raycast's two corrections alone leave up to
    2e-3 m."""
    m = sphere_raycast_map()
    P, cos, _ = sphere_rays()
    ys, xs = np.mgrid[0:SPH_H, 0:SPH_W]
    z, ok = m.surface_depth_f64(xs.ravel(), ys.ravel(), SPH_K, np.eye(4))
    z, ok = z.reshape(SPH_H, SPH_W), ok.reshape(SPH_H, SPH_W)
    sel = cos >= 0.35
    assert sel.sum() > 5000 and ok[sel].all()
    err = np.abs(z[sel] - P[..., 2][sel])
    bound = 2.5e-5 / cos[sel] + 1e-6
    assert np.all(err <= bound), f"refined depth off the sphere by {np.max(err / bound)} x its bound"
    v = m.raycast(SPH_W, SPH_H, SPH_K, colmajor(np.eye(4)))
    assert np.nanmax(np.abs(v[..., 2][sel] - P[..., 2][sel])) > 1e-3
    # rays passing clear of the sphere have no depth
    _, _, approach = sphere_rays()
    assert not ok[approach > SPH_R + 2 * SPH_VS].any()
