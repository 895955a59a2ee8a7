"""The oracle's pins catch planted mistakes.

Each case below is one plausible mistake in the oracle's restatement of
the reference -- a dropped term, a wrong comparison, a border moved by one,
a reversed operand -- planted as a text change in a copy of
oracle/vm_oracle.cpp, compiled into a temporary library. The pins named
with it are then run against that library (tests/test_oracle.py, through
the VM_ORACLE_LIB override of oracle/oracle.py) and must fail; the same
pins pass on the real oracle (tests/test_oracle.py itself). Runs on the
CPU only.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
SRC = os.path.join(ORACLE_DIR, "vm_oracle.cpp")
CXXFLAGS = ["-O1", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math", "-fopenmp", "-shared",
            f"-I{ORACLE_DIR}"]

# name: (text in vm_oracle.cpp, replacement, pins that must fail)
MUTANTS = {
    "raycast_second_correction_dropped": (
        """                sdf = read_sdf_interp(m, res, &valid);
                step = sdf * rtStep;
                res = V3{res.x + step * dir.x, res.y + step * dir.y, res.z + step * dir.z};
                if (valid) found = true;""",
        """                sdf = read_sdf_interp(m, res, &valid);
                if (valid) found = true;""",
        ["test_raycast_sphere_depth"]),
    "raycast_first_correction_dropped": (
        """            if (sdf <= 0.0f) {
                step = sdf * rtStep;
                res = V3{res.x + step * dir.x, res.y + step * dir.y, res.z + step * dir.z};
                sdf = read_sdf_interp(m, res, &valid);""",
        """            if (sdf <= 0.0f) {
                sdf = read_sdf_interp(m, res, &valid);""",
        ["test_raycast_sphere_depth"]),
    "visibility_last_column_excluded": (
        "u > (float)(cols - 1)", "u >= (float)(cols - 1)",
        ["test_visibility_border_and_depth_limits[0]"]),
    "visibility_column_border_at_cols": (
        "u > (float)(cols - 1)", "u > (float)cols",
        ["test_visibility_border_and_depth_limits[1]"]),
    "visibility_first_column_excluded": (
        "bool out = u < 0 ||", "bool out = u <= 0 ||",
        ["test_visibility_border_and_depth_limits[2]"]),
    "visibility_row_border_at_rows": (
        "v > (float)(rows - 1)", "v > (float)rows",
        ["test_visibility_border_and_depth_limits[5]"]),
    "visibility_min_depth_excluded": (
        "p.z < m->dmin || p.z > m->dmax", "p.z <= m->dmin || p.z > m->dmax",
        ["test_visibility_border_and_depth_limits[7]"]),
    "visibility_max_depth_excluded": (
        "p.z < m->dmin || p.z > m->dmax", "p.z < m->dmin || p.z >= m->dmax",
        ["test_visibility_border_and_depth_limits[8]"]),
    "hash_wrong_prime": (
        "((uint32_t)y * 19349669u)", "((uint32_t)y * 19349663u)",
        ["test_hash_known_answers"]),
    "allocation_full_truncation": (
        "const float truncHalf = (float)((double)m->trunc * 0.5);", "const float truncHalf = m->trunc;",
        ["test_allocation_half_truncation_band"]),
    "integration_unclamped_sdf": (
        "sdf = std::fmin(1.0f, sdf / tr);", "sdf = sdf / tr;",
        ["test_integration_plane_first_frame"]),
    "mesh_winding_not_reversed": (
        """            V3 p0{c.x * vs, c.y * vs, c.z * vs};
            V3 p1{b.x * vs, b.y * vs, b.z * vs};
            V3 p2{a.x * vs, a.y * vs, a.z * vs};""",
        """            V3 p0{a.x * vs, a.y * vs, a.z * vs};
            V3 p1{b.x * vs, b.y * vs, b.z * vs};
            V3 p2{c.x * vs, c.y * vs, c.z * vs};""",
        ["test_polygonize_single_cube_cases"]),
    "shading_specular_dropped": (
        """            float I = Ax * Ka * Dx + Lx * Kd * Dx * std::fmax(0.f, nl) +
                      Lx * Ks * Sx * std::pow(std::fmax(0.f, dot3(R, Vv)), n);""",
        """            float I = Ax * Ka * Dx + Lx * Kd * Dx * std::fmax(0.f, nl);""",
        ["test_render_shading_closed_form"]),
    "refinement_newton_steps_skipped": (
        "        const S t = -v / den;", "        const S t = 0 * v / den;",
        ["test_surface_depth_f64_sphere"]),
    "rendering_block_floor_for_ceil": (
        "if ((float)lrx < std::ceil(px)) lrx = (short)f2i_sat(std::ceil(px));",
        "if ((float)lrx < std::floor(px)) lrx = (short)f2i_sat(std::floor(px));",
        ["test_rendering_blocks_closed_form"]),
}


@pytest.fixture(scope="module")
def mutant_libs(tmp_path_factory):
    src = open(SRC).read()
    out = tmp_path_factory.mktemp("oracle_mutants")
    jobs = {}
    for name, (old, new, _) in MUTANTS.items():
        assert src.count(old) >= 1, f"{name}: planted text not found in vm_oracle.cpp"
        path = out / f"{name}.cpp"
        path.write_text(src.replace(old, new, 1))
        jobs[name] = (str(path), str(out / f"lib{name}.so"))

    def build(item):
        name, (cpp, so) = item
        r = subprocess.run(["g++", *CXXFLAGS, cpp, "-o", so], capture_output=True, text=True)
        return name, r.returncode, r.stderr

    with ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
        for name, rc, err in ex.map(build, jobs.items()):
            assert rc == 0, f"{name} does not compile:\n{err}"
    return {name: so for name, (_, so) in jobs.items()}


@pytest.mark.parametrize("name", sorted(MUTANTS))
def test_pins_catch_planted_mistake(mutant_libs, name):
    pins = MUTANTS[name][2]
    env = dict(os.environ, VM_ORACLE_LIB=mutant_libs[name])
    ids = [f"tests/test_oracle.py::{p}" for p in pins]
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-p", "no:cacheprovider", *ids], cwd=ROOT, env=env,
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 1, f"{name}: pins {pins} did not fail (rc {r.returncode})\n{r.stdout[-2000:]}"
    assert f"{len(pins)} failed" in r.stdout, f"{name}: not every pin failed\n{r.stdout[-2000:]}"
