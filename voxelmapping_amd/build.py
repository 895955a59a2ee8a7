"""Build the in-tree native libraries.

* ``voxelmapping_amd/libvmapping_amd.so`` -- the gfx950 HIP kernels and the
  C ABI of ``include/vmapping_capi.h`` plus the C++ drop-in API of
  ``include/vmapping.h`` / ``map_struct.h`` / ``image_proc.h``;
* ``voxelmapping_amd/libvm_synth.so`` -- the seeded synthetic scene generator.

Run ``python -m voxelmapping_amd.build`` (or ``__graft_entry__.build()``).
hipcc cross-compiles gfx950 code objects without a GPU present.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
INCLUDE = os.path.join(ROOT, "include")
BUILD = os.path.join(ROOT, "_build")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("VM_OFFLOAD_ARCH", "gfx950")

# -ffp-contract=off: no fused multiply-add, so every float expression rounds as
# written (parity with the CPU oracle); IEEE division and sqrt are requested
# explicitly; packed fp32 instructions are disabled.
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    "-ffp-contract=off", "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
    f"-I{INCLUDE}", f"-I{CSRC}",
]

HIP_SOURCES = ["vm_fusion.hip", "vm_raycast.hip", "vm_mesh.hip", "vm_batch.hip", "vm_adjoint.hip", "vm_pose.hip",
               "vm_capi.hip"]
CXX_SOURCES = ["vm_facade.cpp"]

LIB = os.path.join(PKG, "libvmapping_amd.so")
SYNTH = os.path.join(PKG, "libvm_synth.so")


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"build step failed ({r.returncode}): {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")


def build_hip_library(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    jobs = []
    objs = []
    for src in HIP_SOURCES:
        obj = os.path.join(BUILD, src + ".o")
        objs.append(obj)
        jobs.append([HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
    for src in CXX_SOURCES:
        path = os.path.join(CSRC, src)
        if not os.path.exists(path):
            continue
        obj = os.path.join(BUILD, src + ".o")
        objs.append(obj)
        jobs.append([HIPCC, "-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", f"-I{INCLUDE}", f"-I{CSRC}",
                     "-c", path, "-o", obj])
    workers = min(len(jobs), int(os.environ.get("MAX_JOBS", "8")))
    with ThreadPoolExecutor(max_workers=max(1, workers)) as ex:
        for _ in ex.map(_run, jobs):
            pass
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


def build_synth(verbose: bool = False) -> str:
    tmp = SYNTH + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-fopenmp",
          os.path.join(CSRC, "vm_synth.cpp"), "-o", tmp])
    os.replace(tmp, SYNTH)
    if verbose:
        print("built", SYNTH)
    return SYNTH


CPP_TESTS = os.path.join(ROOT, "tests", "cpp")
DEMO_SRC = os.path.join(CPP_TESTS, "dropin_demo.cpp")
DEMO_BIN = os.path.join(CPP_TESTS, "dropin_demo")
# the same API driven with the reference's OpenCV / Eigen argument types,
# against the stand-in headers of tests/cpp/stub
REFTYPES_SRC = os.path.join(CPP_TESTS, "dropin_reference_types.cpp")
REFTYPES_BIN = os.path.join(CPP_TESTS, "dropin_reference_types")
STUB_INCLUDE = os.path.join(CPP_TESTS, "stub")


def build_cpp_program(src: str, out: str, extra_includes=(), verbose: bool = False) -> str:
    """C++ program using the drop-in headers, linked against the library."""
    tmp = out + ".tmp"
    incs = [f"-I{d}" for d in extra_includes]
    _run([HIPCC, "-O2", "-std=c++17", *incs, f"-I{INCLUDE}", src, "-o", tmp, f"-L{PKG}", "-lvmapping_amd",
          "-Wl,-rpath,$ORIGIN/../../voxelmapping_amd"])
    os.replace(tmp, out)
    if verbose:
        print("built", out)
    return out


def build_dropin_demo(verbose: bool = False) -> str:
    return build_cpp_program(DEMO_SRC, DEMO_BIN, verbose=verbose)


def build_reference_types_demo(verbose: bool = False) -> str:
    return build_cpp_program(REFTYPES_SRC, REFTYPES_BIN, [STUB_INCLUDE], verbose=verbose)


def build_all(verbose: bool = False) -> None:
    build_synth(verbose)
    build_hip_library(verbose)
    build_dropin_demo(verbose)
    build_reference_types_demo(verbose)


if __name__ == "__main__":
    build_all(verbose=True)
    sys.exit(0)
