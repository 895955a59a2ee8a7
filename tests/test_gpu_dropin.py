"""The C++ drop-in API (vmapping.h, map_struct.h, device_functions.h,
image_proc.h) run end to end by compiled programs on a closed-form scene:
once with the library's own value types, once with the reference's OpenCV /
Eigen argument types (stand-in headers, tests/cpp/stub)."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_cpp_dropin_program():
    demo = os.path.join(ROOT, "tests", "cpp", "dropin_demo")
    print("running", demo, flush=True)
    r = subprocess.run([demo], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0 and "DROPIN OK" in r.stdout


def test_cpp_reference_types_program():
    prog = os.path.join(ROOT, "tests", "cpp", "dropin_reference_types")
    print("running", prog, flush=True)
    r = subprocess.run([prog], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0 and "REFTYPES OK" in r.stdout
