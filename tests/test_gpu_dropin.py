"""The C++ drop-in API (vmapping.h, map_struct.h, device_functions.h,
image_proc.h) run end to end by a compiled program on a closed-form scene."""
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
