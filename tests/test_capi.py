"""The C-ABI library loads and exports every function include/vmapping_capi.h
declares, and the Python binding covers them (no GPU calls)."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(vm_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    from voxelmapping_amd import _capi
    lib = ctypes.CDLL(_capi.LIB_PATH)
    names = declared("vmapping_capi.h")
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_signatures_cover_header():
    from voxelmapping_amd import _capi
    assert set(declared("vmapping_capi.h")) == set(_capi.SIGNATURES)


def test_error_reporting_without_gpu_work():
    from voxelmapping_amd import _capi
    L = _capi.lib()
    assert L.vm_map_create(None, None) == -1
    assert b"NULL" in L.vm_last_error()
    assert L.vm_map_destroy(None) == 0
    assert L.vm_version().startswith(b"vmapping-amd")


def test_struct_sizes_match_header():
    from voxelmapping_amd import _capi
    assert ctypes.sizeof(_capi.MapDesc) == 36
    assert ctypes.sizeof(_capi.MapStats) == 32
    assert ctypes.sizeof(_capi.MapArrays) == 9 * 8


def test_cpp_dropin_api_exported_and_linked():
    import subprocess
    from voxelmapping_amd import _capi
    out = subprocess.run(["nm", "-DC", "--defined-only", _capi.LIB_PATH], capture_output=True, text=True).stdout
    for sym in ("vmap::VoxelMapping::VoxelMapping(", "vmap::VoxelMapping::CreateMap(", "vmap::VoxelMapping::FuseDepth(",
                "vmap::VoxelMapping::FuseDepthAndImage(", "vmap::VoxelMapping::RenderScene(",
                "vmap::VoxelMapping::Polygonize(", "vmap::VoxelMapping::SurfacePoints(",
                "vmap::VoxelMapping::FuseDepthAndImageHost(", "vmap::VoxelMapping::reset("):
        assert sym in out, sym
    demo = os.path.join(ROOT, "tests", "cpp", "dropin_demo")
    assert os.path.exists(demo), "run the build first"
    ldd = subprocess.run(["ldd", demo], capture_output=True, text=True).stdout
    assert "libvmapping_amd.so" in ldd and "not found" not in ldd.split("libvmapping_amd.so")[1].split("\n")[0]


def test_reference_type_overloads_compile():
    """The drop-in headers with OpenCV and Eigen present (the stand-in
    headers of tests/cpp/stub define CV_8UC3 and its kind as macros, as OpenCV
    does): the reference's own signatures -- cv::cuda::GpuMat& outputs,
    FuseDepthAndImage(cv::Mat, cv::Mat, ...), std::vector<Eigen::Vector3f>
    GetSurfacePoints() -- compile, and the program built from them links."""
    import subprocess
    from voxelmapping_amd import build as B
    r = subprocess.run([B.HIPCC, "-std=c++17", "-fsyntax-only", f"-I{B.STUB_INCLUDE}", f"-I{B.INCLUDE}",
                        B.REFTYPES_SRC], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    # without the stand-ins the same headers compile on their own types
    r = subprocess.run([B.HIPCC, "-std=c++17", "-fsyntax-only", f"-I{B.INCLUDE}", B.DEMO_SRC],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    assert os.path.exists(B.REFTYPES_BIN), "run the build first"
    ldd = subprocess.run(["ldd", B.REFTYPES_BIN], capture_output=True, text=True).stdout
    assert "libvmapping_amd.so" in ldd
