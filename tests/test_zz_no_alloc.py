"""vm_polygonize_noalloc (marching cubes with the library's fixed mesh work
array, no device allocation) through the C ABI, against vm_polygonize on
the same maps: two calls in a row, the second on a smaller map than the
first (the shared work array then holds the first call's contents), each
with the same block count, triangle count and bit-identical triangle
buffers; and the rejection of a map with more allocated blocks than the
array holds."""
import ctypes as C

import numpy as np
import pytest

from parity import fuse_both, make_pair

pytestmark = pytest.mark.gpu

CAP = 1_500_000


def _mesh(m, noalloc):
    import torch
    from voxelmapping_amd import _capi
    L = _capi.lib()
    verts = torch.full((CAP, 9), 7.0, dtype=torch.float32, device="cuda")
    norms = torch.full_like(verts, 7.0)
    b, t = C.c_uint32(0), C.c_uint32(0)
    fn = L.vm_polygonize_noalloc if noalloc else L.vm_polygonize
    rc = fn(m.handle, C.c_void_p(verts.data_ptr()), C.c_void_p(norms.data_ptr()), CAP, None, C.byref(b),
            C.byref(t))
    assert rc == 0, L.vm_last_error()
    torch.cuda.synchronize()
    return b.value, t.value, verts, norms


def test_noalloc_mesh_equals_polygonize(oracle_threads):
    import torch
    from voxelmapping_amd import synth
    w = synth.room_small(voxel_format=2)
    maps = []
    for frames in (12, 3):  # the second map has fewer blocks
        g, o = make_pair(w, fmt=2)
        fuse_both(w, g, o, range(frames), rgb=True)
        maps.append(g)
    expect = [_mesh(m, noalloc=False) for m in maps]
    assert expect[1][0] < expect[0][0] and expect[1][1] < expect[0][1]
    for m, (eb, et, ev, en) in zip(maps, expect):  # two calls in a row, larger map first
        b, t, v, n = _mesh(m, noalloc=True)
        print(f"noalloc: {b} blocks, {t} triangles (expected {eb}, {et})", flush=True)
        assert (b, t) == (eb, et) and t > 1000
        assert torch.equal(v, ev) and torch.equal(n, en)
    for m in maps:
        m.release()


def test_noalloc_rejects_more_blocks_than_it_holds():
    """A map whose allocated count exceeds VM_NOALLOC_MAX_BLOCKS (2^19) is
    refused before any launch, with an error naming the limit; the same
    map's vm_polygonize is unaffected."""
    import torch
    from voxelmapping_amd import DeviceMap, VMError
    m = DeviceMap((1 << 20) + (1 << 16), 1 << 20, 1 << 20, 0.01, 0.04, 0)
    c = np.zeros(16, np.int32)
    c[0] = (1 << 20) - 1 - ((1 << 19) + 1)  # heap pointer of 2^19 + 1 allocated blocks
    m.upload(3, c)
    verts = torch.zeros((16, 9), dtype=torch.float32, device="cuda")
    norms = torch.zeros_like(verts)
    with pytest.raises(VMError, match="VM_NOALLOC_MAX_BLOCKS"):
        m.polygonize_noalloc(verts, norms)
    c[0] = (1 << 20) - 1 - 3  # three (dead, never published) blocks: accepted, no triangles
    m.upload(3, c)
    assert m.polygonize_noalloc(verts, norms) == (3, 0)
    m.release()
