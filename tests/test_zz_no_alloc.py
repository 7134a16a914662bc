"""pm_model_depth_ranges_noalloc, the entry point that allocates nothing:
device-resident inputs, the library's fixed device array as scratch.  Two
calls in a row (the second with fewer images and points, so stale centres
or a box left from the first call would show) give the same ranges, bit for
bit, as pm_model_depth_ranges on the same models; sizes beyond the array are
rejected before any device work."""
import ctypes as C

import numpy as np
import pytest

from sensemap_amd import abi, scene
from sensemap_amd import patch_match as pmm

pytestmark = pytest.mark.gpu


def ranges_allocating(eng, m):
    out = np.zeros((m.num_images, 2), np.float32)
    abi.check(eng.lib.pm_model_depth_ranges(eng.handle, C.byref(m.to_c()), out.ctypes.data_as(C.POINTER(C.c_float))))
    return out


def _hip():
    """The HIP runtime the engine library is linked against (already loaded
    with it), for the test's own device buffers."""
    hip = C.CDLL("libamdhip64.so.7")
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipFree.argtypes = [C.c_void_p]
    return hip


def ranges_noalloc(eng, m):
    hip = _hip()
    bufs = []

    def upload(a):
        a = np.ascontiguousarray(a)
        p = C.c_void_p()
        assert hip.hipMalloc(C.byref(p), max(a.nbytes, 4)) == 0
        bufs.append(p)
        if a.nbytes:
            assert hip.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0  # host to device
        return p

    try:
        R, T = upload(m.R), upload(m.T)
        xyz = upload(m.xyz) if m.num_points else None
        out = upload(np.full((m.num_images, 2), -7.0, np.float32))
        abi.check(eng.lib.pm_model_depth_ranges_noalloc(eng.handle, m.num_images, R, T, m.num_points, xyz, out))
        got = np.empty((m.num_images, 2), np.float32)
        assert hip.hipMemcpy(got.ctypes.data, out, got.nbytes, 2) == 0  # device to host
        return got
    finally:
        for p in bufs:
            hip.hipFree(p)


def test_noalloc_depth_ranges_equal_the_allocating_entry_point():
    eng = pmm.Engine(0)
    big = scene.sparse_scene(seed=41, num_cams=60, num_points=8000, num_blind=2).model()
    small = scene.sparse_scene(seed=42, num_cams=9, num_points=300, num_blind=0).model()
    for m in (big, small):
        want = ranges_allocating(eng, m)
        got = ranges_noalloc(eng, m)
        print("[noalloc] images", m.num_images, "points", m.num_points, "equal",
              np.array_equal(got.view(np.uint32), want.view(np.uint32)), flush=True)
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


def test_noalloc_rejects_sizes_beyond_the_array():
    eng = pmm.Engine(0)
    n = abi.PM_DEPTH_RANGES_NOALLOC_MAX_IMAGES + 1
    rc = eng.lib.pm_model_depth_ranges_noalloc(eng.handle, n, 1, 1, 0, None, 1)  # rejected before any device work
    assert rc == abi.PM_ERR_INVALID_ARG
    assert b"fixed device array" in eng.lib.pm_last_error()
