"""ilbm_box_sum_noalloc (the allocation-free box sum) against ilbm_box_sum
through the C ABI: two calls in a row, the second on a smaller box, each
bit-identical to the allocating entry point; a box above the limit is
rejected before any device work."""
import ctypes as C

import numpy as np
import pytest

from ilbmrex_amd import native


def _fab(ptr, lo, hi, ncomp):
    f = native.Fab()
    f.p = ptr
    f.lo = native.I3(*lo)
    f.hi = native.I3(*hi)
    f.ncomp = ncomp
    return f


def test_noalloc_limit_is_rejected_without_device_work():
    L = native.load()
    cap = C.c_int64(0)
    assert L.ilbm_box_sum_noalloc_max_cells(C.byref(cap)) == 0
    assert cap.value == (131072 - 1) * 256
    n = 4096
    # a fab descriptor over (n+1) x n x n cells; the pointer is never touched
    f = _fab(C.c_void_p(0x1000), (0, 0, 0), (n, n - 1, n - 1), 1)
    out = C.c_double(0.0)
    rc = L.ilbm_box_sum_noalloc(C.byref(f), 0, native.I3(0, 0, 0), native.I3(n, n - 1, n - 1), C.byref(out), None)
    assert rc == 1 and b"limit" in L.ilbm_last_error()


@pytest.mark.gpu
def test_noalloc_box_sum_matches_allocating_sum():
    L = native.load()
    rng = np.random.default_rng(20261203)
    shapes = [((3, 70, 90, 130), (-3, -3, -3), (120, 80, 60)),    # first call: larger box
              ((2, 20, 30, 50), (-1, -2, -3), (40, 20, 12))]       # second call: smaller box
    for (nc, nz, ny, nx), lo, top in shapes:
        a = rng.normal(size=(nc, nz, ny, nx))
        dev = native.DeviceArray(a.shape)
        dev.upload(a)
        hi = (lo[0] + nx - 1, lo[1] + ny - 1, lo[2] + nz - 1)
        f = _fab(dev.ptr, lo, hi, nc)
        blo = native.I3(lo[0] + 1, lo[1] + 2, lo[2] + 1)
        bhi = native.I3(*top)
        for comp in range(nc):
            want, got = C.c_double(0.0), C.c_double(0.0)
            assert L.ilbm_box_sum(C.byref(f), comp, blo, bhi, C.byref(want), None) == 0
            assert L.ilbm_box_sum_noalloc(C.byref(f), comp, blo, bhi, C.byref(got), None) == 0
            print(f"box {tuple(blo)}..{tuple(bhi)} comp {comp}: {got.value!r} vs {want.value!r}", flush=True)
            assert got.value == want.value
            sub = a[comp, blo[2] - lo[2]:bhi[2] - lo[2] + 1, blo[1] - lo[1]:bhi[1] - lo[1] + 1,
                    blo[0] - lo[0]:bhi[0] - lo[0] + 1]
            assert abs(got.value - sub.sum()) <= 1e-12 * np.abs(sub).sum()
        dev.free()
