"""The overlapped IB step (far cells on a second stream during the IB
forcing; levels of at least 2^20 cells) gives the same populations, macros
and force records as the single-stream step; Sum(Rho) is summed in another
order."""
import numpy as np
import pytest

from ilbmrex_amd import config as cf

pytestmark = pytest.mark.gpu


def _pair(case, init, nsteps):
    from ilbmrex_amd.native import Level
    out = []
    for ov in (True, False):
        lev = Level(case, init)
        lev.set_overlap(ov)
        lev.advance(nsteps, want_state=True)
        out.append((lev.f(), lev.macro(), lev.force_history(), lev.mass_history()))
        lev.close()
    return out


@pytest.mark.parametrize("storage", [cf.STORE_F64, cf.STORE_F32])
def test_overlap_matches_single_stream_cylinder(storage):
    c = cf.cylinder_2d(nx=1024, ny=1024, D=24.0, xc=200.0, yc=512.0, U=0.05, Re=100.0, storage=storage)
    print(f"[overlap] cylinder 1024x1024 storage={storage}", flush=True)
    (f1, m1, h1, s1), (f2, m2, h2, s2) = _pair(c, cf.uniform_init(c, u=(0.05, 0.0, 0.0)), 40)
    assert np.array_equal(f1, f2) and np.array_equal(m1, m2) and np.array_equal(h1, h2)
    np.testing.assert_allclose(s1, s2, rtol=1e-13)


def test_overlap_matches_single_stream_sphere_near_wall():
    # the band reaches the inflow face: the near box is clipped there
    c = cf.sphere_3d(nx=128, ny=96, nz=96, D=8.0, center=(6.0, 48.0, 48.0), U=0.04, Re=40.0,
                     storage=cf.STORE_F32)
    print("[overlap] sphere 128x96x96 FP32 near the inflow", flush=True)
    (f1, m1, h1, s1), (f2, m2, h2, s2) = _pair(c, cf.uniform_init(c, u=(0.04, 0.0, 0.0)), 20)
    assert np.array_equal(f1, f2) and np.array_equal(m1, m2) and np.array_equal(h1, h2)
    np.testing.assert_allclose(s1, s2, rtol=1e-12)


def test_overlap_band_across_periodic_boundary():
    # body at the periodic y edge: the near box spans the whole y extent
    c = cf.cylinder_2d(nx=1024, ny=1024, D=10.0, xc=30.0, yc=2.0, U=0.05, Re=60.0, storage=cf.STORE_F64)
    print("[overlap] cylinder across the periodic y boundary", flush=True)
    (f1, m1, h1, s1), (f2, m2, h2, s2) = _pair(c, cf.uniform_init(c, u=(0.05, 0.0, 0.0)), 20)
    assert np.array_equal(f1, f2) and np.array_equal(m1, m2) and np.array_equal(h1, h2)


def test_overlap_sphere_off_walls_split_advances():
    """Body away from the walls: the next step's ghost slabs are filled right
    after the far cells; advances of 7 + 1 + 6 steps (the slab fill queued
    ahead and the alternating mass partials across advance calls) against one
    single-stream advance of 14."""
    from ilbmrex_amd.native import Level
    c = cf.sphere_3d(nx=128, ny=96, nz=96, D=10.0, center=(48.0, 47.6, 48.3), U=0.04, Re=50.0,
                     storage=cf.STORE_F64)
    init = cf.uniform_init(c, u=(0.04, 0.0, 0.0))
    print("[overlap] sphere 128x96x96 FP64 off the walls, 7 + 1 + 6 steps", flush=True)
    a = Level(c, init)
    a.set_overlap(True)
    a.advance(7)
    a.advance(1)
    a.advance(6, want_state=True)
    b = Level(c, init)
    b.set_overlap(False)
    b.advance(14, want_state=True)
    assert np.array_equal(a.f(), b.f()) and np.array_equal(a.macro(), b.macro())
    assert np.array_equal(a.force_history(), b.force_history())
    ma, mb = a.mass_history(), b.mass_history()
    assert len(ma) == len(mb) == 14
    np.testing.assert_allclose(ma, mb, rtol=1e-13)
    a.close()
    b.close()


@pytest.mark.parametrize("regions", ["three", "two", "three_after_interp", "three_after_export",
                                     "three_persistent", "three_separate_slabs", "three_ring_on_far",
                                     "three_ring_on_near", "three_ring_own_stream"])
def test_overlap_long_run_crosses_mass_ring_halves(regions, monkeypatch):
    """70 overlapped steps (the Sum(Rho) ring halves fill and are reduced
    twice, the alternating far / ring signals run through every parity)
    against the single-stream step: bit-identical state and force history.
    "three": far cells, near box and the ring between them as three
    launches, the far cells at once; "two": ILBM_OVERLAP_TWO_REGIONS=1
    keeps far + near; the far cells after the step's interpolation or
    export (ILBM_FAR_AFTER), as a persistent launch of 4
    blocks per CU (ILBM_FAR_BLOCKS_PER_CU=4), and with the separate ghost
    slab kernel (ILBM_GHOST_SEPARATE=1); the ring on the far cells', the
    near box's or its own stream (ILBM_RING_STREAM)."""
    from ilbmrex_amd.native import Level
    for v in ("ILBM_OVERLAP_TWO_REGIONS", "ILBM_FAR_AFTER", "ILBM_FAR_BLOCKS_PER_CU", "ILBM_GHOST_SEPARATE",
              "ILBM_RING_STREAM"):
        monkeypatch.delenv(v, raising=False)
    if regions == "two":
        monkeypatch.setenv("ILBM_OVERLAP_TWO_REGIONS", "1")
    if regions == "three_after_interp":
        monkeypatch.setenv("ILBM_FAR_AFTER", "interp")
    if regions == "three_after_export":
        monkeypatch.setenv("ILBM_FAR_AFTER", "export")
    if regions == "three_persistent":
        monkeypatch.setenv("ILBM_FAR_BLOCKS_PER_CU", "4")
    if regions == "three_separate_slabs":
        monkeypatch.setenv("ILBM_GHOST_SEPARATE", "1")
    if regions == "three_ring_on_far":
        monkeypatch.setenv("ILBM_RING_STREAM", "far")
    if regions == "three_ring_on_near":
        monkeypatch.setenv("ILBM_RING_STREAM", "near")
    if regions == "three_ring_own_stream":
        monkeypatch.setenv("ILBM_RING_STREAM", "own")
    c = cf.sphere_3d(nx=128, ny=96, nz=96, D=10.0, center=(48.0, 47.6, 48.3), U=0.04, Re=50.0,
                     storage=cf.STORE_F32)
    init = cf.uniform_init(c, u=(0.04, 0.0, 0.0))
    print(f"[overlap] sphere 128x96x96 FP32, 70 steps, {regions} regions", flush=True)
    (f1, m1, h1, s1), (f2, m2, h2, s2) = _pair(c, init, 70)
    assert np.array_equal(f1, f2) and np.array_equal(m1, m2) and np.array_equal(h1, h2)
    assert len(s1) == 70
    np.testing.assert_allclose(s1, s2, rtol=1e-12)
