"""AMR hierarchy on the GPU against the oracle's hierarchy (FP64).

Every level's populations and macros must be bit-identical to the oracle's
after the coarse steps (same coarse-fine interpolation, equilibrium ghost
cells, subcycling, average-down and derived-field schedule); Sum(Rho) and
the finest level's force record are compared at 1e-12 relative (summation
order).  Reduced storage is compared with the FP64 hierarchy by Cd.
"""
import numpy as np
import pytest

from ilbmrex_amd import config as cf
import refcpu

pytestmark = pytest.mark.gpu


def cyl_case(storage=cf.STORE_F64):
    return cf.cylinder_2d(nx=64, ny=48, D=8.0, xc=20.0, yc=24.0, U=0.05, Re=40.0, storage=storage)


TWO_LEVEL = [((24, 28, 0), (71, 67, 0))]
THREE_LEVEL = [((16, 16, 0), (79, 79, 0)), ((48, 48, 0), (127, 127, 0))]


def run_pair(case, boxes, nsteps, regrid_int=2, regrid=None):
    from ilbmrex_amd.native import AmrHierarchy
    print(f"[amr] {case.name} {case.lattice.name} levels={1 + len(boxes)} storage={case.storage} "
          f"steps={nsteps} regrid={regrid}", flush=True)
    u0 = case.bc_lo_val[0][1]
    state = (1.0, u0, 0.0, 0.0)
    gpu = AmrHierarchy(case, boxes, regrid_int, state)
    orc = refcpu.AmrOracle(case, boxes, regrid_int, cf.uniform_init(case, u=(u0, 0.0, 0.0)))
    if regrid is not None:
        gpu.set_regrid(**regrid)
        orc.set_regrid(**regrid)
    gpu.advance(nsteps, want_state=True)
    orc.advance(nsteps)
    return gpu, orc


def check_levels(gpu, orc):
    for l in range(gpu.nlev):
        for what, a, b in (("f", gpu.f(l), orc.f(l)), ("macro", gpu.macro(l), orc.macro(l))):
            if not np.array_equal(a, b):
                d = np.abs(a - b)
                raise AssertionError(f"level {l} {what}: {np.count_nonzero(a != b)} differ, max {d.max():.3e}")
    mh = gpu.mass_history()
    assert abs(mh[-1] - orc.mass_sum()) <= 1e-12 * abs(orc.mass_sum())


def check_record(gpu, orc):
    fh = gpu.force_history()
    rec = orc.force_record()
    assert fh.shape[1] == len(rec)
    floor = 1e-10 * np.abs(rec).max()
    assert np.all(np.abs(fh[-1] - rec) <= 1e-12 * np.abs(rec) + floor), (fh[-1], rec)


def test_two_level_cylinder_bitexact():
    c = cyl_case()
    gpu, orc = run_pair(c, TWO_LEVEL, 12)
    assert gpu.num_markers() == orc.num_markers() == 51
    check_levels(gpu, orc)
    check_record(gpu, orc)
    info = gpu.info()
    assert info["steps"] == 12 and info["time"] == 12.0


def test_three_level_cylinder_bitexact():
    c = cyl_case()
    gpu, orc = run_pair(c, THREE_LEVEL, 6, regrid_int=3)
    check_levels(gpu, orc)
    check_record(gpu, orc)


def test_two_level_sphere_bitexact():
    c = cf.sphere_3d(nx=32, ny=24, nz=24, D=6.0, center=(10.0, 12.0, 12.0), U=0.04, Re=30.0)
    gpu, orc = run_pair(c, [((8, 10, 10), (35, 37, 37))], 4)
    check_levels(gpu, orc)
    check_record(gpu, orc)


def test_two_level_fp32_cd_close_to_fp64():
    from ilbmrex_amd.native import AmrHierarchy
    res = {}
    for st in (cf.STORE_F64, cf.STORE_F32):
        c = cyl_case(st)
        h = AmrHierarchy(c, TWO_LEVEL, 2, (1.0, 0.05, 0.0, 0.0))
        h.advance(300)
        res[st] = h.force_history()[-100:, 4].mean()
        h.close()
    assert np.isfinite(res[cf.STORE_F32])
    assert abs(res[cf.STORE_F32] - res[cf.STORE_F64]) <= 1e-3 * abs(res[cf.STORE_F64])


def test_two_level_pitching_section_bitexact():
    """Moving body on the finest level: device kinematics and gather
    spreading against the oracle's host kinematics and serial scatter."""
    c = cf.cylinder_2d(nx=64, ny=48, D=10.0, xc=20.0, yc=24.0, U=0.03, Re=100.0).replace(
        geometry_type=cf.GEOM_USER, airfoil_chord=10.0, n_marker=96, pitch_amplitude=0.2,
        pitch_frequency=0.01, pivot_x=2.5, force_eval_method=cf.FORCE_EVAL_MARKER, name="pitching_amr")
    gpu, orc = run_pair(c, [((28, 28, 0), (75, 67, 0))], 8)
    check_levels(gpu, orc)
    check_record(gpu, orc)


def test_pitching_naca_three_levels_runs():
    from ilbmrex_amd.native import AmrHierarchy
    c, boxes = cf.pitching_naca_amr(Re=100.0)
    h = AmrHierarchy(c, boxes, 10, (1.0, c.U0, 0.0, 0.0))
    assert h.num_markers() == 256
    h.advance(20)
    fh = h.force_history()
    assert fh.shape == (20, 23) and np.isfinite(fh).all()
    info = h.info()
    assert info["cells"] == [512 * 256, 128 * 96, 160 * 128]


# ---------------------------------------------------------------------------
# Regrid: the GPU's tags, boxes and remade levels against the oracle's
# ---------------------------------------------------------------------------

def check_boxes(gpu, orc):
    boxes, n = gpu.boxes()
    assert boxes == orc.boxes, (boxes, orc.boxes)
    assert n == orc.regrids()
    return n


def test_two_level_cylinder_regrid_bitexact():
    c = cyl_case()
    gpu, orc = run_pair(c, TWO_LEVEL, 12, regrid_int=2,
                        regrid=dict(threshold_ratio=(0.5, 0.5), blocking_factor=4))
    assert check_boxes(gpu, orc) >= 1
    check_levels(gpu, orc)
    check_record(gpu, orc)


def test_three_level_cylinder_regrid_bitexact():
    """regrid(0) re-boxes levels 1 and 2; regrid(1) between coarse steps
    re-boxes level 2 alone (istep[1] % regrid_int == 0)."""
    c = cyl_case()
    gpu, orc = run_pair(c, THREE_LEVEL, 8, regrid_int=3,
                        regrid=dict(threshold_ratio=(0.5, 0.6, 0.6), blocking_factor=4))
    assert check_boxes(gpu, orc) >= 1
    check_levels(gpu, orc)
    check_record(gpu, orc)


def test_two_level_pitching_section_regrid_bitexact():
    c = cf.cylinder_2d(nx=64, ny=48, D=10.0, xc=20.0, yc=24.0, U=0.03, Re=100.0).replace(
        geometry_type=cf.GEOM_USER, airfoil_chord=10.0, n_marker=96, pitch_amplitude=0.2,
        pitch_frequency=0.01, pivot_x=2.5, force_eval_method=cf.FORCE_EVAL_MARKER, name="pitching_amr")
    gpu, orc = run_pair(c, [((28, 28, 0), (75, 67, 0))], 10, regrid_int=2,
                        regrid=dict(threshold_ratio=(0.7, 0.7), blocking_factor=4))
    assert check_boxes(gpu, orc) >= 1
    check_levels(gpu, orc)
    check_record(gpu, orc)


def test_two_level_sphere_regrid_bitexact():
    c = cf.sphere_3d(nx=32, ny=24, nz=24, D=6.0, center=(10.0, 12.0, 12.0), U=0.04, Re=30.0)
    gpu, orc = run_pair(c, [((8, 10, 10), (35, 37, 37))], 4, regrid_int=2,
                        regrid=dict(threshold_ratio=(0.5, 0.5), blocking_factor=4))
    assert check_boxes(gpu, orc) >= 1
    check_levels(gpu, orc)
    check_record(gpu, orc)
