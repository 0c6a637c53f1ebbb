"""Per-box entry points (ilbm_box_*) on reference-layout device arrays.

AdvancePhiAtLevel composed box call by box call (ilbmrex_amd.boxpath) must
reproduce the oracle bit for bit for f_new and macro_new, for every BC type;
the derived fields are checked on the Taylor-Green field, whose vorticity
and pressure are known in closed form.
"""
import math

import numpy as np
import pytest

from ilbmrex_amd import config as cf
import refcpu

pytestmark = pytest.mark.gpu


def _pair(case, init, nsteps, force=None):
    from ilbmrex_amd.boxpath import BoxLevel
    print(f"[box] {case.name} {case.lattice.name} n={case.n} steps={nsteps}", flush=True)
    b = BoxLevel(case, init, force=force)
    b.advance(nsteps)
    o = refcpu.Oracle(case, init)
    o.advance(nsteps)
    return b, o


def _same(a, b, what):
    if not np.array_equal(a, b):
        d = np.abs(a - b)
        raise AssertionError(f"{what}: {np.count_nonzero(a != b)} differ, max {d.max():.3e}")


@pytest.mark.parametrize("name", ["periodic_forced", "channel", "walls_user1", "duct_user2"])
def test_box_path_matches_oracle(name):
    if name == "periodic_forced":
        c = cf.periodic_box(cf.D2Q9, (40, 24, 1), nu=0.03, force=(1e-5, 2e-6, 0.0))
        init, force, n = cf.random_init(c, seed=4), (1e-5, 2e-6, 0.0), 15
    elif name == "channel":
        c = cf.cylinder_2d(nx=48, ny=32).replace(use_ib=False, name="channel")
        init, force, n = cf.random_init(c, seed=5, du=0.02), None, 15
    elif name == "walls_user1":
        z4 = (1.0, 0.0, 0.0, 0.0)
        c = cf.Case(lattice=cf.D2Q9, n=(20, 16, 1), nu=0.05, periodic=(1, 0, 1),
                    bc_lo=(0, cf.USER_1, 0), bc_hi=(0, cf.USER_1, 0),
                    bc_lo_val=((0.0,) * 4, z4, (0.0,) * 4), bc_hi_val=((0.0,) * 4, z4, (0.0,) * 4),
                    prescribed_force=(1e-5, 0.0, 0.0), name="walls_user1")
        init, force, n = cf.random_init(c, seed=6, du=0.01), (1e-5, 0.0, 0.0), 12
    else:
        c = cf.square_duct().replace(n=(6, 12, 12), name="duct_user2")
        init, force, n = cf.random_init(c, seed=7, du=0.01), c.prescribed_force, 8
    b, o = _pair(c, init, n, force)
    _same(b.f(), o.f(), "f_new")
    _same(b.macro(), o.macro(), "macro_new")
    assert abs(b.mass_sum() - o.mass_sum()) <= 1e-12 * abs(o.mass_sum())


def test_box_derived_fields_taylor_green():
    from ilbmrex_amd.boxpath import BoxLevel
    c = cf.periodic_box(cf.D2Q9, (64, 64, 1), nu=1.0 / 6.0)
    init = cf.taylor_green_init(c)
    b = BoxLevel(c, init)
    vort, P = b.derived()
    k = 2 * math.pi / 64
    x = 0.5 + np.arange(64)
    X, Y = np.meshgrid(x, x)
    # u = -U cos kx sin ky, v = U sin kx cos ky: omega = dv/dx - du/dy = 2 U k cos kx cos ky
    w_exact = 2 * 0.02 * k * np.cos(k * X) * np.cos(k * Y)
    # centred differences: exact up to sin(k)/k
    assert np.abs(vort.reshape(64, 64) - w_exact * math.sin(k) / k).max() < 1e-15 * 64
    assert np.array_equal(P, (init[0] - 1.0) / 3.0)


@pytest.mark.parametrize("which", ["cylinder", "duct3d"])
def test_level_derived_fields_match_oracle(which):
    """UpdateDerivedMacroFields of a resident level against the oracle's
    (a one-level hierarchy refreshes them every step with regrid_int = 1)."""
    from ilbmrex_amd.native import Level
    if which == "cylinder":
        c = cf.cylinder_2d(nx=64, ny=40, D=8.0, xc=20.0, yc=21.0, U=0.05, Re=50.0)
        init, n = cf.uniform_init(c, u=(0.05, 0.0, 0.0)), 20
    else:
        c = cf.square_duct().replace(n=(6, 12, 12))
        init, n = cf.random_init(c, seed=9, du=0.01), 6
    lev = Level(c, init)
    lev.advance(n, want_state=True)
    got = lev.derived()
    orc = refcpu.AmrOracle(c, [], 1, init)
    orc.advance(n)
    ref = orc.macro(0, all_comps=True)[4:6]
    _same(got, ref, "vorticity / pressure")


# ---------------------------------------------------------------------------
# IBMarkerDF through the per-box ABI (ilbm_ib_*), strided fabs, and the
# coarse-fine entry points driving a two-level hierarchy
# ---------------------------------------------------------------------------

def _ib_same(b, o):
    _same(b.f(), o.f(), "f_new")
    _same(b.macro(), o.macro(), "macro_new")
    _same(b.forcing(), o.forcing(), "forcing")
    xyz, uvw, fxyz = b.markers()
    oxyz, ouvw, ofxyz = o.markers()
    _same(xyz, oxyz, "marker positions")
    _same(uvw, ouvw, "marker U")
    _same(fxyz, ofxyz, "marker F")
    rec = o.force_record()
    assert b.record.shape == rec.shape and b.record[0] == rec[0] and b.record[1] == rec[1]
    floor = 1e-10 * np.abs(rec[2:]).max()
    assert np.all(np.abs(b.record[2:] - rec[2:]) <= 1e-12 * np.abs(rec[2:]) + floor), (b.record, rec)


@pytest.mark.parametrize("name", ["cylinder", "cylinder_diag_loop2", "sphere", "pitching", "cavity_lid"])
def test_box_path_ib_matches_oracle(name):
    if name == "cylinder":
        c = cf.cylinder_2d(nx=96, ny=64, D=16.0, xc=32.0, yc=31.9, U=0.05, Re=100.0)
        n = 12
    elif name == "cylinder_diag_loop2":
        c = cf.cylinder_2d(nx=64, ny=48, D=12.0, xc=24.0, yc=24.3, coupling=cf.COUPLING_EXPLICIT_DIAG,
                           loop_ns=2, delta_type=3)
        n = 10
    elif name == "sphere":
        c = cf.sphere_3d(nx=40, ny=28, nz=28, D=8.0, center=(14.0, 14.2, 13.9), U=0.04, Re=50.0)
        n = 5
    elif name == "pitching":
        c = cf.cylinder_2d(nx=96, ny=64, D=20.0, xc=30.0, yc=32.0, U=0.03, Re=200.0).replace(
            geometry_type=cf.GEOM_USER, airfoil_chord=20.0, pitch_frequency=0.01, pivot_x=5.0,
            force_eval_method=cf.FORCE_EVAL_MARKER, name="pitching")
        n = 8
    else:
        c = cf.cavity_ibm_box(lid_only=1, coupling=cf.COUPLING_EXPLICIT)
        n = 8
    init = cf.uniform_init(c, u=(c.bc_lo_val[0][1], 0.0, 0.0))
    b, o = _pair(c, init, n)
    _ib_same(b, o)
    b.close()


def test_box_path_ivc_within_tolerance():
    c = cf.cylinder_2d(nx=80, ny=64, D=16.0, xc=28.0, yc=32.2, coupling=cf.COUPLING_IVC)
    b, o = _pair(c, cf.uniform_init(c, u=(0.05, 0.0, 0.0)), 8)
    assert np.abs(b.f() - o.f()).max() <= 1e-12 * np.abs(o.f()).max()
    _, _, fx = b.markers()
    _, _, ofx = o.markers()
    assert np.abs(fx - ofx).max() <= 1e-10 * np.abs(ofx).max()
    b.close()


def test_box_path_padded_strides():
    """Every array with padded rows and planes, described by explicit
    strides, and the force scratch used through component aliases."""
    from ilbmrex_amd.boxpath import BoxLevel
    c = cf.cylinder_2d(nx=70, ny=40, D=10.0, xc=22.0, yc=20.3, U=0.05, Re=60.0)
    init = cf.uniform_init(c, u=(0.05, 0.0, 0.0))
    b = BoxLevel(c, init, pad=True)
    assert b.f_new.strides[0] > b.f_new.n[0]
    b.advance(10)
    o = refcpu.Oracle(c, init)
    o.advance(10)
    _ib_same(b, o)
    b.close()


@pytest.mark.parametrize("pad", [False, True])
def test_box_hierarchy_two_levels_matches_oracle(pad):
    """A two-level hierarchy built and stepped only through the per-box ABI
    (ilbm_box_interp_from_coarse, ilbm_box_cf_fill, ilbm_box_average_down,
    ilbm_ib_update_forcing on the fine box) against the oracle's hierarchy:
    both levels bit-exact, the force record at 1e-12."""
    from ilbmrex_amd.boxpath import BoxHierarchy
    c = cf.cylinder_2d(nx=64, ny=48, D=8.0, xc=20.0, yc=24.0, U=0.05, Re=40.0)
    box = ((24, 28, 0), (71, 67, 0))
    n = 6
    print(f"[box] two-level cylinder pad={pad}, {n} coarse steps", flush=True)
    h = BoxHierarchy(c, box, regrid_int=2, pad=pad)
    h.advance(n)
    o = refcpu.AmrOracle(c, [box], 2, cf.uniform_init(c, u=(0.05, 0.0, 0.0)))
    o.advance(n)
    for lev, bl in enumerate((h.c0, h.c1)):
        _same(bl.f(), o.f(lev), f"level {lev} f")
        _same(bl.macro(6), o.macro(lev, all_comps=True), f"level {lev} macro")
    rec = o.force_record()
    floor = 1e-10 * np.abs(rec).max()
    assert np.all(np.abs(h.c1.record - rec) <= 1e-12 * np.abs(rec) + floor), (h.c1.record, rec)
    h.close()
