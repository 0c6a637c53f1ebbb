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
