"""Pins of the oracle's ext_dir walls (applyExtDirMesoscopic,
Source/BC/bc_fill.H:159-204, and macroBcFill, :263-301) by plane Couette
flow, and of the pitching-section hierarchy's stability limit, on the CPU.

Couette: walls at y-lo (at rest) and y-hi (ux = U), both ext_dir as in
Examples/Couette_flow/inputs (amrbc.bc_lo/hi = 2 3, bc_hi_1_ux_val = 0.1,
nu = 0.064), x periodic.  The scheme imposes the wall values on the
boundary cells: the steady state is exactly linear with u(0) = 0 and
u(ny-1) = U, and the start-up transient follows the series of the
reference's notebook (Examples/Couette_flow/Visualization.ipynb cell 4)
with the channel height ny - 1 between the boundary-cell centres.  The
notebook's own error table comes from a three-level AMR run with
foextrap x faces; its walls sit elsewhere, so its digits are not
reproduced here.  A sign error in the non-equilibrium term
(bc_fill.H:202) breaks the linear profile.
"""
import numpy as np
import pytest

from ilbmrex_amd import config as cf
import refcpu


def couette_case(nx, ny, nu, U=0.1):
    lo, hi = (1.0, 0.0, 0.0, 0.0), (1.0, U, 0.0, 0.0)
    z = (0.0,) * 4
    return cf.Case(lattice=cf.D2Q9, n=(nx, ny, 1), nu=nu, periodic=(1, 0, 1),
                   bc_lo=(cf.INT_DIR, cf.EXT_DIR, cf.INT_DIR), bc_hi=(cf.INT_DIR, cf.EXT_DIR, cf.INT_DIR),
                   bc_lo_val=(z, lo, z), bc_hi_val=(z, hi, z), name="couette")


def exact_couette(y, t, U, H, nu, terms=600):
    """Start-up Couette series, u(0, t) = 0, u(H, t) = U, u(y, 0) = 0."""
    n = np.arange(1, terms + 1, dtype=float)[:, None]
    series = ((2.0 * U / np.pi) * ((-1.0) ** n / n) * np.sin(n * np.pi * y[None, :] / H)
              * np.exp(-(n * n) * (np.pi ** 2) * nu * t / (H * H))).sum(axis=0)
    return U * y / H + series


def _steady_profile_error(nsteps=8000):
    ny, U = 16, 0.1
    c = couette_case(4, ny, 1.0 / 6.0, U)
    o = refcpu.Oracle(c, cf.uniform_init(c))
    o.advance(nsteps)
    m = o.macro().reshape(4, ny, 4)
    o.close()
    j = np.arange(ny)
    lin = np.abs(m[1] - (U * j / (ny - 1))[:, None]).max()
    return lin, np.abs(m[2]).max(), np.abs(m[0] - 1.0).max()


def test_couette_steady_profile_is_exactly_linear():
    lin, uy, drho = _steady_profile_error()
    assert lin < 1e-14 and uy < 1e-14 and drho < 1e-13


def test_couette_pin_fails_with_wrong_nonequilibrium_sign():
    with refcpu.mutation(refcpu.mutation.EXTDIR_SIGN):
        lin, _, _ = _steady_profile_error()
    assert lin > 1e-3


def test_couette_transient_follows_series():
    """Examples/Couette_flow parameters (64 cells, nu = 0.064, U = 0.1):
    relative L2 of the profile against the series with H = ny - 1."""
    ny, U, nu = 64, 0.1, 0.064
    c = couette_case(4, ny, nu, U)
    o = refcpu.Oracle(c, cf.uniform_init(c))
    y = np.arange(ny, dtype=float)
    t = 0
    # error bound at each time: the discrete start-up converges to the series
    for T, tol in ((25, 2.0e-2), (100, 5.0e-3), (1000, 5.0e-4)):
        o.advance(T - t)
        t = T
        m = o.macro().reshape(4, ny, 4)
        assert np.ptp(m[1], axis=1).max() == 0.0          # x-invariant
        ex = exact_couette(y, T, U, ny - 1.0, nu)
        err = np.sqrt(np.mean((m[1][:, 0] - ex) ** 2)) / np.sqrt(np.mean(ex ** 2))
        assert err < tol, (T, err)
    o.close()


def test_pitching_section_hierarchy_diverges_at_re1000():
    """The pitching NACA 0012 hierarchy (coarse 512x256, chord 64 finest
    cells, +-10 deg, k = 1) with BGK at Re = 1000 (tau 0.5019 coarse,
    0.5077 finest; 62 lattice Reynolds number per coarse cell) leaves the
    physical range within 20 coarse steps in the FP64 oracle; at Re = 100
    the same hierarchy stays near the inflow speed."""
    peak = {}
    for Re in (1000.0, 100.0):
        c, boxes = cf.pitching_naca_amr(storage=cf.STORE_F64, Re=Re)
        A = refcpu.AmrOracle(c, boxes, 1 << 30, cf.uniform_init(c, u=(0.04, 0.0, 0.0)))
        A.advance(20)
        m = A.macro(2)
        peak[Re] = float(np.nanmax(np.abs(m[1:3]))) if np.isfinite(m).all() else float("inf")
        A.close()
    assert peak[1000.0] > 1.0          # 25x the inflow speed: diverged
    assert peak[100.0] < 2.0 * 0.04
