"""Pins of the oracle's immersed-boundary half (IBMarkerDF::update_forcing,
Source/IBM/IBMarkerDF.cpp:611-1917, and ComputeIBForce,
Source/Core/AmrCoreLBM.cpp:456-701) by what the mathematics fixes, on the
CPU.  Every case uses a cell size of 0.5 (so dt = 0.5, dV = 0.25 or 0.125),
so a dropped volume or time-step factor changes the result.

* interpolation reproduces uniform and linear velocity fields exactly
  (partition of unity and zero first moment of the Peskin kernels);
* spreading conserves the force and its first moment:
  sum EulerForce dV = sum F dv, sum x EulerForce dV = sum X F dv, also
  through a periodic boundary (SumBoundary of the ghost images);
* the stored moments are r x (F dv) about the body centre (brute force);
* one explicit iteration on uniform flow gives F = (U_target - U) / dt;
* with rho = 1 the exported Eulerian force equals the marker force sum,
  and Cd = Fx / (1/2 U0^2 D) (2D) or Fx / (1/2 U0^2 pi R^2) (3D);
* momentum balance of the forced collision in a periodic box with a body:
  M(n+1) - M(n) = S(n-1) / (2 tau) + (1 - 1 / (2 tau)) S(n), M the
  post-stream population momentum, S(n) the exported force of step n.

Each pin is also run with a planted mistake in the oracle (the
``refcpu.mutation`` switch) and must then fail.
"""
import math

import numpy as np
import pytest

from ilbmrex_amd import config as cf
import refcpu


def _cyl2d(**kw):
    """Cylinder R = 3 at (12.0, 12.1) in a 32 x 24 domain of 0.5 cells."""
    c = cf.cylinder_2d(nx=64, ny=48, D=6.0, xc=12.0, yc=12.1, U=0.04, Re=40.0, **kw)
    return c.replace(dx=(0.5, 0.5, 1.0))


def _sph3d(**kw):
    c = cf.sphere_3d(nx=40, ny=32, nz=32, D=5.0, center=(8.0, 8.1, 7.9), U=0.04, Re=20.0, **kw)
    return c.replace(dx=(0.5, 0.5, 0.5))


def _centres(case, grown):
    """Cell-centre coordinates per dimension (over the 2-ghost grown box)."""
    out = []
    for d in range(3):
        g = 2 if (grown and d < case.lattice.dim) else 0
        n = case.n[d]
        out.append(case.plo[d] + (np.arange(-g, n + g) + 0.5) * case.dx[d])
    return out


def _linear_field(case, coef):
    """EulerVel over the grown box: comp c = a + b x + c y (+ d z)."""
    x, y, z = _centres(case, True)
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    ev = np.zeros((3,) + X.shape)
    for c in range(3):
        a, bx, by, bz = coef[c]
        ev[c] = a + bx * X + by * Y + (bz * Z if case.lattice.dim == 3 else 0.0)
    if case.lattice.dim == 2:
        ev[2] = 0.0
    return ev


COEF = [(0.03, 1.0e-3, -2.0e-3, 5.0e-4), (-0.02, 4.0e-4, 1.5e-3, -1.0e-3), (0.01, -6.0e-4, 2.0e-4, 8.0e-4)]


def _interp_errors(case, coef):
    o = refcpu.Oracle(case, cf.uniform_init(case))
    xyz, _ = refcpu.build_markers(case)
    U = o.interp_probe(0.0, _linear_field(case, coef))
    o.close()
    ndim = case.lattice.dim
    err = 0.0
    for c in range(ndim):
        a, bx, by, bz = coef[c]
        want = a + bx * xyz[:, 0] + by * xyz[:, 1] + (bz * xyz[:, 2] if ndim == 3 else 0.0)
        err = max(err, np.abs(U[:, c] - want).max())
    return err


@pytest.mark.parametrize("make,delta", [(_cyl2d, 4), (_cyl2d, 3), (_sph3d, 4)])
def test_interpolation_reproduces_linear_fields(make, delta):
    case = make(delta_type=delta)
    assert _interp_errors(case, COEF) < 1e-15
    # uniform field: partition of unity alone
    const = [(c[0], 0.0, 0.0, 0.0) for c in COEF]
    assert _interp_errors(case, const) < 1e-16


def test_interpolation_pin_fails_without_cell_volume():
    with refcpu.mutation(refcpu.mutation.INTERP_NO_DV):
        assert _interp_errors(_cyl2d(), COEF) > 1e-3


def _spread(case, seed=3):
    o = refcpu.Oracle(case, cf.uniform_init(case))
    nm = o.num_markers()
    rng = np.random.default_rng(seed)
    F = rng.normal(size=(nm, 3))
    if case.lattice.dim == 2:
        F[:, 2] = 0.0
    efi, mom = o.spread_probe(0.0, F)
    dv = o.dv()
    o.close()
    xyz, _ = refcpu.build_markers(case)
    return F, efi, mom, dv, xyz


def _moment_errors(case, F, efi, dv, xyz, first=True):
    dV = float(np.prod(case.dx[:case.lattice.dim]))
    ndim = case.lattice.dim
    x, y, z = _centres(case, False)
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    pos = [X.reshape(-1), Y.reshape(-1), Z.reshape(-1)]
    err = 0.0
    for c in range(ndim):
        s0 = efi[c].sum() * dV
        w0 = (F[:, c] * dv).sum()
        err = max(err, abs(s0 - w0) / np.abs(F[:, c] * dv).sum())
        if first:
            for d in range(ndim):
                s1 = (pos[d] * efi[c]).sum() * dV
                w1 = (xyz[:, d] * F[:, c] * dv).sum()
                err = max(err, abs(s1 - w1) / np.abs(xyz[:, d] * F[:, c] * dv).sum())
    return err


@pytest.mark.parametrize("make", [_cyl2d, _sph3d])
def test_spreading_conserves_force_and_first_moment(make):
    case = make()
    F, efi, mom, dv, xyz = _spread(case)
    assert _moment_errors(case, F, efi, dv, xyz) < 1e-13


def test_spreading_conserves_force_through_periodic_boundary():
    """A body whose stencils wrap through the periodic y boundary: the ghost
    images are summed into the valid cells (SumBoundary), so the force is
    still conserved."""
    case = cf.cylinder_2d(nx=64, ny=32, D=12.0, xc=12.0, yc=6.3, U=0.04, Re=60.0).replace(dx=(0.5, 0.5, 1.0))
    F, efi, mom, dv, xyz = _spread(case)
    assert xyz[:, 1].min() >= 0.0              # markers inside the domain,
    assert (xyz[:, 1] < 1.0).any()             # their stencils reach the ghost images
    assert _moment_errors(case, F, efi, dv, xyz, first=False) < 1e-13


def test_spreading_pin_fails_without_marker_volume():
    case = _cyl2d()
    with refcpu.mutation(refcpu.mutation.SPREAD_NO_DV):
        F, efi, mom, dv, xyz = _spread(case)
    assert _moment_errors(case, F, efi, dv, xyz) > 1e-2


def _torque_error(case):
    F, efi, mom, dv, xyz = _spread(case, seed=9)
    ctr = np.array([case.x0, case.y0, case.z0 if case.lattice.dim == 3 else 0.0])
    r = xyz - ctr
    want = np.cross(r, F * dv)                 # brute force, per marker
    return np.abs(mom - want).max() / np.abs(want).max()


@pytest.mark.parametrize("make", [_cyl2d, _sph3d])
def test_moments_are_r_cross_f_dv(make):
    assert _torque_error(make()) < 1e-14


def test_moment_pin_fails_with_swapped_cross_product():
    with refcpu.mutation(refcpu.mutation.MOMENT_SIGN):
        assert _torque_error(_cyl2d()) > 1.0


def _one_step_uniform(case, u):
    o = refcpu.Oracle(case, cf.uniform_init(case, u=u))
    o.advance(1)
    _, _, F = o.markers()
    rec = o.force_record()
    dv = o.dv()
    o.close()
    return F, rec, dv


@pytest.mark.parametrize("make", [_cyl2d, _sph3d])
def test_one_explicit_iteration_on_uniform_flow(make):
    """Fixed body in uniform flow u: U interpolates to u, so the one
    iteration gives F = (0 - u) / dt on every marker, and (rho = 1) the
    exported Eulerian force equals the marker force sum."""
    case = make()
    u = (0.04, 0.01, -0.02 if case.lattice.dim == 3 else 0.0)
    dt = case.dx[0]
    F, rec, dv = _one_step_uniform(case, u)
    for c in range(case.lattice.dim):
        assert np.abs(F[:, c] + u[c] / dt).max() < 1e-14 * abs(u[c]) / dt
    nm = F.shape[0]
    fx_eul, fy_eul, fx_m, fy_m = rec[10], rec[11], rec[14], rec[15]
    assert abs(fx_m - nm * dv * u[0] / dt) < 1e-13 * abs(fx_m)
    assert abs(fx_eul - fx_m) < 1e-12 * abs(fx_m)
    assert abs(fy_eul - fy_m) < 1e-12 * abs(fy_m)
    if case.lattice.dim == 3:
        assert abs(rec[27] - rec[29]) < 1e-12 * abs(rec[29])     # Fz_eul, Fz_marker
    # drag coefficient of the Eulerian force (column Cd_eul)
    area = math.pi * case.R ** 2 if case.lattice.dim == 3 else 2.0 * case.R
    assert abs(rec[12] * (0.5 * case.U0 ** 2 * area) - fx_eul) < 1e-14 * abs(fx_eul)


@pytest.mark.parametrize("which", ["FORCE_SIGN", "EXPORT_NO_DT", "CD_NO_HALF"])
def test_uniform_flow_pins_fail_on_planted_mistakes(which):
    with refcpu.mutation(getattr(refcpu.mutation, which)):
        with pytest.raises(AssertionError):
            test_one_explicit_iteration_on_uniform_flow(_cyl2d)


def _periodic_body(dim):
    if dim == 2:
        c = cf.periodic_box(cf.D2Q9, (48, 40, 1), nu=0.02, name="periodic_body2d")
        return c.replace(dx=(0.5, 0.5, 1.0), nu=0.01, use_ib=True, x0=12.0, y0=10.1, R=3.0,
                         coupling_method=cf.COUPLING_EXPLICIT, loop_ns=1, delta_type=4, U0=0.04)
    c = cf.periodic_box(cf.D3Q19, (24, 20, 20), nu=0.02, name="periodic_body3d")
    return c.replace(dx=(0.5, 0.5, 0.5), nu=0.01, use_ib=True, x0=6.0, y0=5.1, z0=4.9, R=2.0,
                     coupling_method=cf.COUPLING_EXPLICIT, loop_ns=1, delta_type=4, U0=0.04)


def _momentum_balance_error(dim, nsteps=6):
    case = _periodic_body(dim)
    o = refcpu.Oracle(case, cf.random_init(case, seed=5, du=0.03))
    tau = o.tau()
    ndim = case.lattice.dim
    prev_M = prev_S = None
    err, scale = 0.0, 0.0
    for n in range(nsteps):
        o.advance(1)
        m = o.macro()
        S = o.forcing().sum(axis=1)[:ndim]
        P = (m[0] * m[1:1 + ndim]).sum(axis=1)
        M = P - 0.5 * S                 # post-stream population momentum
        if prev_M is not None:
            expect = prev_S / (2.0 * tau) + (1.0 - 1.0 / (2.0 * tau)) * S
            err = max(err, np.abs((M - prev_M) - expect).max())
            scale = max(scale, np.abs(S).max())
        prev_M, prev_S = M, S
    o.close()
    return err / scale


@pytest.mark.parametrize("dim", [2, 3])
def test_momentum_balance_of_ib_forcing(dim):
    assert _momentum_balance_error(dim) < 1e-10


def test_momentum_balance_fails_without_half_force_in_macros():
    with refcpu.mutation(refcpu.mutation.GUO_NO_HALF):
        assert _momentum_balance_error(2) > 1e-3
