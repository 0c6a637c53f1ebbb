"""GPU runs of the cases the oracle pins by closed forms, against the oracle
and against those closed forms:

* the immersed-box cavity of Examples/Cavity_flow/inputs_ibm_box (IVC, full
  box) and its lid-only variant with explicit forcing;
* the body torque (sum of r x (F dv), IBMarkerDF.cpp:1500-1503) against a
  brute-force sum over the GPU's own marker state;
* plane Couette flow between ext_dir walls (Examples/Couette_flow wall
  conditions): bit-exact against the oracle, steady profile exactly linear;
* Taylor-Green decay (Examples/Taylor_Green_Vortex/Prob.H) on a periodic
  level: bit-exact against the oracle and the analytic decay rate;
* momentum balance of the exported IB force in a periodic box.
"""
import numpy as np
import pytest

from ilbmrex_amd import config as cf
import refcpu
from test_gpu_parity import _ib_checks, assert_bitexact, run_pair

pytestmark = pytest.mark.gpu


def test_cavity_ibm_box_ivc_within_tolerance():
    """Full marker box (256 markers, lid at ux = 0.05) with IVC: the LU is
    applied as an explicit inverse on the device (1e-10 of the oracle)."""
    c = cf.cavity_ibm_box()
    lev, orc = run_pair(c, cf.uniform_init(c), 20)
    assert lev.num_markers() == orc.num_markers() == 256
    xyz, _, fx = lev.markers()
    oxyz, _, ofx = orc.markers()
    assert_bitexact(xyz, oxyz, "marker positions")
    assert np.abs(fx - ofx).max() <= 1e-10 * np.abs(ofx).max()
    f, fo = lev.f(), orc.f()
    assert np.abs(f - fo).max() <= 1e-12 * np.abs(fo).max()
    m, mo = lev.macro(), orc.macro()
    assert np.abs(m[1:3] - mo[1:3]).max() <= 1e-10 * np.abs(mo[1:3]).max()
    # the lid drives the flow to the right under the top side
    assert mo[1].reshape(96, 64)[63, 32] > 0.01


def test_cavity_ibm_lid_only_explicit_bitexact():
    c = cf.cavity_ibm_box(lid_only=1, coupling=cf.COUPLING_EXPLICIT)
    lev, orc = run_pair(c, cf.uniform_init(c), 20)
    assert lev.num_markers() == orc.num_markers() == 65
    _ib_checks(lev, orc)


@pytest.mark.parametrize("kind", ["rotating_cylinder", "sphere"])
def test_marker_torque_matches_brute_force(kind):
    if kind == "rotating_cylinder":
        c = cf.cylinder_2d(nx=96, ny=64, D=16.0, xc=32.0, yc=31.7, U=0.05, Re=100.0).replace(om=(0.0, 0.0, 0.004))
        ctr = np.array([c.x0, c.y0, 0.0])
    else:
        c = cf.sphere_3d(nx=48, ny=32, nz=32, D=10.0, center=(16.0, 16.2, 15.9), U=0.04, Re=50.0).replace(
            om=(0.001, -0.002, 0.003))
        ctr = np.array([c.x0, c.y0, c.z0])
    lev, orc = run_pair(c, cf.uniform_init(c, u=(c.U0, 0.0, 0.0)), 12)
    _ib_checks(lev, orc)
    xyz, _, F = lev.markers()
    want = np.cross(xyz - ctr, F * lev.dv()).sum(axis=0)
    got = lev.torque()
    scale = np.abs(np.cross(xyz - ctr, F * lev.dv())).sum(axis=0).max()
    assert np.abs(got - want).max() <= 1e-13 * scale, (got, want)
    assert np.abs(want).max() > 1e-3 * scale          # a torque well above rounding


def test_couette_bitexact_and_linear():
    c = cf.couette_channel(nx=8, ny=32, nu=0.064, U=0.1)
    lev, orc = run_pair(c, cf.uniform_init(c), 400)
    assert_bitexact(lev.f(), orc.f(), "f")
    assert_bitexact(lev.macro(), orc.macro(), "macro")
    s = cf.couette_channel(nx=4, ny=16, nu=1.0 / 6.0, U=0.1)
    from ilbmrex_amd.native import Level
    lev = Level(s, cf.uniform_init(s))
    lev.advance(8000, want_state=True)
    u = lev.macro()[1].reshape(16, 4)
    j = np.arange(16)
    assert np.abs(u - (0.1 * j / 15.0)[:, None]).max() < 1e-14


def test_taylor_green_decay():
    """Periodic 64^2 Taylor-Green vortex, nu = 1/6 (tau = 1): bit-exact
    against the oracle, and the kinetic energy decays at the analytic rate
    exp(-2 nu (kx^2 + ky^2) t) within 0.2 %."""
    c = cf.periodic_box(cf.D2Q9, (64, 64, 1), nu=1.0 / 6.0, name="taylor_green")
    init = cf.taylor_green_init(c, u_max=0.02)
    lev, orc = run_pair(c, init, 200)
    assert_bitexact(lev.f(), orc.f(), "f")
    assert_bitexact(lev.macro(), orc.macro(), "macro")
    m = lev.macro()
    e0 = (init[1] ** 2 + init[2] ** 2).sum()
    e = (m[1] ** 2 + m[2] ** 2).sum()
    k2 = 2.0 * (2.0 * np.pi / 64.0) ** 2
    want = np.exp(-2.0 * (1.0 / 6.0) * k2 * 200)
    assert abs(e / e0 / want - 1.0) < 2e-3


def test_ib_momentum_balance():
    """All-periodic D2Q9 box with a fixed cylinder: the fluid momentum of
    each step changes by S(n-1)/(2 tau) + (1 - 1/(2 tau)) S(n), S the
    exported force (BGK relaxes to the forced velocity of the last step,
    the Guo term adds the rest)."""
    c = cf.periodic_box(cf.D2Q9, (96, 64, 1), nu=0.02, name="periodic_body").replace(
        use_ib=True, x0=30.0, y0=31.6, R=8.0, coupling_method=cf.COUPLING_EXPLICIT, loop_ns=1,
        delta_type=4, U0=0.04)
    from ilbmrex_amd.native import Level
    lev = Level(c, cf.random_init(c, seed=5, du=0.03))
    tau = lev.info()["tau"]
    prev = None
    worst = 0.0
    for n in range(6):
        lev.advance(1, want_state=True)
        m = lev.macro()
        S = lev.forcing().sum(axis=1)[:2]
        M = (m[0] * m[1:3]).sum(axis=1) - 0.5 * S
        if prev is not None:
            pM, pS = prev
            err = np.abs((M - pM) - (pS / (2 * tau) + (1 - 1 / (2 * tau)) * S)).max()
            worst = max(worst, err / np.abs(S).max())
        prev = (M, S)
    assert worst < 1e-10
