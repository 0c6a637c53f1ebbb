"""Pins of the oracle's AMR pieces (CPU).

* cell_cons_interp: linear fields reproduced exactly, the 2^d children
  average to their parent, no new extrema (children stay within the
  parent's neighbour range), flat children at an extremum.  The AMReX
  version the reference builds against is unknown: the limiter's exact
  form is parity unpinned beyond these properties.
* A uniform flow is a fixed point of a two- and three-level hierarchy.
* After a coarse step, covered coarse cells hold the average of their fine
  children (AverageDownTo).
* Taylor-Green vortex (Examples/Taylor_Green_Vortex/Prob.H) with a refined
  patch decays at the analytic viscous rate.
* The immersed body lives on the finest level: markers are built with the
  fine spacing.
"""
import math

import numpy as np
import pytest

from ilbmrex_amd import config as cf
import refcpu


def _lin(shape, coef):
    nz, ny, nx = shape
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    return coef[0] + coef[1] * (i + 0.5) + coef[2] * (j + 0.5) + coef[3] * (k + 0.5)


@pytest.mark.parametrize("dim", [2, 3])
def test_cons_interp_reproduces_linear_fields(dim):
    nz = 5 if dim == 3 else 1
    coefs = [(1.0, 0.3, -0.2, 0.1 if dim == 3 else 0.0), (0.05, -0.01, 0.02, 0.003 if dim == 3 else 0.0)]
    crse = np.stack([_lin((nz, 5, 5), c) for c in coefs])
    clo = (10, 20, 30 if dim == 3 else 0)
    ic = (12, 22, 32 if dim == 3 else 0)
    for di in range(2):
        for dj in range(2):
            for dk in range(2 if dim == 3 else 1):
                fiv = (2 * ic[0] + di, 2 * ic[1] + dj, (2 * ic[2] + dk) if dim == 3 else 0)
                got = refcpu.cons_interp(crse, clo, dim, fiv)
                # fine cell centre in coarse-cell units, relative to clo
                xc = [(fiv[d] + 0.5) / 2.0 - clo[d] for d in range(dim)] + [0.5] * (3 - dim)
                for n, c in enumerate(coefs):
                    exp = c[0] + c[1] * xc[0] + c[2] * xc[1] + c[3] * xc[2]
                    assert abs(got[n] - exp) < 1e-14 * max(1.0, abs(exp))


@pytest.mark.parametrize("dim", [2, 3])
def test_cons_interp_conserves_and_adds_no_extrema(dim):
    rng = np.random.default_rng(5)
    nz = 3 if dim == 3 else 1
    for trial in range(50):
        crse = 1.0 + rng.uniform(-0.2, 0.2, size=(3, nz, 3, 3))
        clo = (4, 7, 9 if dim == 3 else 0)
        ic = (5, 8, 10 if dim == 3 else 0)
        kids = []
        for di in range(2):
            for dj in range(2):
                for dk in range(2 if dim == 3 else 1):
                    fiv = (2 * ic[0] + di, 2 * ic[1] + dj, (2 * ic[2] + dk) if dim == 3 else 0)
                    kids.append(refcpu.cons_interp(crse, clo, dim, fiv))
        kids = np.array(kids)
        centre = crse[:, nz // 2, 1, 1]
        assert np.allclose(kids.mean(axis=0), centre, rtol=0, atol=4e-16 * 2)
        # the limited slope keeps children within the face-neighbour range
        for n in range(3):
            nb = [crse[n, nz // 2, 1, 0], crse[n, nz // 2, 1, 2], crse[n, nz // 2, 0, 1], crse[n, nz // 2, 2, 1]]
            if dim == 3:
                nb += [crse[n, 0, 1, 1], crse[n, 2, 1, 1]]
            lo, hi = min(nb + [centre[n]]), max(nb + [centre[n]])
            span = sum(abs(x - centre[n]) for x in nb)
            assert kids[:, n].min() >= lo - span and kids[:, n].max() <= hi + span


def test_cons_interp_flat_at_extremum():
    crse = np.ones((1, 1, 3, 3))
    crse[0, 0, 1, 1] = 2.0            # a maximum in both directions
    kids = [refcpu.cons_interp(crse, (0, 0, 0), 2, (2 + di, 2 + dj, 0))[0] for di in range(2) for dj in range(2)]
    assert kids == [2.0] * 4


def _uniform_periodic(dim, n, u):
    lat = cf.D2Q9 if dim == 2 else cf.D3Q19
    c = cf.periodic_box(lat, n, nu=0.05)
    return c, cf.uniform_init(c, rho=1.0, u=u)


def test_uniform_flow_is_fixed_point_of_hierarchy():
    c, init = _uniform_periodic(2, (32, 24, 1), (0.05, 0.02, 0.0))
    boxes = [((16, 12, 0), (47, 35, 0)), ((40, 32, 0), (63, 55, 0))]
    A = refcpu.AmrOracle(c, boxes, regrid_int=2, init_macro=init)
    A.advance(6)
    for l in range(3):
        m = A.macro(l)
        assert np.abs(m[0] - 1.0).max() < 1e-14
        assert np.abs(m[1] - 0.05).max() < 1e-15 and np.abs(m[2] - 0.02).max() < 1e-15


def test_covered_cells_hold_average_of_children():
    c = cf.periodic_box(cf.D2Q9, (24, 20, 1), nu=0.04)
    init = cf.random_init(c, seed=8, du=0.02)
    box = ((12, 10, 0), (31, 27, 0))
    A = refcpu.AmrOracle(c, [box], regrid_int=3, init_macro=init)
    A.advance(3)
    f0, f1 = A.f(0).reshape(9, 20, 24), A.f(1).reshape(9, 18, 20)
    m0, m1 = A.macro(0, True).reshape(6, 20, 24), A.macro(1, True).reshape(6, 18, 20)
    for jc in range(5, 14):
        for ic in range(6, 16):
            fj, fi = 2 * jc - 10, 2 * ic - 12
            kids = f1[:, fj:fj + 2, fi:fi + 2]
            c_sum = ((kids[:, 0, 0] + kids[:, 0, 1]) + kids[:, 1, 0]) + kids[:, 1, 1]
            assert np.array_equal(f0[:, jc, ic], 0.25 * c_sum)
            mk = m1[:4, fj:fj + 2, fi:fi + 2]
            m_sum = ((mk[:, 0, 0] + mk[:, 0, 1]) + mk[:, 1, 0]) + mk[:, 1, 1]
            assert np.array_equal(m0[:4, jc, ic], 0.25 * m_sum)
    # derived fields of level 0 were refreshed after the average-down:
    # pressure = (rho - 1) / 3
    assert np.array_equal(m0[5], (m0[0] - 1.0) / 3.0)


def _tgv_amplitude(m, case):
    nx, ny = case.n[0], case.n[1]
    kx, ky = 2 * math.pi / nx, 2 * math.pi / ny
    x = 0.5 + np.arange(nx)
    y = 0.5 + np.arange(ny)
    X, Y = np.meshgrid(x, y)
    shape = -np.cos(kx * X) * np.sin(ky * Y)
    u = m[1].reshape(ny, nx)
    return float((u * shape).sum() / (shape * shape).sum())


@pytest.mark.slow
def test_taylor_green_decay_with_refined_patch():
    c = cf.periodic_box(cf.D2Q9, (64, 64, 1), nu=1.0 / 6.0)
    init = cf.taylor_green_init(c)
    box = ((32, 32, 0), (95, 95, 0))      # the middle quarter refined
    A = refcpu.AmrOracle(c, [box], regrid_int=2, init_macro=init)
    a0 = _tgv_amplitude(A.macro(0), c)
    steps = 150
    A.advance(steps)
    a1 = _tgv_amplitude(A.macro(0), c)
    k2 = 2 * (2 * math.pi / 64) ** 2
    expect = math.exp(-(1.0 / 6.0) * k2 * steps)
    assert abs(a0 - 0.02) < 1e-4
    assert abs(a1 / a0 - expect) < 0.005 * expect   # 0.14 % here; single level 0.003 %


def test_body_markers_use_finest_spacing():
    c = cf.cylinder_2d(nx=64, ny=48, D=8.0, xc=24.0, yc=24.0, U=0.05, Re=40.0)
    init = cf.uniform_init(c, u=(0.05, 0.0, 0.0))
    box = ((28, 28, 0), (67, 67, 0))
    A = refcpu.AmrOracle(c, [box], regrid_int=2, init_macro=init)
    assert A.num_markers() == math.ceil(2 * math.pi * 4.0 / 0.5)
    A.advance(4)
    rec = A.force_record()
    assert len(rec) == 23 and rec[0] == 4 and rec[1] == 4.0
    assert np.isfinite(rec).all() and rec[4] > 0.0


# ---------------------------------------------------------------------------
# Regrid (ErrorEst + one box per level + RemakeLevel)
# ---------------------------------------------------------------------------

def _cyl():
    return cf.cylinder_2d(nx=64, ny=48, D=8.0, xc=20.0, yc=24.0, U=0.05, Re=40.0)


def _feq(case, m):
    lat = case.lattice
    r = np.maximum(m[0], 1.0e-12)
    u, v, w = m[1], m[2], m[3]
    u2 = u * u + v * v + w * w
    out = []
    for q in range(lat.q):
        cu = float(lat.cx[q]) * u + float(lat.cy[q]) * v + float(lat.cz[q]) * w
        out.append(lat.w[q] * r * (1.0 + 3.0 * cu + 4.5 * cu * cu - 1.5 * u2))
    return np.stack(out)


def _expected_region_box(case, bf, buf=1):
    """Level-1 box from the body refinement region alone: cells whose centre
    lies in [c - R - pad, c + R + pad] (pad = max(6 dx, 2)), grown by the
    error buffer, clipped 3 cells inside the domain, aligned outward to
    blocking_factor / 2 coarse cells; refined by 2."""
    pad = max(6.0, 2.0)
    centre = (case.x0, case.y0)
    lo, hi = [], []
    bu = bf // 2
    for d in range(2):
        a = math.ceil(centre[d] - case.R - pad - 0.5) - buf
        b = math.floor(centre[d] + case.R + pad - 0.5) + buf
        a, b = max(a, 3), min(b, case.n[d] - 1 - 3)
        a, b = (a // bu) * bu, (b // bu + 1) * bu - 1
        lo.append(2 * a)
        hi.append(2 * b + 1)
    return (lo[0], lo[1], 0), (hi[0], hi[1], 0)


def test_regrid_box_of_body_region():
    c = _cyl()
    o = refcpu.AmrOracle(c, [((24, 28, 0), (71, 67, 0))], 2, cf.uniform_init(c, u=(0.05, 0.0, 0.0)))
    o.advance(4)
    # a threshold far above the vorticity maximum leaves the body region
    # (and the level-set band inside it) as the only tags
    o.set_regrid(threshold_ratio=(1.0e30, 1.0e30), blocking_factor=8)
    o.regrid_now(0)
    assert o.boxes[0] == _expected_region_box(c, 8)
    assert o.regrids() == 1
    o.close()


def test_remake_keeps_old_data_interpolates_new_cells_and_sets_feq():
    c = _cyl()
    box = ((24, 28, 0), (71, 67, 0))
    o = refcpu.AmrOracle(c, [box], 2, cf.uniform_init(c, u=(0.05, 0.0, 0.0)))
    o.advance(6)
    old = o.macro(1, all_comps=True).reshape(c.nmac, box[1][1] - box[0][1] + 1, box[1][0] - box[0][0] + 1)
    crse = o.macro(0, all_comps=True).reshape(c.nmac, c.n[1], c.n[0])
    o.set_regrid(threshold_ratio=(0.5, 0.5), blocking_factor=4)
    o.regrid_now(0)
    (nlo, nhi), = o.boxes
    assert (nlo, nhi) != box
    ny, nx = nhi[1] - nlo[1] + 1, nhi[0] - nlo[0] + 1
    new = o.macro(1).reshape(4, ny, nx)
    f = o.f(1).reshape(c.lattice.q, ny, nx)
    nkept = ninterp = 0
    for j in range(nlo[1], nhi[1] + 1):
        for i in range(nlo[0], nhi[0] + 1):
            got = new[:, j - nlo[1], i - nlo[0]]
            if box[0][0] <= i <= box[1][0] and box[0][1] <= j <= box[1][1]:
                assert np.array_equal(got, old[:4, j - box[0][1], i - box[0][0]])
                nkept += 1
            else:
                ic, jc = i // 2, j // 2
                blk = crse[:, jc - 1:jc + 2, ic - 1:ic + 2][:, None]
                want = refcpu.cons_interp(blk, (ic - 1, jc - 1, 0), 2, (i, j, 0))
                assert np.array_equal(got, want[:4]), (i, j)
                ninterp += 1
    assert nkept > 0 and ninterp > 0
    np.testing.assert_allclose(f, _feq(c, new), rtol=1e-15, atol=0.0)
    o.close()


def test_regrid_schedule_counts_and_nesting_three_levels():
    c = cf.cylinder_2d(nx=64, ny=48, D=8.0, xc=20.0, yc=24.0, U=0.05, Re=40.0)
    boxes = [((16, 16, 0), (79, 79, 0)), ((48, 48, 0), (127, 127, 0))]
    o = refcpu.AmrOracle(c, boxes, 2, cf.uniform_init(c, u=(0.05, 0.0, 0.0)))
    o.set_regrid(threshold_ratio=(0.5, 0.5, 0.5), blocking_factor=4)
    o.advance(6)
    assert o.regrids() >= 1
    (b1lo, b1hi), (b2lo, b2hi) = o.boxes
    for d in range(2):
        # whole coarse cells, blocking-aligned, 3 cells of nesting margin
        assert b1lo[d] % 4 == 0 and (b1hi[d] + 1) % 4 == 0
        assert b2lo[d] % 4 == 0 and (b2hi[d] + 1) % 4 == 0
        assert b1lo[d] // 2 >= 3 and b1hi[d] // 2 <= c.n[d] - 1 - 3
        assert b2lo[d] // 2 - 3 >= b1lo[d] and b2hi[d] // 2 + 3 <= b1hi[d]
    assert np.isfinite(o.force_record()).all()
    o.close()
