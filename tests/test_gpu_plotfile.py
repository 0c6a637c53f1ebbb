"""Plotfiles and derived fields of GPU levels and hierarchies against the
oracle (FP64): the written rho / u / f planes are the oracle's, the derived
fields of level 0 equal the oracle's UpdateDerivedMacroFields after the
average-down, and the file reads back."""
import numpy as np
import pytest

from ilbmrex_amd import config as cf
from ilbmrex_amd import plotfile as pf
import refcpu

pytestmark = pytest.mark.gpu


def _cyl():
    return cf.cylinder_2d(nx=64, ny=48, D=8.0, xc=20.0, yc=24.0, U=0.05, Re=40.0)


def test_amr_derived_level0_matches_oracle_and_plotfile_round_trip(tmp_path):
    from ilbmrex_amd.native import AmrHierarchy
    c = _cyl()
    box = [((24, 28, 0), (71, 67, 0))]
    print("[plotfile] two-level cylinder, 4 coarse steps", flush=True)
    gpu = AmrHierarchy(c, box, 2, (1.0, 0.05, 0.0, 0.0))
    gpu.advance(4, want_state=True)
    orc = refcpu.AmrOracle(c, box, 2, cf.uniform_init(c, u=(0.05, 0.0, 0.0)))
    orc.advance(4)
    m0 = orc.macro(0, all_comps=True)
    assert np.array_equal(gpu.derived(0), m0[4:6])
    # the finest level's derived fields were last updated after its last
    # sub-step ((step + 1) % regrid_int == 0); its interior cells do not
    # read the coarse-fine ghosts
    d1 = gpu.derived(1).reshape(2, 40, 48)
    o1 = orc.macro(1, all_comps=True).reshape(6, 40, 48)
    assert np.array_equal(d1[:, 1:-1, 1:-1], o1[4:6, 1:-1, 1:-1])
    path = tmp_path / pf.plotfile_name("plt", 4)
    pf.amr_plotfile(path, gpu, 4)
    r = pf.read_plotfile(path)
    names = pf.var_names(c)
    assert r["names"] == names and len(r["levels"]) == 2
    for l in range(2):
        data = r["levels"][l]["data"]
        shp = data.shape[1:]
        rho = orc.macro(l)[0].reshape(shp)
        assert np.array_equal(data[0], rho)
        f = orc.f(l).reshape((c.lattice.q,) + shp)
        assert np.array_equal(data[6:6 + c.lattice.q], f)
    assert r["levels"][1]["lo"] == (24, 28, 0) and r["levels"][1]["hi"] == (71, 67, 0)
    gpu.close()
    orc.close()


def test_level_plotfile_matches_oracle(tmp_path):
    from ilbmrex_amd.native import Level
    c = _cyl()
    init = cf.uniform_init(c, u=(0.05, 0.0, 0.0))
    print("[plotfile] single-level cylinder, 10 steps", flush=True)
    lev = Level(c, init)
    lev.advance(10, want_state=True)
    orc = refcpu.Oracle(c, init)
    orc.advance(10)
    path = tmp_path / pf.plotfile_name("plt", 10)
    pf.level_plotfile(path, lev, 10)
    r = pf.read_plotfile(path)
    data = r["levels"][0]["data"]
    shp = data.shape[1:]
    om = orc.macro()
    assert np.array_equal(data[0], om[0].reshape(shp))
    assert np.array_equal(data[6:15], orc.f().reshape((9,) + shp))
    inside = data[-1] < 0
    assert np.array_equal(data[1][~inside], om[1].reshape(shp)[~inside])
    assert np.all(data[1][inside] == 0.0)
    lev.close()
    orc.close()
