"""AMReX plotfile writer (CPU): header structure, FAB layout, round trip.

AMReX and yt are absent, so the byte layout is checked against the
published structure of HyperCLaw-V1.1 headers and VisMF version-1 headers
(line by line for a small case) and by reading back what was written;
compatibility with AMReX's own reader is parity unpinned.
"""
import math

import numpy as np

from ilbmrex_amd import config as cf
from ilbmrex_amd import plotfile as pf


def _case2d():
    return cf.cylinder_2d(nx=16, ny=8, D=4.0, xc=6.0, yc=4.0, U=0.05, Re=20.0)


def test_var_names_follow_plotfile_var_names():
    c = _case2d()
    names = pf.var_names(c)
    assert names[:6] == ["rho", "ux", "uy", "uz", "vor", "Pressure"]
    assert names[6:15] == [f"f_new_{i}" for i in range(9)]
    assert names[-1] == "phi" and len(names) == 6 + 9 + 1
    box = c.replace(geometry_type=cf.GEOM_BOX)
    assert "phi" not in pf.var_names(box)


def test_level_data_masks_inside_the_body_and_writes_phi():
    c = _case2d().replace(ub=(0.01, 0.0, 0.0))
    n = c.n[0] * c.n[1]
    rng = np.random.default_rng(3)
    macro, der, f = rng.random((4, n)), rng.random((2, n)), rng.random((9, n))
    d = pf.level_data(c, 0, (0, 0, 0), (15, 7, 0), macro, der, f)
    assert d.shape == (16, 1, 8, 16)
    j, i = np.meshgrid(np.arange(8), np.arange(16), indexing="ij")
    phi = np.sqrt((i + 0.5 - c.x0) ** 2 + (j + 0.5 - c.y0) ** 2) - c.R
    np.testing.assert_array_equal(d[-1, 0], phi)
    inside = phi < 0
    assert inside.any() and (~inside).any()
    assert np.all(d[1, 0][inside] == 0.01) and np.all(d[4, 0][inside] == 0.0)
    np.testing.assert_array_equal(d[1, 0][~inside], macro[1].reshape(8, 16)[~inside])
    np.testing.assert_array_equal(d[6:15, 0], f.reshape(9, 8, 16))
    np.testing.assert_array_equal(d[0, 0], macro[0].reshape(8, 16))


def test_header_and_cell_header_layout(tmp_path):
    c = _case2d().replace(use_ib=False)
    names = pf.var_names(c)
    nc = len(names)
    d0 = np.arange(nc * 8 * 16, dtype=float).reshape(nc, 1, 8, 16)
    d1 = -np.arange(nc * 4 * 8, dtype=float).reshape(nc, 1, 4, 8)
    path = tmp_path / pf.plotfile_name("plt", 20)
    assert path.name == "plt00020"
    pf.write_plotfile(path, c, [((0, 0, 0), (15, 7, 0), d0), ((8, 4, 0), (15, 7, 0), d1)], 20.0, [20, 40])
    h = (path / "Header").read_text().split("\n")
    assert h[0] == "HyperCLaw-V1.1" and h[1] == str(nc) and h[2:2 + nc] == names
    k = 2 + nc
    assert h[k:k + 3] == ["2", "20", "1"]
    assert h[k + 3] == "0 0 " and h[k + 4] == "16 8 "
    assert h[k + 5] == "2 "
    assert h[k + 6] == "((0,0) (15,7) (0,0)) ((0,0) (31,15) (0,0)) "
    assert h[k + 7] == "20 40 "
    assert h[k + 8] == "1 1 " and h[k + 9] == "0.5 0.5 "
    assert h[k + 10:k + 12] == ["0", "0"]
    assert h[k + 12:k + 17] == ["0 1 20", "20", "0 16", "0 8", "Level_0/Cell"]
    assert h[k + 17:k + 22] == ["1 1 20", "40", "4 8", "2 4", "Level_1/Cell"]
    ch = (path / "Level_1" / "Cell_H").read_text().split("\n")
    assert ch[:10] == ["1", "1", str(nc), "0", "(1 0", "((8,4) (15,7) (0,0))", ")", "1",
                       "FabOnDisk: Cell_D_00000 0", ""]
    assert ch[10] == f"1,{nc}"
    mins = ch[11].split(",")
    assert len(mins) == nc + 1 and mins[-1] == "" and float(mins[0]) == d1[0].min()
    raw = (path / "Level_1" / "Cell_D_00000").read_bytes()
    head = raw[:raw.index(b"\n")].decode()
    assert head == f"FAB ((8, (64 11 52 0 1 12 0 1023)),(8, (8 7 6 5 4 3 2 1)))((8,4) (15,7) (0,0)) {nc}"
    body = np.frombuffer(raw[len(head) + 1:], dtype="<f8")
    np.testing.assert_array_equal(body, d1.reshape(-1))


def test_round_trip_3d(tmp_path):
    c = cf.sphere_3d(nx=8, ny=6, nz=4, D=2.0, center=(3.0, 3.0, 2.0), U=0.04, Re=10.0)
    nc = len(pf.var_names(c))
    rng = np.random.default_rng(11)
    d0 = rng.random((nc, 4, 6, 8))
    path = tmp_path / "plt00003"
    pf.write_plotfile(path, c, [((0, 0, 0), (7, 5, 3), d0)], 0.1, [3])
    r = pf.read_plotfile(path)
    assert r["dim"] == 3 and r["names"] == pf.var_names(c)
    assert r["time"] == 0.1 and r["level_steps"] == [3]
    lev = r["levels"][0]
    assert lev["lo"] == (0, 0, 0) and lev["hi"] == (7, 5, 3)
    np.testing.assert_array_equal(lev["data"], d0)
    assert lev["realbox"] == [[0.0, 8.0], [0.0, 6.0], [0.0, 4.0]]
    assert (path / "Header").read_text().split("\n")[2 + nc + 1] == format(0.1, ".17g")
    assert math.isclose(float(format(0.1, ".17g")), 0.1)
