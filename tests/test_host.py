"""Host-side checks that need no GPU: the C-ABI library loads and exports
every entry point include/ilbm.h declares, the binding's struct layout
matches the compiled one, and the inputs reader maps reference-style input
files onto the case description the way ReadParameters does."""
import ctypes as C
import os
import re
import subprocess
import sys

import numpy as np
import pytest

from ilbmrex_amd import config as cf

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "ilbm.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ilbm_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    from ilbmrex_amd import native
    lib = native.load()
    names = declared_functions()
    assert "ilbm_level_create" in names and len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(native.EXPORTED) <= set(names)
    assert lib.ilbm_abi_version() == 3


def test_case_desc_layout_matches_binding():
    from ilbmrex_amd import native
    lib = native.load()
    assert lib.ilbm_case_desc_size() == C.sizeof(native.CaseDesc)


def test_status_strings():
    from ilbmrex_amd import native
    lib = native.load()
    assert lib.ilbm_status_string(0) == b"ok"
    assert lib.ilbm_status_string(2) == b"unsupported"


CYLINDER_INPUTS = """
max_step = 100
geometry.prob_lo = 0.0 0.0 0.0
geometry.prob_hi = 256.0 128.0 8.0
geometry.is_periodic = 0 1 1
amr.n_cell = 256 128 8
amr.max_level = 0
lbmPhysicalParameters.nu = 0.016  # Re = 100
lbmPhysicalParameters.nmac = 6
lbmPhysicalParameters.U0 = 0.05
lbm2d.ndir = 9
lbm2d.wis = 4.0 1.0 1.0 1.0 1.0 1.0 1.0 1.0 1.0
lbm2d.wim = 9.0 9.0 9.0 9.0 9.0 36.0 36.0 36.0 36.0
lbm2d.dirx = 0.0 1.0 0.0 -1.0 0.0 1.0 -1.0 -1.0 1.0
lbm2d.diry = 0.0 0.0 1.0 0.0 -1.0 1.0 1.0 -1.0 -1.0
lbm2d.dirz = 0.0 0.0 0.0 0.0 0.0 0.0 0.0 0.0 0.0
amrbc.bc_lo = 3 0
amrbc.bc_hi = 2 0
amrbc.bc_lo_0_rho_val = 1.0
amrbc.bc_lo_0_ux_val = 0.05
ibm.use_cylinder = 1
ibm.method = iamr_marker
ibm.x0 = 64.0
ibm.y0 = 64.0
ibm.R = 16.0
ibm.delta_type = 4
ibm.loop_ns = 1
ibm.coupling_method = explicit
ibm.force_eval_method = marker
"""


def test_inputs_reader_maps_reference_keys():
    c = cf.case_from_inputs(CYLINDER_INPUTS)
    assert c.lattice is cf.D2Q9 and c.n == (256, 128, 1)
    assert c.periodic[:2] == (0, 1)
    assert c.bc_lo[0] == cf.EXT_DIR and c.bc_hi[0] == cf.FOEXTRAP
    assert c.bc_lo_val[0][:2] == (1.0, 0.05)
    assert c.use_ib and c.R == 16.0 and c.coupling_method == cf.COUPLING_EXPLICIT
    assert c.force_eval_method == cf.FORCE_EVAL_MARKER
    assert abs(c.tau - (3 * 0.016 + 0.5)) < 1e-15
    # command-line style overrides win, as with ParmParse
    c2 = cf.case_from_inputs(CYLINDER_INPUTS, overrides=["ibm.coupling_method = wu_shu",
                                                         "lbmPhysicalParameters.nu = 0.032"])
    assert c2.coupling_method == cf.COUPLING_IVC and c2.nu == 0.032


def test_inputs_reader_rejects_what_the_reference_aborts_on():
    with pytest.raises(ValueError):
        cf.case_from_inputs(CYLINDER_INPUTS, overrides=["ibm.method = sharp"])
    with pytest.raises(ValueError):
        cf.case_from_inputs(CYLINDER_INPUTS, overrides=["ibm.method = bogus"])
    with pytest.raises(ValueError):   # lattice table not one of the reference's
        cf.case_from_inputs(CYLINDER_INPUTS, overrides=["lbm2d.wim = 9 9 9 9 9 9 9 9 9"])


def test_lattice_tables():
    for lat in (cf.D2Q9, cf.D3Q19, cf.D3Q27):
        w = lat.w
        assert abs(sum(w) - 1.0) < 1e-15
        opp = lat.opposite()
        assert all(opp[opp[q]] == q for q in range(lat.q))
        # second moment isotropy: sum w c_a c_b = delta_ab / 3
        for a in (lat.cx, lat.cy, lat.cz)[:lat.dim]:
            assert abs(sum(wi * ca * ca for wi, ca in zip(w, a)) - 1.0 / 3.0) < 1e-15


def test_bench_refuses_several_gpus():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True)
    assert r.returncode == 2 and "one GPU" in r.stderr


def test_footprint_and_argument_checks_without_gpu():
    """Entry points that reject their arguments on the host return a status
    before any device work."""
    from ilbmrex_amd import native
    lib = native.load()
    c = cf.cylinder_2d(nx=256, ny=128, storage=cf.STORE_F32)
    b100 = native.Batch.footprint(c, 100)
    b1000 = native.Batch.footprint(c, 1000)
    big = native.Batch.footprint(c.replace(n=(512, 256, 1)), 100)
    assert c.ncells * 9 * 4 * 2 < b100 < b1000 < big
    assert native.Batch.footprint(c.replace(storage=cf.STORE_F64), 100) > b100
    d = native.to_desc(c)
    lo = (C.c_int * 3)(0, 0, 0)
    hi = (C.c_int * 3)(255, 127, 0)
    assert lib.ilbm_box_collide(C.byref(d), None, None, None, lo, hi, None) == 1
    f = native.Fab()
    f.p = 0
    assert lib.ilbm_box_fill_physical(C.byref(d), C.byref(f), 1, None) == 1
    assert lib.ilbm_level_create(None, None, None) == 1
    assert lib.ilbm_amr_create(C.byref(d), 2, None, None, 2, None, None) == 1
    assert lib.ilbm_batch_create(0, 1, 8, None) == 1


def test_serving_workloads_are_seeded_and_in_range():
    from ilbmrex_amd import serving
    a = serving.ensemble_cases(64, seed=3)
    b = serving.ensemble_cases(64, seed=3)
    assert [c.n for c in a] == [c.n for c in b]
    for c in a:
        nx, ny = c.n[0], c.n[1]
        assert nx == 2 * ny and 64 <= ny <= 512
        D = 2 * c.R
        assert 8.0 <= D <= max(8.0, ny / 4.0) + 1e-9 or D >= 0.08 * 20
        assert c.n_marker == round(3.141592653589793 * D)
        assert c.nu >= 0.004 - 1e-12 and c.storage == cf.STORE_F32
    reqs = serving.serving_requests(400, rate=10.0, seed=5)
    t = np.array([r.arrival for r in reqs])
    assert np.all(np.diff(t) > 0) and abs(len(reqs) / t[-1] - 10.0) < 2.0
    kinds = [r.case.lattice.dim for r in reqs]
    assert 0.6 < kinds.count(2) / len(kinds) < 0.8
    assert all(2000 <= r.nsteps <= 20000 for r in reqs)
    for r in reqs:
        if r.case.lattice.dim == 3:
            assert 32 <= r.case.n[0] <= 192 and r.case.n[0] == r.case.n[1] == r.case.n[2]
        else:
            assert 64 <= r.case.n[0] <= 1024 and r.case.n[0] == r.case.n[1]


def test_fab_descriptor_checks_without_gpu():
    """Per-box entry points refuse non-FP64 fabs and overlapping strides
    before any device work."""
    from ilbmrex_amd import native
    L = native.load()
    f = native.make_fab(0x1000, (0, 0, 0), (7, 5, 0), 2)
    f.dtype = native.FAB_F32
    assert L.ilbm_box_setval(C.byref(f), 0.0, None) == 2          # unsupported
    g = native.make_fab(0x1000, (0, 0, 0), (7, 5, 0), 2, strides=(4, 0, 0))
    assert L.ilbm_box_setval(C.byref(g), 0.0, None) == 1          # rows overlap
    assert C.sizeof(native.Fab) == 8 + 4 * 8 + 3 * 8
