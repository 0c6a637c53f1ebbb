"""ctypes binding of the CPU oracle (oracle/librefcpu.so).

Test infrastructure: imported by tests/, by ``__graft_entry__.smoke`` as the
checker and by the cpu_baseline leg of bench.py.  The product never imports
this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
LIB_PATH = os.path.join(ORACLE_DIR, "librefcpu.so")

D3 = C.c_double * 3
I3 = C.c_int * 3
D27 = C.c_double * 27


class RcCase(C.Structure):
    _fields_ = [
        ("dim", C.c_int), ("n", I3), ("plo", D3), ("dx", D3), ("periodic", I3),
        ("bc_lo", I3), ("bc_hi", I3),
        ("bc_lo_val", (C.c_double * 4) * 3), ("bc_hi_val", (C.c_double * 4) * 3),
        ("q", C.c_int), ("cx", D27), ("cy", D27), ("cz", D27), ("w", D27),
        ("nu", C.c_double), ("U0", C.c_double), ("nmac", C.c_int),
        ("use_prescribed_force", C.c_int), ("prescribed_force", D3),
        ("use_ib", C.c_int), ("x0", C.c_double), ("y0", C.c_double), ("z0", C.c_double),
        ("R", C.c_double), ("coupling_method", C.c_int), ("delta_type", C.c_int),
        ("loop_ns", C.c_int), ("mdf_relax", C.c_double), ("explicit_diag_eps", C.c_double),
        ("geometry_type", C.c_int), ("n_marker", C.c_int), ("rd", C.c_double),
        ("ubx", C.c_double), ("uby", C.c_double), ("ubz", C.c_double),
        ("omx", C.c_double), ("omy", C.c_double), ("omz", C.c_double),
        ("ivc_diag_reg", C.c_double), ("ivc_rebuild_matrix", C.c_int),
        ("renormalize_delta", C.c_int), ("force_eval_method", C.c_int),
        ("airfoil_chord", C.c_double), ("airfoil_thickness", C.c_double),
        ("pitch_mean", C.c_double), ("pitch_amplitude", C.c_double),
        ("pitch_frequency", C.c_double), ("pitch_phase", C.c_double),
        ("pivot_x", C.c_double),
        ("box_lo", D3), ("box_hi", D3), ("box_ds", C.c_double), ("box_lid_only", C.c_int),
        ("box_lid_u", D3), ("box_wall_tol", C.c_double),
    ]


class RcRegrid(C.Structure):
    _fields_ = [("threshold_ratio", C.c_double * 4), ("no_refine_xlo", C.c_double), ("no_refine_xhi", C.c_double),
                ("refine_upstream", C.c_double), ("refine_downstream", C.c_double), ("refine_cross", C.c_double),
                ("blocking_factor", C.c_int), ("n_error_buf", C.c_int)]


def build() -> None:
    subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        P = C.POINTER
        L.rc_create.restype = C.c_void_p
        L.rc_create.argtypes = [P(RcCase), P(C.c_double)]
        L.rc_destroy.argtypes = [C.c_void_p]
        L.rc_step.argtypes = [C.c_void_p]
        L.rc_advance.argtypes = [C.c_void_p, C.c_int]
        L.rc_mass_sum.restype = C.c_double
        L.rc_mass_sum.argtypes = [C.c_void_p]
        for nm in ("rc_get_f", "rc_get_macro", "rc_get_forcing"):
            getattr(L, nm).argtypes = [C.c_void_p, P(C.c_double)]
        L.rc_force_record.argtypes = [C.c_void_p, P(C.c_double), C.c_int]
        L.rc_num_markers.argtypes = [C.c_void_p]
        L.rc_marker_state.argtypes = [C.c_void_p, P(C.c_double), P(C.c_double), P(C.c_double)]
        L.rc_last_marker_force_sum.argtypes = [C.c_void_p, P(C.c_double)]
        L.rc_marker_dv.restype = C.c_double
        L.rc_marker_dv.argtypes = [C.c_void_p]
        L.rc_tau.restype = C.c_double
        L.rc_tau.argtypes = [C.c_void_p]
        L.rc_feq.restype = C.c_double
        L.rc_feq.argtypes = [C.c_double] * 8 + [C.c_int]
        L.rc_delta.restype = C.c_double
        L.rc_delta.argtypes = [C.c_double, C.c_double, C.c_double, C.c_int]
        L.rc_delta_type_to_kernel.argtypes = [C.c_int]
        L.rc_overlap_1d.restype = C.c_double
        L.rc_overlap_1d.argtypes = [C.c_double, C.c_double, C.c_double, C.c_int,
                                    C.c_double, C.c_int, C.c_int]
        L.rc_lu_factorize.argtypes = [P(C.c_double), P(C.c_int), C.c_int, C.c_double]
        L.rc_lu_solve.argtypes = [P(C.c_double), P(C.c_int), P(C.c_double), P(C.c_double), C.c_int]
        L.rc_build_markers.argtypes = [P(RcCase), C.c_double, P(C.c_double), C.c_int, P(C.c_double)]
        L.rc_collide_forced.argtypes = [P(C.c_double), P(C.c_double), C.c_int, C.c_double,
                                        P(C.c_double), P(C.c_double), P(C.c_double), P(C.c_double)] + \
            [C.c_double] * 7 + [C.c_int]
        L.rc_amr_create.restype = C.c_void_p
        L.rc_amr_create.argtypes = [P(RcCase), C.c_int, P(C.c_int), P(C.c_int), C.c_int, P(C.c_double)]
        L.rc_amr_destroy.argtypes = [C.c_void_p]
        L.rc_amr_step.argtypes = [C.c_void_p]
        L.rc_amr_advance.argtypes = [C.c_void_p, C.c_int]
        L.rc_amr_level.restype = C.c_void_p
        L.rc_amr_level.argtypes = [C.c_void_p, C.c_int]
        L.rc_amr_set_regrid.argtypes = [C.c_void_p, P(RcRegrid)]
        L.rc_amr_boxes.argtypes = [C.c_void_p, P(C.c_int), P(C.c_int)]
        L.rc_amr_regrids.argtypes = [C.c_void_p]
        L.rc_amr_regrid_now.argtypes = [C.c_void_p, C.c_int]
        L.rc_get_macro_all.argtypes = [C.c_void_p, P(C.c_double)]
        L.rc_cons_interp_test.argtypes = [P(C.c_double), P(C.c_int), P(C.c_int), C.c_int, C.c_int,
                                          P(C.c_int), P(C.c_double)]
        L.rc_ib_interp_probe.argtypes = [C.c_void_p, C.c_double, P(C.c_double), P(C.c_double)]
        L.rc_ib_spread_probe.argtypes = [C.c_void_p, C.c_double, P(C.c_double), P(C.c_double), P(C.c_double)]
        L.rc_set_mutation.argtypes = [C.c_int]
        _lib = L
    return _lib


def dptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_double))


def to_rc_case(case) -> RcCase:
    """Maps an ilbmrex_amd.config.Case onto the oracle's struct."""
    c = RcCase()
    lat = case.lattice
    c.dim = lat.dim
    c.n = I3(*case.n)
    c.plo = D3(*case.plo)
    c.dx = D3(*case.dx)
    c.periodic = I3(*case.periodic)
    c.bc_lo = I3(*case.bc_lo)
    c.bc_hi = I3(*case.bc_hi)
    for d in range(3):
        for k in range(4):
            c.bc_lo_val[d][k] = case.bc_lo_val[d][k]
            c.bc_hi_val[d][k] = case.bc_hi_val[d][k]
    c.q = lat.q
    w = lat.w
    for q in range(lat.q):
        c.cx[q], c.cy[q], c.cz[q], c.w[q] = float(lat.cx[q]), float(lat.cy[q]), float(lat.cz[q]), w[q]
    c.nu, c.U0, c.nmac = case.nu, case.U0, case.nmac
    if case.prescribed_force is not None:
        c.use_prescribed_force = 1
        c.prescribed_force = D3(*case.prescribed_force)
    c.use_ib = int(case.use_ib)
    c.x0, c.y0, c.z0, c.R = case.x0, case.y0, case.z0, case.R
    c.coupling_method, c.delta_type, c.loop_ns = case.coupling_method, case.delta_type, case.loop_ns
    c.mdf_relax, c.explicit_diag_eps = case.mdf_relax, case.explicit_diag_eps
    c.geometry_type, c.n_marker, c.rd = case.geometry_type, case.n_marker, case.rd
    c.ubx, c.uby, c.ubz = case.ub
    c.omx, c.omy, c.omz = case.om
    c.ivc_diag_reg, c.ivc_rebuild_matrix = case.ivc_diag_reg, case.ivc_rebuild_matrix
    c.renormalize_delta, c.force_eval_method = case.renormalize_delta, case.force_eval_method
    c.airfoil_chord, c.airfoil_thickness = case.airfoil_chord, case.airfoil_thickness
    c.pitch_mean, c.pitch_amplitude = case.pitch_mean, case.pitch_amplitude
    c.pitch_frequency, c.pitch_phase, c.pivot_x = case.pitch_frequency, case.pitch_phase, case.pivot_x
    c.box_lo, c.box_hi = D3(*case.box_lo), D3(*case.box_hi)
    c.box_ds, c.box_lid_only = case.box_ds, case.box_lid_only
    c.box_lid_u, c.box_wall_tol = D3(*case.box_lid_u), case.box_wall_tol
    return c


class Oracle:
    """FP64 CPU level (AdvancePhiAtLevel + Evolve hooks) on one box."""

    def __init__(self, case, init_macro: np.ndarray):
        self.case = case
        self._c = to_rc_case(case)
        m = np.ascontiguousarray(init_macro, dtype=np.float64).reshape(4, -1)
        self._h = lib().rc_create(C.byref(self._c), dptr(m))

    def close(self):
        if self._h:
            lib().rc_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def advance(self, n: int):
        lib().rc_advance(self._h, int(n))

    def mass_sum(self) -> float:
        return lib().rc_mass_sum(self._h)

    def f(self) -> np.ndarray:
        out = np.empty(self.case.lattice.q * self.case.ncells)
        lib().rc_get_f(self._h, dptr(out))
        return out.reshape(self.case.lattice.q, -1)

    def macro(self) -> np.ndarray:
        out = np.empty(4 * self.case.ncells)
        lib().rc_get_macro(self._h, dptr(out))
        return out.reshape(4, -1)

    def forcing(self) -> np.ndarray:
        out = np.empty(3 * self.case.ncells)
        lib().rc_get_forcing(self._h, dptr(out))
        return out.reshape(3, -1)

    def force_record(self) -> np.ndarray:
        out = np.zeros(40)
        n = lib().rc_force_record(self._h, dptr(out), 40)
        return out[:n]

    def num_markers(self) -> int:
        return lib().rc_num_markers(self._h)

    def markers(self):
        n = self.num_markers()
        xyz, uvw, fxyz = np.zeros(3 * n), np.zeros(3 * n), np.zeros(3 * n)
        lib().rc_marker_state(self._h, dptr(xyz), dptr(uvw), dptr(fxyz))
        return xyz.reshape(n, 3), uvw.reshape(n, 3), fxyz.reshape(n, 3)

    def last_marker_force_sum(self) -> np.ndarray:
        out = np.zeros(3)
        lib().rc_last_marker_force_sum(self._h, dptr(out))
        return out

    def dv(self) -> float:
        return lib().rc_marker_dv(self._h)

    def tau(self) -> float:
        return lib().rc_tau(self._h)

    def grown_shape(self):
        """(nz, ny, nx) of the valid box grown by the 2 ghost cells the IB
        stages read."""
        n = self.case.n
        g = [2, 2, 2 if self.case.lattice.dim == 3 else 0]
        return (n[2] + 2 * g[2], n[1] + 2 * g[1], n[0] + 2 * g[0])

    def interp_probe(self, time: float, ev: np.ndarray) -> np.ndarray:
        """Marker U [nm][3] interpolated from EulerVel ev [3][nz][ny][nx]
        over the valid box grown by 2 (rc_ib_interp_probe)."""
        ev = np.ascontiguousarray(ev, dtype=np.float64)
        assert ev.shape == (3,) + self.grown_shape()
        out = np.zeros(3 * self.num_markers())
        lib().rc_ib_interp_probe(self._h, float(time), dptr(ev.reshape(-1)), dptr(out))
        return out.reshape(-1, 3)

    def spread_probe(self, time: float, F: np.ndarray):
        """EulerForce [3][cells] and moments [nm][3] of marker forces F [nm][3]
        (rc_ib_spread_probe)."""
        F = np.ascontiguousarray(F, dtype=np.float64)
        nm = self.num_markers()
        assert F.shape == (nm, 3)
        efi = np.zeros(3 * self.case.ncells)
        mom = np.zeros(3 * nm)
        lib().rc_ib_spread_probe(self._h, float(time), dptr(F.reshape(-1)), dptr(efi), dptr(mom))
        return efi.reshape(3, -1), mom.reshape(nm, 3)


class AmrOracle:
    """FP64 CPU AMR hierarchy: level 0 = the case's domain, level l >= 1 one
    box (level-l cell indices, inclusive) at refinement ratio 2."""

    def __init__(self, case, boxes, regrid_int: int, init_macro: np.ndarray):
        self.case = case
        self.boxes = [tuple(map(tuple, b)) for b in boxes]
        self.nlev = 1 + len(boxes)
        self._c = to_rc_case(case)
        lo = (C.c_int * max(3 * len(boxes), 1))()
        hi = (C.c_int * max(3 * len(boxes), 1))()
        for l, (blo, bhi) in enumerate(self.boxes):
            for d in range(3):
                lo[3 * l + d], hi[3 * l + d] = blo[d], bhi[d]
        m = np.ascontiguousarray(init_macro, dtype=np.float64).reshape(4, -1)
        self._h = lib().rc_amr_create(C.byref(self._c), self.nlev, lo, hi, int(regrid_int), dptr(m))
        if not self._h:
            raise ValueError("rc_amr_create rejected the hierarchy")

    def close(self):
        if self._h:
            lib().rc_amr_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def advance(self, n: int):
        lib().rc_amr_advance(self._h, int(n))
        self.boxes = self.current_boxes()

    def set_regrid(self, threshold_ratio=(1.0, 1.0, 1.0, 1.0), blocking_factor=8, n_error_buf=1,
                   no_refine_xlo=0.0, no_refine_xhi=0.0, refine=(-1.0, -1.0, -1.0)):
        r = RcRegrid()
        tr = list(threshold_ratio) + [threshold_ratio[-1]] * (4 - len(threshold_ratio))
        for i in range(4):
            r.threshold_ratio[i] = float(tr[i])
        r.no_refine_xlo, r.no_refine_xhi = float(no_refine_xlo), float(no_refine_xhi)
        r.refine_upstream, r.refine_downstream, r.refine_cross = map(float, refine)
        r.blocking_factor, r.n_error_buf = int(blocking_factor), int(n_error_buf)
        lib().rc_amr_set_regrid(self._h, C.byref(r))

    def current_boxes(self):
        n = max(self.nlev - 1, 1)
        lo, hi = (C.c_int * (3 * n))(), (C.c_int * (3 * n))()
        lib().rc_amr_boxes(self._h, lo, hi)
        return [(tuple(lo[3 * l:3 * l + 3]), tuple(hi[3 * l:3 * l + 3])) for l in range(self.nlev - 1)]

    def regrid_now(self, lbase: int = 0):
        lib().rc_amr_regrid_now(self._h, int(lbase))
        self.boxes = self.current_boxes()

    def regrids(self) -> int:
        return lib().rc_amr_regrids(self._h)

    def ncells(self, l):
        if l == 0:
            return self.case.ncells
        lo, hi = self.boxes[l - 1]
        dim = self.case.lattice.dim
        return int(np.prod([hi[d] - lo[d] + 1 for d in range(dim)]))

    def _lev(self, l):
        return lib().rc_amr_level(self._h, int(l))

    def f(self, l) -> np.ndarray:
        out = np.empty(self.case.lattice.q * self.ncells(l))
        lib().rc_get_f(self._lev(l), dptr(out))
        return out.reshape(self.case.lattice.q, -1)

    def macro(self, l, all_comps=False) -> np.ndarray:
        nc = self.case.nmac if all_comps else 4
        out = np.empty(nc * self.ncells(l))
        if all_comps:
            lib().rc_get_macro_all(self._lev(l), dptr(out))
        else:
            lib().rc_get_macro(self._lev(l), dptr(out))
        return out.reshape(nc, -1)

    def forcing(self, l) -> np.ndarray:
        out = np.empty(3 * self.ncells(l))
        lib().rc_get_forcing(self._lev(l), dptr(out))
        return out.reshape(3, -1)

    def mass_sum(self) -> float:
        return lib().rc_mass_sum(self._lev(0))

    def force_record(self) -> np.ndarray:
        out = np.zeros(40)
        n = lib().rc_force_record(self._lev(self.nlev - 1), dptr(out), 40)
        return out[:n]

    def num_markers(self) -> int:
        return lib().rc_num_markers(self._lev(self.nlev - 1))


def cons_interp(crse: np.ndarray, clo, dim: int, fiv) -> np.ndarray:
    """cell_cons_interp of the oracle at fine cell fiv from a coarse block
    crse[ncomp][nz][ny][nx] whose first cell is clo."""
    crse = np.ascontiguousarray(crse, dtype=np.float64)
    nc = crse.shape[0]
    shp = crse.shape[1:][::-1]
    lo = (C.c_int * 3)(*clo)
    hi = (C.c_int * 3)(*[clo[d] + shp[d] - 1 for d in range(3)])
    f = (C.c_int * 3)(*fiv)
    out = np.zeros(nc)
    lib().rc_cons_interp_test(dptr(crse.reshape(-1)), lo, hi, dim, nc, f, dptr(out))
    return out


class mutation:
    """Context manager planting one of the oracle's test mistakes
    (rc_set_mutation); the restatement is restored on exit."""
    EXTDIR_SIGN, SPREAD_NO_DV, EXPORT_NO_DT, INTERP_NO_DV = 1, 2, 3, 4
    CD_NO_HALF, FORCE_SIGN, MOMENT_SIGN, GUO_NO_HALF = 5, 6, 7, 8

    def __init__(self, which: int):
        self.which = int(which)

    def __enter__(self):
        lib().rc_set_mutation(self.which)
        return self

    def __exit__(self, *a):
        lib().rc_set_mutation(0)


def build_markers(case, time: float = 0.0):
    c = to_rc_case(case)
    buf = np.zeros(3 * 20000)
    dv = C.c_double(0.0)
    n = lib().rc_build_markers(C.byref(c), time, dptr(buf), 20000, C.byref(dv))
    return buf[:3 * n].reshape(n, 3), dv.value
