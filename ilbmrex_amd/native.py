"""ctypes binding of libilbm_hip.so (the C ABI declared in include/ilbm.h).

The HIP library is the only compute path: there is no CPU fallback.  If the
library has not been built, importing ``Level`` raises immediately.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional

import numpy as np

from . import config as cf

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG_DIR, "_lib", "libilbm_hip.so")

D3 = C.c_double * 3
I3 = C.c_int * 3
D27 = C.c_double * 27


class CaseDesc(C.Structure):
    """ilbm_case_desc (include/ilbm.h)."""
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
        ("storage", C.c_int),
    ]


EXPORTED = [
    "ilbm_level_create", "ilbm_level_destroy", "ilbm_level_advance", "ilbm_level_synchronize",
    "ilbm_level_get_f", "ilbm_level_get_macro", "ilbm_level_get_forcing",
    "ilbm_level_mass_history", "ilbm_level_force_history", "ilbm_level_num_markers",
    "ilbm_level_marker_dv", "ilbm_level_marker_state", "ilbm_level_marker_force_sum",
    "ilbm_level_marker_torque", "ilbm_level_info", "ilbm_level_kernel_time",
    "ilbm_level_set_timing", "ilbm_level_set_overlap", "ilbm_copy_probe", "ilbm_status_string", "ilbm_last_error",
    "ilbm_abi_version", "ilbm_case_desc_size",
    "ilbm_batch_create", "ilbm_batch_destroy", "ilbm_batch_footprint", "ilbm_batch_add",
    "ilbm_batch_advance", "ilbm_batch_finished", "ilbm_batch_progress",
    "ilbm_batch_force_history", "ilbm_batch_mass_history", "ilbm_batch_remove",
    "ilbm_batch_case_tile", "ilbm_batch_info", "ilbm_device_memory",
    "ilbm_amr_create", "ilbm_amr_destroy", "ilbm_amr_advance", "ilbm_amr_level_cells",
    "ilbm_amr_get_f", "ilbm_amr_get_macro", "ilbm_amr_mass_history", "ilbm_amr_force_history",
    "ilbm_amr_num_markers", "ilbm_amr_info", "ilbm_amr_set_regrid", "ilbm_amr_boxes", "ilbm_amr_get_derived",
    "ilbm_box_equilibrium", "ilbm_box_fill_periodic", "ilbm_box_fill_physical", "ilbm_box_collide",
    "ilbm_box_bounce_back", "ilbm_box_stream_macro", "ilbm_box_derived", "ilbm_box_sum",
    "ilbm_level_checkpoint_size", "ilbm_level_checkpoint", "ilbm_level_restore", "ilbm_level_get_derived",
    "ilbm_batch_checkpoint_size", "ilbm_batch_checkpoint", "ilbm_batch_add_from_checkpoint",
    "ilbm_box_copy", "ilbm_box_setval", "ilbm_device_alloc", "ilbm_device_free", "ilbm_copy_to_device",
    "ilbm_copy_to_host", "ilbm_device_synchronize", "ilbm_box_sum_noalloc", "ilbm_box_sum_noalloc_max_cells",
    "ilbm_box_cf_fill", "ilbm_box_interp_from_coarse", "ilbm_box_average_down",
    "ilbm_ib_create", "ilbm_ib_destroy", "ilbm_ib_update_forcing", "ilbm_ib_num_markers", "ilbm_ib_force_sum",
    "ilbm_ib_torque", "ilbm_ib_marker_state",
]

ERR_FULL = 6

FAB_F64, FAB_F32 = 0, 1


class Fab(C.Structure):
    """ilbm_fab (include/ilbm.h): an Array4 in device memory; strides of 0
    are the dense strides of the box."""
    _fields_ = [("p", C.c_void_p), ("lo", I3), ("hi", I3), ("ncomp", C.c_int), ("dtype", C.c_int),
                ("jstride", C.c_int64), ("kstride", C.c_int64), ("nstride", C.c_int64)]


def make_fab(ptr: int, lo, hi, ncomp: int, strides=(0, 0, 0), comp: int = 0) -> Fab:
    """A descriptor of device memory at `ptr` (FP64).  With comp > 0 the fab
    is the alias of component comp of a larger array (MultiFab make_alias):
    the pointer moves by comp component strides."""
    f = Fab()
    nx, ny, nz = (hi[d] - lo[d] + 1 for d in range(3))
    js = strides[0] or nx
    ks = strides[1] or js * ny
    ns = strides[2] or ks * nz
    f.p = int(ptr) + 8 * comp * ns
    f.lo = I3(*lo)
    f.hi = I3(*hi)
    f.ncomp = int(ncomp)
    f.dtype = FAB_F64
    f.jstride, f.kstride, f.nstride = (int(x) for x in strides)
    return f


_lib = None


class IlbmError(RuntimeError):
    pass


def load():
    """Loads the in-tree HIP library; raises if it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise IlbmError(f"HIP library not built: {LIB_PATH} (run __graft_entry__.build())")
    L = C.CDLL(LIB_PATH)
    P = C.POINTER
    vp = C.c_void_p
    L.ilbm_level_create.argtypes = [P(CaseDesc), P(C.c_double), P(vp)]
    L.ilbm_level_destroy.argtypes = [vp]
    L.ilbm_level_advance.argtypes = [vp, C.c_int, C.c_int]
    L.ilbm_level_synchronize.argtypes = [vp]
    for nm in ("ilbm_level_get_f", "ilbm_level_get_macro", "ilbm_level_get_forcing"):
        getattr(L, nm).argtypes = [vp, P(C.c_double)]
    L.ilbm_level_mass_history.argtypes = [vp, P(C.c_double), C.c_int, P(C.c_int)]
    L.ilbm_level_force_history.argtypes = [vp, P(C.c_double), C.c_int, P(C.c_int), P(C.c_int)]
    L.ilbm_level_num_markers.argtypes = [vp, P(C.c_int)]
    L.ilbm_level_marker_dv.argtypes = [vp, P(C.c_double)]
    L.ilbm_level_marker_state.argtypes = [vp, P(C.c_double), P(C.c_double), P(C.c_double)]
    L.ilbm_level_marker_force_sum.argtypes = [vp, P(C.c_double)]
    L.ilbm_level_marker_torque.argtypes = [vp, P(C.c_double)]
    L.ilbm_level_info.argtypes = [vp, P(C.c_double)]
    L.ilbm_level_kernel_time.argtypes = [vp, P(C.c_double), P(C.c_int)]
    L.ilbm_level_set_timing.argtypes = [vp, C.c_int]
    L.ilbm_level_set_overlap.argtypes = [vp, C.c_int]
    L.ilbm_batch_create.argtypes = [C.c_int64, C.c_int, C.c_int, P(vp)]
    L.ilbm_batch_destroy.argtypes = [vp]
    L.ilbm_batch_footprint.argtypes = [P(CaseDesc), C.c_int, P(C.c_int64)]
    L.ilbm_batch_add.argtypes = [vp, P(CaseDesc), P(C.c_double), C.c_int, P(C.c_int)]
    L.ilbm_batch_advance.argtypes = [vp, C.c_int, C.c_int, P(C.c_int)]
    L.ilbm_batch_finished.argtypes = [vp, P(C.c_int), C.c_int, P(C.c_int)]
    L.ilbm_batch_progress.argtypes = [vp, C.c_int, P(C.c_int), P(C.c_int)]
    L.ilbm_batch_force_history.argtypes = [vp, C.c_int, P(C.c_double), C.c_int, P(C.c_int), P(C.c_int)]
    L.ilbm_batch_mass_history.argtypes = [vp, C.c_int, P(C.c_double), C.c_int, P(C.c_int)]
    L.ilbm_batch_remove.argtypes = [vp, C.c_int]
    L.ilbm_batch_case_tile.argtypes = [vp, C.c_int, P(C.c_int)]
    L.ilbm_batch_info.argtypes = [vp, P(C.c_double)]
    L.ilbm_device_memory.argtypes = [P(C.c_int64), P(C.c_int64)]
    L.ilbm_amr_create.argtypes = [P(CaseDesc), C.c_int, P(C.c_int), P(C.c_int), C.c_int, P(C.c_double), P(vp)]
    L.ilbm_amr_destroy.argtypes = [vp]
    L.ilbm_amr_advance.argtypes = [vp, C.c_int, C.c_int]
    L.ilbm_amr_level_cells.argtypes = [vp, C.c_int, P(C.c_int64)]
    L.ilbm_amr_get_f.argtypes = [vp, C.c_int, P(C.c_double)]
    L.ilbm_amr_get_macro.argtypes = [vp, C.c_int, P(C.c_double)]
    L.ilbm_amr_mass_history.argtypes = [vp, P(C.c_double), C.c_int, P(C.c_int)]
    L.ilbm_amr_force_history.argtypes = [vp, P(C.c_double), C.c_int, P(C.c_int), P(C.c_int)]
    L.ilbm_amr_num_markers.argtypes = [vp, P(C.c_int)]
    L.ilbm_amr_info.argtypes = [vp, P(C.c_double)]
    L.ilbm_amr_set_regrid.argtypes = [vp, P(RegridParams)]
    L.ilbm_amr_get_derived.argtypes = [vp, C.c_int, P(C.c_double)]
    L.ilbm_amr_boxes.argtypes = [vp, P(C.c_int), P(C.c_int), P(C.c_int)]
    L.ilbm_level_get_derived.argtypes = [vp, P(C.c_double)]
    L.ilbm_level_checkpoint_size.argtypes = [vp, P(C.c_int64)]
    L.ilbm_level_checkpoint.argtypes = [vp, C.c_void_p, C.c_int64]
    L.ilbm_level_restore.argtypes = [vp, C.c_void_p, C.c_int64]
    L.ilbm_batch_checkpoint_size.argtypes = [vp, C.c_int, P(C.c_int64)]
    L.ilbm_batch_checkpoint.argtypes = [vp, C.c_int, C.c_void_p, C.c_int64]
    L.ilbm_batch_add_from_checkpoint.argtypes = [vp, P(CaseDesc), C.c_void_p, C.c_int64, C.c_int, P(C.c_int)]
    PF = P(Fab)
    L.ilbm_box_equilibrium.argtypes = [P(CaseDesc), PF, PF, P(C.c_int), P(C.c_int), vp]
    L.ilbm_box_fill_periodic.argtypes = [P(CaseDesc), PF, P(C.c_int), P(C.c_int), vp]
    L.ilbm_box_fill_physical.argtypes = [P(CaseDesc), PF, C.c_int, vp]
    L.ilbm_box_collide.argtypes = [P(CaseDesc), PF, PF, PF, P(C.c_int), P(C.c_int), vp]
    L.ilbm_box_bounce_back.argtypes = [P(CaseDesc), PF, P(C.c_int), P(C.c_int), vp]
    L.ilbm_box_stream_macro.argtypes = [P(CaseDesc), PF, PF, PF, PF, P(C.c_int), P(C.c_int), vp]
    L.ilbm_box_derived.argtypes = [P(CaseDesc), PF, PF, P(C.c_int), P(C.c_int), vp]
    L.ilbm_box_sum.argtypes = [PF, C.c_int, P(C.c_int), P(C.c_int), P(C.c_double), vp]
    L.ilbm_box_sum_noalloc.argtypes = [PF, C.c_int, P(C.c_int), P(C.c_int), P(C.c_double), vp]
    L.ilbm_box_sum_noalloc_max_cells.argtypes = [P(C.c_int64)]
    L.ilbm_box_cf_fill.argtypes = [P(CaseDesc), PF, PF, C.c_double, C.c_double, C.c_double, PF, PF, C.c_int,
                                   P(C.c_int), P(C.c_int), vp]
    L.ilbm_box_interp_from_coarse.argtypes = [P(CaseDesc), PF, PF, C.c_int, P(C.c_int), P(C.c_int), vp]
    L.ilbm_box_average_down.argtypes = [P(CaseDesc), PF, PF, P(C.c_int), P(C.c_int), PF, PF, C.c_int, vp]
    L.ilbm_ib_create.argtypes = [P(CaseDesc), P(C.c_int), P(C.c_int), P(vp)]
    L.ilbm_ib_destroy.argtypes = [vp]
    L.ilbm_ib_update_forcing.argtypes = [vp, PF, PF, PF, PF, PF, PF, PF, C.c_double, C.c_double, vp]
    L.ilbm_ib_num_markers.argtypes = [vp, P(C.c_int64)]
    L.ilbm_ib_force_sum.argtypes = [vp, P(C.c_double)]
    L.ilbm_ib_torque.argtypes = [vp, P(C.c_double)]
    L.ilbm_ib_marker_state.argtypes = [vp, P(C.c_double), P(C.c_double), P(C.c_double)]
    L.ilbm_box_copy.argtypes = [PF, PF, P(C.c_int), P(C.c_int), C.c_int, vp]
    L.ilbm_box_setval.argtypes = [PF, C.c_double, vp]
    L.ilbm_device_alloc.argtypes = [C.c_int64, P(vp)]
    L.ilbm_device_free.argtypes = [vp]
    L.ilbm_copy_to_device.argtypes = [vp, vp, C.c_int64]
    L.ilbm_copy_to_host.argtypes = [vp, vp, C.c_int64]
    L.ilbm_status_string.restype = C.c_char_p
    L.ilbm_status_string.argtypes = [C.c_int]
    L.ilbm_last_error.restype = C.c_char_p
    _lib = L
    return L


def _dp(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _check(rc: int, what: str):
    if rc != 0:
        L = load()
        msg = L.ilbm_status_string(rc).decode()
        detail = L.ilbm_last_error().decode()
        raise IlbmError(f"{what}: {msg} ({detail})")


def to_desc(case: "cf.Case") -> CaseDesc:
    d = CaseDesc()
    lat = case.lattice
    d.dim = lat.dim
    d.n = I3(*case.n)
    d.plo = D3(*case.plo)
    d.dx = D3(*case.dx)
    d.periodic = I3(*case.periodic)
    d.bc_lo = I3(*case.bc_lo)
    d.bc_hi = I3(*case.bc_hi)
    for k in range(3):
        for m in range(4):
            d.bc_lo_val[k][m] = case.bc_lo_val[k][m]
            d.bc_hi_val[k][m] = case.bc_hi_val[k][m]
    d.q = lat.q
    w = lat.w
    for q in range(lat.q):
        d.cx[q], d.cy[q], d.cz[q], d.w[q] = float(lat.cx[q]), float(lat.cy[q]), float(lat.cz[q]), w[q]
    d.nu, d.U0, d.nmac = case.nu, case.U0, case.nmac
    if case.prescribed_force is not None:
        d.use_prescribed_force = 1
        d.prescribed_force = D3(*case.prescribed_force)
    d.use_ib = int(case.use_ib)
    d.x0, d.y0, d.z0, d.R = case.x0, case.y0, case.z0, case.R
    d.coupling_method, d.delta_type, d.loop_ns = case.coupling_method, case.delta_type, case.loop_ns
    d.mdf_relax, d.explicit_diag_eps = case.mdf_relax, case.explicit_diag_eps
    d.geometry_type, d.n_marker, d.rd = case.geometry_type, case.n_marker, case.rd
    d.ubx, d.uby, d.ubz = case.ub
    d.omx, d.omy, d.omz = case.om
    d.ivc_diag_reg, d.ivc_rebuild_matrix = case.ivc_diag_reg, case.ivc_rebuild_matrix
    d.renormalize_delta, d.force_eval_method = case.renormalize_delta, case.force_eval_method
    d.airfoil_chord, d.airfoil_thickness = case.airfoil_chord, case.airfoil_thickness
    d.pitch_mean, d.pitch_amplitude = case.pitch_mean, case.pitch_amplitude
    d.pitch_frequency, d.pitch_phase, d.pivot_x = case.pitch_frequency, case.pitch_phase, case.pivot_x
    d.box_lo, d.box_hi = D3(*case.box_lo), D3(*case.box_hi)
    d.box_ds, d.box_lid_only = case.box_ds, case.box_lid_only
    d.box_lid_u, d.box_wall_tol = D3(*case.box_lid_u), case.box_wall_tol
    d.storage = case.storage
    return d


class Level:
    """One resident level on the GPU (AmrCoreLBM level 0 + IBMarkerDF)."""

    def __init__(self, case: "cf.Case", init_macro: Optional[np.ndarray] = None):
        case.validate()
        self.case = case
        L = load()
        if init_macro is None:
            init_macro = cf.uniform_init(case)
        m = np.ascontiguousarray(init_macro, dtype=np.float64).reshape(4, -1)
        if m.shape[1] != case.ncells:
            raise ValueError("init_macro must be [4][cells]")
        self._desc = to_desc(case)
        h = C.c_void_p()
        _check(L.ilbm_level_create(C.byref(self._desc), _dp(m), C.byref(h)), "ilbm_level_create")
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            load().ilbm_level_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def advance(self, nsteps: int, want_state: bool = False):
        _check(load().ilbm_level_advance(self._h, int(nsteps), int(want_state)), "advance")

    def synchronize(self):
        _check(load().ilbm_level_synchronize(self._h), "synchronize")

    def f(self) -> np.ndarray:
        out = np.empty(self.case.lattice.q * self.case.ncells)
        _check(load().ilbm_level_get_f(self._h, _dp(out)), "get_f")
        return out.reshape(self.case.lattice.q, -1)

    def macro(self) -> np.ndarray:
        out = np.empty(4 * self.case.ncells)
        _check(load().ilbm_level_get_macro(self._h, _dp(out)), "get_macro")
        return out.reshape(4, -1)

    def derived(self) -> np.ndarray:
        """vorticity and pressure of the exported state, [2][cells]."""
        out = np.empty(2 * self.case.ncells)
        _check(load().ilbm_level_get_derived(self._h, _dp(out)), "get_derived")
        return out.reshape(2, -1)

    def forcing(self) -> np.ndarray:
        out = np.empty(3 * self.case.ncells)
        _check(load().ilbm_level_get_forcing(self._h, _dp(out)), "get_forcing")
        return out.reshape(3, -1)

    def mass_history(self) -> np.ndarray:
        n = C.c_int(0)
        _check(load().ilbm_level_mass_history(self._h, None, 0, C.byref(n)), "mass_history")
        out = np.empty(max(n.value, 1))
        _check(load().ilbm_level_mass_history(self._h, _dp(out), n.value, C.byref(n)), "mass_history")
        return out[:n.value]

    def force_history(self) -> np.ndarray:
        nr, nc = C.c_int(0), C.c_int(0)
        _check(load().ilbm_level_force_history(self._h, None, 0, C.byref(nr), C.byref(nc)), "force_history")
        out = np.empty(max(nr.value * nc.value, 1))
        _check(load().ilbm_level_force_history(self._h, _dp(out), nr.value, C.byref(nr), C.byref(nc)),
               "force_history")
        return out[:nr.value * nc.value].reshape(nr.value, max(nc.value, 1))

    def num_markers(self) -> int:
        n = C.c_int(0)
        _check(load().ilbm_level_num_markers(self._h, C.byref(n)), "num_markers")
        return n.value

    def dv(self) -> float:
        v = C.c_double(0)
        _check(load().ilbm_level_marker_dv(self._h, C.byref(v)), "marker_dv")
        return v.value

    def markers(self):
        n = self.num_markers()
        xyz, uvw, f = np.zeros(3 * n), np.zeros(3 * n), np.zeros(3 * n)
        _check(load().ilbm_level_marker_state(self._h, _dp(xyz), _dp(uvw), _dp(f)), "marker_state")
        return xyz.reshape(n, 3), uvw.reshape(n, 3), f.reshape(n, 3)

    def last_marker_force_sum(self) -> np.ndarray:
        out = np.zeros(3)
        _check(load().ilbm_level_marker_force_sum(self._h, _dp(out)), "marker_force_sum")
        return out

    def torque(self) -> np.ndarray:
        out = np.zeros(3)
        _check(load().ilbm_level_marker_torque(self._h, _dp(out)), "marker_torque")
        return out

    def info(self) -> dict:
        out = np.zeros(8)
        _check(load().ilbm_level_info(self._h, _dp(out)), "info")
        return {"tau": out[0], "dt": out[1], "time": out[2], "steps": int(out[3]),
                "band_cells": int(out[4]), "pop_bytes": int(out[5]), "padded": bool(out[6]),
                "addressing": ("wrap", "wall64", "wall32", "wall32x", "wall32x")[int(out[6])],
                "ghosts_fused": int(out[6]) == 4, "blocks": int(out[7])}

    def set_timing(self, enable: bool):
        _check(load().ilbm_level_set_timing(self._h, int(enable)), "set_timing")

    def set_overlap(self, enable: bool):
        """IB level: far cells stream-collide on a second stream during the IB forcing."""
        _check(load().ilbm_level_set_overlap(self._h, int(enable)), "set_overlap")

    def checkpoint(self) -> bytes:
        """The level's dynamic state (ilbm_level_checkpoint)."""
        L = load()
        n = C.c_int64(0)
        _check(L.ilbm_level_checkpoint_size(self._h, C.byref(n)), "checkpoint_size")
        buf = C.create_string_buffer(n.value)
        _check(L.ilbm_level_checkpoint(self._h, buf, n.value), "checkpoint")
        return buf.raw

    def restore(self, data: bytes):
        buf = C.create_string_buffer(data, len(data))
        _check(load().ilbm_level_restore(self._h, buf, len(data)), "restore")

    def kernel_time(self):
        ms, n = C.c_double(0), C.c_int(0)
        _check(load().ilbm_level_kernel_time(self._h, C.byref(ms), C.byref(n)), "kernel_time")
        return ms.value, n.value


class BatchFull(IlbmError):
    """Admission refused: the HBM budget or the case slots are exhausted."""


class Batch:
    """Many independent cases resident on the GPU, advanced together
    (ilbm_batch_* in include/ilbm.h).  A case is admitted while others run,
    advances in lock step with them and stops at its own step count."""

    def __init__(self, hbm_budget_bytes: int, max_cases: int = 1024, graph_steps: int = 8):
        L = load()
        h = C.c_void_p()
        _check(L.ilbm_batch_create(int(hbm_budget_bytes), int(max_cases), int(graph_steps), C.byref(h)),
               "ilbm_batch_create")
        self._h = h
        self._cases = {}

    def close(self):
        if getattr(self, "_h", None):
            load().ilbm_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @staticmethod
    def footprint(case: "cf.Case", nsteps: int) -> int:
        b = C.c_int64(0)
        d = to_desc(case)
        _check(load().ilbm_batch_footprint(C.byref(d), int(nsteps), C.byref(b)), "footprint")
        return b.value

    def add(self, case: "cf.Case", init_macro: Optional[np.ndarray], nsteps: int) -> int:
        """Admits a case; raises BatchFull when it does not fit."""
        case.validate()
        if init_macro is None:
            init_macro = cf.uniform_init(case)
        m = np.ascontiguousarray(init_macro, dtype=np.float64).reshape(4, -1)
        if m.shape[1] != case.ncells:
            raise ValueError("init_macro must be [4][cells]")
        d = to_desc(case)
        cid = C.c_int(-1)
        rc = load().ilbm_batch_add(self._h, C.byref(d), _dp(m), int(nsteps), C.byref(cid))
        if rc == ERR_FULL:
            raise BatchFull("batch full")
        _check(rc, "ilbm_batch_add")
        self._cases[cid.value] = case
        return cid.value

    def advance(self, max_steps: int, use_graph: bool = True) -> int:
        n = C.c_int(0)
        _check(load().ilbm_batch_advance(self._h, int(max_steps), int(use_graph), C.byref(n)), "advance")
        return n.value

    def finished(self) -> list:
        n = C.c_int(0)
        _check(load().ilbm_batch_finished(self._h, None, 0, C.byref(n)), "finished")
        ids = (C.c_int * max(n.value, 1))()
        _check(load().ilbm_batch_finished(self._h, ids, n.value, C.byref(n)), "finished")
        return [ids[i] for i in range(n.value)]

    def progress(self, cid: int):
        d, t = C.c_int(0), C.c_int(0)
        _check(load().ilbm_batch_progress(self._h, int(cid), C.byref(d), C.byref(t)), "progress")
        return d.value, t.value

    def force_history(self, cid: int) -> np.ndarray:
        L = load()
        nr, nc = C.c_int(0), C.c_int(0)
        _check(L.ilbm_batch_force_history(self._h, int(cid), None, 0, C.byref(nr), C.byref(nc)), "force_history")
        out = np.empty(max(nr.value * nc.value, 1))
        _check(L.ilbm_batch_force_history(self._h, int(cid), _dp(out), nr.value, C.byref(nr), C.byref(nc)),
               "force_history")
        return out[:nr.value * nc.value].reshape(nr.value, max(nc.value, 1))

    def mass_history(self, cid: int) -> np.ndarray:
        L = load()
        n = C.c_int(0)
        _check(L.ilbm_batch_mass_history(self._h, int(cid), None, 0, C.byref(n)), "mass_history")
        out = np.empty(max(n.value, 1))
        _check(L.ilbm_batch_mass_history(self._h, int(cid), _dp(out), n.value, C.byref(n)), "mass_history")
        return out[:n.value]

    def checkpoint(self, cid: int) -> bytes:
        L = load()
        n = C.c_int64(0)
        _check(L.ilbm_batch_checkpoint_size(self._h, int(cid), C.byref(n)), "checkpoint_size")
        buf = C.create_string_buffer(n.value)
        _check(L.ilbm_batch_checkpoint(self._h, int(cid), buf, n.value), "checkpoint")
        return buf.raw

    def add_from_checkpoint(self, case: "cf.Case", data: bytes, nsteps: int) -> int:
        d = to_desc(case)
        buf = C.create_string_buffer(data, len(data))
        cid = C.c_int(-1)
        rc = load().ilbm_batch_add_from_checkpoint(self._h, C.byref(d), buf, len(data), int(nsteps), C.byref(cid))
        if rc == ERR_FULL:
            raise BatchFull("batch full")
        _check(rc, "add_from_checkpoint")
        self._cases[cid.value] = case
        return cid.value

    def remove(self, cid: int):
        _check(load().ilbm_batch_remove(self._h, int(cid)), "remove")
        self._cases.pop(cid, None)

    def tile(self, cid: int) -> int:
        t = C.c_int(0)
        _check(load().ilbm_batch_case_tile(self._h, int(cid), C.byref(t)), "case_tile")
        return t.value

    def info(self) -> dict:
        out = np.zeros(8)
        _check(load().ilbm_batch_info(self._h, _dp(out)), "info")
        return {"resident": int(out[0]), "active": int(out[1]), "bytes": int(out[2]), "budget": int(out[3]),
                "graph_launches": int(out[4]), "captures": int(out[5]), "eager_steps": int(out[6]),
                "lups": float(out[7])}


class RegridParams(C.Structure):
    """ilbm_regrid_params (include/ilbm.h)."""
    _fields_ = [("threshold_ratio", C.c_double * 4), ("no_refine_xlo", C.c_double), ("no_refine_xhi", C.c_double),
                ("refine_upstream", C.c_double), ("refine_downstream", C.c_double), ("refine_cross", C.c_double),
                ("blocking_factor", C.c_int), ("n_error_buf", C.c_int)]


class AmrHierarchy:
    """Levels 0..nlev-1 at refinement ratio 2 (ilbm_amr_* in include/ilbm.h):
    level 0 is the case's domain, each box in ``boxes`` (level-l cell
    indices, inclusive lo / hi) one refined level, advanced with subcycling.
    The initial state (rho, ux, uy, uz) is uniform."""

    def __init__(self, case: "cf.Case", boxes, regrid_int: int, state=(1.0, 0.0, 0.0, 0.0)):
        case.validate()
        self.case = case
        self.nlev = 1 + len(boxes)
        L = load()
        lo = (C.c_int * max(3 * len(boxes), 1))()
        hi = (C.c_int * max(3 * len(boxes), 1))()
        for l, (blo, bhi) in enumerate(boxes):
            for d in range(3):
                lo[3 * l + d], hi[3 * l + d] = int(blo[d]), int(bhi[d])
        self._desc = to_desc(case)
        st = (C.c_double * 4)(*[float(x) for x in state])
        h = C.c_void_p()
        _check(L.ilbm_amr_create(C.byref(self._desc), self.nlev, lo, hi, int(regrid_int), st, C.byref(h)),
               "ilbm_amr_create")
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            load().ilbm_amr_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def advance(self, nsteps: int, want_state: bool = False):
        _check(load().ilbm_amr_advance(self._h, int(nsteps), int(want_state)), "amr_advance")

    def set_regrid(self, threshold_ratio=(1.0,), blocking_factor: int = 8, n_error_buf: int = 1,
                   no_refine_xlo: float = 0.0, no_refine_xhi: float = 0.0, refine=(-1.0, -1.0, -1.0)):
        """Switch regridding on every regrid_int steps (amrvorth.thresholdRatio,
        amr.blocking_factor, amr.n_error_buf, amrvorth.no_refine_xlo/hi,
        ibm.refine_upstream/downstream/cross)."""
        p = RegridParams()
        tr = list(threshold_ratio) + [threshold_ratio[-1]] * (4 - len(threshold_ratio))
        for i in range(4):
            p.threshold_ratio[i] = float(tr[i])
        p.no_refine_xlo, p.no_refine_xhi = float(no_refine_xlo), float(no_refine_xhi)
        p.refine_upstream, p.refine_downstream, p.refine_cross = (float(x) for x in refine)
        p.blocking_factor, p.n_error_buf = int(blocking_factor), int(n_error_buf)
        _check(load().ilbm_amr_set_regrid(self._h, C.byref(p)), "amr_set_regrid")

    def boxes(self):
        """current boxes of levels 1.. as ((lo), (hi)) and the regrid count"""
        n = max(self.nlev - 1, 1)
        lo, hi, nr = (C.c_int * (3 * n))(), (C.c_int * (3 * n))(), C.c_int(0)
        _check(load().ilbm_amr_boxes(self._h, lo, hi, C.byref(nr)), "amr_boxes")
        return [(tuple(lo[3 * l:3 * l + 3]), tuple(hi[3 * l:3 * l + 3])) for l in range(self.nlev - 1)], nr.value

    def cells(self, lev: int) -> int:
        n = C.c_int64(0)
        _check(load().ilbm_amr_level_cells(self._h, int(lev), C.byref(n)), "amr_level_cells")
        return n.value

    def f(self, lev: int) -> np.ndarray:
        out = np.empty(self.case.lattice.q * self.cells(lev))
        _check(load().ilbm_amr_get_f(self._h, int(lev), _dp(out)), "amr_get_f")
        return out.reshape(self.case.lattice.q, -1)

    def macro(self, lev: int) -> np.ndarray:
        out = np.empty(4 * self.cells(lev))
        _check(load().ilbm_amr_get_macro(self._h, int(lev), _dp(out)), "amr_get_macro")
        return out.reshape(4, -1)

    def derived(self, lev: int) -> np.ndarray:
        """vorticity and pressure of an exported level, [2][cells]"""
        out = np.empty(2 * self.cells(lev))
        _check(load().ilbm_amr_get_derived(self._h, int(lev), _dp(out)), "amr_get_derived")
        return out.reshape(2, -1)

    def mass_history(self) -> np.ndarray:
        n = C.c_int(0)
        _check(load().ilbm_amr_mass_history(self._h, None, 0, C.byref(n)), "amr_mass_history")
        out = np.empty(max(n.value, 1))
        _check(load().ilbm_amr_mass_history(self._h, _dp(out), n.value, C.byref(n)), "amr_mass_history")
        return out[:n.value]

    def force_history(self) -> np.ndarray:
        nr, nc = C.c_int(0), C.c_int(0)
        L = load()
        _check(L.ilbm_amr_force_history(self._h, None, 0, C.byref(nr), C.byref(nc)), "amr_force_history")
        out = np.empty(max(nr.value * nc.value, 1))
        _check(L.ilbm_amr_force_history(self._h, _dp(out), nr.value, C.byref(nr), C.byref(nc)),
               "amr_force_history")
        return out[:nr.value * nc.value].reshape(nr.value, max(nc.value, 1))

    def num_markers(self) -> int:
        n = C.c_int(0)
        _check(load().ilbm_amr_num_markers(self._h, C.byref(n)), "amr_num_markers")
        return n.value

    def info(self) -> dict:
        out = np.zeros(8)
        _check(load().ilbm_amr_info(self._h, _dp(out)), "amr_info")
        return {"levels": int(out[0]), "steps": int(out[1]), "time": out[2],
                "cells": [int(x) for x in out[3:3 + int(out[0])]], "tau_finest": out[6], "tau0": out[7]}


class DeviceArray:
    """FP64 device buffer of the library's allocator, shaped [ncomp][nz][ny][nx]."""

    def __init__(self, shape):
        self.shape = tuple(int(x) for x in shape)
        self.nbytes = 8 * int(np.prod(self.shape))
        p = C.c_void_p()
        _check(load().ilbm_device_alloc(self.nbytes, C.byref(p)), "device_alloc")
        self.ptr = p.value

    def upload(self, a: np.ndarray):
        a = np.ascontiguousarray(a, dtype=np.float64)
        assert a.size * 8 == self.nbytes
        _check(load().ilbm_copy_to_device(self.ptr, a.ctypes.data, self.nbytes), "copy_to_device")

    def download(self) -> np.ndarray:
        out = np.empty(self.shape)
        _check(load().ilbm_copy_to_host(out.ctypes.data, self.ptr, self.nbytes), "copy_to_host")
        return out

    def free(self):
        if self.ptr:
            load().ilbm_device_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def device_memory():
    """(free, total) device bytes."""
    f, t = C.c_int64(0), C.c_int64(0)
    _check(load().ilbm_device_memory(C.byref(f), C.byref(t)), "device_memory")
    return f.value, t.value


def load_probe(nbytes: int, iters: int) -> float:
    """Average milliseconds of a plain device copy of nbytes (ilbm_copy_probe)."""
    L = load()
    L.ilbm_copy_probe.argtypes = [C.c_int64, C.c_int, C.POINTER(C.c_double)]
    ms = C.c_double(0.0)
    _check(L.ilbm_copy_probe(int(nbytes), int(iters), C.byref(ms)), "copy_probe")
    return ms.value
