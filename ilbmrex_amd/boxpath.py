"""AdvancePhiAtLevel driven box by box through the per-box C ABI.

BoxLevel keeps the reference's own state on the device -- f_new / f_old
(ndir components) and macro_new / macro_old (nmac components: rho, ux, uy,
uz, vorticity, pressure), AMReX Array4 layout with 3 ghost cells
(AmrCoreLBM.H:175-180, 276) -- in buffers of the library's allocator, and
steps it with one ilbm_box_* / ilbm_ib_* call per stage in the order of
Source/Core/AdvanceLBMAtLevel.cpp:20-386 for a level of one box:

    swap f_new/f_old, macro_new/macro_old
    FillPatch(f_old) -> sfSborder       copy + periodic images + mesoscopicBcFill
                                        (lev > 0: equilibrium of the coarse
                                        interpolation in the ghost cells)
    FillPatch(macro_old) -> smSborder   copy + periodic images + macroBcFill
                                        (lev > 0: cell_cons_interp of the coarse)
    force scratch: prescribed force or 0; IBMarkerDF::update_forcing on the
                   aliases rho, u, v, w of smSborder; FillBoundary
    collide(sfSborder; smSborder, force)
    FillBoundary(sfSborder), bounce-back ghosts
    stream + forced macros -> f_new, macro_new

Sum(Rho) and ComputeIBForce (AmrCoreLBM.cpp:366-414, 456-701) are composed
from ilbm_box_sum of the exported force and the marker force sum.
BoxHierarchy couples two such levels as timeStepWithSubcycling does
(AmrCoreLBM.cpp:1688-1754): the coarse step, two fine substeps, AverageDownTo,
and the derived fields (UpdateDerivedMacroFields) on the reference's
schedule.  With ``pad`` the arrays are allocated with padded rows and planes
and described by explicit strides.

It is the reference-layout counterpart of the fused level (ilbm_level_*)
and hierarchy (ilbm_amr_*): the same FP64 results, with the intermediate
arrays of the reference visible.
"""
from __future__ import annotations

import ctypes as C
import math
from typing import Optional

import numpy as np

from . import config as cf
from . import native

NG = 3


def _i3(v):
    return (C.c_int * 3)(*[int(x) for x in v])


class _Arr:
    """A device array [ncomp][nz][ny][nx] of a fab box, optionally with
    padded rows / planes (explicit strides)."""

    def __init__(self, ncomp, lo, hi, pad):
        self.ncomp = ncomp
        self.lo, self.hi = tuple(lo), tuple(hi)
        n = [hi[d] - lo[d] + 1 for d in range(3)]
        self.n = n
        if pad:
            js = (n[0] + 7) // 8 * 8 + 8
            ks = js * (n[1] + 1)
            ns = ks * n[2] + 64
        else:
            js, ks, ns = n[0], n[0] * n[1], n[0] * n[1] * n[2]
        self.strides = (js, ks, ns) if pad else (0, 0, 0)
        self.js, self.ks, self.ns = js, ks, ns
        self.dev = native.DeviceArray((ncomp * ns,))
        native._check(native.load().ilbm_box_setval(self.ref(), 0.0, None), "ilbm_box_setval")

    def fab(self, comp=0, ncomp=None):
        return native.make_fab(self.dev.ptr, self.lo, self.hi, self.ncomp - comp if ncomp is None else ncomp,
                               self.strides, comp)

    def ref(self, comp=0, ncomp=None):
        return C.byref(self.fab(comp, ncomp))

    def _view(self, flat):
        return np.lib.stride_tricks.as_strided(flat, shape=(self.ncomp, self.n[2], self.n[1], self.n[0]),
                                               strides=(8 * self.ns, 8 * self.ks, 8 * self.js, 8))

    def upload(self, a):
        """a: [ncomp][nz][ny][nx] over the fab box"""
        flat = np.zeros(self.ncomp * self.ns)
        self._view(flat)[...] = a
        self.dev.upload(flat)

    def download(self):
        return self._view(self.dev.download()).copy()

    def free(self):
        self.dev.free()


class BoxLevel:
    """One level of one box in the reference's layout.  ``vlo``/``vhi``: the
    box in level indices (default the whole domain of ``case``, whose ``n``,
    ``dx`` and BCs describe the level)."""

    def __init__(self, case: "cf.Case", init_macro: Optional[np.ndarray] = None, force: Optional[tuple] = None,
                 vlo=None, vhi=None, pad: bool = False, nmac: int = 4):
        case.validate()
        self.case = case
        self.L = native.load()
        self.desc = native.to_desc(case)
        dim = case.lattice.dim
        self.dim = dim
        self.q = case.lattice.q
        self.nmac = nmac
        self.vlo = tuple(vlo) if vlo is not None else (0, 0, 0)
        self.vhi = tuple(vhi) if vhi is not None else tuple(case.n[d] - 1 if d < dim else 0 for d in range(3))
        g = [NG if d < dim else 0 for d in range(3)]
        self.flo = tuple(self.vlo[d] - g[d] for d in range(3))
        self.fhi = tuple(self.vhi[d] + g[d] for d in range(3))
        self.nv = tuple(self.vhi[d] - self.vlo[d] + 1 for d in range(3))
        arr = lambda nc: _Arr(nc, self.flo, self.fhi, pad)
        self.f_new, self.f_old, self.sfS = arr(self.q), arr(self.q), arr(self.q)
        self.m_new, self.m_old, self.smS = arr(nmac), arr(nmac), arr(nmac)
        self.fx = arr(3)
        self.uniform_force = tuple(force) if force is not None else (0.0, 0.0, 0.0)
        self.has_force = case.use_ib or any(x != 0.0 for x in self.uniform_force)
        self.tau = 3.0 * case.nu / case.dx[0] + 0.5
        self.dt = case.dx[0]
        self.t_new, self.t_old, self.istep = 0.0, -1.0e200, 0
        self.ib = None
        if case.use_ib:
            h = C.c_void_p()
            native._check(self.L.ilbm_ib_create(C.byref(self.desc), _i3(self.vlo), _i3(self.vhi), C.byref(h)),
                          "ilbm_ib_create")
            self.ib = h
        self.have_prev = False
        self.prev_body = (0.0, 0.0, 0.0)
        self.record = None
        if init_macro is not None:
            m = np.zeros((nmac,) + self._shape())
            m[(slice(0, 4),) + self._valid()] = np.asarray(init_macro, dtype=np.float64).reshape(
                (4,) + tuple(self.nv[d] for d in (2, 1, 0)))
            self.m_new.upload(m)
            self.m_old.upload(m)
            self.equilibrium()

    def _shape(self):
        return tuple(self.fhi[d] - self.flo[d] + 1 for d in (2, 1, 0))

    def _valid(self):
        return tuple(slice(self.vlo[d] - self.flo[d], self.vhi[d] - self.flo[d] + 1) for d in (2, 1, 0))

    def _check(self, rc, what):
        native._check(rc, f"ilbm_{what}")

    def equilibrium(self):
        """InitEquilibrium: f_new = f_old = feq(macro_new) on the valid box."""
        d = C.byref(self.desc)
        self._check(self.L.ilbm_box_equilibrium(d, self.m_new.ref(), self.f_new.ref(), _i3(self.vlo), _i3(self.vhi),
                                                None), "box_equilibrium")
        self._check(self.L.ilbm_box_copy(self.f_new.ref(), self.f_old.ref(), _i3(self.vlo), _i3(self.vhi), self.q,
                                         None), "box_copy")

    def close(self):
        if self.ib:
            self.L.ilbm_ib_destroy(self.ib)
            self.ib = None
        for a in (self.f_new, self.f_old, self.sfS, self.m_new, self.m_old, self.smS, self.fx):
            a.free()

    # -- FillPatch ---------------------------------------------------------
    def _fill_patch_single(self, dst: _Arr, src: _Arr, meso: bool):
        d = C.byref(self.desc)
        nc = dst.ncomp
        self._check(self.L.ilbm_box_setval(dst.ref(), 0.0, None), "box_setval")
        self._check(self.L.ilbm_box_copy(src.ref(), dst.ref(), _i3(self.vlo), _i3(self.vhi), nc, None),
                    "box_copy")                                              # ParallelCopy
        self._check(self.L.ilbm_box_fill_periodic(d, dst.ref(), _i3(self.vlo), _i3(self.vhi), None),
                    "box_fill_periodic")
        self._check(self.L.ilbm_box_fill_physical(d, dst.ref(), int(meso), None), "box_fill_physical")

    def fill_patch(self, time, coarse: Optional["BoxLevel"] = None):
        """sfSborder / smSborder of a step starting at `time`."""
        if coarse is None:
            self._fill_patch_single(self.sfS, self.f_old, True)
            self._fill_patch_single(self.smS, self.m_old, False)
            return
        # FillPatchMacro / FillPatchMesoscopic for lev > 0: own valid cells,
        # ghosts from the coarse level at `time`
        for dst, src, nc in ((self.smS, self.m_old, self.nmac), (self.sfS, self.f_old, self.q)):
            self._check(self.L.ilbm_box_copy(src.ref(), dst.ref(), _i3(self.vlo), _i3(self.vhi), nc, None),
                        "box_copy")
        self._check(self.L.ilbm_box_cf_fill(C.byref(self.desc), coarse.m_old.ref(), coarse.m_new.ref(), coarse.t_old,
                                            coarse.t_new, time, self.sfS.ref(), self.smS.ref(), self.nmac,
                                            _i3(self.vlo), _i3(self.vhi), None), "box_cf_fill")

    # -- one AdvancePhiAtLevel ------------------------------------------------
    def advance_level(self, time, coarse: Optional["BoxLevel"] = None):
        self.f_new, self.f_old = self.f_old, self.f_new
        self.m_new, self.m_old = self.m_old, self.m_new
        self.fill_patch(time, coarse)
        d = C.byref(self.desc)
        lo, hi = _i3(self.vlo), _i3(self.vhi)
        force = None
        if self.has_force:
            for c in range(3):
                self._check(self.L.ilbm_box_setval(self.fx.ref(c, 1), float(self.uniform_force[c]), None),
                            "box_setval")
            if self.ib:
                s = self.smS
                self._check(self.L.ilbm_ib_update_forcing(
                    self.ib, s.ref(0, 1), s.ref(1, 1), s.ref(2, 1), s.ref(3, 1),
                    self.fx.ref(0, 1), self.fx.ref(1, 1), self.fx.ref(2, 1), self.dt, time, None), "ib_update_forcing")
            self._check(self.L.ilbm_box_fill_periodic(d, self.fx.ref(), lo, hi, None), "box_fill_periodic")
            force = self.fx.ref()
        self._check(self.L.ilbm_box_collide(d, self.sfS.ref(), self.smS.ref(), force, lo, hi, None), "box_collide")
        self._check(self.L.ilbm_box_fill_periodic(d, self.sfS.ref(), lo, hi, None), "box_fill_periodic")
        self._check(self.L.ilbm_box_bounce_back(d, self.sfS.ref(), lo, hi, None), "box_bounce_back")
        force = self.fx.ref() if self.has_force else None
        self._check(self.L.ilbm_box_stream_macro(d, self.sfS.ref(), self.f_new.ref(), self.m_new.ref(), force, lo, hi,
                                                 None), "box_stream_macro")

    def step(self):
        """timeStepWithSubcycling at max_level = 0, then ComputeIBForce."""
        time = self.t_new
        self.t_old = self.t_new
        self.t_new += self.dt
        self.advance_level(time)
        self.istep += 1
        if self.ib:
            self.compute_ib_force(time + self.dt, self.istep)

    def advance(self, n: int):
        for _ in range(n):
            self.step()

    # -- diagnostics -----------------------------------------------------------
    def box_sum(self, arr: _Arr, comp: int) -> float:
        out = C.c_double(0.0)
        self._check(self.L.ilbm_box_sum(arr.ref(), comp, _i3(self.vlo), _i3(self.vhi), C.byref(out), None), "box_sum")
        return out.value

    def compute_ib_force(self, time, step):
        """ComputeIBForce (AmrCoreLBM.cpp:456-701) of this box: the Eulerian
        force -sum(F dV)/dt of the exported force, the marker force
        -sum(F dv), time-centred body force, coefficients."""
        c = self.case
        dV = c.dx[0] * c.dx[1] * (c.dx[2] if self.dim == 3 else 1.0)
        s = [self.box_sum(self.fx, k) for k in range(3)]
        if self.dim != 3:
            s[2] = 0.0
        eul = tuple(-(x * dV) / self.dt for x in s)
        ms = (C.c_double * 3)()
        self._check(self.L.ilbm_ib_force_sum(self.ib, ms), "ib_force_sum")
        mark = tuple(-ms[k] for k in range(3))
        body = mark if c.force_eval_method == cf.FORCE_EVAL_MARKER else eul
        tc = tuple(0.5 * (p + b) for p, b in zip(self.prev_body, body)) if self.have_prev else body
        self.prev_body, self.have_prev = body, True
        D = 2.0 * c.R
        denom = 0.5 * 1.0 * c.U0 * c.U0 * (math.pi * c.R * c.R if self.dim == 3 else D)
        co = (lambda f: f / denom) if denom > 0.0 else (lambda f: 0.0)
        row = [float(step), time, tc[0], tc[1], co(tc[0]), co(tc[1]), body[0], body[1], co(body[0]), co(body[1]),
               eul[0], eul[1], co(eul[0]), co(eul[1]), mark[0], mark[1], co(mark[0]), co(mark[1]),
               eul[0], eul[1], co(eul[0]), co(eul[1]), (time * c.U0 / D) if D > 0.0 else 0.0]
        if self.dim == 3:
            row += [tc[2], co(tc[2]), body[2], co(body[2]), eul[2], co(eul[2]), mark[2], co(mark[2]), eul[2],
                    co(eul[2])]
        self.record = np.array(row)

    def markers(self):
        n = C.c_int64(0)
        self._check(self.L.ilbm_ib_num_markers(self.ib, C.byref(n)), "ib_num_markers")
        xyz, uvw, f = np.zeros(3 * n.value), np.zeros(3 * n.value), np.zeros(3 * n.value)
        dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
        self._check(self.L.ilbm_ib_marker_state(self.ib, dp(xyz), dp(uvw), dp(f)), "ib_marker_state")
        return xyz.reshape(-1, 3), uvw.reshape(-1, 3), f.reshape(-1, 3)

    def _valid_of(self, a: _Arr, nc):
        native._check(self.L.ilbm_device_synchronize(), "synchronize")
        return a.download()[(slice(0, nc),) + self._valid()].reshape(nc, -1)

    def f(self) -> np.ndarray:
        return self._valid_of(self.f_new, self.q)

    def macro(self, ncomp: int = 4) -> np.ndarray:
        return self._valid_of(self.m_new, ncomp)

    def forcing(self) -> np.ndarray:
        return self._valid_of(self.fx, 3)

    def mass_sum(self) -> float:
        return self.box_sum(self.m_new, 0)

    def derived(self) -> np.ndarray:
        """vorticity and pressure of macro_new (UpdateDerivedMacroFields)."""
        out = _Arr(6, self.flo, self.fhi, False)
        self._check(self.L.ilbm_box_setval(out.ref(), 0.0, None), "box_setval")
        self.update_derived(self.t_new, None, out)
        res = self._valid_of(out, 6)[4:6]
        out.free()
        return res

    def update_derived(self, time, coarse: Optional["BoxLevel"] = None, target: Optional[_Arr] = None):
        """UpdateDerivedMacroFields (AmrCoreLBM.cpp:150-168): FillPatchMacro
        of comps 0..3 with one ghost at `time` (the level's macro_new), then
        vorticity and pressure into comps 4, 5 of macro_new."""
        g = [1 if d < self.dim else 0 for d in range(3)]
        plo = tuple(self.vlo[d] - g[d] for d in range(3))
        phi = tuple(self.vhi[d] + g[d] for d in range(3))
        patch = _Arr(4, plo, phi, False)
        self._check(self.L.ilbm_box_copy(self.m_new.ref(0, 4), patch.ref(), _i3(self.vlo), _i3(self.vhi), 4, None),
                    "box_copy")
        if coarse is None:
            self._check(self.L.ilbm_box_fill_periodic(C.byref(self.desc), patch.ref(), _i3(self.vlo), _i3(self.vhi),
                                                      None), "box_fill_periodic")
            self._check(self.L.ilbm_box_fill_physical(C.byref(self.desc), patch.ref(), 0, None), "box_fill_physical")
        else:
            self._check(self.L.ilbm_box_cf_fill(C.byref(self.desc), coarse.m_old.ref(), coarse.m_new.ref(),
                                                coarse.t_old, coarse.t_new, time, None, patch.ref(), 4,
                                                _i3(self.vlo), _i3(self.vhi), None), "box_cf_fill")
        dst = target if target is not None else self.m_new
        self._check(self.L.ilbm_box_derived(C.byref(self.desc), patch.ref(), dst.ref(), _i3(self.vlo), _i3(self.vhi),
                                            None), "box_derived")
        patch.free()


class BoxHierarchy:
    """Two levels at refinement ratio 2, each one box, stepped through the
    per-box ABI as timeStepWithSubcycling does (AmrCoreLBM.cpp:1688-1754);
    the immersed body lives on the fine level.  The boxes stay fixed;
    derived fields follow the reference's schedule (every regrid_int steps
    of a level and after each AverageDownTo)."""

    def __init__(self, case: "cf.Case", fine_box, regrid_int: int, pad: bool = False):
        case.validate()
        self.case = case
        dim = case.lattice.dim
        self.regrid_int = regrid_int
        self.nmac = case.nmac
        L = native.load()
        self.L = L
        coarse_case = case.replace(use_ib=False)
        u0 = case.bc_lo_val[0][1:4] if not case.periodic[0] else (0.0, 0.0, 0.0)
        self.c0 = BoxLevel(coarse_case, cf.uniform_init(coarse_case, u=tuple(u0)), pad=pad, nmac=self.nmac)
        # MakeNewLevelFromScratch: initdata, derived fields, macro_old = macro_new
        self.c0.update_derived(0.0)
        self.c0._check(L.ilbm_box_copy(self.c0.m_new.ref(), self.c0.m_old.ref(), _i3(self.c0.flo), _i3(self.c0.fhi),
                                       self.nmac, None), "box_copy")
        fine_case = case.replace(n=tuple(case.n[d] * 2 if d < dim else 1 for d in range(3)),
                                 dx=tuple(case.dx[d] / 2.0 if d < dim else case.dx[d] for d in range(3)))
        flo, fhi = fine_box
        self.c1 = BoxLevel(fine_case, None, vlo=flo, vhi=fhi, pad=pad, nmac=self.nmac)
        # MakeNewLevelFromCoarse: InterpFromCoarseLevel of the macros, f = feq
        self.c1._check(L.ilbm_box_interp_from_coarse(C.byref(self.c1.desc), self.c0.m_new.ref(), self.c1.m_new.ref(),
                                                     self.nmac, _i3(self.c1.vlo), _i3(self.c1.vhi), None),
                       "box_interp_from_coarse")
        self.c1.equilibrium()
        self.c1._check(L.ilbm_box_copy(self.c1.m_new.ref(), self.c1.m_old.ref(), _i3(self.c1.vlo), _i3(self.c1.vhi),
                                       self.nmac, None), "box_copy")
        self.c1.update_derived(0.0, self.c0)
        self.istep = 0

    def _substep(self, lev: BoxLevel, time, coarse: Optional[BoxLevel]):
        lev.t_old = lev.t_new
        lev.t_new += lev.dt
        lev.advance_level(time, coarse)
        if self.regrid_int > 0 and (lev.istep + 1) % self.regrid_int == 0:
            lev.update_derived(lev.t_new, coarse)
        lev.istep += 1

    def step(self):
        c0, c1 = self.c0, self.c1
        cur = c0.t_new
        self._substep(c0, cur, None)
        for i in (1, 2):
            self._substep(c1, cur + (i - 1) * c1.dt, c0)
        # AverageDownTo(0), then the coarse derived fields
        c0._check(self.L.ilbm_box_average_down(C.byref(c1.desc), c1.f_new.ref(), c1.m_new.ref(), _i3(c1.vlo),
                                               _i3(c1.vhi), c0.f_new.ref(), c0.m_new.ref(), self.nmac, None),
                  "box_average_down")
        c0.update_derived(c0.t_new)
        cur += c0.dt
        self.istep += 1
        if c1.ib:
            c1.compute_ib_force(cur, self.istep)
        c0.t_new = c1.t_new = cur

    def advance(self, n: int):
        for _ in range(n):
            self.step()

    def close(self):
        self.c0.close()
        self.c1.close()
