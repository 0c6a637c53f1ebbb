"""AdvancePhiAtLevel driven box by box through the per-box C ABI.

BoxLevel keeps the reference's own state on the device -- f_new / f_old
(ndir components) and macro_new / macro_old (rho, ux, uy, uz), AMReX Array4
layout with 3 ghost cells (AmrCoreLBM.H:175-180, 276) -- in buffers of the
library's allocator, and steps it with one ilbm_box_* call per stage in the
order of Source/Core/AdvanceLBMAtLevel.cpp:20-386 for a level of one box:

    swap f_new/f_old, macro_new/macro_old
    FillPatch(f_old) -> sfSborder   copy + periodic images + mesoscopicBcFill
    FillPatch(macro_old) -> smSborder   copy + periodic images + macroBcFill
    collide(sfSborder; smSborder, force)
    FillBoundary(sfSborder), bounce-back ghosts
    stream + forced macros -> f_new, macro_new

It is the reference-layout counterpart of the fused level (ilbm_level_*): the
same FP64 results, with the intermediate arrays of the reference visible.
"""
from __future__ import annotations

import ctypes as C
from typing import Optional

import numpy as np

from . import config as cf
from . import native

NG = 3


class BoxLevel:
    def __init__(self, case: "cf.Case", init_macro: np.ndarray, force: Optional[tuple] = None):
        case.validate()
        self.case = case
        self.L = native.load()
        self.desc = native.to_desc(case)
        dim = case.lattice.dim
        self.dim = dim
        self.q = case.lattice.q
        self.vlo = (0, 0, 0)
        self.vhi = tuple(case.n[d] - 1 if d < dim else 0 for d in range(3))
        g = [NG if d < dim else 0 for d in range(3)]
        self.flo = tuple(self.vlo[d] - g[d] for d in range(3))
        self.fhi = tuple(self.vhi[d] + g[d] for d in range(3))
        self.shape3 = tuple(self.fhi[d] - self.flo[d] + 1 for d in (2, 1, 0))
        arr = lambda nc: native.DeviceArray((nc,) + self.shape3)
        self.f_new, self.f_old, self.sfS = arr(self.q), arr(self.q), arr(self.q)
        self.m_new, self.m_old, self.smS = arr(4), arr(4), arr(4)
        self.force = None
        if force is not None and any(x != 0.0 for x in force):
            self.force = arr(3)
            h = np.zeros((3,) + self.shape3)
            for c in range(3):
                h[c] = float(force[c])
            self.force.upload(h)
        m = np.zeros((4,) + self.shape3)
        m[(slice(None),) + self._valid()] = np.asarray(init_macro, dtype=np.float64).reshape(
            (4,) + tuple(case.n[d] if d < dim else 1 for d in (2, 1, 0)))
        self.m_new.upload(m)
        self.m_old.upload(m)
        self._check(self.L.ilbm_box_equilibrium(C.byref(self.desc), self._fab(self.m_new), self._fab(self.f_new),
                                                self._i3(self.vlo), self._i3(self.vhi), None), "equilibrium")
        self._check(self.L.ilbm_box_copy(self._fab(self.f_new), self._fab(self.f_old), self._i3(self.vlo),
                                         self._i3(self.vhi), self.q, None), "copy")

    def _valid(self):
        return tuple(slice(self.vlo[d] - self.flo[d], self.vhi[d] - self.flo[d] + 1) for d in (2, 1, 0))

    @staticmethod
    def _i3(v):
        return (C.c_int * 3)(*v)

    def _fab(self, a):
        if a is None:
            return None
        f = native.Fab()
        f.p = a.ptr
        f.lo = native.I3(*self.flo)
        f.hi = native.I3(*self.fhi)
        f.ncomp = a.shape[0]
        return C.byref(f)

    def _check(self, rc, what):
        native._check(rc, f"ilbm_box_{what}")

    def _fill_patch(self, dst, src, meso):
        d = C.byref(self.desc)
        self._check(self.L.ilbm_box_setval(self._fab(dst), 0.0, None), "setval")
        self._check(self.L.ilbm_box_copy(self._fab(src), self._fab(dst), self._i3(self.vlo), self._i3(self.vhi),
                                         dst.shape[0], None), "copy")              # ParallelCopy
        self._check(self.L.ilbm_box_fill_periodic(d, self._fab(dst), self._i3(self.vlo), self._i3(self.vhi), None),
                    "fill_periodic")
        self._check(self.L.ilbm_box_fill_physical(d, self._fab(dst), int(meso), None), "fill_physical")

    def step(self):
        self.f_new, self.f_old = self.f_old, self.f_new
        self.m_new, self.m_old = self.m_old, self.m_new
        self._fill_patch(self.sfS, self.f_old, True)
        self._fill_patch(self.smS, self.m_old, False)
        d = C.byref(self.desc)
        lo, hi = self._i3(self.vlo), self._i3(self.vhi)
        self._check(self.L.ilbm_box_collide(d, self._fab(self.sfS), self._fab(self.smS), self._fab(self.force),
                                            lo, hi, None), "collide")
        self._check(self.L.ilbm_box_fill_periodic(d, self._fab(self.sfS), lo, hi, None), "fill_periodic")
        self._check(self.L.ilbm_box_bounce_back(d, self._fab(self.sfS), lo, hi, None), "bounce_back")
        self._check(self.L.ilbm_box_stream_macro(d, self._fab(self.sfS), self._fab(self.f_new), self._fab(self.m_new),
                                                 self._fab(self.force), lo, hi, None), "stream_macro")

    def advance(self, n: int):
        for _ in range(n):
            self.step()

    def _valid_of(self, a, nc):
        native._check(self.L.ilbm_device_synchronize(), "synchronize")
        return a.download()[(slice(0, nc),) + self._valid()].reshape(nc, -1)

    def f(self) -> np.ndarray:
        return self._valid_of(self.f_new, self.q)

    def macro(self) -> np.ndarray:
        return self._valid_of(self.m_new, 4)

    def mass_sum(self) -> float:
        out = C.c_double(0.0)
        self._check(self.L.ilbm_box_sum(self._fab(self.m_new), 0, self._i3(self.vlo), self._i3(self.vhi),
                                        C.byref(out), None), "sum")
        return out.value

    def derived(self) -> np.ndarray:
        """vorticity and pressure of macro_new (UpdateDerivedMacroFields)."""
        patch = native.DeviceArray((4,) + self.shape3)
        self._fill_patch(patch, self.m_new, False)
        out = native.DeviceArray((6,) + self.shape3)
        self._check(self.L.ilbm_box_setval(self._fab(out), 0.0, None), "setval")
        d = C.byref(self.desc)
        self._check(self.L.ilbm_box_derived(d, self._fab(patch), self._fab(out), self._i3(self.vlo),
                                            self._i3(self.vhi), None), "derived")
        res = self._valid_of(out, 6)[4:6]
        patch.free()
        out.free()
        return res
