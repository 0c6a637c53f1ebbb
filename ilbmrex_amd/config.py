"""Case description: lattice tables, boundary conditions, IB parameters.

This mirrors what ``AmrCoreLBM::ReadParameters`` and the AmrCoreLBM
constructor assemble from an ``inputs`` file
(Source/Core/AmrCoreLBM.cpp:190-356, 1188-1507) for one level, plus
``MarkerIBParams`` (Source/IBM/MarkerIBParams.H:8-132).

The lattice tables are the ones the reference's example inputs define, in
their direction order, with ``w = wis / wim`` evaluated in double precision
(AmrCoreLBM.cpp:1218-1221):

* D2Q9  -- Examples/Cylinder_flow/inputs:35-40
* D3Q19 -- Examples/SquareDuct_flow/inputs:36-41
* D3Q27 -- Examples/Sphere_flow/inputs:34-39
"""
from __future__ import annotations

import dataclasses
import math
import re
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence

# AMReX BCType codes used by amrbc.bc_lo / amrbc.bc_hi.
INT_DIR, FOEXTRAP, EXT_DIR, USER_1, USER_2 = 0, 2, 3, 1001, 1002
SUPPORTED_BC = (INT_DIR, FOEXTRAP, EXT_DIR, USER_1, USER_2)

GEOM_CYLINDER, GEOM_BOX, GEOM_USER = 0, 1, 2
COUPLING_EXPLICIT, COUPLING_IVC, COUPLING_EXPLICIT_DIAG = 0, 1, 2
FORCE_EVAL_EULERIAN, FORCE_EVAL_MARKER, FORCE_EVAL_ME = 0, 1, 2

# Population storage formats of the MI355X path.
STORE_F64, STORE_F32, STORE_F16S = 0, 1, 2
STORAGE_NAMES = {"f64": STORE_F64, "f32": STORE_F32, "f16s": STORE_F16S}


@dataclass(frozen=True)
class Lattice:
    name: str
    dim: int
    cx: tuple
    cy: tuple
    cz: tuple
    wis: tuple
    wim: tuple

    @property
    def q(self) -> int:
        return len(self.cx)

    @property
    def w(self) -> List[float]:
        return [float(a) / float(b) for a, b in zip(self.wis, self.wim)]

    def opposite(self) -> List[int]:
        out = []
        for q in range(self.q):
            r = (-self.cx[q], -self.cy[q], -self.cz[q])
            out.append(next(p for p in range(self.q)
                            if (self.cx[p], self.cy[p], self.cz[p]) == r))
        return out


D2Q9 = Lattice(
    "D2Q9", 2,
    (0, 1, 0, -1, 0, 1, -1, -1, 1),
    (0, 0, 1, 0, -1, 1, 1, -1, -1),
    (0,) * 9,
    (4.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0),
    (9.0, 9.0, 9.0, 9.0, 9.0, 36.0, 36.0, 36.0, 36.0),
)
D3Q19 = Lattice(
    "D3Q19", 3,
    (0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0),
    (0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1),
    (0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1),
    (12.0, 2.0, 2.0, 2.0, 2.0, 2.0, 2.0) + (1.0,) * 12,
    (36.0,) * 19,
)
D3Q27 = Lattice(
    "D3Q27", 3,
    (0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0, 1, -1, 1, -1, 1, -1, -1, 1),
    (0, 0, 0, 1, -1, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 1, -1, 1, -1, 1, -1, -1, 1, 1, -1),
    (0, 0, 0, 0, 0, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0, -1, 1, -1, 1, 1, -1, -1, 1, 1, -1, 1, -1),
    (8.0,) + (2.0,) * 6 + (1.0,) * 20,
    (27.0,) * 7 + (54.0,) * 12 + (216.0,) * 8,
)
LATTICES = {"D2Q9": D2Q9, "D3Q19": D3Q19, "D3Q27": D3Q27}


@dataclass
class Case:
    """One level of an iLBMReX run (fields follow the reference's names)."""

    lattice: Lattice = D2Q9
    n: tuple = (64, 64, 1)
    plo: tuple = (0.0, 0.0, 0.0)
    dx: tuple = (1.0, 1.0, 1.0)
    periodic: tuple = (1, 1, 1)
    bc_lo: tuple = (INT_DIR, INT_DIR, INT_DIR)
    bc_hi: tuple = (INT_DIR, INT_DIR, INT_DIR)
    bc_lo_val: tuple = ((0.0,) * 4,) * 3   # per dim: rho, ux, uy, uz
    bc_hi_val: tuple = ((0.0,) * 4,) * 3
    nu: float = 0.01
    U0: float = 0.1                         # AmrCoreLBM.cpp:1243 default
    nmac: int = 6
    prescribed_force: Optional[tuple] = None
    # immersed boundary (ibm.*)
    use_ib: bool = False
    x0: float = 0.0
    y0: float = 0.0
    z0: float = 0.0
    R: float = 0.0
    coupling_method: int = COUPLING_IVC     # MarkerIBParams.H:26
    delta_type: int = 0
    loop_ns: int = 2
    mdf_relax: float = 1.0
    explicit_diag_eps: float = 1.0e-8
    geometry_type: int = GEOM_CYLINDER
    n_marker: int = 0
    rd: float = 0.0
    ub: tuple = (0.0, 0.0, 0.0)
    om: tuple = (0.0, 0.0, 0.0)
    ivc_diag_reg: float = 1.0e-12
    ivc_rebuild_matrix: int = 0
    renormalize_delta: int = 0
    force_eval_method: int = FORCE_EVAL_ME
    # user-defined pitching section (IBMUserDefinedGeometry.H constants)
    airfoil_chord: float = 20.0
    airfoil_thickness: float = 0.12
    pitch_mean: float = 0.0
    pitch_amplitude: float = 0.2617993877991494
    pitch_frequency: float = 0.0015
    pitch_phase: float = 0.0
    pivot_x: float = 0.0
    # 2D marker box
    box_lo: tuple = (0.0, 0.0, 0.0)
    box_hi: tuple = (0.0, 0.0, 0.0)
    box_ds: float = 0.0
    box_lid_only: int = 0
    box_lid_u: tuple = (0.0, 0.0, 0.0)
    box_wall_tol: float = 0.0
    # MI355X path only
    storage: int = STORE_F64
    name: str = "case"

    @property
    def dim(self) -> int:
        return self.lattice.dim

    @property
    def ncells(self) -> int:
        return int(self.n[0]) * int(self.n[1]) * int(self.n[2])

    @property
    def tau(self) -> float:
        # AmrCoreLBM.cpp:225
        return 3.0 * self.nu / self.dx[0] + 0.5

    def replace(self, **kw) -> "Case":
        return dataclasses.replace(self, **kw)

    def validate(self) -> None:
        if self.dim == 2 and self.n[2] != 1:
            raise ValueError("2D case needs n[2] == 1")
        for d in range(self.dim):
            for b in (self.bc_lo[d], self.bc_hi[d]):
                if b not in SUPPORTED_BC:
                    raise ValueError(f"unsupported BC type {b} in dim {d}")
            if self.periodic[d] and (self.bc_lo[d] != INT_DIR or self.bc_hi[d] != INT_DIR):
                raise ValueError(f"periodic dim {d} must use int_dir (0)")
            if not self.periodic[d] and (self.bc_lo[d] == INT_DIR or self.bc_hi[d] == INT_DIR):
                raise ValueError(f"non-periodic dim {d} needs a physical BC")


# ---------------------------------------------------------------------------
# ParmParse subset: "key = v1 v2 ..." lines, '#' comments, later keys win.
# ---------------------------------------------------------------------------
def parse_inputs(text: str, overrides: Sequence[str] = ()) -> Dict[str, List[str]]:
    out: Dict[str, List[str]] = {}
    lines = list(text.splitlines()) + list(overrides)
    for raw in lines:
        line = raw.split("#", 1)[0].strip()
        if not line or "=" not in line:
            continue
        k, v = line.split("=", 1)
        toks = [t.strip("'\"") for t in v.split()]
        out[k.strip()] = toks
    return out


def _get(pp, key, default, conv=float):
    if key not in pp:
        return default
    vals = pp[key]
    return conv(vals[0]) if vals else default


def _arr(pp, key, conv=float):
    return [conv(x) for x in pp.get(key, [])]


def case_from_inputs(text: str, overrides: Sequence[str] = (), storage: str = "f64") -> Case:
    """Builds a level-0 ``Case`` the way ReadParameters reads an inputs file
    (Source/Core/AmrCoreLBM.cpp:1188-1507, boundary values :258-325).
    AMR keys (amr.max_level > 0) are accepted but only level 0 is built."""
    pp = parse_inputs(text, overrides)
    ncell = [int(x) for x in pp["amr.n_cell"]]
    plo = [float(x) for x in pp.get("geometry.prob_lo", ["0", "0", "0"])]
    phi = [float(x) for x in pp["geometry.prob_hi"]]
    per = [int(x) for x in pp.get("geometry.is_periodic", ["0", "0", "0"])]
    if "lbm3d.ndir" in pp and "lbm2d.ndir" not in pp:
        dim = 3
    elif "lbm2d.ndir" in pp and "lbm3d.ndir" not in pp:
        dim = 2
    else:
        dim = 2 if ncell[2] <= 8 and "lbm2d.ndir" in pp else 3
    pre = "lbm2d" if dim == 2 else "lbm3d"
    q = int(_get(pp, f"{pre}.ndir", 9 if dim == 2 else 19, int))
    lat = None
    wis, wim = _arr(pp, f"{pre}.wis"), _arr(pp, f"{pre}.wim")
    dx_, dy_, dz_ = (_arr(pp, f"{pre}.dirx"), _arr(pp, f"{pre}.diry"), _arr(pp, f"{pre}.dirz"))
    for L in LATTICES.values():
        if L.dim == dim and L.q == q and list(map(float, L.cx)) == dx_ and \
                list(map(float, L.cy)) == dy_ and list(map(float, L.cz)) == dz_ and \
                list(L.wis) == wis and list(L.wim) == wim:
            lat = L
    if lat is None:
        raise ValueError(f"lattice table in inputs is not one of {list(LATTICES)}")
    n = (ncell[0], ncell[1], ncell[2] if dim == 3 else 1)
    dx = tuple((phi[d] - plo[d]) / ncell[d] for d in range(3))
    if dim == 2:
        dx = (dx[0], dx[1], 1.0)
    allper = all(per[d] for d in range(dim))
    bc_lo = [INT_DIR] * 3
    bc_hi = [INT_DIR] * 3
    lo_val = [[0.0] * 4 for _ in range(3)]
    hi_val = [[0.0] * 4 for _ in range(3)]
    if not allper:
        blo = [int(x) for x in pp["amrbc.bc_lo"]]
        bhi = [int(x) for x in pp["amrbc.bc_hi"]]
        for d in range(dim):
            bc_lo[d], bc_hi[d] = blo[d], bhi[d]
            for side, arr, code in (("lo", lo_val, blo[d]), ("hi", hi_val, bhi[d])):
                if code in (EXT_DIR, USER_1, USER_2):
                    for c_i, nm in enumerate(("rho", "ux", "uy", "uz")):
                        arr[d][c_i] = _get(pp, f"amrbc.bc_{side}_{d}_{nm}_val", 0.0)
    pf = _arr(pp, "lbm.prescribed_force")
    prescribed = None
    if len(pf) >= 2:
        prescribed = (pf[0], pf[1], pf[2] if len(pf) > 2 else 0.0)
    use_cyl = int(_get(pp, "ibm.use_cylinder", 0, int))
    method = pp.get("ibm.method", ["none"])[0]
    if method in ("0", "none", ""):
        use_ib = False
    elif method in ("1", "marker", "iamr_marker"):
        use_ib = True
    elif method in ("diffuse", "sharp"):
        raise ValueError(f"ibm.method = {method} is not enabled (reference aborts)")
    else:
        raise ValueError(f"Unknown ibm.method = {method}")
    coupling = pp.get("ibm.coupling_method", pp.get("ibm.marker_df.coupling_method", ["ivc"]))[0]
    cm = {"ivc": COUPLING_IVC, "wu_shu": COUPLING_IVC, "explicit_diag": COUPLING_EXPLICIT_DIAG,
          "diag_ivc": COUPLING_EXPLICIT_DIAG}.get(coupling, COUPLING_EXPLICIT)
    geom = pp.get("ibm.marker_geometry", ["cylinder"])[0]
    gt = {"box": GEOM_BOX, "cavity_box": GEOM_BOX, "user_defined": GEOM_USER,
          "custom": GEOM_USER}.get(geom, GEOM_CYLINDER)
    fem = {"eulerian": FORCE_EVAL_EULERIAN, "marker": FORCE_EVAL_MARKER}.get(
        pp.get("ibm.force_eval_method", ["momentum_exchange"])[0], FORCE_EVAL_ME)
    renorm = int(_get(pp, "ibm.delta_renorm", _get(pp, "ibm.renormalize_delta", 0, int), int))
    c = Case(
        lattice=lat, n=n, plo=tuple(plo), dx=dx, periodic=tuple(per[:3]) if dim == 3 else (per[0], per[1], 1),
        bc_lo=tuple(bc_lo), bc_hi=tuple(bc_hi),
        bc_lo_val=tuple(tuple(v) for v in lo_val), bc_hi_val=tuple(tuple(v) for v in hi_val),
        nu=_get(pp, "lbmPhysicalParameters.nu", 0.01),
        U0=_get(pp, "lbmPhysicalParameters.U0", 0.1),
        nmac=int(_get(pp, "lbmPhysicalParameters.nmac", 6, int)),
        prescribed_force=prescribed,
        use_ib=bool(use_cyl) and use_ib,
        x0=_get(pp, "ibm.x0", 0.0), y0=_get(pp, "ibm.y0", 0.0), z0=_get(pp, "ibm.z0", 0.0),
        R=_get(pp, "ibm.R", 0.0),
        coupling_method=cm,
        delta_type=int(_get(pp, "ibm.delta_type", 0, int)),
        loop_ns=int(_get(pp, "ibm.loop_ns", 2, int)),
        mdf_relax=_get(pp, "ibm.mdf_relax", 1.0),
        explicit_diag_eps=_get(pp, "ibm.explicit_diag_eps", 1.0e-8),
        geometry_type=gt,
        n_marker=int(_get(pp, "ibm.n_marker", 0, int)),
        rd=_get(pp, "ibm.rd", 0.0),
        ub=(_get(pp, "ibm.ubx", 0.0), _get(pp, "ibm.uby", 0.0), _get(pp, "ibm.ubz", 0.0)),
        om=(_get(pp, "ibm.omx", 0.0), _get(pp, "ibm.omy", 0.0), _get(pp, "ibm.omz", 0.0)),
        ivc_diag_reg=_get(pp, "ibm.ivc_diag_reg", 1.0e-12),
        ivc_rebuild_matrix=int(_get(pp, "ibm.ivc_rebuild_matrix", 0, int)),
        renormalize_delta=renorm,
        force_eval_method=fem,
        # ibm.box_* (AmrCoreLBM.cpp:1317-1328)
        box_lo=(_get(pp, "ibm.box_xlo", 0.0), _get(pp, "ibm.box_ylo", 0.0), _get(pp, "ibm.box_zlo", 0.0)),
        box_hi=(_get(pp, "ibm.box_xhi", 0.0), _get(pp, "ibm.box_yhi", 0.0), _get(pp, "ibm.box_zhi", 0.0)),
        box_ds=_get(pp, "ibm.box_ds", 0.0),
        box_lid_only=int(_get(pp, "ibm.box_lid_only", 0, int)),
        box_lid_u=(_get(pp, "ibm.box_lid_ux", 0.0), _get(pp, "ibm.box_lid_uy", 0.0), _get(pp, "ibm.box_lid_uz", 0.0)),
        box_wall_tol=_get(pp, "ibm.box_wall_tol", 0.0),
        storage=STORAGE_NAMES[storage],
    )
    c.validate()
    return c


# ---------------------------------------------------------------------------
# Presets for the measured workloads
# ---------------------------------------------------------------------------
def periodic_box(lattice: Lattice, n: Sequence[int], nu: float, force=None,
                 storage: int = STORE_F64, name: str = "periodic") -> Case:
    n3 = (int(n[0]), int(n[1]), int(n[2]) if lattice.dim == 3 else 1)
    return Case(lattice=lattice, n=n3, nu=nu, prescribed_force=force, storage=storage,
                periodic=(1, 1, 1), name=name)


def force_validation(storage: int = STORE_F64) -> Case:
    """Examples/ForceValidation/inputs: periodic 64x64 D2Q9, nu=0.01, F=1e-6."""
    return periodic_box(D2Q9, (64, 64, 1), 0.01, force=(1.0e-6, 0.0, 0.0),
                        storage=storage, name="force_validation")


def square_duct(storage: int = STORE_F64) -> Case:
    """Examples/SquareDuct_flow/inputs: 8x32x32 D3Q19, user_2 walls in y,z."""
    wall = (1.0, 0.0, 0.0, 0.0)
    return Case(lattice=D3Q19, n=(8, 32, 32), plo=(0.0, -16.0, -16.0), nu=0.06,
                periodic=(1, 0, 0), bc_lo=(INT_DIR, USER_2, USER_2),
                bc_hi=(INT_DIR, USER_2, USER_2),
                bc_lo_val=((0.0,) * 4, wall, wall), bc_hi_val=((0.0,) * 4, wall, wall),
                prescribed_force=(1.0e-5, 0.0, 0.0), storage=storage, name="square_duct")


def cylinder_2d(nx=512, ny=256, D=32.0, xc=128.0, yc=128.0, U=0.05, Re=100.0,
                coupling=COUPLING_EXPLICIT, loop_ns=1, delta_type=4,
                storage: int = STORE_F64, uy_bias: float = 0.0) -> Case:
    """Flow past a fixed cylinder: ext_dir inflow at x-lo, foextrap outflow
    at x-hi, periodic in y (Examples/Cylinder_flow/inputs layout)."""
    nu = U * D / Re
    inflow = (1.0, U, uy_bias, 0.0)
    return Case(lattice=D2Q9, n=(nx, ny, 1), nu=nu, U0=U, periodic=(0, 1, 1),
                bc_lo=(EXT_DIR, INT_DIR, INT_DIR), bc_hi=(FOEXTRAP, INT_DIR, INT_DIR),
                bc_lo_val=(inflow, (0.0,) * 4, (0.0,) * 4),
                use_ib=True, x0=xc, y0=yc, R=0.5 * D, coupling_method=coupling,
                loop_ns=loop_ns, delta_type=delta_type, ivc_diag_reg=1.3,
                force_eval_method=FORCE_EVAL_ME, storage=storage, name="cylinder2d")


def sphere_3d(nx=256, ny=128, nz=128, D=24.0, center=(64.0, 64.0, 64.0), U=0.04,
              Re=100.0, coupling=COUPLING_EXPLICIT, loop_ns=1, delta_type=4,
              storage: int = STORE_F64) -> Case:
    """Flow past a fixed sphere on a uniform lattice (Examples/Sphere_flow
    layout with the D3Q19 table)."""
    nu = U * D / Re
    inflow = (1.0, U, 0.0, 0.0)
    return Case(lattice=D3Q19, n=(nx, ny, nz), nu=nu, U0=U, periodic=(0, 1, 1),
                bc_lo=(EXT_DIR, INT_DIR, INT_DIR), bc_hi=(FOEXTRAP, INT_DIR, INT_DIR),
                bc_lo_val=(inflow, (0.0,) * 4, (0.0,) * 4),
                use_ib=True, x0=center[0], y0=center[1], z0=center[2], R=0.5 * D,
                coupling_method=coupling, loop_ns=loop_ns, delta_type=delta_type,
                ivc_diag_reg=1.3, force_eval_method=FORCE_EVAL_ME, storage=storage,
                name="sphere3d")


def couette_channel(nx=4, ny=64, nu=0.064, U=0.1, storage: int = STORE_F64) -> Case:
    """Plane Couette flow: ext_dir walls at y-lo (at rest) and y-hi
    (ux = U), the wall conditions of Examples/Couette_flow/inputs, with x
    periodic."""
    z = (0.0,) * 4
    return Case(lattice=D2Q9, n=(nx, ny, 1), nu=nu, U0=U, periodic=(1, 0, 1),
                bc_lo=(INT_DIR, EXT_DIR, INT_DIR), bc_hi=(INT_DIR, EXT_DIR, INT_DIR),
                bc_lo_val=(z, (1.0, 0.0, 0.0, 0.0), z), bc_hi_val=(z, (1.0, U, 0.0, 0.0), z),
                storage=storage, name="couette")


def cavity_ibm_box(lid_only: int = 0, coupling: int = COUPLING_IVC, storage: int = STORE_F64) -> Case:
    """Examples/Cavity_flow/inputs_ibm_box: 64x96 D2Q9 with ext_dir walls at
    rest, nu = 0.02, an immersed box of markers over [0, 64]^2 (clipped
    inside the domain, spacing 1) whose top side moves at ux = 0.05, IVC
    with ivc_diag_reg = 20 and one forcing iteration."""
    w = (1.0, 0.0, 0.0, 0.0)
    return Case(lattice=D2Q9, n=(64, 96, 1), nu=0.02, U0=0.1, periodic=(0, 0, 1),
                bc_lo=(EXT_DIR, EXT_DIR, INT_DIR), bc_hi=(EXT_DIR, EXT_DIR, INT_DIR),
                bc_lo_val=(w, w, (0.0,) * 4), bc_hi_val=(w, w, (0.0,) * 4),
                use_ib=True, x0=32.0, y0=32.0, R=8.0, coupling_method=coupling, loop_ns=1,
                delta_type=4, ivc_diag_reg=20.0, geometry_type=GEOM_BOX,
                box_lo=(0.0, 0.0, 0.0), box_hi=(64.0, 64.0, 0.0), box_ds=1.0, box_lid_only=lid_only,
                box_lid_u=(0.05, 0.0, 0.0), force_eval_method=FORCE_EVAL_ME, storage=storage,
                name="cavity_ibm_box")


def uniform_init(case: Case, rho: float = 1.0, u=(0.0, 0.0, 0.0)):
    """initdata of the uniform-flow examples (e.g. Examples/Cylinder_flow/Prob.H:12-30)."""
    import numpy as np
    n = case.ncells
    m = np.empty((4, n), dtype=np.float64)
    m[0] = rho
    m[1], m[2], m[3] = u
    return m


def random_init(case: Case, seed: int = 20261202, drho: float = 0.01, du: float = 0.05):
    """Seeded perturbed state used for the roofline workload:
    rho = 1 + drho*xi, u = du*eta with xi, eta ~ U(-1, 1)."""
    import numpy as np
    rng = np.random.default_rng(seed)
    n = case.ncells
    m = np.empty((4, n), dtype=np.float64)
    m[0] = 1.0 + drho * rng.uniform(-1.0, 1.0, n)
    for c in (1, 2, 3):
        m[c] = du * rng.uniform(-1.0, 1.0, n)
    if case.dim == 2:
        m[3] = 0.0
    return m


def taylor_green_init(case: Case, u_max: float = 0.02, rho0: float = 1.0):
    """Examples/Taylor_Green_Vortex/Prob.H:9-54 at t = 0 (2D)."""
    import numpy as np
    nx, ny = case.n[0], case.n[1]
    Lx, Ly = nx * case.dx[0], ny * case.dx[1]
    kx, ky = 2.0 * np.pi / Lx, 2.0 * np.pi / Ly
    x = case.plo[0] + (0.5 + np.arange(nx)) * case.dx[0]
    y = case.plo[1] + (0.5 + np.arange(ny)) * case.dx[1]
    X, Y = np.meshgrid(x, y)          # [j][i]
    m = np.zeros((4, ny, nx))
    m[1] = -u_max * np.sqrt(ky / kx) * np.cos(kx * X) * np.sin(ky * Y)
    m[2] = u_max * np.sqrt(kx / ky) * np.sin(kx * X) * np.cos(ky * Y)
    p = -0.25 * rho0 * u_max * u_max * ((ky / kx) * np.cos(2 * kx * X) + (kx / ky) * np.cos(2 * ky * Y))
    m[0] = rho0 + 3.0 * p
    return m.reshape(4, -1)


def pitching_naca_amr(storage: int = STORE_F32, U: float = 0.04, Re: float = 1000.0, k_red: float = 1.0):
    """Pitching NACA 0012 on a three-level hierarchy (ratio 2): coarse
    512x256, chord 64 finest cells (16 coarse), 256 markers, +-10 deg about
    the quarter chord at reduced frequency k = pi f c / U, explicit direct
    forcing on the finest level.  Returns (case, boxes); the boxes cover the
    swept section with 3-cell nesting margins and stay fixed.

    At Re = 1000 (tau = 0.5019 coarse, 0.5077 finest) this BGK hierarchy
    diverges within about 12 coarse steps, in FP64 on the CPU as on the
    GPU, with explicit or IVC forcing; Re = 100 stays finite."""
    import math
    chord = 16.0
    nu = U * chord / Re
    f = k_red * U / (math.pi * chord)
    inflow = (1.0, U, 0.0, 0.0)
    c = Case(lattice=D2Q9, n=(512, 256, 1), nu=nu, U0=U, periodic=(0, 1, 1),
             bc_lo=(EXT_DIR, INT_DIR, INT_DIR), bc_hi=(FOEXTRAP, INT_DIR, INT_DIR),
             bc_lo_val=(inflow, (0.0,) * 4, (0.0,) * 4),
             use_ib=True, x0=128.0, y0=128.0, R=0.5 * chord, coupling_method=COUPLING_EXPLICIT,
             loop_ns=1, delta_type=4, geometry_type=GEOM_USER, n_marker=256, airfoil_chord=chord,
             airfoil_thickness=0.12, pitch_mean=0.0, pitch_amplitude=math.radians(10.0),
             pitch_frequency=f, pitch_phase=0.0, pivot_x=0.25 * chord,
             force_eval_method=FORCE_EVAL_MARKER, storage=storage, name="pitching_naca_amr")
    boxes = [((224, 208, 0), (351, 303, 0)), ((480, 448, 0), (639, 575, 0))]
    return c, boxes

