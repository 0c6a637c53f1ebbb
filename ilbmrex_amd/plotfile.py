"""AMReX plotfiles of a level or a hierarchy (WritePlotFile,
Source/Core/AmrCoreLBM.cpp:1914-1930).

Variables (PlotFileMF / PlotFileVarNames, AmrCoreLBM.cpp:1760-1911): rho,
ux, uy, uz, vor, Pressure (macro_new), f_new_0 .. f_new_{Q-1}, and for the
cylinder / sphere marker geometry the level set phi = r - R, with the
velocity of cells inside the body (phi < 0) replaced by the body velocity
(ubx, uby, ubz) and their vorticity by 0.  A plotfile is named
plot_file + the coarse step in five digits (PlotFileName, :1756-1758).

The layout is AMReX's WriteMultiLevelPlotfile with one box per level:

    pltNNNNN/Header                  HyperCLaw-V1.1 text header
    pltNNNNN/Level_L/Cell_H          VisMF header (version 1, NFiles)
    pltNNNNN/Level_L/Cell_D_00000    "FAB <real descriptor><box> <ncomp>"
                                     line, then the FP64 data, component
                                     planes with x fastest (little endian)

written from the published format of AMReX's plotfile and VisMF headers.
AMReX and yt are absent here, so compatibility with their readers is
parity unpinned; read_plotfile parses what write_plotfile writes and the
tests check the round trip and the header structure.
"""
from __future__ import annotations

import os
from pathlib import Path
from typing import List, Sequence, Tuple

import numpy as np

REAL_DESC = "((8, (64 11 52 0 1 12 0 1023)),(8, (8 7 6 5 4 3 2 1)))"   # IEEE FP64, little endian


def plotfile_name(prefix: str, step: int) -> str:
    """amrex::Concatenate(plot_file, istep[0], 5)"""
    return f"{prefix}{int(step):05d}"


def _g(x: float) -> str:
    """std::ostream << double at precision 17 (printf %.17g)"""
    return format(float(x), ".17g")


def _iv(v, dim) -> str:
    return "(" + ",".join(str(int(v[d])) for d in range(dim)) + ")"


def _box(lo, hi, dim) -> str:
    return "(" + _iv(lo, dim) + " " + _iv(hi, dim) + " " + _iv((0, 0, 0), dim) + ")"


def var_names(case) -> List[str]:
    """PlotFileVarNames (AmrCoreLBM.cpp:1860-1911)."""
    names = ["rho", "ux", "uy", "uz", "vor", "Pressure"][:case.nmac]
    names += [f"macro_{m}" for m in range(6, case.nmac)]
    names += [f"f_new_{i}" for i in range(case.lattice.q)]
    if _has_phi(case):
        names.append("phi")
    return names


def _has_phi(case) -> bool:
    # marker_levelset_ibm: the level set exists for the cylinder / sphere geometry
    return bool(case.use_ib) and case.geometry_type == 0


def level_data(case, lev: int, box_lo, box_hi, macro: np.ndarray, derived: np.ndarray, f: np.ndarray) -> np.ndarray:
    """PlotFileMF of one level: [ncomp][nz][ny][nx] over the level's box.

    macro: [4][cells] rho, ux, uy, uz; derived: [2][cells] vorticity,
    pressure (or None when nmac = 4); f: [Q][cells]; cells x fastest."""
    dim = case.lattice.dim
    shape = tuple(box_hi[d] - box_lo[d] + 1 if d < dim else 1 for d in (2, 1, 0))
    comps = [np.asarray(macro[c], dtype=np.float64).reshape(shape) for c in range(4)]
    if case.nmac > 4:
        if derived is None:
            raise ValueError("nmac = 6 needs the derived fields")
        comps += [np.asarray(derived[c], dtype=np.float64).reshape(shape) for c in range(2)]
    comps += [np.zeros(shape)] * max(case.nmac - 6, 0)
    comps += [np.asarray(f[q], dtype=np.float64).reshape(shape) for q in range(case.lattice.q)]
    out = np.stack(comps)
    if _has_phi(case):
        fac = 2 ** lev
        dx = [case.dx[d] / fac for d in range(3)]
        ax = [case.plo[d] + (np.arange(box_lo[d], box_hi[d] + 1) + 0.5) * dx[d] if d < dim else np.zeros(1)
              for d in range(3)]
        z, y, x = np.meshgrid(ax[2], ax[1], ax[0], indexing="ij")
        r2 = (x - case.x0) ** 2 + (y - case.y0) ** 2
        if dim == 3:
            r2 = r2 + (z - case.z0) ** 2
        phi = np.sqrt(r2) - case.R
        inside = phi < 0.0
        out[1][inside] = case.ub[0]
        out[2][inside] = case.ub[1]
        out[3][inside] = case.ub[2]
        if case.nmac > 4:
            out[4][inside] = 0.0
        out = np.concatenate([out, phi[None]])
    return out


def write_plotfile(path, case, levels: Sequence[Tuple[Sequence[int], Sequence[int], np.ndarray]], time: float,
                   level_steps: Sequence[int], ref_ratio: int = 2) -> None:
    """WriteMultiLevelPlotfile for one box per level.  levels[l] =
    (box lo, box hi in level-l indices, data [ncomp][nz][ny][nx])."""
    dim = case.lattice.dim
    names = var_names(case)
    root = Path(path)
    root.mkdir(parents=True, exist_ok=True)
    finest = len(levels) - 1
    dom_hi0 = [case.n[d] - 1 for d in range(3)]
    prob_hi = [case.plo[d] + case.dx[d] * case.n[d] for d in range(3)]
    h = ["HyperCLaw-V1.1", str(len(names))] + names
    h += [str(dim), _g(time), str(finest)]
    h.append("".join(_g(case.plo[d]) + " " for d in range(dim)))
    h.append("".join(_g(prob_hi[d]) + " " for d in range(dim)))
    h.append("".join(f"{ref_ratio} " for _ in range(finest)))
    h.append("".join(_box((0, 0, 0), [(dom_hi0[d] + 1) * ref_ratio ** l - 1 for d in range(3)], dim) + " "
                     for l in range(finest + 1)))
    h.append("".join(f"{int(level_steps[l])} " for l in range(finest + 1)))
    for l in range(finest + 1):
        h.append("".join(_g(case.dx[d] / ref_ratio ** l) + " " for d in range(dim)))
    h += ["0", "0"]   # Cartesian coordinates, boundary width
    for l, (lo, hi, data) in enumerate(levels):
        data = np.ascontiguousarray(data, dtype="<f8")
        if data.shape[0] != len(names):
            raise ValueError(f"level {l}: {data.shape[0]} components, {len(names)} names")
        h.append(f"{l} 1 {_g(time)}")
        h.append(str(int(level_steps[l])))
        dxl = [case.dx[d] / ref_ratio ** l for d in range(3)]
        for d in range(dim):
            h.append(f"{_g(case.plo[d] + lo[d] * dxl[d])} {_g(case.plo[d] + (hi[d] + 1) * dxl[d])}")
        h.append(f"Level_{l}/Cell")
        _write_level(root / f"Level_{l}", lo, hi, data, dim)
    (root / "Header").write_text("\n".join(h) + "\n")


def _sci(x: float) -> str:
    """ostream in scientific notation at precision 16"""
    return format(float(x), ".16e")


def _write_level(d: Path, lo, hi, data: np.ndarray, dim: int) -> None:
    d.mkdir(parents=True, exist_ok=True)
    nc = data.shape[0]
    fab_head = f"FAB {REAL_DESC}{_box(lo, hi, dim)} {nc}\n".encode("ascii")
    with open(d / "Cell_D_00000", "wb") as fh:
        fh.write(fab_head)
        fh.write(data.tobytes(order="C"))
    mins = "".join(_sci(data[c].min()) + "," for c in range(nc))
    maxs = "".join(_sci(data[c].max()) + "," for c in range(nc))
    hdr = ["1", "1", str(nc), "0", f"(1 0", _box(lo, hi, dim), ")", "1", "FabOnDisk: Cell_D_00000 0", "",
           f"1,{nc}", mins, "", f"1,{nc}", maxs, ""]
    (d / "Cell_H").write_text("\n".join(hdr) + "\n")


def _parse_box(s: str, dim: int):
    s = s.strip()
    parts = s.replace("(", " ").replace(")", " ").replace(",", " ").split()
    v = [int(p) for p in parts]
    lo = tuple(v[0:dim]) + (0,) * (3 - dim)
    hi = tuple(v[dim:2 * dim]) + (0,) * (3 - dim)
    return lo, hi


def read_plotfile(path) -> dict:
    """Parse a plotfile written by write_plotfile (one box per level)."""
    root = Path(path)
    lines = (root / "Header").read_text().split("\n")
    it = iter(lines)
    out = {"version": next(it)}
    nv = int(next(it))
    out["names"] = [next(it) for _ in range(nv)]
    dim = int(next(it))
    out["dim"] = dim
    out["time"] = float(next(it))
    finest = int(next(it))
    out["prob_lo"] = [float(x) for x in next(it).split()]
    out["prob_hi"] = [float(x) for x in next(it).split()]
    out["ref_ratio"] = [int(x) for x in next(it).split()]
    dom = next(it)
    out["domains"] = dom
    out["level_steps"] = [int(x) for x in next(it).split()]
    out["dx"] = [[float(x) for x in next(it).split()] for _ in range(finest + 1)]
    out["coord"] = int(next(it))
    out["bwidth"] = int(next(it))
    levels = []
    for _ in range(finest + 1):
        lev, ngrid, _t = next(it).split()
        next(it)   # level step
        rb = [[float(x) for x in next(it).split()] for _ in range(dim)]
        cell = next(it)
        levels.append(_read_level(root / os.path.dirname(cell), dim, rb))
    out["levels"] = levels
    return out


def _read_level(d: Path, dim: int, realbox) -> dict:
    hl = (d / "Cell_H").read_text().split("\n")
    nc = int(hl[2])
    lo, hi = _parse_box(hl[5], dim)
    name, off = hl[8].split()[1], int(hl[8].split()[2])
    raw = (d / name).read_bytes()[off:]
    nl = raw.index(b"\n")
    head = raw[:nl].decode("ascii")
    if not head.startswith("FAB " + REAL_DESC):
        raise ValueError("unexpected FAB header: " + head)
    shape = tuple(hi[k] - lo[k] + 1 for k in (2, 1, 0))
    data = np.frombuffer(raw[nl + 1:nl + 1 + 8 * nc * int(np.prod(shape))], dtype="<f8").reshape((nc,) + shape)
    return {"lo": lo, "hi": hi, "ncomp": nc, "data": data.copy(), "realbox": realbox, "fab_header": head}


def level_plotfile(path, lev, step: int) -> None:
    """Plotfile of a resident single level (native.Level) after an advance
    with want_state."""
    case = lev.case
    macro = lev.macro()
    derived = lev.derived() if case.nmac > 4 else None
    f = lev.f()
    hi = tuple(case.n[d] - 1 if d < case.lattice.dim else 0 for d in range(3))
    data = level_data(case, 0, (0, 0, 0), hi, macro, derived, f)
    write_plotfile(path, case, [((0, 0, 0), hi, data)], lev.info()["time"], [step])


def amr_plotfile(path, h, step: int) -> None:
    """Plotfile of a hierarchy (native.AmrHierarchy) after an advance with
    want_state; every level's box and its own derived fields."""
    case = h.case
    boxes, _ = h.boxes()
    dim = case.lattice.dim
    levels = []
    steps = []
    for l in range(h.nlev):
        if l == 0:
            lo, hi = (0, 0, 0), tuple(case.n[d] - 1 if d < dim else 0 for d in range(3))
        else:
            lo, hi = boxes[l - 1]
        derived = h.derived(l) if case.nmac > 4 else None
        levels.append((lo, hi, level_data(case, l, lo, hi, h.macro(l), derived, h.f(l))))
        steps.append(step * 2 ** l)
    write_plotfile(path, case, levels, h.info()["time"], steps)
