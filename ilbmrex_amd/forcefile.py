"""force.dat wire format and the Cd / Cl / Strouhal analysis of a force
history.

Writer: the column layout ComputeIBForce appends per step
(Source/Core/AmrCoreLBM.cpp:636-700): comment header, then step, time, the
body force and coefficients time-centred, raw, Eulerian, marker and
momentum-exchange, t_conv, and in 3D ten z-columns; numbers as C++ ostream
prints a double by default (six significant digits).

Reader and estimators: the acceptance analysis of the reference's cylinder
example (Examples/Cylinder_flow/analyze_force.py) is an algorithm whose
steps the Strouhal number depends on, so its steps are kept; the code here is
written from this specification:

* rows: every line that is not blank or a comment is split into numbers
  (tokens that are not numbers are skipped, bytes that are not ASCII and NUL
  characters are ignored); rows of fewer than six numbers are skipped.
  Per row: step = column 0, time = column 1, t_conv = column 22 for rows of
  more than 22 numbers, column 18 for rows of 19-22 numbers (the legacy
  layout), else absent; (Cd, Cl) from the columns of the chosen source
  (primary 4/5, eulerian 12/13, marker 16/17, momentum exchange 20/21), or
  4/5 when the row is too short for them.
* Strouhal from peaks: mean m and population standard deviation s of the
  lift; no estimate if s < 1e-12.  Candidates are interior samples higher
  than the previous one and not lower than the next.  Walking them in time
  order, keep a candidate whose excess over m is at least amp_frac * s and
  whose time is at least min_sep_frac * D / (0.5 U) after the last kept one
  (the shortest period of a Strouhal number of 0.5).  With k >= 3 kept
  peaks, the period is the mean spacing of their times and St = D / (U T).
* Strouhal from the spectrum: at least 16 samples, a positive mean sample
  spacing dt = (t_last - t_first) / (n - 1), a demeaned lift of RMS at
  least 1e-12; the largest real-FFT magnitude among the bins whose St =
  f D / U lies in [st_min, st_max] (the first on ties); none if that is the
  zero-frequency bin.
* Summary: drop the first discard_frac of the rows, take t_conv as the time
  base when every remaining row has it (then D = U = 1), Cd mean, Cl mean
  and RMS about the mean; St from peaks when it falls in the St band, else
  from the spectrum; Cd and St matched against the acceptance bands of the
  Re = 100 cylinder (Cd in [1.20, 1.80], St in [0.14, 0.20], St only counted
  when the lift RMS reaches 0.02).
"""
from __future__ import annotations

import math
from pathlib import Path
from typing import List, Optional, Sequence, Tuple

import numpy as np

COLS_2D = ("step time Fx_body_tc Fy_body_tc Cd Cl Fx_body_raw Fy_body_raw Cd_raw Cl_raw "
           "Fx_eul Fy_eul Cd_eul Cl_eul Fx_marker Fy_marker Cd_marker Cl_marker "
           "Fx_me Fy_me Cd_me Cl_me t_conv").split()
COLS_3D = COLS_2D + ("Fz_body_tc Cz Fz_body_raw Cz_raw Fz_eul Cz_eul Fz_marker Cz_marker "
                     "Fz_me Cz_me").split()

SOURCE_COLS = {"primary": (4, 5), "eulerian": (12, 13), "marker": (16, 17), "me": (20, 21)}

Row = Tuple[int, float, Optional[float], float, float]


def _fmt(x: float) -> str:
    """std::ostream << double with the default precision (printf %g, 6 digits)."""
    return "%g" % x


def force_dat_text(history: np.ndarray, dim: int) -> str:
    hist = np.asarray(history, dtype=float)
    names = COLS_3D if dim == 3 else COLS_2D
    out = []
    if dim == 3:
        out.append("# 3D coefficients are normalized by frontal area A = pi R^2.")
        out.append("# In 3D, the historical Cl columns represent the y-force coefficient.")
        out.append("# Additional z-force and Cz columns are appended after t_conv.")
    else:
        out.append("# 2D coefficients are normalized by reference length D = 2R.")
    out.append("# " + "  ".join(names))
    for row in hist:
        if len(row) != len(names):
            raise ValueError(f"a {dim}D force row has {len(names)} columns, not {len(row)}")
        out.append("  ".join([str(int(row[0]))] + [_fmt(v) for v in row[1:]]))
    return "\n".join(out) + "\n"


def write_force_dat(path, history: np.ndarray, dim: int) -> None:
    Path(path).write_text(force_dat_text(history, dim))


def _numbers(text: str) -> List[float]:
    vals = []
    for tok in text.split():
        try:
            vals.append(float(tok))
        except ValueError:
            continue
    return vals


def _row(vals: List[float], cd_col: int, cl_col: int) -> Row:
    n = len(vals)
    t_conv = vals[22] if n > 22 else vals[18] if n > 18 else None
    if cl_col >= n:
        cd_col, cl_col = 4, 5
    return int(vals[0]), vals[1], t_conv, vals[cd_col], vals[cl_col]


def read_force(path, source: str = "primary") -> List[Row]:
    """(step, time, t_conv or None, Cd, Cl) of every data row."""
    cd_col, cl_col = SOURCE_COLS[source]
    text = Path(path).read_bytes().decode("ascii", errors="ignore")
    rows = []
    for line in text.splitlines():
        line = line.replace("\0", " ").strip()
        if not line or line[0] == "#":
            continue
        vals = _numbers(line)
        if len(vals) >= 6:
            rows.append(_row(vals, cd_col, cl_col))
    return rows


def estimate_st_peak(times: Sequence[float], cls: Sequence[float], d: float, u: float,
                     min_sep_period_frac: float = 0.4, amp_frac: float = 0.4):
    """Strouhal number from the spacing of the filtered lift maxima:
    (St or None, periods used, reason when None)."""
    y = np.asarray(cls, dtype=float)
    t = np.asarray(times, dtype=float)
    n = len(y)
    mean = float(y.sum()) / n
    std = math.sqrt(max(float(((y - mean) ** 2).sum()) / n, 0.0))
    if std < 1.0e-12:
        return None, 0, "no lift oscillation (standard deviation below 1e-12)"
    gap = min_sep_period_frac * d / (0.5 * u)
    height = amp_frac * std
    kept: List[float] = []
    for i in range(1, n - 1):
        if not (y[i] > y[i - 1] and y[i] >= y[i + 1]):
            continue
        if y[i] - mean < height:
            continue
        if kept and t[i] - kept[-1] < gap:
            continue
        kept.append(float(t[i]))
    if len(kept) < 3:
        return None, len(kept), f"only {len(kept)} lift maxima pass the amplitude and spacing filters"
    period = (kept[-1] - kept[0]) / (len(kept) - 1)
    if period <= 0.0:
        return None, len(kept) - 1, "maxima do not advance in time"
    return d / (u * period), len(kept) - 1, ""


def estimate_st_fft(times: Sequence[float], cls: Sequence[float], d: float, u: float,
                    st_min: float = 0.05, st_max: float = 0.50):
    """Strouhal number of the strongest spectral line of the lift inside
    [st_min, st_max]: (St or None, reason when None)."""
    n = len(times)
    if n < 16:
        return None, f"{n} samples; the spectrum needs 16"
    dt = (times[-1] - times[0]) / (n - 1)
    if dt <= 0.0:
        return None, "times do not increase"
    y = np.asarray(cls, dtype=float)
    y = y - y.mean()
    if math.sqrt(float(np.mean(y * y))) < 1.0e-12:
        return None, "no lift oscillation (RMS below 1e-12)"
    mag = np.abs(np.fft.rfft(y))
    st = np.fft.rfftfreq(n, d=dt) * d / u
    inside = np.flatnonzero((st >= st_min) & (st <= st_max))
    if inside.size == 0:
        return None, f"no spectral bin between St {st_min} and {st_max}"
    k = int(inside[np.argmax(mag[inside])])
    if k == 0:
        return None, "the strongest bin in the band is the mean"
    return float(st[k]), ""


def analyze(rows: Sequence[Row], u0: float, diameter: float, discard_frac: float = 0.25,
            time_basis: str = "auto", cd_range=(1.20, 1.80), st_range=(0.14, 0.20),
            cl_rms_min_periodic=2.0e-2, st_band=(0.05, 0.50)) -> dict:
    """Cd mean, Cl statistics and St of a force history, matched against the
    acceptance bands of the Re = 100 cylinder."""
    if len(rows) < 20:
        raise ValueError(f"{len(rows)} force rows; the analysis needs 20")
    tail = list(rows[int(len(rows) * discard_frac):])
    conv = all(r[2] is not None for r in tail)
    if time_basis == "conv" and not conv:
        raise ValueError("t_conv time base requested but a row has no t_conv column")
    if time_basis == "conv" or (time_basis == "auto" and conv):
        times, d_eff, u_eff = [r[2] for r in tail], 1.0, 1.0
    else:
        times, d_eff, u_eff = [r[1] for r in tail], diameter, u0
    cd = np.array([r[3] for r in tail], dtype=float)
    cl = np.array([r[4] for r in tail], dtype=float)
    cd_mean = float(cd.mean())
    cl_mean = float(cl.mean())
    cl_rms = float(np.sqrt(np.mean((cl - cl_mean) ** 2)))
    st_peak, periods, _ = estimate_st_peak(times, cl, d_eff, u_eff)
    st_fft, _ = estimate_st_fft(times, cl, d_eff, u_eff, *st_band)
    if st_peak is not None and st_band[0] <= st_peak <= st_band[1]:
        st, src = st_peak, "peak"
    elif st_fft is not None:
        st, src = st_fft, "fft_fallback"
    else:
        st, src = None, "none"
    periodic = cl_rms >= cl_rms_min_periodic
    return {"cd_mean": cd_mean, "cl_mean": cl_mean, "cl_rms": cl_rms, "st": st, "st_source": src,
            "periods": periods, "periodic_lift": periodic,
            "match_cd": cd_range[0] <= cd_mean <= cd_range[1],
            "match_st": bool(st is not None and periodic and st_range[0] <= st <= st_range[1])}
