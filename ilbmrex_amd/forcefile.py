"""force.dat wire format and the Cd / St analysis of a force history.

* write_force_dat / read_force: the column layout ComputeIBForce writes
  (Source/Core/AmrCoreLBM.cpp:636-700: header comments, step, time, then
  body force and coefficients time-centred, raw, Eulerian, marker and
  momentum-exchange, t_conv; in 3D ten z-columns appended) with C++
  ostream's default 6-significant-digit formatting, and the reader of
  Examples/Cylinder_flow/analyze_force.py:17-53 (current 23/33-column and
  legacy 19-column rows).
* estimate_st_peak / estimate_st_fft / analyze: the Strouhal estimators and
  acceptance summary of analyze_force.py:56-253 -- mean-crossing peak
  periods with amplitude and separation filters, else the dominant FFT bin
  inside a Strouhal band -- restated; pinned by synthetic sinusoids of known
  frequency.
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


def _fmt(x: float) -> str:
    """std::ostream << double with the default precision (printf %g, 6 digits)."""
    return "%g" % x


def force_dat_text(history: np.ndarray, dim: int) -> str:
    hist = np.asarray(history, dtype=float)
    out = []
    if dim == 3:
        out.append("# 3D coefficients are normalized by frontal area A = pi R^2.")
        out.append("# In 3D, the historical Cl columns represent the y-force coefficient.")
        out.append("# Additional z-force and Cz columns are appended after t_conv.")
    else:
        out.append("# 2D coefficients are normalized by reference length D = 2R.")
    out.append("# " + "  ".join(COLS_3D if dim == 3 else COLS_2D))
    ncol = len(COLS_3D if dim == 3 else COLS_2D)
    for row in hist:
        if len(row) != ncol:
            raise ValueError(f"expected {ncol} columns, got {len(row)}")
        out.append("  ".join([str(int(row[0]))] + [_fmt(v) for v in row[1:]]))
    return "\n".join(out) + "\n"


def write_force_dat(path, history: np.ndarray, dim: int) -> None:
    Path(path).write_text(force_dat_text(history, dim))


def _floats(line: str) -> List[float]:
    vals = []
    for tok in line.split():
        try:
            vals.append(float(tok))
        except ValueError:
            pass
    return vals


def read_force(path, source: str = "primary") -> List[Tuple[int, float, Optional[float], float, float]]:
    """(step, time, t_conv or None, Cd, Cl) per row."""
    c_cd, c_cl = SOURCE_COLS[source]
    rows = []
    for raw in Path(path).open("rb"):
        s = raw.decode("ascii", errors="ignore").replace("\x00", " ").strip()
        if not s or s.startswith("#"):
            continue
        v = _floats(s)
        if len(v) < 6:
            continue
        t_conv = v[22] if len(v) > 22 else (v[18] if len(v) > 18 else None)
        cd, cl = (v[c_cd], v[c_cl]) if c_cl < len(v) else (v[4], v[5])
        rows.append((int(v[0]), v[1], t_conv, cd, cl))
    return rows


def _local_maxima(y: Sequence[float]) -> List[int]:
    return [i for i in range(1, len(y) - 1) if y[i] > y[i - 1] and y[i] >= y[i + 1]]


def estimate_st_peak(times, cls, d, u, min_sep_period_frac=0.4, amp_frac=0.4):
    """St from the mean period between filtered lift maxima; (St or None,
    periods used, reason)."""
    n = len(cls)
    cmean = sum(cls) / n
    cstd = math.sqrt(max(sum((x - cmean) ** 2 for x in cls) / n, 0.0))
    if cstd < 1.0e-12:
        return None, 0, "lift variance too small"
    min_sep = min_sep_period_frac * (d / (0.5 * u))  # shortest period of St <= 0.5
    thr = amp_frac * cstd
    peaks, last = [], -1.0e100
    for i in _local_maxima(cls):
        if cls[i] - cmean < thr or times[i] - last < min_sep:
            continue
        peaks.append(i)
        last = times[i]
    if len(peaks) < 3:
        return None, len(peaks), "insufficient filtered peaks"
    periods = [times[peaks[k + 1]] - times[peaks[k]] for k in range(len(peaks) - 1)]
    tmean = sum(periods) / len(periods)
    if tmean <= 0:
        return None, len(periods), "non-positive period"
    return (1.0 / tmean) * d / u, len(periods), ""


def estimate_st_fft(times, cls, d, u, st_min=0.05, st_max=0.50):
    """St of the largest rfft bin of the demeaned lift inside [st_min, st_max]."""
    if len(times) < 16:
        return None, "too few points for fft"
    dt = (times[-1] - times[0]) / max(len(times) - 1, 1)
    if dt <= 0:
        return None, "non-positive dt"
    x = np.asarray(cls, dtype=float)
    x = x - x.mean()
    if float(np.sqrt(np.mean(x * x))) < 1.0e-12:
        return None, "lift variance too small"
    f = np.fft.rfftfreq(len(x), d=dt)
    mag = np.abs(np.fft.rfft(x))
    if len(mag) < 3:
        return None, "insufficient fft bins"
    st = f * d / u
    band = (st >= st_min) & (st <= st_max)
    if not np.any(band):
        return None, "no fft bins in requested St band"
    idx = np.where(band)[0]
    j = int(idx[np.argmax(mag[idx])])
    if j <= 0 or j >= len(f):
        return None, "invalid dominant bin in St band"
    return float(st[j]), ""


def analyze(rows, u0: float, diameter: float, discard_frac: float = 0.25, time_basis: str = "auto",
            cd_range=(1.20, 1.80), st_range=(0.14, 0.20), cl_rms_min_periodic=2.0e-2,
            st_band=(0.05, 0.50)) -> dict:
    """Cd mean, Cl statistics and St of a force history, with the reference
    acceptance bands of the Re = 100 cylinder."""
    if len(rows) < 20:
        raise ValueError("not enough samples")
    sub = rows[int(len(rows) * discard_frac):]
    has_conv = all(r[2] is not None for r in sub)
    use_conv = time_basis == "conv" or (time_basis == "auto" and has_conv)
    if use_conv:
        times, d_eff, u_eff = [r[2] for r in sub], 1.0, 1.0
    else:
        times, d_eff, u_eff = [r[1] for r in sub], diameter, u0
    cds = [r[3] for r in sub]
    cls = [r[4] for r in sub]
    cd_mean = sum(cds) / len(cds)
    cl_mean = sum(cls) / len(cls)
    cl_rms = math.sqrt(sum((x - cl_mean) ** 2 for x in cls) / len(cls))
    st_peak, nper, _ = estimate_st_peak(times, cls, d_eff, u_eff)
    st_fft, _ = estimate_st_fft(times, cls, d_eff, u_eff, *st_band)
    if st_peak is not None and st_band[0] <= st_peak <= st_band[1]:
        st, src = st_peak, "peak"
    elif st_fft is not None:
        st, src = st_fft, "fft_fallback"
    else:
        st, src = None, "none"
    periodic = cl_rms >= cl_rms_min_periodic
    return {"cd_mean": cd_mean, "cl_mean": cl_mean, "cl_rms": cl_rms, "st": st, "st_source": src,
            "periods": nper, "periodic_lift": periodic,
            "match_cd": cd_range[0] <= cd_mean <= cd_range[1],
            "match_st": bool(st is not None and periodic and st_range[0] <= st <= st_range[1])}
