"""Writes tests/golden/st_vectors.json: Strouhal-estimator vectors of the
reference's Examples/Cylinder_flow/analyze_force.py.

Run in a container that holds the reference checkout (it imports
analyze_force.py from there and runs its main() on force.dat files written
by ilbmrex_amd.forcefile).  The signals are synthetic lift histories made by
signal() below from the parameters stored with each vector, so the fixture
holds parameters and the reference's outputs only.  Vectors named
"recorded_*" replay a force history stored in tests/golden/ as .npy.

    python tests/golden/make_st_vectors.py /root/reference
"""
import contextlib
import io
import json
import math
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))


def lcg_noise(n, seed):
    """Deterministic uniform noise in [-0.5, 0.5) from a 32-bit LCG."""
    out = np.empty(n)
    s = seed & 0xFFFFFFFF
    for k in range(n):
        s = (1664525 * s + 1013904223) & 0xFFFFFFFF
        out[k] = s / 4294967296.0 - 0.5
    return out


def signal(p):
    """Force history [n][23] of a synthetic lift signal described by p."""
    n, D, U = p["n"], p["D"], p["U"]
    t = np.arange(1, n + 1, dtype=float) * p.get("dt", 1.0)
    f = p["st"] * U / D
    cl = p["amp"] * np.sin(2 * math.pi * f * t + p.get("phase", 0.0))
    if p.get("harm", 0.0):
        cl += p["harm"] * p["amp"] * np.sin(4 * math.pi * f * t + 0.3)
    if p.get("st2", 0.0):
        cl += p.get("amp2", 0.0) * np.sin(2 * math.pi * (p["st2"] * U / D) * t)
    if p.get("growth", 0.0):
        cl *= 1.0 - np.exp(-t / p["growth"])
    if p.get("noise", 0.0):
        cl += p["noise"] * lcg_noise(n, p.get("seed", 1))
    cl += p.get("offset", 0.0)
    h = np.zeros((n, 23))
    h[:, 0] = np.arange(1, n + 1)
    h[:, 1] = t
    h[:, 4] = p.get("cd", 1.35) + 0.02 * np.sin(4 * math.pi * f * t)
    h[:, 5] = cl
    h[:, 22] = t * U / D
    return h


def recorded(fn):
    """Force history [n][23] of a recorded series stored as [n][5] columns
    step, time, Cd, Cl, t_conv."""
    r = np.load(os.path.join(HERE, fn))
    h = np.zeros((len(r), 23))
    h[:, 0], h[:, 1], h[:, 4], h[:, 5], h[:, 22] = r[:, 0], r[:, 1], r[:, 2], r[:, 3], r[:, 4]
    return h


CASES = [
    dict(name="sinusoid_st0165", n=20000, D=32.0, U=0.05, st=0.165, amp=0.3, harm=0.05),
    dict(name="noisy_st0180", n=12000, D=32.0, U=0.05, st=0.18, amp=0.25, noise=0.08, seed=7),
    dict(name="growing_st0170", n=20000, D=32.0, U=0.05, st=0.17, amp=0.4, growth=4000.0, noise=0.01, seed=3),
    dict(name="beat_st0150", n=16000, D=24.0, U=0.04, st=0.15, amp=0.2, st2=0.21, amp2=0.12),
    dict(name="few_periods", n=1500, D=32.0, U=0.05, st=0.165, amp=0.3),
    dict(name="steady_lift", n=4000, D=32.0, U=0.05, st=0.165, amp=0.0),
    dict(name="tiny_amplitude", n=4000, D=32.0, U=0.05, st=0.165, amp=0.01, offset=0.002),
    dict(name="band_edge_high", n=8000, D=16.0, U=0.05, st=0.48, amp=0.3),
    dict(name="band_edge_low", n=40000, D=32.0, U=0.05, st=0.06, amp=0.3),
    dict(name="coarse_sampling", n=3000, D=32.0, U=0.05, st=0.2, amp=0.3, dt=5.0),
    dict(name="short_series", n=24, D=32.0, U=0.05, st=0.165, amp=0.3, dt=50.0),
]


def run_reference_main(af, path, u0, diameter, extra=()):
    argv = sys.argv
    sys.argv = ["analyze_force.py", "--force", path, "--u0", repr(u0), "--diameter", repr(diameter), *extra]
    buf = io.StringIO()
    try:
        with contextlib.redirect_stdout(buf):
            af.main()
    finally:
        sys.argv = argv
    out = {}
    for tok in buf.getvalue().split():
        if "=" in tok:
            k, v = tok.split("=", 1)
            out[k] = v
    return out, buf.getvalue()


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    sys.path.insert(0, os.path.join(ref, "Examples", "Cylinder_flow"))
    sys.path.insert(0, ROOT)
    import analyze_force as af          # the reference's estimators
    from ilbmrex_amd import forcefile as ff
    vectors = []
    cases = list(CASES)
    for fn in sorted(os.listdir(HERE)):
        if fn.startswith("recorded_") and fn.endswith(".npy"):
            cases.append(dict(name=fn[:-4], recorded=fn))
    tmp = os.path.join(HERE, "_tmp_force.dat")
    for p in cases:
        if "recorded" in p:
            h = recorded(p["recorded"])
            with open(os.path.join(HERE, p["recorded"][:-4] + ".json")) as f:
                p.update(json.load(f))
        else:
            h = signal(p)
        ff.write_force_dat(tmp, h, 2)
        rows = af.read_force(__import__("pathlib").Path(tmp), "primary")
        sub = rows[int(len(rows) * 0.25):]
        times = [r[1] for r in sub]
        cls = [r[4] for r in sub]
        st_peak, nper, msg_peak = af.estimate_st_peak(times, cls, p["D"], p["U"], 0.4, 0.4)
        st_fft, msg_fft = af.estimate_st_fft(times, cls, p["D"], p["U"], 0.05, 0.5)
        summary, text = ({}, "") if len(rows) < 20 else run_reference_main(af, tmp, p["U"], p["D"])
        vectors.append({"params": p,
                        "rows": len(rows),
                        "st_peak": st_peak, "periods": nper, "peak_ok": msg_peak == "",
                        "st_fft": st_fft, "fft_ok": msg_fft == "",
                        "summary": summary})
        print(p["name"], st_peak, nper, st_fft, summary.get("St"), summary.get("match_St"), flush=True)
    os.remove(tmp)
    with open(os.path.join(HERE, "st_vectors.json"), "w") as f:
        json.dump({"source": "Examples/Cylinder_flow/analyze_force.py (estimate_st_peak, estimate_st_fft, "
                             "read_force, main) on force.dat files written by ilbmrex_amd.forcefile; lbm time "
                             "basis for the estimator calls (first 25 % dropped), the script's defaults for main",
                   "vectors": vectors}, f, indent=1)


if __name__ == "__main__":
    main()
