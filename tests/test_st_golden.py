"""The Strouhal estimators and the force.dat reader of ilbmrex_amd.forcefile
against vectors of the reference's Examples/Cylinder_flow/analyze_force.py
(written by tests/golden/make_st_vectors.py, which imports that script):
the peak and spectral estimates, the periods used, and the summary its
main() prints (Cd mean, Cl mean / RMS, St, the acceptance matches), on
synthetic lift signals and on a recorded GPU force history of the 512x256
cylinder."""
import importlib.util
import json
import os

import pytest

from ilbmrex_amd import forcefile as ff

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _golden():
    spec = importlib.util.spec_from_file_location("make_st_vectors", os.path.join(HERE, "make_st_vectors.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    with open(os.path.join(HERE, "st_vectors.json")) as f:
        return mod, json.load(f)["vectors"]


MOD, VECTORS = _golden()


def _history(p):
    if p.get("recorded"):
        return MOD.recorded(p["recorded"])
    return MOD.signal(p)


def _close(a, b, rel=1e-12):
    if a is None or b is None:
        return a is None and b is None
    return abs(a - b) <= rel * max(abs(b), 1e-300)


@pytest.mark.parametrize("vec", VECTORS, ids=[v["params"]["name"] for v in VECTORS])
def test_estimators_match_reference_vectors(vec, tmp_path):
    p = vec["params"]
    path = tmp_path / "force.dat"
    ff.write_force_dat(path, _history(p), 2)
    rows = ff.read_force(path)
    assert len(rows) == vec["rows"]
    sub = rows[int(len(rows) * 0.25):]
    times = [r[1] for r in sub]
    cls = [r[4] for r in sub]
    st, periods, why = ff.estimate_st_peak(times, cls, p["D"], p["U"], 0.4, 0.4)
    assert _close(st, vec["st_peak"]) and periods == vec["periods"] and (why == "") == vec["peak_ok"]
    st_f, why_f = ff.estimate_st_fft(times, cls, p["D"], p["U"], 0.05, 0.5)
    assert _close(st_f, vec["st_fft"]) and (why_f == "") == vec["fft_ok"]
    s = vec["summary"]
    if not s:
        with pytest.raises(ValueError):
            ff.analyze(rows, p["U"], p["D"])
        return
    a = ff.analyze(rows, p["U"], p["D"])
    # main() prints six decimals
    for key, ref in (("cd_mean", "Cd_mean"), ("cl_mean", "Cl_mean"), ("cl_rms", "Cl_rms")):
        assert abs(a[key] - float(s[ref])) <= 5.1e-7, (key, a[key], s[ref])
    if s["St"] == "NaN":
        assert a["st"] is None
    else:
        assert abs(a["st"] - float(s["St"])) <= 5.1e-7
    assert a["match_cd"] == (s["match_Cd"] == "True")
    assert a["match_st"] == (s["match_St"] == "True")
    assert a["periodic_lift"] == (s["periodic_lift"] == "True")
