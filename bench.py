#!/usr/bin/env python3
"""Benchmark of the per-timestep IB-LBM path on one MI355X.

Headline workload (``value``): D3Q19 BGK flow past a fixed sphere on a
256x128x128 lattice, D = 24 cells at (64, 64, 64), 1810 markers, Re = 100,
U = 0.04, explicit direct forcing with one iteration and the 4-point
kernel, FP32 population storage with FP32 collision and FP64 IB / force
reductions.  A step is one coarse step of the reference (IB forcing,
collide, stream, macros, Sum(Rho), ComputeIBForce).  ``value`` is MLUPS =
cells * steps / wall seconds of the timed region, state resident in HBM.

``roofline`` is for the dominant kernel (the fused stream-collide k_step):
algorithmic bytes per launch = cells * 2 * Q * sizeof(population), divided
by its average duration from HIP events that the kernel's dispatch records
(hipExtLaunchKernelGGL) on the level's stream.

Also reported:
- the same case with FP64 and FP16-shifted storage;
- ``accuracy``: 10000 steps of the sphere in each storage from the same state,
  with the FP32 and FP16-shifted velocity field (relative L2), mean Cd and Cl
  against the FP64 run;
- the periodic D3Q19 256^3 roofline case (F = 0 and F = 1e-6; FP32, FP64,
  FP16-shifted; 200 warm-up and 1000 timed steps);
- the 512x256 D2Q9 cylinder (FP64 explicit and IVC, FP32 explicit, and FP64
  as a one-case graph-replayed batch; ``--workload cylinder`` runs only the
  first three);
- the three-level pitching NACA 0012 hierarchy (FP32, Re = 100, regrid every
  10 coarse steps, 10,000 coarse steps) with force statistics;
- a plain copy kernel's bandwidth;
- the CPU oracle (FP64, OpenMP), timed on a bounded sample of the headline
  case (60 steps), with the host CPU's model string.

``batched``: the 256-case D2Q9 cylinder ensemble (FP32, 5000 steps, one
launch per kernel for all cases of a lane, two lanes on two streams) against
the plain per-case FP64 path, with the Cd error of FP32 against FP64 on a
sample; the continuous-batching server on Poisson arrivals of mixed
cylinder / sphere requests at 2, 8 and 32 requests/s (p50 / p99 latency,
aggregate MLUPS); and one stream of that mix (2,000-20,000 steps per
request) served by both the batched FP32 server and the plain per-case FP64
path, with the throughput ratio and each response's Cd error.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

from ilbmrex_amd import config as cf  # noqa: E402

HBM_PEAK_GBS = 8000.0
METRIC = "MLUPS & %HBM roofline, FP32 vs FP64; batched p50/p99 latency at Cd err<0.5%"


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def sphere_case(storage):
    return cf.sphere_3d(nx=256, ny=128, nz=128, D=24.0, center=(64.0, 64.0, 64.0), U=0.04,
                        Re=100.0, coupling=cf.COUPLING_EXPLICIT, loop_ns=1, delta_type=4,
                        storage=storage)


def periodic_case(storage, force):
    return cf.periodic_box(cf.D3Q19, (256, 256, 256), nu=(0.6 - 0.5) / 3.0, force=force,
                           storage=storage, name="periodic256")


OVERLAP = True


def run_level(case, init, warmup, steps, want_forces=False):
    """Wall-clock MLUPS of `steps` steps (an IB level overlaps its far cells
    with the IB forcing), then the stream-collide kernel's average duration
    from HIP events on a further min(steps, 200) steps run without the
    overlap, so that the kernel covers the whole level."""
    from ilbmrex_amd.native import Level
    lev = Level(case, init)
    lev.set_overlap(OVERLAP)
    lev.advance(warmup)
    lev.synchronize()
    t0 = time.perf_counter()
    lev.advance(steps)
    lev.synchronize()
    wall = time.perf_counter() - t0
    fh = lev.force_history() if want_forces else None
    lev.set_timing(True)
    lev.advance(min(steps, 200))
    lev.synchronize()
    kms, nl = lev.kernel_time()
    info = lev.info()
    lev.close()
    cells = case.ncells
    q = case.lattice.q
    bpl = 2 * q * info["pop_bytes"]
    mlups = cells * steps / wall / 1e6
    kgbs = cells * bpl / (kms * 1e-3) / 1e9 if kms > 0 else 0.0
    return {"mlups": mlups, "ms_per_step": 1e3 * wall / steps, "kernel_ms": kms,
            "kernel_launches": nl, "bytes_per_lup": bpl, "kernel_gbs": kgbs,
            "kernel_frac": kgbs / HBM_PEAK_GBS, "bytes_per_launch": cells * bpl}, fh


CYLINDER_CASES = (("cylinder512_f64_explicit", cf.COUPLING_EXPLICIT, cf.STORE_F64),
                  ("cylinder512_f64_ivc", cf.COUPLING_IVC, cf.STORE_F64),
                  ("cylinder512_f32_explicit", cf.COUPLING_EXPLICIT, cf.STORE_F32))


def cylinder_variants(steps, only=None):
    """The reference's CPU-runnable case, D2Q9 cylinder 512x256, D = 32 at
    (128, 128), 101 markers, Re = 100: explicit and IVC forcing in FP64 and
    explicit forcing in FP32, each on one stream (below 2^20 cells)."""
    out = {}
    for name, coup, st in CYLINDER_CASES:
        if only and name not in only:
            continue
        case = cf.cylinder_2d(nx=512, ny=256, D=32.0, xc=128.0, yc=128.0, U=0.05, Re=100.0,
                              coupling=coup, storage=st)
        log(f"[bench] {name}: {steps} steps")
        r, fh = run_level(case, cf.uniform_init(case, u=(0.05, 0.0, 0.0)), 100, steps, want_forces=True)
        r["us_per_step"] = round(1e3 * r["ms_per_step"], 2)
        r["cd_last"] = float(fh[-1, 4])
        out[name] = r
    return out


def accuracy(nsteps):
    """The sphere in FP64, FP32 and FP16-shifted storage for nsteps steps
    from the same state: relative L2 error of the velocity field at the end,
    relative error of mean Cd and mean |Cl - Cl_64| over the last 1000
    steps, against the FP64 run."""
    from ilbmrex_amd.native import Level
    out = {"steps": nsteps}
    res = {}
    for key, st in (("f64", cf.STORE_F64), ("f32", cf.STORE_F32), ("f16s", cf.STORE_F16S)):
        case = sphere_case(st)
        log(f"[bench] accuracy: sphere {key}, {nsteps} steps")
        lev = Level(case, cf.uniform_init(case, u=(0.04, 0.0, 0.0)))
        lev.set_overlap(OVERLAP)
        lev.advance(nsteps, want_state=True)
        res[key] = (lev.macro()[1:4].copy(), lev.force_history())
        lev.close()
    u64, fh64 = res["f64"]
    tail = slice(max(0, nsteps - 1000), nsteps)
    cd64 = fh64[tail, 4].mean()
    out["cd_f64_mean"] = float(cd64)
    out["cl_f64_rms"] = float(np.sqrt(np.mean(fh64[tail, 5] ** 2)))
    for key in ("f32", "f16s"):
        u, fh = res[key]
        out[f"vel_rel_l2_{key}_vs_f64"] = float(np.linalg.norm(u - u64) / np.linalg.norm(u64))
        out[f"cd_rel_err_{key}_vs_f64"] = float(abs(fh[tail, 4].mean() - cd64) / abs(cd64))
        out[f"cl_mean_abs_diff_{key}_vs_f64"] = float(np.mean(np.abs(fh[tail, 5] - fh64[tail, 5])))
        out[f"finite_{key}"] = bool(np.all(np.isfinite(u)) and np.all(np.isfinite(fh)))
    return out


def cpu_baseline(steps):
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import refcpu
    refcpu.build()
    case = sphere_case(cf.STORE_F64)
    init = cf.uniform_init(case, u=(0.04, 0.0, 0.0))
    orc = refcpu.Oracle(case, init)
    orc.advance(1)
    t0 = time.perf_counter()
    orc.advance(steps)
    wall = time.perf_counter() - t0
    orc.close()
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    model = "unknown CPU"
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"value": case.ncells * steps / wall / 1e6, "unit": "MLUPS", "cores": threads,
            "kind": "port", "cpu": model, "wall_s": round(wall, 2),
            "sample": f"{steps} steps (after 1 untimed) of the same 256x128x128 D3Q19 sphere case, "
                      f"FP64, oracle/refcpu.c with OpenMP ({threads} threads; collision, streaming "
                      f"and ghost fills parallel, the IB stages serial)"}


def amr_variant(steps):
    from ilbmrex_amd.native import AmrHierarchy
    # Re = 1000 at this resolution diverges under BGK (the FP64 oracle of
    # the same hierarchy blows up within 12 coarse steps, explicit or IVC
    # forcing).  Of Re = 100, 110, 125, 150, ..., 1000 over 10,000 coarse
    # steps, 100 is the largest whose forces stay physical (110 and 125
    # stay finite with |Cl| >> 1, 150 and above go non-finite).
    case, boxes = cf.pitching_naca_amr(storage=cf.STORE_F32, Re=100.0)
    log(f"[bench] amr3_pitching_naca_f32: {steps} coarse steps ({4 * steps} finest)")
    h = AmrHierarchy(case, boxes, 10, (1.0, case.U0, 0.0, 0.0))
    # regrid every 10 coarse steps, vorticity tagging with the thresholds of
    # the reference's pitching-airfoil inputs (amrvorth.thresholdRatio)
    h.set_regrid(threshold_ratio=(0.85, 0.90, 0.90), blocking_factor=8)
    h.advance(20)
    info = h.info()
    t0 = time.perf_counter()
    h.advance(steps)
    wall = time.perf_counter() - t0
    fh = h.force_history()
    boxes_end, nregrid = h.boxes()
    info_end = h.info()
    h.close()
    cells = info_end["cells"]
    updates = cells[0] + 2 * cells[1] + 4 * cells[2]
    tail = fh[-min(2000, len(fh)):]
    forces = {"window_coarse_steps": len(tail), "cd_mean": float(np.mean(tail[:, 4])),
              "cd_std": float(np.std(tail[:, 4])), "cl_mean": float(np.mean(tail[:, 5])),
              "cl_rms": float(np.sqrt(np.mean((tail[:, 5] - np.mean(tail[:, 5])) ** 2)))}
    return {"mlups_all_levels": updates * steps / wall / 1e6, "ms_per_coarse_step": 1e3 * wall / steps,
            "coarse_steps": steps, "finest_steps": 4 * steps, "forces_last_window": forces,
            "finest_steps_per_s": 4 * steps / wall, "cells_per_level": cells,
            "cells_per_level_start": info["cells"], "boxes_end": boxes_end, "regrids_changing_boxes": nregrid,
            "markers": 256, "cd_mean_last_100": float(np.mean(fh[-100:, 4])),
            "finite": bool(np.isfinite(fh).all()),
            "workload": "coarse 512x256 D2Q9, levels at ratio 2 (128x96, 160x128 boxes), chord 64 finest "
                        "cells, 256 markers, +-10 deg about the quarter chord, k=1.0, Re=100 (Re=1000 diverges under "
                        "BGK at this resolution, FP64 oracle included), U=0.04, explicit direct forcing, FP32, "
                        "regrid every 10 coarse steps (vorticity tagging, thresholdRatio 0.85/0.90, blocking "
                        "factor 8, one box per level)"}


def batched_section(args):
    from ilbmrex_amd import serving
    from ilbmrex_amd.native import device_memory
    free, total = device_memory()
    budget = int(min(free - (16 << 30), 288e9))
    out = {"hbm_budget_bytes": budget}
    cases = serving.ensemble_cases(256, seed=20261202)
    ns = args.ensemble_steps
    log(f"[bench] ensemble: 256 cylinder cases FP32, {ns} steps")
    wall, lups, _, info = serving.run_ensemble(cases, ns, budget)
    ens = {"cases": 256, "steps": ns, "storage": "f32", "cells": sum(c.ncells for c in cases),
           "wall_s": round(wall, 4), "mlups": round(lups / wall / 1e6, 1),
           "graph_launches": info["graph_launches"], "eager_steps": info["eager_steps"]}
    sample = cases[::32]
    log("[bench] ensemble: plain per-case FP64 path on 8 of the cases")
    ens["plain_fp64_mlups"] = round(serving.plain_rate(sample, 500, storage=cf.STORE_F64), 1)
    ens["plain_fp64_sample"] = "8 of the 256 cases (every 32nd), 500 steps each, one Level at a time"
    ens["speedup_vs_plain_fp64"] = round(ens["mlups"] / ens["plain_fp64_mlups"], 2)
    log("[bench] ensemble: Cd of FP32 against FP64 on 8 cases")
    _, _, h32, _ = serving.run_ensemble(sample, ns, budget, keep=True)
    _, _, h64, _ = serving.run_ensemble([c.replace(storage=cf.STORE_F64) for c in sample], ns, budget, keep=True)
    tail = max(1, ns // 5)
    errs = [abs(a[-tail:, 4].mean() - b[-tail:, 4].mean()) / abs(b[-tail:, 4].mean()) for a, b in zip(h32, h64)]
    ens["cd_err_f32_vs_f64_max"] = float(max(errs))
    ens["cd_err_note"] = f"relative error of mean Cd over the last {tail} steps, 8 cases"
    out["ensemble"] = ens
    sweep = []
    for rate in args.rates:
        reqs = serving.serving_requests(args.requests, rate=rate, seed=20261202)
        log(f"[bench] serving: {args.requests} requests at {rate}/s (FP32, continuous batching)")
        st = serving.serve(reqs, budget, chunk=args.chunk)
        st["rate_per_s"] = rate
        sweep.append({k: (round(v, 4) if isinstance(v, float) else v) for k, v in st.items()})
    out["serving_f32"] = sweep
    # config 5's own stream (2k-20k steps, 70/30 cylinder / sphere) through
    # both paths: the batched FP32 server and the plain per-case FP64 path
    rate = args.plain_rate
    log(f"[bench] serving: {args.requests} requests at {rate}/s, batched FP32 vs plain FP64 on the same stream")
    reqs_b = serving.serving_requests(args.requests, rate=rate, seed=20261202)
    b = serving.serve(reqs_b, budget, chunk=args.chunk, keep_history=True)
    reqs_p = serving.serving_requests(args.requests, rate=rate, seed=20261202, storage=cf.STORE_F64)
    p = serving.serve_plain(reqs_p, storage=cf.STORE_F64, keep_history=True)
    cd_err, shedding = [], []
    for rb, rp in zip(reqs_b, reqs_p):
        tail = max(1, rb.nsteps // 5)
        c32, c64 = rb.history[-tail:, 4].mean(), rp.history[-tail:, 4].mean()
        cd_err.append(abs(c32 - c64) / abs(c64))
        # a request whose FP64 lift oscillates sheds vortices; its shedding
        # phase (seeded by rounding on a symmetric start) differs between the
        # storages, so a window mean of Cd differs by more than precision
        shedding.append(bool(rp.history[-tail:, 5].std() > 1e-3))
    steady = [e for e, sh in zip(cd_err, shedding) if not sh]
    out["serving_config5_stream"] = {
        "requests": args.requests, "rate_per_s": rate, "steps": [2000, 20000], "mix": "70% D2Q9 cylinders, 30% D3Q19 spheres",
        "batched_f32": {k: (round(v, 4) if isinstance(v, float) else v) for k, v in b.items()},
        "plain_f64": {k: (round(v, 4) if isinstance(v, float) else v) for k, v in p.items()},
        "throughput_ratio": round(b["mlups"] / p["mlups"], 3),
        "cd_err_f32_vs_f64_max": float(max(cd_err)),
        "cd_err_f32_vs_f64_max_steady": float(max(steady)) if steady else None,
        "requests_within_0p5pct": int(sum(e < 5e-3 for e in cd_err)),
        "shedding_requests": [{"rid": r.rid, "cd_err": round(e, 5)} for r, e, sh in zip(reqs_b, cd_err, shedding) if sh],
        "cd_err_note": "relative error of each response's mean Cd over its last 20 % of steps, batched FP32 "
                       "against the same request on the plain FP64 path; requests whose FP64 lift oscillates "
                       "(std > 1e-3) shed vortices with a storage-dependent phase and are listed apart"}
    return out


def read_traffic():
    p = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(p):
        return None
    try:
        with open(p) as f:
            return json.load(f).get("k_step_sphere_fp32_bytes_per_launch")
    except Exception:
        return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--variant-steps", type=int, default=1000, help="timed steps of the periodic 256^3 cases")
    ap.add_argument("--roofline-warmup", type=int, default=200)
    ap.add_argument("--accuracy-steps", type=int, default=10000,
                    help="steps of the sphere in FP64 / FP32 / FP16-shifted storage for the error report")
    ap.add_argument("--cpu-steps", type=int, default=60)
    ap.add_argument("--no-variants", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--amr-steps", type=int, default=10000,
                    help="timed coarse steps of the pitching hierarchy (4 finest steps each)")
    ap.add_argument("--ensemble-steps", type=int, default=5000)
    ap.add_argument("--requests", type=int, default=32)
    ap.add_argument("--rates", type=float, nargs="+", default=[2.0, 8.0, 32.0])
    ap.add_argument("--chunk", type=int, default=200)
    ap.add_argument("--plain-rate", type=float, default=8.0,
                    help="arrival rate of the config-5 stream served by both the batched and the plain FP64 path")
    ap.add_argument("--no-overlap", action="store_true",
                    help="IB levels step on one stream (profiler runs of the whole-level kernel)")
    ap.add_argument("--cylinder-steps", type=int, default=2000,
                    help="timed steps of each 512x256 cylinder case")
    ap.add_argument("--cylinder-cases", nargs="+", choices=[n for n, _, _ in CYLINDER_CASES], default=None,
                    help="--workload cylinder: run only these cases")
    ap.add_argument("--workload", choices=["sphere", "ensemble", "amr", "cylinder"], default="sphere",
                    help="ensemble / amr / cylinder: time only that workload (profiler runs) and print its summary")
    ap.add_argument("--probe", action="store_true",
                    help="also run the 1 GiB copy kernel (profiler counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only; --gpus {args.gpus} is not supported "
              f"(the multi-GPU sharding is a design, not a build).", file=sys.stderr)
        sys.exit(2)
    from ilbmrex_amd import native
    native.load()
    global OVERLAP
    OVERLAP = not args.no_overlap
    if args.workload == "ensemble":
        from ilbmrex_amd import serving
        if args.probe:
            ms = native.load_probe(1 << 30, 3)
            log(f"[bench] copy probe 1 GiB: {ms:.3f} ms ({2 * (1 << 30) / (ms * 1e-3) / 1e9:.0f} GB/s)")
        cases = serving.ensemble_cases(256, seed=20261202)
        wall, lups, _, info = serving.run_ensemble(cases, args.ensemble_steps, 100 << 30)
        print(json.dumps({"workload": "ensemble of 256 D2Q9 cylinder cases, FP32", "steps": args.ensemble_steps,
                          "mlups": round(lups / wall / 1e6, 1), "wall_s": round(wall, 4)}), flush=True)
        return
    if args.workload == "amr":
        print(json.dumps(amr_variant(args.amr_steps)), flush=True)
        return
    if args.workload == "cylinder":
        print(json.dumps(cylinder_variants(args.cylinder_steps, args.cylinder_cases)), flush=True)
        return

    # headline: sphere, FP32 storage
    c32 = sphere_case(cf.STORE_F32)
    init = cf.uniform_init(c32, u=(0.04, 0.0, 0.0))
    log(f"[bench] sphere 256x128x128 FP32: warmup {args.warmup}, timed {args.steps}")
    head, _ = run_level(c32, init, args.warmup, args.steps)
    log(f"[bench] sphere FP32: {head['mlups']:.1f} MLUPS, k_step {head['kernel_ms']*1e3:.1f} us "
        f"({head['kernel_gbs']:.0f} GB/s)")

    variants = {}
    if not args.no_variants:
        # the roofline cases first, on a GPU not yet warmed by the long runs below
        vw, vs = args.roofline_warmup, args.variant_steps
        for name, st, force in (("periodic256_f32_F0", cf.STORE_F32, None),
                                ("periodic256_f32_F1e-6", cf.STORE_F32, (1e-6, 0.0, 0.0)),
                                ("periodic256_f64_F0", cf.STORE_F64, None),
                                ("periodic256_f16s_F0", cf.STORE_F16S, None)):
            case = periodic_case(st, force)
            pinit = cf.random_init(case, seed=20261202)
            log(f"[bench] {name}: warmup {vw}, timed {vs}")
            r, _ = run_level(case, pinit, vw, vs)
            variants[name] = r
        for name, case in (("sphere_f64", sphere_case(cf.STORE_F64)),
                           ("sphere_f16s", sphere_case(cf.STORE_F16S))):
            log(f"[bench] {name}: warmup {args.warmup}, timed {args.steps}")
            r, _ = run_level(case, init, args.warmup, args.steps)
            variants[name] = r
        variants["accuracy"] = accuracy(args.accuracy_steps)
        # the reference's CPU-runnable case: D2Q9 cylinder 512x256
        variants.update(cylinder_variants(args.cylinder_steps))
        # the same case as a one-case batch: hipGraph replay of the step sequence
        case = cf.cylinder_2d(nx=512, ny=256, D=32.0, xc=128.0, yc=128.0, U=0.05, Re=100.0,
                              storage=cf.STORE_F64)
        from ilbmrex_amd import serving
        log("[bench] cylinder512_f64_explicit_graph: 2000 steps")
        wall, lups, _, _ = serving.run_ensemble([case], 2000, 8 << 30)
        variants["cylinder512_f64_explicit_graph"] = {"mlups": lups / wall / 1e6,
                                                      "ms_per_step": 1e3 * wall / 2000}
        # pitching NACA 0012 on three levels, FP32
        variants["amr3_pitching_naca_f32"] = amr_variant(args.amr_steps)
        copy_ms = native.load_probe(2 << 30, 20)
        variants["copy_probe_gbs"] = 2 * (2 << 30) / (copy_ms * 1e-3) / 1e9

    if args.probe:
        ms = native.load_probe(1 << 30, 3)
        log(f"[bench] copy probe 1 GiB: {ms:.3f} ms ({2 * (1 << 30) / (ms * 1e-3) / 1e9:.0f} GB/s)")

    batched = None
    if not args.no_batch:
        batched = batched_section(args)

    cpu = None
    if not args.no_cpu:
        log(f"[bench] CPU oracle: {args.cpu_steps} steps")
        cpu = cpu_baseline(args.cpu_steps)

    out = {
        "metric": METRIC,
        "value": round(head["mlups"], 2),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(head["ms_per_step"], 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (uniform inflow initial state, no dataset)",
        "config": {
            "workload": "D3Q19 BGK flow past a fixed sphere, 256x128x128 uniform lattice, D=24 at "
                        "(64,64,64), 1810 markers, Re=100, U=0.04, explicit direct forcing (loop_ns=1, "
                        "4-point delta), Guo forcing, ext_dir inflow / foextrap outflow, periodic y,z; "
                        "FP32 population storage, FP32 collision, FP64 IB and force sums",
            "cells": c32.ncells,
            "markers": 1810,
            "parallelism": "single GPU",
        },
        "roofline": {
            "bound": "hbm",
            "kernel": "k_step<D3Q19,StF32,slabs> (fused pull-stream + macros + BGK)",
            "achieved": round(head["kernel_gbs"], 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(head["kernel_frac"], 4),
            "traffic": read_traffic(),
            "algorithmic_bytes_per_launch": head["bytes_per_launch"],
            "avg_launch_ms": round(head["kernel_ms"], 5),
        },
        "cpu_baseline": cpu,
        "variants": variants,
        "batched": batched,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
