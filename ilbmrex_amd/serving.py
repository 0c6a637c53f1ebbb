"""Batched ensembles and a continuous-batching request loop over ilbm_batch.

Workloads (seeded, synthetic):

* ensemble_cases: independent D2Q9 flow-past-cylinder cases with grid
  log-uniform from 128x64 to 1024x512 (aspect 2), D ~ U[8, 48], Re ~ U[20, 200],
  cylinder centre offset ~ U(-10 %, +10 %) of mid-height, markers = round(pi D),
  U = 0.05, explicit direct forcing.  Two limits keep every case inside the
  stable range of BGK and the channel: D <= ny / 4 and nu = U D / Re >= 0.004
  (D >= 0.08 Re).
* serving_requests: Poisson arrivals; 70 % D2Q9 cylinders on n x n grids with n
  log-uniform in [64, 1024], 30 % D3Q19 spheres on n^3 grids with n
  log-uniform in [32, 192] (D = n / 5, markers ~ pi D^2 from the shell
  formula), step counts ~ U[2000, 20000].

serve() admits requests in arrival order while their device footprint fits
the budget, advances the batch in chunks, retires finished cases (their
force histories are the responses) and admits newly arrived requests between
chunks, so arrivals overlap with running work.
"""
from __future__ import annotations

import math
import time
from collections import deque
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from . import config as cf


def _clip_d(D: float, Re: float, limit: float, U: float = 0.05) -> float:
    lo = max(8.0, 0.004 * Re / U)
    return float(min(max(D, lo), max(lo, limit)))


def ensemble_cases(n: int = 256, seed: int = 20261202, storage: int = cf.STORE_F32) -> List[cf.Case]:
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n):
        ny = int(round(math.exp(rng.uniform(math.log(64), math.log(512)))))
        nx = 2 * ny
        D = rng.uniform(8.0, 48.0)
        Re = rng.uniform(20.0, 200.0)
        off = rng.uniform(-0.1, 0.1)
        D = _clip_d(D, Re, ny / 4.0)
        c = cf.cylinder_2d(nx=nx, ny=ny, D=D, xc=nx / 4.0, yc=ny / 2.0 * (1.0 + off), U=0.05, Re=Re,
                           storage=storage)
        out.append(c.replace(n_marker=int(round(math.pi * D)), name=f"ens{k}"))
    return out


@dataclass
class Request:
    rid: int
    arrival: float
    case: cf.Case
    nsteps: int
    admit: Optional[float] = None
    finish: Optional[float] = None
    history: Optional[np.ndarray] = field(default=None, repr=False)

    @property
    def latency(self) -> float:
        return self.finish - self.arrival


def serving_case(kind: str, n: int, Re: float, storage: int) -> cf.Case:
    if kind == "cyl":
        D = _clip_d(n / 8.0, Re, n / 4.0)
        c = cf.cylinder_2d(nx=n, ny=n, D=D, xc=n / 4.0, yc=n / 2.0, U=0.05, Re=Re, storage=storage)
        return c.replace(n_marker=int(round(math.pi * D)))
    D = max(6.0, n / 5.0)
    Re = min(Re, 0.04 * D / 0.004)
    return cf.sphere_3d(nx=n, ny=n, nz=n, D=D, center=(n / 3.0, n / 2.0, n / 2.0), U=0.04, Re=Re,
                        storage=storage)


def serving_requests(count: int, rate: float, seed: int = 20261202, storage: int = cf.STORE_F32,
                     steps=(2000, 20000), cyl_n=(64, 1024), sph_n=(32, 192)) -> List[Request]:
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for k in range(count):
        t += rng.exponential(1.0 / rate)
        if rng.uniform() < 0.7:
            n = int(round(math.exp(rng.uniform(math.log(cyl_n[0]), math.log(cyl_n[1])))))
            case = serving_case("cyl", n, rng.uniform(20.0, 200.0), storage)
        else:
            n = int(round(math.exp(rng.uniform(math.log(sph_n[0]), math.log(sph_n[1])))))
            case = serving_case("sph", n, rng.uniform(20.0, 200.0), storage)
        nsteps = int(rng.integers(steps[0], steps[1] + 1))
        out.append(Request(k, t, case.replace(name=f"req{k}"), nsteps))
    return out


def initial_state(case: cf.Case) -> np.ndarray:
    return cf.uniform_init(case, u=(case.bc_lo_val[0][1], case.bc_lo_val[0][2], case.bc_lo_val[0][3]))


def run_ensemble(cases: List[cf.Case], nsteps: int, budget: int, chunk: int = 1 << 30,
                 graph_steps: int = 8, keep: bool = False):
    """Admits all cases at once and advances them to nsteps; returns
    (wall seconds of the advance, lattice updates, histories or None)."""
    from .native import Batch
    with Batch(budget, max_cases=len(cases) + 1, graph_steps=graph_steps) as B:
        ids = [B.add(c, initial_state(c), nsteps) for c in cases]
        t0 = time.perf_counter()
        done = 0
        while done < nsteps:
            done += B.advance(min(chunk, nsteps - done))
        wall = time.perf_counter() - t0
        hist = [B.force_history(i) for i in ids] if keep else None
        info = B.info()
    return wall, sum(c.ncells for c in cases) * nsteps, hist, info


def serve(requests: List[Request], budget: int, chunk: int = 200, graph_steps: int = 8,
          max_cases: int = 4096, keep_history: bool = False, clock=time.perf_counter, min_chunk: int = 16):
    """Continuous batching over arrival times given in seconds from the start.

    Between advances the loop retires finished cases and admits every
    arrived request that fits the budget.  An advance runs at most `chunk`
    steps, and fewer (down to `min_chunk`) when the next request is due
    sooner: the steps until its arrival are estimated from the measured
    time per step of the resident set, so an arrival waits about one short
    chunk instead of a whole one."""
    from .native import Batch, BatchFull
    pending = deque(sorted(requests, key=lambda r: r.arrival))
    waiting: deque = deque()
    active = {}
    lups = 0.0
    step_s = None          # seconds per step of the current resident set
    with Batch(budget, max_cases=max_cases, graph_steps=graph_steps) as B:
        t0 = clock()
        while pending or waiting or active:
            now = clock() - t0
            while pending and pending[0].arrival <= now:
                waiting.append(pending.popleft())
            admitted = False
            while waiting:
                r = waiting[0]
                try:
                    cid = B.add(r.case, initial_state(r.case), r.nsteps)
                except BatchFull:
                    if not active:
                        raise
                    break
                waiting.popleft()
                r.admit = clock() - t0
                active[cid] = r
                admitted = True
            if admitted:
                step_s = None
            if not active:
                if pending:
                    time.sleep(max(0.0, min(pending[0].arrival - (clock() - t0), 0.01)))
                continue
            n = chunk
            if pending and step_s:
                due = pending[0].arrival - (clock() - t0)
                n = int(min(chunk, max(min_chunk, due / step_s)))
            ta = clock()
            done = B.advance(n)
            if done > 0:
                step_s = (clock() - ta) / done
            for cid in B.finished():
                r = active.pop(cid)
                r.finish = clock() - t0
                if keep_history:
                    r.history = B.force_history(cid)
                lups += r.case.ncells * r.nsteps
                B.remove(cid)
        makespan = clock() - t0
        info = B.info()
    lat = np.array([r.latency for r in requests])
    return {"p50_s": float(np.percentile(lat, 50)), "p99_s": float(np.percentile(lat, 99)),
            "mean_s": float(lat.mean()), "makespan_s": makespan, "mlups": lups / makespan / 1e6,
            "requests": len(requests), "graph_launches": info["graph_launches"],
            "captures": info["captures"]}


def serve_plain(requests: List[Request], storage: int = cf.STORE_F64, clock=time.perf_counter,
                keep_history: bool = False):
    """The plain per-case path as a server: requests run one at a time, in
    arrival order, each on its own Level with the given storage."""
    from .native import Level
    reqs = sorted(requests, key=lambda r: r.arrival)
    lups = 0.0
    t0 = clock()
    for r in reqs:
        wait = r.arrival - (clock() - t0)
        if wait > 0:
            time.sleep(wait)
        c = r.case.replace(storage=storage)
        r.admit = clock() - t0
        lev = Level(c, initial_state(c))
        lev.advance(r.nsteps)
        lev.synchronize()
        r.history = lev.force_history() if keep_history else None
        lev.close()
        r.finish = clock() - t0
        lups += c.ncells * r.nsteps
    makespan = clock() - t0
    lat = np.array([r.latency for r in reqs])
    return {"p50_s": float(np.percentile(lat, 50)), "p99_s": float(np.percentile(lat, 99)),
            "mean_s": float(lat.mean()), "makespan_s": makespan, "mlups": lups / makespan / 1e6,
            "requests": len(reqs)}


def plain_rate(cases: List[cf.Case], nsteps: int, storage: int = cf.STORE_F64) -> float:
    """MLUPS of the plain per-case path: each case alone on its own Level,
    one after another."""
    from .native import Level
    wall = 0.0
    lups = 0.0
    for c in cases:
        c2 = c.replace(storage=storage)
        lev = Level(c2, initial_state(c2))
        lev.advance(2)
        lev.synchronize()
        t0 = time.perf_counter()
        lev.advance(nsteps)
        lev.synchronize()
        wall += time.perf_counter() - t0
        lups += c2.ncells * nsteps
        lev.close()
    return lups / wall / 1e6
