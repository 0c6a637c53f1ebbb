"""Batched cases (ilbm_batch_*) on the GPU.

* FP64 batch of mixed cases (cylinders of different sizes, IVC, 2-iteration
  explicit-diag, a sphere, a periodic box) against the oracle and against
  the single-level path: force records bit-identical to the single level
  and within 1e-12 of the oracle (summation order), mass within 1e-12.
* The same case alone and admitted mid-run into a mix (other parity, other
  classes sharing the launches) gives bit-identical histories.
* hipGraph replays (with eager steps and synchronisations between them)
  against an all-eager run: bit-identical.
* Admission: footprint budget and slot limit refuse with BatchFull; removal
  frees the budget.
"""
import numpy as np
import pytest

from ilbmrex_amd import config as cf
import refcpu

pytestmark = pytest.mark.gpu

BUDGET = 8 << 30


def _init(c):
    u = c.bc_lo_val[0][1:4] if not c.periodic[0] else (0.0, 0.0, 0.0)
    return cf.uniform_init(c, u=tuple(u))


def mixed_cases(storage):
    return [
        cf.cylinder_2d(nx=96, ny=64, D=16.0, xc=30.0, yc=33.0, U=0.05, Re=100.0, storage=storage),
        cf.cylinder_2d(nx=150, ny=70, D=10.0, xc=40.0, yc=34.0, U=0.05, Re=60.0, storage=storage),
        cf.cylinder_2d(nx=80, ny=48, D=12.0, xc=24.0, yc=24.0, U=0.05, Re=50.0,
                       coupling=cf.COUPLING_IVC, storage=storage),
        cf.cylinder_2d(nx=72, ny=40, D=8.0, xc=20.0, yc=21.0, U=0.05, Re=40.0,
                       coupling=cf.COUPLING_EXPLICIT, loop_ns=2, delta_type=3, storage=storage),
        cf.sphere_3d(nx=40, ny=24, nz=24, D=8.0, center=(12.0, 12.0, 12.5), U=0.04, Re=50.0,
                     storage=storage),
        cf.periodic_box(cf.D2Q9, (64, 32, 1), nu=0.02, force=(1e-5, 0.0, 0.0), storage=storage),
    ]


def test_batch_fp64_matches_oracle_and_level():
    from ilbmrex_amd.native import Batch, Level
    # explicit-diag forcing of this body goes unstable after ~40 steps in the
    # oracle as well; 30 steps stay in the stable range
    diag = cf.cylinder_2d(nx=72, ny=40, D=8.0, xc=20.0, yc=21.0, U=0.05, Re=40.0,
                          coupling=cf.COUPLING_EXPLICIT_DIAG, loop_ns=2, delta_type=3)
    cases = mixed_cases(cf.STORE_F64) + [diag]
    nsteps = 30
    print(f"[batch] FP64 mix of {len(cases)} cases, {nsteps} steps", flush=True)
    with Batch(BUDGET, graph_steps=8) as B:
        ids = [B.add(c, _init(c), nsteps) for c in cases]
        done = 0
        while done < nsteps:
            done += B.advance(nsteps - done)
        for c, cid in zip(cases, ids):
            fh_b = B.force_history(cid)
            mh_b = B.mass_history(cid)
            assert len(mh_b) == nsteps
            lev = Level(c, _init(c))
            lev.advance(nsteps)
            fh_l, mh_l = lev.force_history(), lev.mass_history()
            lev.close()
            orc = refcpu.Oracle(c, _init(c))
            orc.advance(nsteps)
            assert abs(mh_b[-1] - orc.mass_sum()) <= 1e-12 * abs(orc.mass_sum()), c.name
            assert np.allclose(mh_b, mh_l, rtol=1e-13, atol=0.0), c.name
            if c.use_ib:
                assert fh_b.shape == fh_l.shape == (nsteps, 23 if c.lattice.dim == 2 else 33)
                assert np.array_equal(fh_b, fh_l), f"{c.name}: batch and single level differ"
                rec = orc.force_record()
                tol = 1e-10 if c.coupling_method == cf.COUPLING_IVC else 1e-12
                floor = 1e-10 * np.abs(rec).max()
                assert np.all(np.abs(fh_b[-1] - rec) <= tol * np.abs(rec) + floor), c.name


def _run(cases_steps, order, graph=True, extra_first=0):
    """Admits cases at the given advance-call indices; returns histories."""
    from ilbmrex_amd.native import Batch
    hist = {}
    with Batch(BUDGET, graph_steps=8) as B:
        ids = {}
        call = 0
        while True:
            for k, (c, n, at) in enumerate(cases_steps):
                if at == call and k not in ids:
                    ids[k] = B.add(c, _init(c), n)
            if call == 0 and extra_first:
                B.advance(extra_first, use_graph=graph)
            else:
                B.advance(order[call % len(order)], use_graph=graph)
            call += 1
            if len(ids) == len(cases_steps) and all(B.progress(i)[0] >= B.progress(i)[1] for i in ids.values()):
                break
            assert call < 500
        for k, cid in ids.items():
            hist[k] = (B.force_history(cid), B.mass_history(cid))
    return hist


def test_alone_and_in_mix_identical_fp32():
    cases = mixed_cases(cf.STORE_F32)
    target = cases[0]
    n = 40
    print("[batch] alone vs mix, FP32", flush=True)
    alone = _run([(target, n, 0)], order=[7, 8, 5])
    # admitted at the second call, after the others already took 7 steps
    # (other ping-pong parity), with other lattices / classes in the launches
    mix = [(cases[1], 200, 0), (cases[4], 120, 0), (cases[5], 90, 0), (target, n, 1), (cases[3], 60, 1)]
    mixed = _run(mix, order=[7, 8, 5, 16, 3])
    fa, ma = alone[0]
    fm, mm = mixed[3]
    assert fa.shape == fm.shape == (n, 23)
    assert np.array_equal(fa, fm), "force history changes with the mix"
    assert np.array_equal(ma, mm), "mass history changes with the mix"
    assert np.all(np.isfinite(fa))


def test_graph_replays_match_eager():
    cases = mixed_cases(cf.STORE_F32)[:4] + [mixed_cases(cf.STORE_F16S)[4]]
    n = 64
    print("[batch] graph replays vs eager", flush=True)
    spec = [(c, n, 0) for c in cases]
    eager = _run(spec, order=[5], graph=False)
    graph = _run(spec, order=[16, 3, 8, 9, 24], graph=True)
    for k in eager:
        assert np.all(np.isfinite(eager[k][1])), f"case {k}: unstable"
        assert np.array_equal(eager[k][0], graph[k][0]), f"case {k}: force history"
        assert np.array_equal(eager[k][1], graph[k][1]), f"case {k}: mass history"


def test_admission_budget_and_slots():
    from ilbmrex_amd.native import Batch, BatchFull
    c = mixed_cases(cf.STORE_F32)[0]
    need = Batch.footprint(c, 100)
    assert need > c.ncells * 9 * 4 * 2
    with Batch(int(need * 2.5), max_cases=8, graph_steps=8) as B:
        a = B.add(c, _init(c), 100)
        b = B.add(c, _init(c), 50)
        with pytest.raises(BatchFull):
            B.add(c, _init(c), 100)
        assert B.info()["bytes"] <= B.info()["budget"]
        assert B.advance(1000) == 50
        assert B.finished() == [b]
        B.remove(b)
        d = B.add(c, _init(c), 10)
        assert B.advance(1000) == 10
        assert sorted(B.finished()) == [d]
        assert B.progress(a) == (60, 100)
    with Batch(1 << 40, max_cases=1, graph_steps=8) as B:
        B.add(c, _init(c), 5)
        with pytest.raises(BatchFull):
            B.add(c, _init(c), 5)


def test_ensemble_fp32_cd_within_half_percent_of_fp64():
    """A few ensemble draws: FP32 batched Cd against the FP64 path (which is
    the oracle's arithmetic, see the parity tests) over 2000 steps."""
    from ilbmrex_amd import serving
    cases = serving.ensemble_cases(6, seed=11)
    cases = [c for c in cases if c.ncells <= 200_000][:4]
    n = 2000
    print(f"[batch] ensemble Cd FP32 vs FP64, {len(cases)} cases, {n} steps", flush=True)
    _, _, h32, _ = serving.run_ensemble(cases, n, BUDGET, keep=True)
    _, _, h64, _ = serving.run_ensemble([c.replace(storage=cf.STORE_F64) for c in cases], n, BUDGET, keep=True)
    for a, b in zip(h32, h64):
        cd32, cd64 = a[-500:, 4].mean(), b[-500:, 4].mean()
        assert np.isfinite(cd32) and abs(cd32 - cd64) <= 5e-3 * abs(cd64), (cd32, cd64)


def test_batch_large_3d_slice_layout_matches_level():
    """A 3D case of 2^20 cells or more (z-slice population layout) next to a
    2D case in one batch: histories identical to the single level's, the
    record within 1e-12 of the oracle."""
    from ilbmrex_amd.native import Batch, Level
    big = cf.sphere_3d(nx=128, ny=96, nz=96, D=12.0, center=(40.0, 48.3, 47.8), U=0.04, Re=50.0)
    small = mixed_cases(cf.STORE_F64)[0]
    nsteps = 4
    print("[batch] 128x96x96 sphere + 96x64 cylinder, FP64", flush=True)
    with Batch(BUDGET, graph_steps=2) as B:
        ids = [B.add(c, _init(c), nsteps) for c in (big, small)]
        done = 0
        while done < nsteps:
            done += B.advance(nsteps - done)
        for c, cid in zip((big, small), ids):
            fh_b = B.force_history(cid)
            lev = Level(c, _init(c))
            lev.advance(nsteps)
            assert np.array_equal(fh_b, lev.force_history()), c.name
            lev.close()
            if c is big:
                fh_big = fh_b
    orc = refcpu.Oracle(big, _init(big))
    orc.advance(nsteps)
    rec = orc.force_record()
    assert np.all(np.abs(fh_big[-1] - rec) <= 1e-12 * np.abs(rec) + 1e-10 * np.abs(rec).max())


@pytest.mark.parametrize("tile,nc", [(0, 1), (64, 1), (128, 1), (0, 2), (64, 2), (128, 2)])
def test_batch_tilings_match_level(tile, nc, monkeypatch):
    """Each stream-collide tiling of the batched launch (0: linear tiles of
    256 cells with no row padding; 64 / 128: 2D tiles), with one or two
    cells per thread, gives the single level's state: checkpoint bytes and
    force histories identical, mass within 1e-13 (block partials differ).
    Rows of 150, 80 and 72 cells are not multiples of the tile widths; the
    3D case spans slices in one tile."""
    from ilbmrex_amd.native import Batch, Level
    monkeypatch.setenv("ILBM_BATCH_TILE", str(tile))
    monkeypatch.setenv("ILBM_BATCH_NC", str(nc))
    cases = mixed_cases(cf.STORE_F32)
    nsteps = 12
    print(f"[batch] tiling {tile}, {nc} cells per thread: {len(cases)} cases, {nsteps} steps", flush=True)
    with Batch(BUDGET, graph_steps=4) as B:
        ids = [B.add(c, _init(c), nsteps) for c in cases]
        assert all(B.tile(cid) == tile for cid in ids)
        done = 0
        while done < nsteps:
            done += B.advance(nsteps - done)
        for c, cid in zip(cases, ids):
            lev = Level(c, _init(c))
            lev.advance(nsteps)
            assert B.checkpoint(cid) == lev.checkpoint(), f"{c.name}: state differs"
            if c.use_ib:
                assert np.array_equal(B.force_history(cid), lev.force_history()), c.name
            np.testing.assert_allclose(B.mass_history(cid), lev.mass_history(), rtol=1e-13, atol=0.0)
            lev.close()


@pytest.mark.parametrize("phase", ["staggered", "together"])
def test_two_lane_batch_matches_levels(phase, monkeypatch):
    """A batch of at least 2^22 active cells splits each class into two lanes
    on two streams (advanced eagerly), half a step apart or in phase
    (ILBM_BATCH_PHASE); every case's state and force history equal the
    single level's, mass within 1e-13."""
    from ilbmrex_amd.native import Batch, Level
    monkeypatch.setenv("ILBM_BATCH_PHASE", phase)
    st = cf.STORE_F32
    cases = [
        cf.cylinder_2d(nx=1024, ny=1024, D=24.0, xc=200.0, yc=512.0, U=0.05, Re=100.0, storage=st),
        cf.cylinder_2d(nx=1024, ny=768, D=20.0, xc=180.0, yc=380.0, U=0.05, Re=80.0, storage=st),
        cf.cylinder_2d(nx=1024, ny=1024, D=16.0, xc=150.0, yc=500.0, U=0.05, Re=60.0, storage=st),
        cf.cylinder_2d(nx=512, ny=512, D=12.0, xc=100.0, yc=256.0, U=0.05, Re=50.0, storage=st),
        cf.periodic_box(cf.D2Q9, (1024, 1024, 1), nu=0.02, force=(1e-6, 0.0, 0.0), storage=st),
        cf.sphere_3d(nx=64, ny=48, nz=48, D=10.0, center=(20.0, 24.0, 24.2), U=0.04, Re=50.0, storage=st),
    ]
    assert sum(c.ncells for c in cases) >= 1 << 22
    nsteps = 10
    print(f"[batch] two lanes: {len(cases)} cases, {nsteps} steps", flush=True)
    with Batch(BUDGET, graph_steps=4) as B:
        ids = [B.add(c, _init(c), nsteps) for c in cases]
        done = 0
        while done < nsteps:
            done += B.advance(nsteps - done)
        info = B.info()
        assert info["graph_launches"] == 0 and info["eager_steps"] == nsteps
        for c, cid in zip(cases, ids):
            lev = Level(c, _init(c))
            lev.set_overlap(False)
            lev.advance(nsteps)
            assert B.checkpoint(cid) == lev.checkpoint(), f"{c.name}: state differs"
            if c.use_ib:
                assert np.array_equal(B.force_history(cid), lev.force_history()), c.name
            np.testing.assert_allclose(B.mass_history(cid), lev.mass_history(), rtol=1e-13, atol=0.0)
            lev.close()
