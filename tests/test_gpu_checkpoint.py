"""Checkpoint / restart of a level and preemption of a batched case:
resuming from the dump must continue bit for bit (fields, force history,
mass history) as the uninterrupted run."""
import numpy as np
import pytest

from ilbmrex_amd import config as cf

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("storage", [cf.STORE_F64, cf.STORE_F32])
def test_level_restart_continues_bitexact(storage):
    from ilbmrex_amd.native import Level
    c = cf.cylinder_2d(nx=96, ny=64, D=12.0, xc=30.0, yc=33.0, U=0.05, Re=80.0, storage=storage)
    init = cf.uniform_init(c, u=(0.05, 0.0, 0.0))
    print(f"[ckpt] level storage={storage}", flush=True)
    a = Level(c, init)
    a.advance(30)
    ck = a.checkpoint()
    a.advance(20, want_state=True)
    b = Level(c, init)
    b.restore(ck)
    b.advance(20, want_state=True)
    assert np.array_equal(a.f(), b.f()) and np.array_equal(a.macro(), b.macro())
    assert np.array_equal(a.force_history()[-20:], b.force_history())
    assert np.array_equal(a.mass_history()[-20:], b.mass_history())
    assert b.info()["steps"] == 50 and b.info()["time"] == 50.0
    # a dump does not restore into a different case
    other = Level(c.replace(n=(96, 66, 1)), cf.uniform_init(c.replace(n=(96, 66, 1))))
    with pytest.raises(Exception):
        other.restore(ck)


def test_batch_preemption_resumes_bitexact():
    from ilbmrex_amd.native import Batch
    c = cf.cylinder_2d(nx=80, ny=48, D=10.0, xc=24.0, yc=25.0, U=0.05, Re=60.0, storage=cf.STORE_F32)
    d = cf.cylinder_2d(nx=128, ny=64, D=12.0, xc=30.0, yc=32.0, U=0.05, Re=90.0, storage=cf.STORE_F32)
    init_c, init_d = cf.uniform_init(c, u=(0.05, 0, 0)), cf.uniform_init(d, u=(0.05, 0, 0))
    print("[ckpt] batch preemption", flush=True)
    with Batch(4 << 30) as B:
        x = B.add(c, init_c, 40)
        B.add(d, init_d, 200)
        while B.progress(x)[0] < 40:
            B.advance(40)
        ref_f, ref_m = B.force_history(x), B.mass_history(x)
    with Batch(4 << 30) as B:
        x = B.add(c, init_c, 25)
        y = B.add(d, init_d, 200)
        B.advance(25)
        first_f, first_m = B.force_history(x), B.mass_history(x)
        ck = B.checkpoint(x)
        B.remove(x)                       # preempted
        B.advance(13)                     # the other case keeps running
        x2 = B.add_from_checkpoint(c, ck, 15)
        while B.progress(x2)[0] < 15:
            B.advance(15)
        f2, m2 = B.force_history(x2), B.mass_history(x2)
        assert B.progress(y)[0] == 25 + 13 + 15
    assert np.array_equal(np.vstack([first_f, f2]), ref_f)
    assert np.array_equal(np.concatenate([first_m, m2]), ref_m)


@pytest.mark.parametrize("shape,storage", [("cylinder", cf.STORE_F32), ("cylinder", cf.STORE_F16S),
                                           ("sphere", cf.STORE_F64)])
@pytest.mark.parametrize("first,second", [("0", "1"), ("1", "0"), ("0", "0")])
@pytest.mark.parametrize("at", [0, 30])
def test_restart_across_fused_and_separate_ghost_slabs(shape, storage, first, second, at, monkeypatch):
    """A level walled in x only whose stream-collide kernel writes the x-face
    ghost slabs (ILBM_GHOST_SEPARATE=0) keeps its boundary layer from the
    initial state only: a checkpoint taken after a step carries the current
    slabs, one taken before the first step the boundary layer.  Restored into
    a level with fused or separate slabs, either continues bit for bit."""
    from ilbmrex_amd.native import Level
    monkeypatch.delenv("ILBM_ADDR_GENERIC", raising=False)
    monkeypatch.delenv("ILBM_FORCE_ADDR64", raising=False)
    if shape == "cylinder":
        c = cf.cylinder_2d(nx=96, ny=64, D=12.0, xc=30.0, yc=33.0, U=0.05, Re=80.0, storage=storage)
    else:
        c = cf.sphere_3d(nx=40, ny=24, nz=24, D=8.0, center=(12.0, 12.0, 12.0), U=0.05, Re=40.0, storage=storage)
    init = cf.uniform_init(c, u=(0.05, 0.0, 0.0))
    print(f"[ckpt] {shape} storage={storage} slabs {first}->{second} at step {at}", flush=True)
    monkeypatch.setenv("ILBM_GHOST_SEPARATE", "0")
    ref = Level(c, init)
    ref.advance(at + 20, want_state=True)
    monkeypatch.setenv("ILBM_GHOST_SEPARATE", first)
    a = Level(c, init)
    assert a.info()["ghosts_fused"] == (first == "0")
    if at:
        a.advance(at)
    ck = a.checkpoint()
    a.close()
    monkeypatch.setenv("ILBM_GHOST_SEPARATE", second)
    b = Level(c, init)
    b.restore(ck)
    b.advance(20, want_state=True)
    assert np.array_equal(ref.f(), b.f()) and np.array_equal(ref.macro(), b.macro())
    assert np.array_equal(ref.force_history()[-20:], b.force_history())
    ref.close()
    b.close()
