"""The incremental cell re-sort (sort_mode 1, the default from a million beads up) against
sorting every key each build (sort_mode 0).

Both produce the order of a stable sort of (cell key, bead index) pairs from index order
(the reference's thrust::sort_by_key over keys and indices, include/cuda/cell.h:254-260),
so every downstream result -- the staging order of the pair kernels, their FP64 sums, the
trajectory -- must be bit-identical.  A state after hundreds of steps is compared bit for
bit: a single out-of-place bead in any cell changes the summation order of its
neighbours' forces and, within a few steps, the positions.  Cases cover the few-percent
movers of an ordinary step, box moves (every bead rescaled, cell grids changed), a state
upload that moves every bead, the captured 8-step graph, batches, and the step phases.
"""
import numpy as np
import pytest

import softmold_amd as sm
from conftest import gpu_available

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason="needs the MI355X")]


def _pair(d, **kw):
    a = sm.Solver.from_system(d, **kw)
    b = sm.Solver.from_system(d, **kw)
    a.tune("sort_mode", 1)
    b.tune("sort_mode", 0)
    assert a.tune_get("sort_mode") == 1 and b.tune_get("sort_mode") == 0
    return a, b


def _same(a, b):
    for x, y in zip(a.download(), b.download()):
        assert np.array_equal(x, y)


_systems = {}


def system(name):
    if name not in _systems:
        m = {"L5000": lambda: sm.Mpd.liposome(1234, 5000, 3.45),
             "L80000": lambda: sm.Mpd.liposome(1234, 80000, 3.45),
             "B4000": lambda: sm.Mpd.bilayer(1234, 4000, 3.45)}[name]()
        _systems[name] = m.system()
        m.close()
    return _systems[name]


@pytest.mark.parametrize("name,kw,steps", [
    ("L5000", {}, 400),
    ("L80000", {}, 200),
    ("B4000", dict(delta_lxy=0.05, tension=1.0, seed=5), 400),   # box moves every 8 steps
])
def test_incremental_equals_full_sort(name, kw, steps):
    d = system(name)
    print(f"case incremental sort {name} {kw} {steps} steps", flush=True)
    a, b = _pair(d, **kw)
    for s in (a, b):
        s.init_forces()
    _same(a, b)
    h = steps // 2
    for s in (a, b):
        s.run(0, h)                 # eager (the fused kick + next first half)
    _same(a, b)
    for s in (a, b):
        s.tune("graph", 1)
        s.run(h, steps - h)         # captured 8-step blocks
    _same(a, b)
    if kw:
        t, acc = a.mc_stats()
        assert (t, acc) == b.mc_stats() and t > 0 and 0 < acc < t


def test_incremental_after_an_upload_that_moves_every_bead():
    d = system("L5000")
    a, b = _pair(d)
    for s in (a, b):
        s.init_forces()
        s.run(0, 20)
    pos, vel, acc, box = a.download()
    # a rigid shift by half a box: every bead changes cell
    shifted = np.mod(pos + 0.5 * box, box)
    for s in (a, b):
        s.upload(pos=shifted, vel=vel, acc=acc)
        s.run(20, 40)
    _same(a, b)


def test_batch_incremental_equals_full_sort():
    # replicas share their constant tables: three bilayers, one with box moves
    ds = [sm.Mpd.bilayer(7, 700, 3.45).system(), sm.Mpd.bilayer(9, 900, 3.45).system(),
          sm.Mpd.bilayer(11, 1500, 3.45).system()]
    ds[1]["delta_lxy"], ds[1]["tension"] = 0.05, 1.0
    a = sm.Solver.batch(ds)
    b = sm.Solver.batch(ds)
    a.tune("sort_mode", 1)
    b.tune("sort_mode", 0)
    for s in (a, b):
        s.init_forces()
        s.run(0, 120)
        s.tune("graph", 1)
        s.run(120, 200)
    _same(a, b)


def test_incremental_with_the_step_phases():
    """sm_step_first twice before the sort (the second call recomputes the keys and the
    mover counts from scratch), then the rest of the step, against the full sort."""
    d = system("L5000")
    a, b = _pair(d)
    for s in (a, b):
        s.init_forces()
        for step in range(1, 30):
            s.step_first(step)
            if step % 7 == 0:
                s.step_first(step)
            s.step_finish(step)
    _same(a, b)


def test_switching_modes_mid_run():
    d = system("L5000")
    a, b = _pair(d)
    for s in (a, b):
        s.init_forces()
    for k in range(4):
        for s in (a, b):
            s.run(50 * k, 50)
        a.tune("sort_mode", k % 2)
    _same(a, b)


def test_autotuner_keeps_the_trajectory():
    """The autotuner may pick either sort mode; the run after it equals an untuned run."""
    d = system("L5000")
    a, b = _pair(d)
    for s in (a, b):
        s.init_forces()
        s.run(0, 16)
    a.autotune(16)
    for s in (a, b):
        s.run(16, 64)
    _same(a, b)

