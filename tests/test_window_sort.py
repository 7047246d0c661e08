"""The window sort (sort_mode 2, the default for one replica of up to 524288 beads) against
sorting every key each build (sort_mode 0).

Keys whose cell lies in a window around the previous build's occupied cells are stored as
their index in the window's (z, y, x) order and sorted in two 9-bit passes; any key outside
it sends the build through the cell indices' three 8-bit passes in the same two launches.
Both must give the order of a stable sort of (cell key, bead index) pairs from index order
(include/cuda/cell.h:254-260), so every downstream result must be bit-identical to the
every-key sort's.  Cases: ordinary steps (the window path), box moves (the grid changes),
uploads that move every bead out of the window (the fallback, then the window again),
the captured 8-step graph, the step phases, tiny systems and mode switches.
"""
import numpy as np
import pytest

import softmold_amd as sm
from conftest import gpu_available

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason="needs the MI355X")]

_systems = {}


def system(name):
    if name not in _systems:
        m = {"L500": lambda: sm.Mpd.liposome(1234, 500, 3.45),
             "L5000": lambda: sm.Mpd.liposome(1234, 5000, 3.45),
             "L80000": lambda: sm.Mpd.liposome(1234, 80000, 3.45),
             "L100000": lambda: sm.Mpd.liposome(1234, 100000, 3.45)}[name]()
        _systems[name] = m.system()
        m.close()
    return _systems[name]


def _pair(d, **kw):
    a = sm.Solver.from_system(d, **kw)
    b = sm.Solver.from_system(d, **kw)
    assert a.tune_get("sort_mode") == 2     # the default at this size
    b.tune("sort_mode", 0)
    return a, b


def _same(a, b):
    for x, y in zip(a.download(), b.download()):
        assert np.array_equal(x, y)


@pytest.mark.parametrize("name,kw,steps", [
    ("L500", {}, 200),                                           # 512-key tiles
    ("L5000", {}, 400),
    ("L80000", {}, 200),                                         # 2048-key tiles, the bench's system
    ("L100000", {}, 100),                                        # a window of over 2^17 cells: 9 + 9 bits
    ("L5000", dict(delta_lxy=0.05, tension=1.0, seed=5), 400),   # box moves every 8 steps
])
def test_window_sort_equals_full_sort(name, kw, steps):
    d = system(name)
    print(f"case window sort {name} {kw} {steps} steps", flush=True)
    a, b = _pair(d, **kw)
    for s in (a, b):
        s.init_forces()
    _same(a, b)
    h = steps // 2
    for s in (a, b):
        s.run(0, h)                 # eager (the fused kick + next first half)
    _same(a, b)
    assert a.tune_get("sort_window_path") == 1
    for s in (a, b):
        s.tune("graph", 1)
        s.run(h, steps - h)         # captured 8-step blocks
    _same(a, b)
    assert a.tune_get("sort_window_path") == 1
    if kw:
        t, acc = a.mc_stats()
        assert (t, acc) == b.mc_stats() and t > 0 and 0 < acc < t


def test_window_sort_falls_back_when_beads_leave_the_window():
    d = system("L5000")
    a, b = _pair(d)
    for s in (a, b):
        s.init_forces()
        s.run(0, 20)
    assert a.tune_get("sort_window_path") == 1
    pos, vel, acc, box = a.download()
    # a rigid shift by a quarter box: every bead leaves the window of the last build
    shifted = np.mod(pos + 0.25 * box, box)
    for s in (a, b):
        s.upload(pos=shifted, vel=vel, acc=acc)
        s.step_first(20)
        s.step_finish(20)
    assert a.tune_get("sort_window_path") == 0     # that build took the fallback
    _same(a, b)
    for s in (a, b):
        s.run(21, 30)
    assert a.tune_get("sort_window_path") == 1     # and the next builds the window again
    _same(a, b)


def test_window_sort_with_the_step_phases():
    """sm_step_first twice before the sort: the second call starts the window's histograms,
    occupied range and fallback flag over."""
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
    assert a.tune_get("sort_window_path") == 1


def test_switching_among_the_three_modes():
    d = system("L5000")
    a, b = _pair(d)
    for s in (a, b):
        s.init_forces()
    for k in range(6):
        for s in (a, b):
            s.run(40 * k, 40)
        a.tune("sort_mode", (k + 2) % 3)
    _same(a, b)


def test_window_sort_refused_for_batches():
    ds = [sm.Mpd.liposome(7, 300, 3.45).system(), sm.Mpd.liposome(9, 400, 3.45).system()]
    s = sm.Solver.batch(ds)
    assert s.tune_get("sort_mode") == 0
    with pytest.raises(sm.SmError):
        s.tune("sort_mode", 2)
