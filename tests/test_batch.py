"""Batched replicas, the captured step and the autotuner against the single-system path.

A replica's beads, cell keys, Philox counters and FP64 sums depend on nothing outside the
replica, so a replica stepped in a batch must match the same system stepped alone
bit for bit; the same holds for the hipGraph replay of the step and for every launch
shape the autotuner may pick.  The single-system path is itself checked against the FP64
oracle in test_gpu_parity.py.
"""
import numpy as np
import pytest

import softmold_amd as sm
from conftest import gpu_available

pytestmark = pytest.mark.gpu


def _need_gpu():
    if not gpu_available():
        pytest.skip("no GPU")


def _systems():
    # replicas share constant tables: bilayers of one generator, uneven sizes
    a = sm.Mpd.bilayer(11, 400, 3.45).system()
    b = sm.Mpd.bilayer(12, 250, 3.45).system()
    c = sm.Mpd.bilayer(13, 700, 3.45).system()
    a.update(temperature=2.7, delta_lxy=0.05, tension=0.5)
    b.update(temperature=3.3, delta_lxy=0.05, tension=1.0)
    c.update(temperature=3.0, delta_lxy=0.0)        # no box moves in this one
    return [a, b, c]


def test_mixed_generators_are_refused():
    """Replicas must share the constant tables (the liposome and bilayer generators'
    tables differ in two Umax entries)."""
    if not gpu_available():
        pytest.skip("no GPU")
    a = sm.Mpd.liposome(11, 400, 3.45).system()
    b = sm.Mpd.bilayer(12, 300, 3.45).system()
    with pytest.raises(sm.SmError, match="constant tables"):
        sm.Solver.batch([a, b])


def test_large_box_replicas_equal_solo_runs():
    """Two 400^3 vesicles (8M cells each): cell bases far apart in the key space."""
    _need_gpu()
    systems = [sm.Mpd.liposome(21, 600, 3.45).system(), sm.Mpd.liposome(22, 900, 3.45).system()]
    systems[1].update(delta_lxy=0.01, tension=0.3)
    bt = sm.Solver.batch(systems)
    bt.init_forces()
    bt.run(0, 17)
    bp, bv, _, bbox = _state(bt)
    for r, d in enumerate(systems):
        s = sm.Solver.from_system(d)
        s.init_forces()
        s.run(0, 17)
        p, v, _, box = _state(s)
        lo, hi = bt.first[r], bt.first[r] + bt.count[r]
        assert np.array_equal(p, bp[lo:hi]) and np.array_equal(v, bv[lo:hi]), r
        assert np.array_equal(box[0], bbox[r]), r
        s.close()


def _state(s):
    pos, vel, acc, box = s.download()
    return pos, vel, acc, np.atleast_2d(box)


def test_batch_replicas_equal_solo_runs():
    _need_gpu()
    systems = _systems()
    steps = 41                                       # 5 MC moves (steps 8..40)
    print("case: batch of 3 vs solo, 41 steps", flush=True)
    bt = sm.Solver.batch(systems)
    assert bt.nrep == 3 and list(bt.count) == [len(d["type"]) for d in systems]
    bt.init_forces()
    bt.run(1, steps)
    bp, bv, ba, bbox = _state(bt)
    bm = bt.measure()
    bk = bt.cell_keys()
    btr, bac = bt.mc_stats()
    for r, d in enumerate(systems):
        s = sm.Solver.from_system(d)
        s.init_forces()
        s.run(1, steps)
        p, v, a, box = _state(s)
        lo, hi = bt.first[r], bt.first[r] + bt.count[r]
        assert np.array_equal(p, bp[lo:hi]), r
        assert np.array_equal(v, bv[lo:hi]), r
        assert np.array_equal(a, ba[lo:hi]), r
        assert np.array_equal(box[0], bbox[r]), r
        assert np.array_equal(s.measure(), bm[r]), r
        assert np.array_equal(s.cell_keys(), bk[lo:hi]), r
        t, acc = s.mc_stats()
        assert (t, acc) == (btr[r], bac[r]), r
        if d["delta_lxy"]:
            assert t == 5
            ha, hr = s.mc_hist()
            ba_, br_ = bt.mc_hist(r)
            assert np.array_equal(ha, ba_) and np.array_equal(hr, br_)
        else:
            assert t == 0
        s.close()


def test_graph_replay_equals_eager_steps():
    _need_gpu()
    systems = _systems()
    out = []
    for graph in (0, 1):
        print(f"case: graph={graph} batch of 3, 50 steps", flush=True)
        s = sm.Solver.batch(systems)
        s.tune("graph", graph)
        s.init_forces()
        s.run(1, 3)       # eager steps 1..3
        s.run(4, 30)      # eager 4..8, graph blocks 9..16, 17..24, 25..32, eager 33
        s.run(34, 16)     # eager 34..40, graph 41..48, eager 49
        out.append(_state(s) + (s.measure(), s.mc_stats()))
        s.close()
    for x, y in zip(out[0][:5], out[1][:5]):
        assert np.array_equal(x, y)
    assert all(np.array_equal(x, y) for x, y in zip(out[0][5], out[1][5]))


def test_launch_shapes_give_identical_forces():
    """Each cell is processed the same way whatever chunk holds it: forces are bit-identical
    across chunk sizes.  The two-phase variant sums the same terms but may round a product
    differently (FMA contraction), so it agrees to FP32 rounding."""
    _need_gpu()
    d = sm.Mpd.liposome(1234, 2000, 3.45).system()
    ref = {}
    for variant in (0, 1):
        for chunk in (4, 8, 16, 64):
            s = sm.Solver.from_system(d)
            s.tune("pair_chunk", chunk)
            s.tune("pair_variant", variant)
            fp, fc = s.compute_forces()
            s.close()
            if variant not in ref:
                ref[variant] = (fp, fc)
            assert np.array_equal(fp, ref[variant][0]), (chunk, variant)
            assert np.array_equal(fc, ref[variant][1]), (chunk, variant)
    a, b = ref[0][0], ref[1][0]
    assert np.linalg.norm(a - b) <= 1e-6 * np.linalg.norm(a)


def test_autotune_restores_state_and_keeps_results():
    _need_gpu()
    systems = _systems()
    a = sm.Solver.batch(systems)
    a.init_forces()
    a.run(1, 10)
    before = _state(a)
    print("case: autotune 16 steps per trial", flush=True)
    ms = a.autotune(16)
    assert ms > 0
    after = _state(a)
    for x, y in zip(before, after):
        assert np.array_equal(x, y)
    chunk, variant, graph = a.tune_get("pair_chunk"), a.tune_get("pair_variant"), a.tune_get("graph")
    assert chunk in (2, 4, 8, 16, 32, 64) and variant in (0, 1) and graph in (0, 1)
    a.run(11, 30)
    b = sm.Solver.batch(systems)
    b.init_forces()
    b.run(1, 10)
    b.run(11, 30)
    for x, y in zip(_state(a), _state(b)):
        assert np.array_equal(x, y)
    assert np.array_equal(a.measure(), b.measure())


def test_knobs_validate_and_round_trip():
    """sm_tune / sm_tune_get: every knob reads back what was set, and an unknown knob or a
    value outside its range is refused with SM_ERR_INVALID, leaving the knob unchanged."""
    if not gpu_available():
        pytest.skip("no GPU")
    s = sm.Solver.from_system(sm.Mpd.bilayer(1234, 250, 3.45).system())
    valid = {"pair_chunk": (1, 8, 4096), "pair_variant": (0, 1), "graph": (0, 1), "sort_mode": (0, 1, 2),
             "mc_fused": (0, 1)}
    invalid = {"pair_chunk": (0, -1, 4097), "pair_variant": (-1, 2), "graph": (2,), "sort_mode": (-1, 3),
               "mc_fused": (-1, 2)}
    for knob, values in valid.items():
        for v in values:
            s.tune(knob, v)
            assert s.tune_get(knob) == v, knob
        kept = s.tune_get(knob)
        for v in invalid[knob]:
            with pytest.raises(sm.SmError) as err:
                s.tune(knob, v)
            assert err.value.code == sm.SM_ERR_INVALID and knob in str(err.value)
            assert s.tune_get(knob) == kept
    for call in (lambda: s.tune("no_such_knob", 1), lambda: s.tune_get("no_such_knob")):
        with pytest.raises(sm.SmError) as err:
            call()
        assert err.value.code == sm.SM_ERR_INVALID and "unknown knob" in str(err.value)
    # the context still steps after the refusals
    s.tune("pair_chunk", 8)
    s.init_forces()
    s.run(0, 8)
    assert np.all(np.isfinite(s.download()[0]))
