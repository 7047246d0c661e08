"""The batched replica sweep at its full size: 64 liposome replicas of 6,072 to 119,553 beads
(2,805,567 in all), lipid counts log-uniform in [2000, 40000] from MT19937 seed 20261202,
tension linspace(0, 1, 8) x temperature 3 linspace(0.9, 1.1, 8), deltaLXY 0.01.

Every replica's cell grid is 200^3 in a 400^3 box; the replicas' dense cell tables hold
584M cells, so the batch's cell keys carry 30 bits and the radix sort runs all four of its
8-bit passes -- the only configuration that does.  The reference has no batch runner: each
replica is one `MD` run of its own (MD.cpp:589-721), so every replica is checked against the
FP64 oracle on its own system.

Bars: cell keys and the force kernel's accepted pair counts bit-exact; step-0 pair forces
within 1e-5 per particle (relative to max(|F_i|, 0.05 rms|F|)) and 1e-6 in L2; FP64 potential
energies to 1e-10 and the kinetic energy to 2^-23 (velocities are held in FP32); the smallest, median and largest replicas of a 600-step batched run (captured 8-step
graphs, chunk 32) bit-identical
to the same replicas run alone; the vesicle area per lipid 4 pi R^2 / N_outer of the
smallest and the median replica within 4 standard errors of independent oracle copies
(noise seeds differ) and within 0.5 %.  Marked gpu: it runs on the MI355X.
"""
import os

import numpy as np
import pytest

import oracle as O
import softmold_amd as sm
from conftest import gpu_available
from softmold_amd import observables as obs
from softmold_amd.sweep import replica_sweep, sweep_counts

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="needs the MI355X")]

HEAD = 2
FORCE_TOL, FORCE_TOL_L2 = 1e-5, 1e-6


@pytest.fixture(scope="module")
def sweep():
    O.set_threads(min(16, os.cpu_count() or 1))
    return replica_sweep()


def oracle_of(d, **kw):
    args = dict(cutoff=d["cutoff"], dt=d["dt"], gamma=d["gamma"], temp=d["temperature"],
                delta_lxy=d["delta_lxy"], tension=d["tension"], seed=d["seed"])
    args.update(kw)
    return O.System(d["pos"], d["type"], d["vel"], d["box"], d["fconst"], d["uconst"], d["n_types"],
                    d["chain"], d["chain_const"], **args)


def picks(sweep):
    n = np.array([len(d["type"]) for d in sweep])
    order = np.argsort(n, kind="stable")
    return int(order[0]), int(order[len(order) // 2]), int(order[-1])


def test_sweep_is_the_configured_workload(sweep):
    n = np.array([len(d["type"]) for d in sweep])
    assert len(sweep) == 64 and n.sum() == 2805567
    assert np.array_equal(n, 3 * sweep_counts())
    assert n.min() == 6072 and n.max() == 119553


def test_every_replica_at_step0_matches_oracle(sweep):
    print(f"case sweep step0 R={len(sweep)} N={sum(len(d['type']) for d in sweep)}", flush=True)
    b = sm.Solver.batch(sweep)
    keys = b.cell_keys()
    fp, fc = b.compute_forces()
    meas = b.measure()
    fcount, _ = b.force_pair_count()
    worst = 0.0
    for r, d in enumerate(sweep):
        lo, hi = b.first[r], b.first[r] + b.count[r]
        o = oracle_of(d)
        np.testing.assert_array_equal(keys[lo:hi], o.cell_keys(), err_msg=f"replica {r}")
        np.testing.assert_array_equal(fcount[lo:hi], o.pair_count_active(), err_msg=f"replica {r}")
        ref = o.pair_force()
        mag = np.linalg.norm(ref, axis=1)
        floor = 0.05 * np.sqrt((mag ** 2).mean())
        per = np.linalg.norm(fp[lo:hi] - ref, axis=1) / np.maximum(mag, floor)
        l2 = np.linalg.norm(fp[lo:hi] - ref) / np.linalg.norm(ref)
        worst = max(worst, per.max())
        assert per.max() < FORCE_TOL and l2 < FORCE_TOL_L2, (r, per.max(), l2)
        np.testing.assert_allclose(fc[lo:hi], o.chain_force(), rtol=1e-9, atol=1e-9)
        assert meas[r, 0] == pytest.approx(o.pair_potential(), rel=1e-10)
        assert meas[r, 1] == pytest.approx(o.chain_potential(), rel=1e-9, abs=1e-12)
        # velocities are stored in FP32: each v^2 carries at most 2^-23 relative error
        assert meas[r, 2] == pytest.approx(o.kinetic(), rel=2.0 ** -23)
    print(f"sweep step-0 pair force: worst per-particle error {worst:.3e} over 64 replicas")
    b.close()


def _radius_series(steps, every, solver):
    """Radius of each replica after every `every` steps of steps 0..steps-1 (MD.cpp's loop
    from t0 = 0), the run cut into blocks of `every` steps."""
    out = []
    for s0 in range(0, steps, every):
        solver.run(s0, every)
        out.append(solver.vesicle_radius(HEAD))
    return np.array(out)


def test_batch_equals_solo_and_area_per_lipid_matches_oracle(sweep):
    steps, every = 600, 24
    small, median, large = picks(sweep)
    print(f"case sweep run steps={steps} picks={small, median, large}", flush=True)
    b = sm.Solver.batch(sweep)
    b.tune("pair_chunk", 32)
    b.tune("graph", 1)
    b.init_forces()
    rad = _radius_series(steps, every, b)                 # [time, replica]
    pos, vel, _, box = b.download()
    trials, accepted = b.mc_stats()
    assert trials.min() == trials.max() == (steps - 1) // 8   # box moves at steps 8, 16, ...
    b.close()

    # the same replicas alone, in the plain solver's defaults (eager, size-chosen chunk)
    for r in (small, median, large):
        s = sm.Solver.from_system(sweep[r])
        s.init_forces()
        s.run(0, steps)
        p, v, _, bx = s.download()
        lo = int(np.sum([len(d["type"]) for d in sweep[:r]]))
        hi = lo + len(sweep[r]["type"])
        np.testing.assert_array_equal(p, pos[lo:hi], err_msg=f"replica {r}")
        np.testing.assert_array_equal(v, vel[lo:hi], err_msg=f"replica {r}")
        np.testing.assert_array_equal(bx, box[r])
        assert s.mc_stats() == (int(trials[r]), int(accepted[r]))
        s.close()

    # area per lipid against independent oracle copies (same replica, other noise seeds)
    copies = 4
    for r in (small, median):
        d = sweep[r]
        n_outer = obs.liposome_leaflets(len(d["type"]) // 3, 3.45)[1]
        a_gpu = obs.vesicle_area_per_lipid(rad[len(rad) // 2:, r], n_outer).mean()
        a_orc = []
        for c in range(copies):
            o = oracle_of(d, seed=d["seed"] + 5000 + c)
            md = O.MD(o, rng_mode=1)
            md.init_forces()
            series = []
            for s0 in range(0, steps, every):
                md.run(s0, s0 + every - 1)
                series.append(O.vesicle_radius(o.pos, d["type"], o.box, HEAD))
            series = np.array(series)
            a_orc.append(obs.vesicle_area_per_lipid(series[len(series) // 2:], n_outer).mean())
        a_orc = np.array(a_orc)
        sd = a_orc.std(ddof=1)
        gap = abs(a_gpu - a_orc.mean())
        print(f"replica {r} (N={len(d['type'])}): area per lipid GPU {a_gpu:.6f} oracle "
              f"{a_orc.mean():.6f} +- {sd:.2e} (copies), gap {gap / a_orc.mean():.2e} relative")
        assert gap <= 4.0 * sd * np.sqrt(1.0 + 1.0 / copies) + 1e-12
        assert gap <= 0.005 * a_orc.mean()
