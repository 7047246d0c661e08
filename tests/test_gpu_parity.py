"""Parity of the gfx950 path (through the C ABI) with the FP64 oracle on the same inputs.

Bars: cell keys and the neighbour pair set bit-exact; FP64 energies and dU to 1e-10
relative; step-0 pair forces (FP32 pair arithmetic on FP64 position differences) within
1e-5 relative, measured two ways -- per particle |dF_i| / max(|F_i|, 0.05 rms|F|) and
globally ||dF||_2 / ||F||_2.
"""
import numpy as np
import pytest

import oracle as O
import softmold_amd as sm
from conftest import gpu_available

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="needs the MI355X")]

FORCE_TOL = 1e-5      # per-particle, floored (the step-0 force bar for FP32 pair arithmetic)
FORCE_TOL_L2 = 1e-6   # global L2


@pytest.fixture(scope="module", autouse=True)
def oracle_threads():
    O.set_threads(8)
    yield


_cache = {}


def system(name):
    if name not in _cache:
        gen = {
            "L500": lambda: sm.Mpd.liposome(1234, 500, 3.45),
            "B400": lambda: sm.Mpd.bilayer(1234, 400, 3.45),
            "L5000": lambda: sm.Mpd.liposome(1234, 5000, 3.45),     # 15,000 beads
            "B4000": lambda: sm.Mpd.bilayer(1234, 4000, 3.45),      # 12,000 beads
            "L80000": lambda: sm.Mpd.liposome(1234, 80000, 3.45),   # 240,000 beads
            "B1M": lambda: sm.Mpd.bilayer(1234, 1000000, 3.45),     # 2,999,094 beads
        }[name]
        _cache[name] = gen().system()
    return _cache[name]


def oracle_of(d, **kw):
    args = dict(cutoff=d["cutoff"], dt=d["dt"], gamma=d["gamma"], temp=d["temperature"],
                delta_lxy=d["delta_lxy"], tension=d["tension"], seed=d["seed"])
    args.update(kw)
    return O.System(d["pos"], d["type"], d["vel"], d["box"], d["fconst"], d["uconst"], d["n_types"],
                    d["chain"], d["chain_const"], **args)


def force_errors(f, ref):
    mag = np.linalg.norm(ref, axis=1)
    floor = 0.05 * np.sqrt((mag ** 2).mean())
    per = np.linalg.norm(f - ref, axis=1) / np.maximum(mag, floor)
    l2 = np.linalg.norm(f - ref) / np.linalg.norm(ref)
    return per, l2


@pytest.mark.parametrize("name", ["L500", "B400", "L5000", "B4000", "L80000", "B1M"])
def test_cell_keys_bit_exact(name):
    d = system(name)
    print(f"case cell_keys {name} N={len(d['type'])}", flush=True)
    s = sm.Solver.from_system(d)
    np.testing.assert_array_equal(s.cell_keys(), oracle_of(d).cell_keys())


@pytest.mark.parametrize("name", ["L500", "B400", "B4000"])
def test_pair_set_bit_exact(name):
    d = system(name)
    print(f"case pair_set {name}", flush=True)
    s = sm.Solver.from_system(d)
    ours = s.pair_list()
    ref = oracle_of(d).pair_list()
    assert len(ours) == len(ref)
    assert set(map(tuple, ours.tolist())) == set(map(tuple, ref.tolist()))


@pytest.mark.parametrize("name", ["L5000", "L80000", "B1M"])
def test_neighbour_counts_bit_exact(name):
    d = system(name)
    print(f"case counts {name}", flush=True)
    s = sm.Solver.from_system(d)
    c, total = s.pair_count()
    rc = oracle_of(d).pair_count()
    np.testing.assert_array_equal(c, rc)
    assert total == rc.sum() // 2


@pytest.mark.parametrize("name", ["L500", "B400", "L5000", "B4000", "L80000", "B1M", "dense"])
def test_force_kernel_pair_set_bit_exact(name):
    """The pair force kernel's own accepted set -- its staging, per-quarter-cell culling and
    FP64 test, counting the pairs whose force term is not identically zero -- equals the
    oracle's count of such pairs over its full FORCE_COUNT traversal, per particle, at the
    default and at a long chunk.  A culling that dropped a pair with a nonzero term would
    show here as a count one short."""
    d = dense_system() if name == "dense" else system(name)
    print(f"case force_pair_set {name} N={len(d['type'])}", flush=True)
    s = sm.Solver.from_system(d)
    ref = oracle_of(d).pair_count_active()
    for chunk in (None, 64):
        if chunk:
            s.tune("pair_chunk", chunk)
        c, total = s.force_pair_count()
        np.testing.assert_array_equal(c, ref)
        assert total == ref.sum() // 2
    # the culling-free count is larger wherever head pairs sit between r_m and rc
    assert s.pair_count()[1] > total


@pytest.mark.parametrize("name", ["L500", "B400", "L5000", "B4000", "L80000", "B1M"])
def test_step0_pair_forces_fp32_vs_fp64(name):
    d = system(name)
    print(f"case forces {name}", flush=True)
    s = sm.Solver.from_system(d)
    fp, _ = s.compute_forces()
    ref = oracle_of(d).pair_force()
    per, l2 = force_errors(fp, ref)
    print(f"{name}: per-particle max {per.max():.3e} p99.9 {np.quantile(per, 0.999):.3e} L2 {l2:.3e}")
    assert l2 < FORCE_TOL_L2
    assert per.max() < FORCE_TOL
    # Newton's third law holds to FP32 rounding: each pair force is exactly antisymmetric
    # (r and the magnitude depend on r^2 only), so what is left is the FP32 rounding of the
    # per-particle sums, at most a few ulps of sum |F_i| in total
    assert np.linalg.norm(fp.sum(0)) < 2.0 * 2.0 ** -24 * np.abs(fp).sum()


def test_chain_forces_on_deformed_chains():
    d = dict(system("B4000"))
    rng = np.random.default_rng(5)
    pos = d["pos"] + rng.normal(0, 0.05, d["pos"].shape)
    pos = np.mod(pos, d["box"])
    d["pos"] = pos
    s = sm.Solver.from_system(d)
    _, fc = s.compute_forces()
    ref = oracle_of(d).chain_force()
    per, l2 = force_errors(fc, ref)
    print(f"chain: per max {per.max():.3e} L2 {l2:.3e}")
    assert l2 < 1e-6 and per.max() < 1e-5


@pytest.mark.parametrize("name", ["L500", "B400", "L5000", "B4000", "B1M"])
def test_energies_fp64(name):
    d = system(name)
    s = sm.Solver.from_system(d)
    U_pair, U_chain, K = s.measure()
    o = oracle_of(d)
    # FP64 sums of ~10^5 terms in different orders (and the oracle's own OpenMP order
    # varies run to run at ~1e-12): 1e-10 relative
    assert U_pair == pytest.approx(o.pair_potential(), rel=1e-10)
    assert U_chain == pytest.approx(o.chain_potential(), rel=1e-6, abs=1e-12)
    # velocities live in FP32 on the device
    assert K == pytest.approx(o.kinetic(), rel=1e-6)


def test_5000_lipid_vesicle_known_answer():
    s = sm.Solver.from_system(system("L5000"))
    U_pair, _, K = s.measure()
    assert U_pair == pytest.approx(-388321.0139620534, rel=1e-10)
    assert K == pytest.approx(1.5 * 15000 * 3.0, rel=1e-6)


@pytest.mark.parametrize("fused", [1, 0])
@pytest.mark.parametrize("name", ["B400", "B4000", "L5000", "B1M"])
def test_potential_delta_fp64(name, fused):
    """dU of the pair and CHAIN terms under two trial scales, from the traversal fused with
    the pair force (what a box-move step runs) and from the dU gather."""
    d = system(name)
    s = sm.Solver.from_system(d)
    s.tune("mc_fused", fused)
    o = oracle_of(d)
    for scale in ([1.0001, 1.0001, 1 / 1.0001 ** 2], [0.999, 0.999, 1 / 0.999 ** 2]):
        dp, dc = s.potential_delta(scale)
        assert dp == pytest.approx(o.pair_dpotential(scale), rel=1e-10, abs=1e-10)
        assert dc == pytest.approx(o.chain_dpotential(scale), rel=1e-8, abs=1e-12)


def dense_system():
    """Cells with more than 64 particles and stencils beyond one LDS pass; particles at
    exactly 0 and L; head and tail types mixed."""
    rng = np.random.default_rng(9)
    n = 3000
    box = np.array([12.0, 12.0, 12.0])
    pos = rng.uniform(0, 1, (n, 3)) * np.array([12.0, 12.0, 3.0])   # packed into z < 3
    pos[0] = [0.0, 0.0, 0.0]
    pos[1] = [12.0, 6.5, 1.25]   # on the upper x face (clamped into the last cell)
    types = rng.integers(2, 4, n).astype(np.int32)
    d = dict(system("L500"))
    d.update(pos=pos, type=types, vel=np.zeros((n, 3)), box=box,
             chain=np.zeros((0, 3), np.int32), chain_const=np.zeros((0, 4)))
    return d


def test_dense_cells_and_boundary_particles():
    d = dense_system()
    s = sm.Solver.from_system(d)
    o = oracle_of(d)
    np.testing.assert_array_equal(s.cell_keys(), o.cell_keys())
    assert set(map(tuple, s.pair_list().tolist())) == set(map(tuple, o.pair_list().tolist()))
    fp, _ = s.compute_forces()
    per, l2 = force_errors(fp, o.pair_force())
    print(f"dense: per max {per.max():.3e} L2 {l2:.3e}")
    assert l2 < FORCE_TOL_L2 and per.max() < FORCE_TOL


def test_deterministic_runs():
    d = system("B4000")
    out = []
    for _ in range(2):
        s = sm.Solver.from_system(d, delta_lxy=0.01, tension=0.5, seed=42)
        s.init_forces()
        s.run(0, 25)
        out.append(s.download())
    for a, b in zip(out[0], out[1]):
        np.testing.assert_array_equal(a, b)


def test_step_phases_equal_run():
    d = system("B400")
    a = sm.Solver.from_system(d, delta_lxy=0.01, tension=1.0, seed=7)
    b = sm.Solver.from_system(d, delta_lxy=0.01, tension=1.0, seed=7)
    a.init_forces()
    b.init_forces()
    a.run(0, 17)
    for i in range(17):
        b.step_first(i)
        b.step_finish(i, measure=(i % 5 == 0))
    for x, y in zip(a.download(), b.download()):
        np.testing.assert_array_equal(x, y)
    assert a.mc_stats() == b.mc_stats()


@pytest.mark.parametrize("name,steps", [("B400", 24), ("L500", 12)])
def test_trajectory_tracks_oracle_with_same_noise(name, steps):
    """Same initial state and the same counter-RNG noise: the FP32 GPU trajectory stays
    within FP32 drift of the FP64 oracle, and every MC decision is the same."""
    d = system(name)
    kw = dict(delta_lxy=0.01, tension=1.0, seed=99) if name == "B400" else dict(seed=99)
    s = sm.Solver.from_system(d, **kw)
    o = oracle_of(d, **kw)
    md = O.MD(o, rng_mode=1)
    s.init_forces()
    md.init_forces()
    s.run(0, steps)
    acc = md.run(0, steps - 1)
    pos, vel, _, box = s.download()
    if name == "B400":
        t, a = s.mc_stats()
        assert t == md.trials() == (steps - 1) // 8 and a == acc
        np.testing.assert_allclose(box, o.box, rtol=1e-14)
    dp = pos - o.pos
    dp -= box * np.round(dp / box)
    print(f"{name}: max |dp| {np.abs(dp).max():.3e}  max |dv| {np.abs(vel - o.vel).max():.3e}")
    assert np.abs(dp).max() < 1e-4
    assert np.abs(vel - o.vel).max() < 1e-2


def test_runaway_bead_is_reported_as_a_state_error():
    """A bead given a velocity that carries it past the box's far side within one step (one
    wrap does not bring it back) trips the key kernel's position check; every index stays
    clamped inside the grid, the step completes, and the next synchronising call reports
    SM_ERR_STATE.  A new context on the same device then runs normally."""
    d = system("L500")
    s = sm.Solver.from_system(d)
    s.init_forces()
    s.run(0, 4)
    s.sync()
    pos, vel, acc, box = s.download()
    vel = vel.copy()
    vel[17] = [1.0e5, 0.0, 0.0]   # 2000 per step at dt 0.02, five box lengths
    s.upload(vel=vel)
    s.run(4, 2)
    with pytest.raises(sm.SmError, match="left the box") as e:
        s.sync()
    assert e.value.code == sm.SM_ERR_STATE
    s.close()
    t = sm.Solver.from_system(d)
    t.init_forces()
    t.run(0, 8)
    t.sync()
    assert np.all(np.isfinite(t.measure()))


@pytest.mark.parametrize("name", ["B4000", "B400"])
def test_fused_box_move_dU_keeps_the_trajectory(name):
    """Box moves every 8 steps (deltaLXY 0.05, tension 1) with the pair dU taken in the
    step's force traversal and in a gather after the kick: the same decisions and the same
    trajectory, bit for bit (dU decides only the accept bit), eager and in captured graphs."""
    d = dict(system(name), delta_lxy=0.05, tension=1.0, seed=11)
    print(f"case fused dU {name}", flush=True)
    a = sm.Solver.from_system(d)
    b = sm.Solver.from_system(d)
    b.tune("mc_fused", 0)
    for s in (a, b):
        s.init_forces()
        s.run(0, 120)
        s.tune("graph", 1)
        s.run(120, 200)
    assert a.mc_stats() == b.mc_stats() and a.mc_stats()[0] == 39
    for x, y in zip(a.download(), b.download()):
        assert np.array_equal(x, y)
    for r in range(2):
        ha, hb = a.mc_hist(), b.mc_hist()
        assert np.array_equal(ha[r], hb[r])



@pytest.mark.parametrize("sort_mode", [0, 1, 2])
def test_systems_smaller_than_a_wave(sort_mode):
    """One bead, two beads, and one 3-bead lipid chain beside a free bead, in a 10^3 box:
    fewer beads than a wave, a key sort of one partial tile, and pair-kernel chunks past
    the last bead.  Step-0 forces against the oracle, then 24 steps of the same noise."""
    base = system("L500")
    cases = {
        "one": (np.array([[5.0, 5.0, 5.0]]), np.array([3], np.int32), np.zeros((0, 3), np.int32)),
        "two": (np.array([[5.0, 5.0, 5.0], [5.6, 5.3, 4.8]]), np.array([3, 2], np.int32),
                np.zeros((0, 3), np.int32)),
        "chain": (np.array([[1.0, 9.5, 0.2], [1.5, 9.8, 9.9], [2.1, 0.1, 9.7], [1.2, 9.2, 0.9]]),
                  np.array([2, 3, 3, 3], np.int32), np.array([[0, 1, 3]], np.int32)),
    }
    for name, (pos, types, chain) in cases.items():
        n = len(types)
        d = dict(base)
        d.update(pos=pos, type=types, vel=np.zeros((n, 3)), box=np.array([10.0, 10.0, 10.0]), chain=chain,
                 chain_const=np.array(base["chain_const"])[: len(chain)])
        print(f"case tiny {name} n={n} sort_mode {sort_mode}", flush=True)
        s = sm.Solver.from_system(d)
        s.tune("sort_mode", sort_mode)
        o = oracle_of(d)
        np.testing.assert_array_equal(s.cell_keys(), o.cell_keys())
        fp, _ = s.compute_forces()
        ref = o.pair_force()
        assert np.all(np.abs(fp - ref) <= 1e-5 * np.maximum(np.abs(ref), 1.0)), (fp, ref)
        md = O.MD(o, rng_mode=1)
        md.init_forces()
        md.run(0, 23)
        s.init_forces()
        s.run(0, 24)
        p, v, _, _ = s.download()
        dp = p - o.pos
        dp -= 10.0 * np.round(dp / 10.0)
        print(f"{name}: max |dp| {np.abs(dp).max():.3e} max |dv| {np.abs(v - o.vel).max():.3e}")
        assert np.abs(dp).max() < 1e-6 and np.abs(v - o.vel).max() < 1e-4
