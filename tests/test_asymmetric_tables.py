"""Two-body tables whose (a, b) and (b, a) rows differ.

The reference evaluates each pair once, with the row of the particle its half-shell cell
traversal makes first (MD.h:811-819, cellOpt.h:712-926): across cells the particle whose
cell has the other's in its forward half shell (cellOpt.h:715), within a cell the lower
particle index (the cell's list runs in descending index, cellOpt.h:572-585).  The device
kernels visit every pair from both sides, so each side picks that row.

The oracle follows the reference's traversal and is pinned on the asymmetric systems A500
and AB400 (L500 and B400 with the (2, 3) and (3, 2) rows replaced) by the compiled
reference's own step-0 forces, energies and MD trajectories (tests/test_oracle.py).  Here
the device path is compared with the oracle on those systems and on larger ones carrying
the same table, to the bars of the symmetric case: pair set bit-exact, FP64 energies and
dU to 1e-10, step-0 forces to 1e-5 per particle and 1e-6 in L2, the same-noise trajectory
with identical box-move decisions, and the adjoint against FP64 differences.
"""
import os
import tempfile

import numpy as np
import pytest

import oracle as O
import softmold_amd as sm
from conftest import gpu_available

FORCE_TOL, FORCE_TOL_L2 = 1e-5, 1e-6


def _fixture_system(golden_traj, name):
    """The system dict of a reference-written input held by the trajectory fixtures."""
    d = tempfile.mkdtemp()
    path = os.path.join(d, name)
    with open(path + ".mpd", "wb") as f:
        f.write(golden_traj[name + "/input_mpd"].tobytes())
    m = sm.Mpd.read(path)
    out = m.system()
    m.close()
    return out


def _with_table(d, src):
    e = dict(d)
    e["fconst"] = np.array(src["fconst"], np.float64, copy=True)
    e["uconst"] = np.array(src["uconst"], np.float64, copy=True)
    return e


def test_fixture_tables_are_asymmetric(golden_traj):
    for name in ("A500", "AB400t"):
        d = _fixture_system(golden_traj, name)
        nT = d["n_types"]
        F = np.asarray(d["fconst"]).reshape(nT, nT, 6)
        U = np.asarray(d["uconst"]).reshape(nT, nT, 6)
        assert not np.array_equal(F[2, 3], F[3, 2]) and not np.array_equal(U[2, 3], U[3, 2])
        # every other pair is symmetric, and r_m >= cutoff / 2 holds for every row
        F2 = F.copy()
        F2[2, 3] = F2[3, 2] = 0
        assert np.array_equal(F2, F2.transpose(1, 0, 2))
        assert np.all(2 * F[:, :, 0] >= d["cutoff"])


def test_short_r_m_is_refused():
    """With r_m below half the cutoff the reference's force reads its branch at row offset
    int(r/r_m)*3 = 6 for r in [2 r_m, cutoff) (MD.h:819), past the row; such tables are
    refused on the host before any device call."""
    d = sm.Mpd.bilayer(1234, 400, 3.45).system()
    F = np.array(d["fconst"], np.float64, copy=True).reshape(-1, 6)
    F[2 * d["n_types"] + 3, 0] = 0.9
    e = dict(d, fconst=F.ravel())
    with pytest.raises(sm.SmError, match="r_m below half the cutoff") as err:
        sm.Solver.from_system(e)
    assert err.value.code == sm.SM_ERR_UNSUPPORTED


# ---- device ----------------------------------------------------------------------------

gpu = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason="needs the MI355X")]

_cache = {}


def system(golden_traj, name):
    if name not in _cache:
        if name in ("A500", "AB400t"):
            _cache[name] = _fixture_system(golden_traj, name)
        else:
            table = system(golden_traj, "A500")
            gen = {"L5000a": lambda: sm.Mpd.liposome(1234, 5000, 3.45),
                   "B4000a": lambda: sm.Mpd.bilayer(1234, 4000, 3.45),
                   "L80000a": lambda: sm.Mpd.liposome(1234, 80000, 3.45)}[name]
            m = gen()
            _cache[name] = _with_table(m.system(), table)
            m.close()
    return _cache[name]


def oracle_of(d, **kw):
    args = dict(cutoff=d["cutoff"], dt=d["dt"], gamma=d["gamma"], temp=d["temperature"],
                delta_lxy=d["delta_lxy"], tension=d["tension"], seed=d["seed"])
    args.update(kw)
    return O.System(d["pos"], d["type"], d["vel"], d["box"], d["fconst"], d["uconst"], d["n_types"],
                    d["chain"], d["chain_const"], **args)


@pytest.fixture(scope="module", autouse=True)
def oracle_threads():
    O.set_threads(8)
    yield


NAMES = ["A500", "AB400t", "L5000a", "B4000a", "L80000a"]


@pytest.mark.parametrize("name", NAMES)
@pytest.mark.gpu
def test_asymmetric_step0_forces(golden_traj, name):
    if not gpu_available():
        pytest.skip("no GPU")
    d = system(golden_traj, name)
    print(f"case asym forces {name} N={len(d['type'])}", flush=True)
    s = sm.Solver.from_system(d)
    o = oracle_of(d)
    ref = o.pair_force()
    for variant in (0, 1):   # one-phase and two-phase pair kernels
        s.tune("pair_variant", variant)
        fp, _ = s.compute_forces()
        mag = np.linalg.norm(ref, axis=1)
        per = np.linalg.norm(fp - ref, axis=1) / np.maximum(mag, 0.05 * np.sqrt((mag ** 2).mean()))
        l2 = np.linalg.norm(fp - ref) / np.linalg.norm(ref)
        print(f"{name} variant {variant}: per-particle max {per.max():.3e} L2 {l2:.3e}")
        assert per.max() < FORCE_TOL and l2 < FORCE_TOL_L2
        # both sides of a pair take the same row, so the pair forces cancel in the sum
        assert np.linalg.norm(fp.sum(0)) < 2.0 * 2.0 ** -24 * np.abs(fp).sum()


@pytest.mark.parametrize("name", NAMES)
@pytest.mark.gpu
def test_asymmetric_force_pair_set(golden_traj, name):
    """The force kernel's own accepted pairs (those with a term not identically zero under
    the first particle's row) equal the oracle's, per particle."""
    if not gpu_available():
        pytest.skip("no GPU")
    d = system(golden_traj, name)
    print(f"case asym force_pair_set {name}", flush=True)
    s = sm.Solver.from_system(d)
    ref = oracle_of(d).pair_count_active()
    for chunk in (None, 64):
        if chunk:
            s.tune("pair_chunk", chunk)
        c, total = s.force_pair_count()
        np.testing.assert_array_equal(c, ref)
        assert total == ref.sum() // 2


@pytest.mark.parametrize("name", NAMES)
@pytest.mark.gpu
def test_asymmetric_energies_and_delta(golden_traj, name):
    if not gpu_available():
        pytest.skip("no GPU")
    d = system(golden_traj, name)
    print(f"case asym energies {name}", flush=True)
    s = sm.Solver.from_system(d)
    o = oracle_of(d)
    U_pair, _, _ = s.measure()
    assert U_pair == pytest.approx(o.pair_potential(), rel=1e-10)
    for scale in ([1.0001, 1.0001, 1 / 1.0001 ** 2], [0.999, 0.999, 1 / 0.999 ** 2]):
        dp, _ = s.potential_delta(scale)
        assert dp == pytest.approx(o.pair_dpotential(scale), rel=1e-10, abs=1e-10)


@pytest.mark.gpu
def test_asymmetric_trajectory_tracks_oracle_with_same_noise(golden_traj):
    """AB400 with box moves (deltaLXY 0.01, tension 1), 24 steps of the same counter-RNG
    noise on both sides, in one solver and as the second replica of a batch."""
    if not gpu_available():
        pytest.skip("no GPU")
    d = system(golden_traj, "AB400t")
    kw = dict(delta_lxy=0.01, tension=1.0, seed=99)
    print("case asym trajectory AB400t 24 steps", flush=True)
    o = oracle_of(d, **kw)
    md = O.MD(o, rng_mode=1)
    md.init_forces()
    acc = md.run(0, 23)
    e = dict(d, **kw)
    s = sm.Solver.from_system(e)
    s.init_forces()
    s.run(0, 24)
    pos, vel, _, box = s.download()
    t, a = s.mc_stats()
    assert t == md.trials() == 2 and a == acc
    np.testing.assert_allclose(box, o.box, rtol=1e-14)
    dp = pos - o.pos
    dp -= box * np.round(dp / box)
    print(f"solo: max |dp| {np.abs(dp).max():.3e}  max |dv| {np.abs(vel - o.vel).max():.3e}")
    assert np.abs(dp).max() < 1e-4 and np.abs(vel - o.vel).max() < 1e-2
    # the same system behind another replica that carries the same table
    other = dict(_with_table(sm.Mpd.bilayer(77, 400, 3.45).system(), d), **kw)
    b = sm.Solver.batch([other, e])
    b.init_forces()
    b.run(0, 24)
    pb = b.download()[0][b.first[1]:b.first[1] + b.count[1]]
    assert np.array_equal(pb, pos)


@pytest.mark.gpu
def test_asymmetric_adjoint_matches_fp64_differences(golden_traj):
    """The adjoint's pair kernel picks the first particle's row as the force kernels do:
    its gradient on AB400 (tail-tail constants and the CHAIN constants; the asymmetric
    head-tail rows stay in place) against FP64 central differences of the oracle's J."""
    if not gpu_available():
        pytest.skip("no GPU")
    from test_adjoint import PAIR, THETA0, _fd64, _oracle_J

    from softmold_amd.adjoint import objective

    d = dict(system(golden_traj, "AB400t"), delta_lxy=0.0)
    S = 24
    print("case asym adjoint AB400 24 steps", flush=True)
    s = sm.Solver.from_system(d)
    J, g = objective(s, PAIR, S, 8)
    J64 = _oracle_J(d, THETA0, S)
    fd64 = _fd64(d, S, 1e-4)
    print(f"J gpu {J:.10g} oracle {J64:.10g}\nadjoint {g}\nFP64 differences {fd64}")
    assert J == pytest.approx(J64, rel=1e-6)
    np.testing.assert_allclose(g, fd64, rtol=1e-3)


@pytest.mark.gpu
def test_asymmetric_md_hip_step0_matches_reference_md(golden_traj, tmp_path):
    """The drop-in executable on the reference-written AB400 input (asymmetric head-tail
    rows, tension 1): its step-0 potential_ and kinetic_ lines equal the compiled reference
    MD's, as MD prints them (6 significant digits).  Later lines differ by design: the
    noise comes from a counter RNG."""
    if not gpu_available():
        pytest.skip("no GPU")
    import subprocess

    cli = os.path.join(os.path.dirname(sm.__file__), "md_hip")
    name = "AB400t"
    (tmp_path / f"{name}.mpd").write_bytes(golden_traj[name + "/input_mpd"].tobytes())
    print("case: md_hip on AB400t", flush=True)
    r = subprocess.run([cli, name], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr

    def first(path):
        with open(path) as f:
            return [float(x) for x in f.readline().split("\t")]

    for dat in ("potential", "kinetic"):
        got = first(tmp_path / f"{dat}_{name}.dat")
        want = golden_traj[f"{name}/{dat}"][0]
        assert got[0] == want[0] == 0.0
        assert float(f"{got[1]:.6g}") == want[1], (dat, got, want)
