"""The constant-fitting adjoint (sm_adjoint, softmold_amd.adjoint).

theta = (Umin, Umax) of the tail-tail pair (type 3 in the reference generators,
generate/include.h:31-32), k_bond, k_bend.  J = mean U after each kick over S steps with
fixed noise.  Checked against central differences of the GPU's own J, against central
differences of the FP64 oracle's J (the FP64 CPU reference gradient), for invariance to
the checkpoint interval, and through torch.autograd.
"""
import os

import numpy as np
import pytest

import softmold_amd as sm
from conftest import gpu_available

TAIL = 3
PAIR = (TAIL, TAIL)


def test_pair_rows_rebuild_the_generator_tables():
    """laradjiRevaleeFC/PC restated: rebuilding the tail-tail rows from Umin=-6, Umax=200
    gives the generator's tables back exactly (liposome.cpp:46-59)."""
    from softmold_amd.adjoint import pair_rows

    d = sm.Mpd.bilayer(1234, 400, 3.45).system()
    F, U = pair_rows(d["fconst"], d["uconst"], d["n_types"], TAIL, TAIL, -6.0, 200.0)
    assert np.array_equal(F, d["fconst"]) and np.array_equal(U, d["uconst"])
    # a default pair (Umin 0, Umax 100) rebuilt from its own values is unchanged too
    F2, U2 = pair_rows(d["fconst"], d["uconst"], d["n_types"], 2, 3, 0.0, 100.0)
    assert np.array_equal(F2, d["fconst"]) and np.array_equal(U2, d["uconst"])


def _system(n_lipids=400):
    d = sm.Mpd.bilayer(1234, n_lipids, 3.45).system()
    d["delta_lxy"] = 0.0   # no box moves: their decisions would make J piecewise constant
    return d


THETA0 = np.array([-6.0, 200.0, 100.0, 100.0])
STEP = np.array([6e-3, 0.2, 0.5, 0.5])   # central-difference steps


def _oracle_J(d, theta, nsteps, decisions=False):
    """The FP64 oracle's J along its own trajectory; with decisions=True also the accept
    bits of its box-move trials."""
    import oracle as O
    from softmold_amd.adjoint import pair_rows

    O.set_threads(min(16, os.cpu_count() or 1))
    F, U = pair_rows(d["fconst"], d["uconst"], d["n_types"], TAIL, TAIL, theta[0], theta[1])
    cc = np.array(d["chain_const"], np.float64, copy=True)
    cc[:, 1] = theta[2]
    cc[:, 3] = theta[3]
    s = O.System(d["pos"], d["type"], d["vel"], d["box"], F, U, d["n_types"], d["chain"], cc,
                 cutoff=d["cutoff"], dt=d["dt"], gamma=d["gamma"], temp=d["temperature"],
                 delta_lxy=d["delta_lxy"], tension=d["tension"], seed=d["seed"],
                 bonds=d.get("bond"), bond_const=d.get("bond_const"), bends=d.get("bend"),
                 bend_const=d.get("bend_const"), balls=d.get("ball"), ball_const=d.get("ball_const"))
    md = O.MD(s, rng_mode=1)
    J = md.objective(nsteps)
    return (J, md.decisions()) if decisions else J


def _fd64(d, S, scale, decisions=None):
    """Central differences of the FP64 oracle's J with steps STEP * scale.  With
    `decisions` (the accept bits at THETA0), every perturbed run must make the same box-move
    decisions: the pathwise derivative holds them fixed, and a flipped decision would make
    the difference jump."""
    fd = np.zeros(4)
    for p in range(4):
        h = STEP[p] * scale
        a, b = THETA0.copy(), THETA0.copy()
        a[p] += h
        b[p] -= h
        if decisions is None:
            fd[p] = (_oracle_J(d, a, S) - _oracle_J(d, b, S)) / (2 * h)
        else:
            (ja, da), (jb, db) = _oracle_J(d, a, S, True), _oracle_J(d, b, S, True)
            assert da == db == decisions, (p, bin(da), bin(db), bin(decisions))
            fd[p] = (ja - jb) / (2 * h)
    return fd


@pytest.mark.gpu
def test_adjoint_gradient_matches_fp64_reference():
    """32 steps: the trajectory is chaotic enough that central differences need small
    steps to stay in the linear regime (at STEP they are off by up to 17 % at 40 steps,
    converging to the adjoint as the step shrinks), so the FP64 oracle, free of FP32 noise,
    supplies the reference gradient with steps STEP * 1e-4.  The adjoint follows the
    FP32-forward trajectory, which leaves the FP64 one at the chaotic rate: against the
    FP64 reference gradient its largest component error was 4e-6 at 24 steps, 2e-5 at 32
    and 8e-4 at 40."""
    if not gpu_available():
        pytest.skip("no GPU")
    from softmold_amd.adjoint import get_params, objective

    d = _system()
    S, k = 32, 10
    s = sm.Solver.from_system(d)
    assert np.allclose(get_params(s, PAIR), THETA0)
    before = s.download()
    print("case: adjoint B400 32 steps, checkpoint 10", flush=True)
    J, g = objective(s, PAIR, S, k)
    after = s.download()
    for x, y in zip(before, after):
        assert np.array_equal(x, y)
    # the same J from the FP64 oracle along its own trajectory
    J64 = _oracle_J(d, THETA0, S)
    fd64 = _fd64(d, S, 1e-4)
    import adjoint_ref as AR
    gref = AR.adjoint_gradient(d, THETA0, PAIR, S)
    print(f"J gpu {J:.10g} oracle {J64:.10g}\nadjoint {g}\nFP64 central differences {fd64}\n"
          f"FP64 reference gradient {gref}")
    assert J == pytest.approx(J64, rel=1e-6)
    np.testing.assert_allclose(g, fd64, rtol=1e-3)
    np.testing.assert_allclose(g, gref, rtol=1e-3)


@pytest.mark.gpu
def test_adjoint_with_box_moves_matches_fp64_differences():
    """Constant-tension box moves on (deltaLXY 0.01, tension 1, a trial every 8 steps):
    accepted moves rescale the positions, which the adjoint carries through x -> S x with
    the decisions held fixed.  32 steps hold 4 trials; the FP64 oracle's central
    differences (steps STEP * 1e-4) make the same decisions at every perturbed theta."""
    if not gpu_available():
        pytest.skip("no GPU")
    from softmold_amd.adjoint import objective

    d = sm.Mpd.bilayer(1234, 400, 3.45).system()
    d["delta_lxy"], d["tension"] = 0.01, 1.0
    S, k = 32, 10
    s = sm.Solver.from_system(d)
    print("case: adjoint B400 with box moves, 32 steps, checkpoint 10", flush=True)
    J, g = objective(s, PAIR, S, k)
    J64, dec = _oracle_J(d, THETA0, S, True)
    assert dec != 0, "no trial accepted: the rescaling branch would not be exercised"
    fd64 = _fd64(d, S, 1e-4, decisions=dec)
    print(f"decisions {dec:04b} J gpu {J:.10g} oracle {J64:.10g}\nadjoint {g}\nFP64 central differences {fd64}")
    assert J == pytest.approx(J64, rel=1e-6)
    np.testing.assert_allclose(g, fd64, rtol=1e-3)


@pytest.mark.gpu
def test_adjoint_on_the_4000_lipid_bilayer():
    """The adjoint's configuration: the 4000-lipid bilayer (12,000 beads) as the generator
    writes it (box moves on, deltaLXY 0.01), 200 steps, checkpoint every 20.

    Central differences of J have a linear regime only over a short horizon: measured with
    the FP64 oracle on this system (Umin steps 6e-4 ... 6e-13), they agree to 1e-5 up to 40
    steps, drift by 10 % at 80 and grow without bound from 120 steps (to 1e11 at 200): the
    trajectory's sensitivity grows as e^(lambda t) with lambda ~ 6 per unit time, and the
    FP32-forward trajectory the adjoint follows leaves the FP64 one at that rate (largest
    component error against the FP64 reference gradient 3e-6 at 24 steps, 4e-4 at 32, 1e-3
    at 40).  So the 200-step run is timed and checked for J, finiteness and checkpoint
    invariance, and the gradient is checked over 24 steps with the same checkpoint interval
    (segments of 20 and 4 steps): with the box moves against FP64 central differences
    that make the same decisions, and without them also against the FP64 CPU reference
    gradient."""
    if not gpu_available():
        pytest.skip("no GPU")
    import time

    import adjoint_ref as AR
    from softmold_amd.adjoint import objective

    d = sm.Mpd.bilayer(1234, 4000, 3.45).system()
    assert len(d["type"]) == 12000 and d["delta_lxy"] == 0.01
    s = sm.Solver.from_system(d)
    print("case: adjoint B4000 200 steps, checkpoint 20", flush=True)
    s.sync()
    t0 = time.perf_counter()
    J, g = objective(s, PAIR, 200, 20)
    wall = time.perf_counter() - t0
    J1, g1 = objective(s, PAIR, 200, 200)
    J64 = _oracle_J(d, THETA0, 200)
    print(f"200 steps: adjoint wall {wall:.3f} s, J gpu {J:.10g} oracle {J64:.10g} "
          f"(relative {abs(J - J64) / abs(J64):.2e}), gradient {g}")
    assert np.all(np.isfinite(g))
    assert J == J1 and np.array_equal(g, g1)
    # the FP32 trajectory decorrelates from the FP64 one in the run's second half (same
    # noise), so the time-averaged energies agree statistically: the potential energy of
    # 12,000 beads fluctuates by ~1e-3 of its mean
    assert abs(J - J64) <= 1e-3 * abs(J64)

    S, k = 24, 20
    print("case: adjoint B4000 with box moves, 24 steps, checkpoint 20", flush=True)
    Jm, gm = objective(s, PAIR, S, k)
    J64m, dec = _oracle_J(d, THETA0, S, True)
    fdm = _fd64(d, S, 1e-4, decisions=dec)
    print(f"decisions {dec:03b} J gpu {Jm:.10g} oracle {J64m:.10g}\nadjoint {gm}\nFP64 central differences {fdm}")
    assert Jm == pytest.approx(J64m, rel=1e-6)
    np.testing.assert_allclose(gm, fdm, rtol=1e-3)

    d0 = dict(d)
    d0["delta_lxy"] = 0.0
    s0 = sm.Solver.from_system(d0)
    print("case: adjoint B4000 24 steps, checkpoint 20", flush=True)
    J0, g0 = objective(s0, PAIR, S, k)
    fd0 = _fd64(d0, S, 1e-4)
    gref = AR.adjoint_gradient(d0, THETA0, PAIR, S)
    print(f"J gpu {J0:.10g}\nadjoint {g0}\nFP64 central differences {fd0}\nFP64 reference gradient {gref}")
    np.testing.assert_allclose(g0, fd0, rtol=1e-3)
    np.testing.assert_allclose(g0, gref, rtol=1e-3)


@pytest.mark.gpu
def test_adjoint_gradient_matches_gpu_differences():
    """3 steps, in the linear regime at STEP: the adjoint against central differences of
    the GPU's own FP32-forward J, and of the oracle's."""
    if not gpu_available():
        pytest.skip("no GPU")
    from softmold_amd.adjoint import objective, set_params

    d = _system()
    S = 3
    s = sm.Solver.from_system(d)
    J, g = objective(s, PAIR, S, 1)
    fd = np.zeros(4)
    for p in range(4):
        Js = []
        for sgn in (1, -1):
            th = THETA0.copy()
            th[p] += sgn * STEP[p]
            set_params(s, PAIR, th)
            Js.append(objective(s, PAIR, S, 1, grad=False)[0])
        fd[p] = (Js[0] - Js[1]) / (2 * STEP[p])
    set_params(s, PAIR, THETA0)
    fd64 = _fd64(d, S, 1.0)
    print(f"adjoint {g}\nGPU central differences {fd}\nFP64 central differences {fd64}")
    np.testing.assert_allclose(g, fd, rtol=1e-3)
    np.testing.assert_allclose(g, fd64, rtol=1e-3)


@pytest.mark.gpu
def test_adjoint_checkpoint_interval_does_not_change_the_gradient():
    if not gpu_available():
        pytest.skip("no GPU")
    from softmold_amd.adjoint import objective

    s = sm.Solver.from_system(_system(300))
    J1, g1 = objective(s, PAIR, 24, 24)
    J2, g2 = objective(s, PAIR, 24, 5)
    assert J1 == J2
    assert np.array_equal(g1, g2)


@pytest.mark.gpu
def test_autograd_function():
    if not gpu_available():
        pytest.skip("no GPU")
    import torch

    from softmold_amd.adjoint import ConstantsObjective, objective

    s = sm.Solver.from_system(_system(300))
    theta = torch.tensor(THETA0, dtype=torch.float64, requires_grad=True)
    J = ConstantsObjective.apply(theta, s, PAIR, 16, 4)
    (2.0 * J).backward()
    Jr, gr = objective(s, PAIR, 16, 4)
    assert float(J.detach()) == Jr
    np.testing.assert_array_equal(theta.grad.numpy(), 2.0 * gr)


def test_fp64_reference_gradient_matches_fine_differences():
    """The FP64 CPU reference gradient (oracle/adjoint_ref.py: the reverse recursion on the
    oracle's trajectory with finite-difference Hessian-vector products) against central
    differences of the oracle's J with small steps."""
    import adjoint_ref as AR

    d = _system(300)
    S = 20
    g = AR.adjoint_gradient(d, THETA0, PAIR, S)
    fd = _fd64(d, S, 1e-4)
    print("reference", g, "differences", fd)
    np.testing.assert_allclose(g, fd, rtol=1e-4)


@pytest.mark.gpu
def test_batch_adjoint_equals_solo_adjoints():
    """A batch of three bilayers (one with box moves) gives, per replica, the J and the
    gradient of that system in a context of its own, bit for bit: the forward replicas are
    independent, the adjoint kernels work per bead, and the per-replica sums are segmented
    at the replica starts."""
    if not gpu_available():
        pytest.skip("no GPU")
    from softmold_amd.adjoint import objective

    ds = [_system(400), _system(250), sm.Mpd.bilayer(77, 300, 3.45).system()]
    ds[2]["delta_lxy"], ds[2]["tension"] = 0.01, 1.0
    print("case: batch adjoint, 3 bilayers, 24 steps", flush=True)
    b = sm.Solver.batch(ds)
    J, g = objective(b, PAIR, 24, 8)
    assert J.shape == (3,) and g.shape == (3, 4)
    for r, d in enumerate(ds):
        s = sm.Solver.from_system(d)
        Js, gs = objective(s, PAIR, 24, 8)
        print(f"replica {r}: J {J[r]:.10g} solo {Js:.10g}; grad {g[r]} solo {gs}")
        assert J[r] == Js
        assert np.array_equal(g[r], gs)


@pytest.mark.gpu
def test_adjoint_with_generic_records_matches_fp64_differences():
    """The 392-lipid bilayer with 60 BOND records between random head pairs, 30 BEND records
    over random head triples and a BALL of 40 members (their own constants, outside
    theta): 24 steps, the adjoint against FP64 central differences of the oracle's J, whose
    forces and energies include the records."""
    if not gpu_available():
        pytest.skip("no GPU")
    from test_generic_bonded import _crosslinked, _with_ball

    from softmold_amd.adjoint import objective

    d = _with_ball(_crosslinked(_system()))
    S = 24
    print("case: adjoint with BOND/BEND records, B400, 24 steps", flush=True)
    s = sm.Solver.from_system(d)
    J, g = objective(s, PAIR, S, 8)
    J64 = _oracle_J(d, THETA0, S)
    fd64 = _fd64(d, S, 1e-4)
    print(f"J gpu {J:.10g} oracle {J64:.10g}\nadjoint {g}\nFP64 differences {fd64}")
    assert J == pytest.approx(J64, rel=1e-6)
    np.testing.assert_allclose(g, fd64, rtol=1e-3)
    # the records change the gradient: the same system without them differs
    _, g0 = objective(sm.Solver.from_system(_system()), PAIR, S, 8)
    assert np.abs(g - g0).max() > 1e-3 * np.abs(g).max()

