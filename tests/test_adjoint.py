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


def _oracle_J(d, theta, nsteps):
    import oracle as O
    from softmold_amd.adjoint import pair_rows

    O.set_threads(min(16, os.cpu_count() or 1))
    F, U = pair_rows(d["fconst"], d["uconst"], d["n_types"], TAIL, TAIL, theta[0], theta[1])
    cc = np.array(d["chain_const"], np.float64, copy=True)
    cc[:, 1] = theta[2]
    cc[:, 3] = theta[3]
    s = O.System(d["pos"], d["type"], d["vel"], d["box"], F, U, d["n_types"], d["chain"], cc,
                 cutoff=d["cutoff"], dt=d["dt"], gamma=d["gamma"], temp=d["temperature"],
                 delta_lxy=d["delta_lxy"], tension=d["tension"], seed=d["seed"])
    return O.MD(s, rng_mode=1).objective(nsteps)


def _fd64(d, S, scale):
    """Central differences of the FP64 oracle's J with steps STEP * scale."""
    fd = np.zeros(4)
    for p in range(4):
        h = STEP[p] * scale
        a, b = THETA0.copy(), THETA0.copy()
        a[p] += h
        b[p] -= h
        fd[p] = (_oracle_J(d, a, S) - _oracle_J(d, b, S)) / (2 * h)
    return fd


@pytest.mark.gpu
def test_adjoint_gradient_matches_fp64_reference():
    """40 steps: the trajectory is chaotic enough that central differences need small
    steps to stay in the linear regime (at STEP they are off by up to 17 %, converging to
    the adjoint as the step shrinks), so the FP64 oracle, free of FP32 noise, supplies
    the reference gradient with steps STEP * 1e-4."""
    if not gpu_available():
        pytest.skip("no GPU")
    from softmold_amd.adjoint import get_params, objective

    d = _system()
    S, k = 40, 10
    s = sm.Solver.from_system(d)
    assert np.allclose(get_params(s, PAIR), THETA0)
    before = s.download()
    print("case: adjoint B400 40 steps, checkpoint 10", flush=True)
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
