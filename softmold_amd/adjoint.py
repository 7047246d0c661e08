"""Constant fitting through the MD trajectory: a torch.autograd.Function over the discrete
adjoint of the step (sm_adjoint in the C ABI).

theta = (Umin, Umax) of one type pair, then k_bond and k_bend of the CHAIN sets.  The
objective is J(theta) = mean over steps 1..S of U_pair + U_chain after each step's kick,
along the trajectory from the solver's current state with the counter RNG (fixed noise).
The forward runs in the solver's FP32 pair arithmetic; the adjoint accumulates in FP64 and
recomputes the trajectory from checkpoints every `checkpoint` steps.

Area per lipid changes only through the Monte-Carlo box moves, whose accept/reject
decisions are piecewise constant in theta: its pathwise gradient is zero, which is what the
adjoint (decisions held fixed) and small central differences both give.  The gradient of
its ensemble average comes from the fluctuation formula in softmold_amd.response.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import Solver, _check, _dp, lib

__all__ = ["pair_rows", "ConstantsObjective", "objective", "set_params", "get_params"]


def pair_rows(fconst, uconst, n_types: int, a: int, b: int, umin: float, umax: float):
    """Copies of the F and U tables with the (a,b), (b,a) rows rebuilt from Umin, Umax by
    laradjiRevaleeFC / laradjiRevaleePC (potentials/laradjiRevalee.h:84-152), keeping each
    row's r_m (C0) and r_c (C3)."""
    F = np.array(fconst, np.float64, copy=True)
    U = np.array(uconst, np.float64, copy=True)
    for row in {6 * (a * n_types + b), 6 * (b * n_types + a)}:
        rm, rc = F[row], F[row + 3]
        F[row + 1] = (2.0 * (umax - umin)) / (rm * rm)
        F[row + 2] = 0.0
        F[row + 4] = (6.0 * umin) / ((rc - rm) * (rc - rm))
        F[row + 5] = (6.0 * umin) / ((rc - rm) * (rc - rm) * (rc - rm))
        U[row:row + 6] = [rm, (umax - umin) / (rm * rm), umin, rc, (3.0 * umin) / ((rc - rm) * (rc - rm)),
                          (2.0 * umin) / ((rc - rm) * (rc - rm) * (rc - rm))]
    return F, U


def set_params(solver: Solver, pair: tuple[int, int], theta) -> None:
    t = np.ascontiguousarray(theta, np.float64)
    _check(lib().sm_set_params(solver.h, int(pair[0]), int(pair[1]), _dp(t)))


def get_params(solver: Solver, pair: tuple[int, int]) -> np.ndarray:
    t = np.zeros(4)
    _check(lib().sm_get_params(solver.h, int(pair[0]), int(pair[1]), _dp(t)))
    return t


def objective(solver: Solver, pair: tuple[int, int], nsteps: int, checkpoint: int, grad: bool = True):
    """(J, dJ/dtheta or None) at the solver's current constants; the solver state is kept.
    A batch gives one J per replica and a [replicas, 4] gradient."""
    R = solver.nrep
    J = np.zeros(R)
    g = np.zeros((R, 4))
    _check(lib().sm_adjoint(solver.h, int(pair[0]), int(pair[1]), int(nsteps), int(checkpoint), _dp(J),
                            _dp(g) if grad else None))
    if R == 1:
        return float(J[0]), (g[0] if grad else None)
    return J, (g if grad else None)


try:
    import torch

    class ConstantsObjective(torch.autograd.Function):
        """J(theta) for a solver, type pair, step count and checkpoint interval.

            J = ConstantsObjective.apply(theta, solver, (TAIL, TAIL), 200, 20)
            J.backward()          # theta.grad = dJ/dtheta from the FP64 adjoint
        """

        @staticmethod
        def forward(ctx, theta, solver, pair, nsteps, checkpoint):
            th = theta.detach().to("cpu", torch.float64).contiguous().numpy()
            set_params(solver, pair, th)
            J, g = objective(solver, pair, nsteps, checkpoint, grad=theta.requires_grad)
            ctx.save_for_backward(torch.as_tensor(g if g is not None else np.zeros(4), dtype=torch.float64))
            return torch.tensor(J, dtype=theta.dtype, device=theta.device)

        @staticmethod
        def backward(ctx, gout):
            (g,) = ctx.saved_tensors
            return gout * g.to(gout.device, gout.dtype), None, None, None, None
except ImportError:  # torch is optional for the C-level adjoint
    ConstantsObjective = None
