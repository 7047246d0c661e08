"""Equilibrium response of ensemble averages to the interaction constants.

The area per lipid changes only through the Monte-Carlo box moves (MD.cpp:589-721), whose
accept/reject decisions are piecewise constant in the constants theta: the pathwise
derivative that the adjoint computes is zero for it, and the pathwise derivative of any
time average grows with the trajectory's chaos (e^(6 t) per unit time at this step).  The
derivative of the *ensemble* average is well defined instead.  The Langevin thermostat
samples exp(-U/T) at fixed box, and the box move accepts with min(1, exp((U_old - U_new +
tension (A_new - A_old)) / T)) from a symmetric proposal at constant volume, so the joint
stationary density is proportional to exp(-(U(x; theta) - tension A) / T) and, for an
observable O,

    d<O>/dtheta = <dO/dtheta> - (1/T) Cov(O, dU/dtheta)

(the likelihood-ratio, or fluctuation, formula).  The area per lipid has no explicit theta
dependence, so d<a>/dtheta = -Cov(a, dU/dtheta) / T; for the potential energy
d<U>/dtheta = <dU/dtheta> - Cov(U, dU/dtheta) / T.  dU/dtheta comes from the device in
FP64 (Solver.param_derivative); covariances are taken over sampled runs of independent
copies of the system (different noise seeds) about the pooled means, and the copies'
scatter gives the standard error.
"""
from __future__ import annotations

import numpy as np

from . import Solver
from .observables import bilayer_area_per_lipid, vesicle_area_per_lipid

__all__ = ["sample", "response", "chain_lipids"]


def chain_lipids(d: dict) -> int:
    """Lipids (CHAIN molecules) of a system dict."""
    return int(sum(int(n) for _, n, _ in np.asarray(d["chain"]).reshape(-1, 3)))


def sample(solver: Solver, pair, first_step: int, nsteps: int, every: int, n_lipids, kind: str = "bilayer",
           n_outer=None, head_type: int = 2) -> dict:
    """Steps first_step .. first_step + nsteps - 1 in blocks of `every`; after each block, per
    replica: the area per lipid (bilayer: 2 Lx Ly / N_lipids; vesicle: 4 pi R^2 / N_outer
    from the device's head radius), U = U_pair + U_chain, and dU/dtheta (4 values).
    Returns arrays a[t, r], U[t, r], D[t, r, 4]."""
    R = solver.nrep
    n_lipids = np.broadcast_to(np.asarray(n_lipids, float), (R,))
    a, U, D = [], [], []
    step = first_step
    while step < first_step + nsteps:
        k = min(every, first_step + nsteps - step)
        solver.run(step, k)
        step += k
        if kind == "bilayer":
            box = np.asarray(solver.box()).reshape(R, 3)
            a.append(bilayer_area_per_lipid(box, 1.0) / n_lipids)
        else:
            rad = np.asarray(solver.vesicle_radius(head_type)).reshape(R)
            a.append(vesicle_area_per_lipid(rad, 1.0) / np.broadcast_to(np.asarray(n_outer, float), (R,)))
        m = np.asarray(solver.measure()).reshape(R, 3)
        U.append(m[:, 0] + m[:, 1])
        D.append(np.asarray(solver.param_derivative(pair)).reshape(R, 4))
    return {"a": np.array(a), "U": np.array(U), "D": np.array(D)}


def response(s: dict, temperature) -> dict:
    """The fluctuation-formula gradients from sampled runs of R independent copies of one
    system at one temperature.  Covariances are taken about the means pooled over all
    copies and samples: a copy's own time mean would bias each covariance by about the
    observable's correlation time over the run length, which is long for the area (it
    relaxes through the box moves).  Each copy's mean product is one estimate, and their
    scatter gives the standard error.

    Returns area (<a>) and its standard error, d_area (d<a>/dtheta, 4) and d_area_se, U,
    d_U and d_U_se, the per-copy estimates, and drift (the second half's <a> minus the
    first half's, over its standard error: a stationary run gives |drift| of order 1)."""
    a, U, D = s["a"], s["U"], s["D"]
    R = a.shape[1]
    T = float(np.mean(temperature))
    da = a - a.mean()
    dU = U - U.mean()
    dD = D - D.mean((0, 1))
    g_a = -(da[:, :, None] * dD).mean(0) / T          # [R, 4]
    g_U = D.mean(0) - (dU[:, :, None] * dD).mean(0) / T
    se = lambda x: x.std(0, ddof=1) / np.sqrt(x.shape[0]) if x.shape[0] > 1 else np.full(x.shape[1:], np.nan)
    half = a.shape[0] // 2
    h1, h2 = a[:half].mean(0), a[half:].mean(0)
    drift = (h2 - h1).mean() / (se(h2 - h1) if R > 1 else np.nan)
    return {"area": float(a.mean()), "area_se": float(se(a.mean(0))), "d_area": g_a.mean(0), "d_area_se": se(g_a),
            "U": float(U.mean()), "d_U": g_U.mean(0), "d_U_se": se(g_U), "per_replica_area": a.mean(0),
            "per_replica_d_area": g_a, "drift": float(drift)}
