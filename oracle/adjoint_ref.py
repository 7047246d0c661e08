"""FP64 CPU reference gradient of the constant-fitting objective -- test infrastructure.

J(theta) = (1/S) sum_{s=1..S} U(x_s) along the oracle's MD loop with the counter RNG
(md_oracle.c orc_md_objective; MD.cpp:333-511 without box moves), theta = (Umin, Umax of
one type pair, k_bond, k_bend).  `adjoint_gradient` runs the reverse recursion of the
velocity-Verlet + Langevin step on the oracle's FP64 trajectory, with Hessian-vector
products and d(force)/d(theta) taken by central differences of the oracle's FP64 forces
(so it shares no code with the device adjoint).  Only tests import this module.
"""
from __future__ import annotations

import numpy as np

import oracle as O


def pair_rows(fconst, uconst, n_types, a, b, umin, umax):
    """laradjiRevaleeFC / laradjiRevaleePC (potentials/laradjiRevalee.h:84-152) for the
    (a,b), (b,a) rows, keeping each row's r_m and r_c."""
    F = np.array(fconst, np.float64, copy=True)
    U = np.array(uconst, np.float64, copy=True)
    for row in {6 * (a * n_types + b), 6 * (b * n_types + a)}:
        rm, rc = F[row], F[row + 3]
        w = rc - rm
        F[row + 1:row + 6] = [2.0 * (umax - umin) / (rm * rm), 0.0, rc, 6.0 * umin / (w * w), 6.0 * umin / (w ** 3)]
        U[row:row + 6] = [rm, (umax - umin) / (rm * rm), umin, rc, 3.0 * umin / (w * w), 2.0 * umin / (w ** 3)]
    return F, U


def _system(d, theta, pair, pos=None):
    F, U = pair_rows(d["fconst"], d["uconst"], d["n_types"], pair[0], pair[1], theta[0], theta[1])
    cc = np.array(d["chain_const"], np.float64, copy=True)
    cc[:, 1] = theta[2]
    cc[:, 3] = theta[3]
    return O.System(d["pos"] if pos is None else pos, d["type"], d["vel"], d["box"], F, U, d["n_types"],
                    d["chain"], cc, cutoff=d["cutoff"], dt=d["dt"], gamma=d["gamma"], temp=d["temperature"],
                    delta_lxy=0.0, tension=0.0, seed=d["seed"])


def objective(d, theta, pair, nsteps):
    return O.MD(_system(d, theta, pair), rng_mode=1).objective(nsteps)


def adjoint_gradient(d, theta, pair, nsteps, eps_x=1e-6, rel_theta=1e-6):
    theta = np.asarray(theta, np.float64)
    s = _system(d, theta, pair)
    md = O.MD(s, rng_mode=1)
    md.init_forces()
    x0 = s.pos.copy()
    traj = []
    for st in range(1, nsteps + 1):
        md.run(st, st)
        traj.append(s.pos.copy())
    box = np.asarray(d["box"], np.float64)

    def force(x, th=theta):
        q = _system(d, th, pair, np.mod(x, box))
        return q.pair_force() + q.chain_force()

    def energy(x, th):
        q = _system(d, th, pair, x)
        return q.pair_potential() + q.chain_potential()

    def dtheta(fun, x, p):
        h = rel_theta * max(1.0, abs(theta[p]))
        a, b = theta.copy(), theta.copy()
        a[p] += h
        b[p] -= h
        return (fun(x, a) - fun(x, b)) / (2 * h)

    moving = (np.asarray(d["type"]) != 0)[:, None]
    dt, gam = d["dt"], d["gamma"]
    c, obj = 0.5 * dt, 1.0 / nsteps
    X, V, A = np.zeros_like(x0), np.zeros_like(x0), np.zeros_like(x0)
    g = np.zeros(4)
    for st in range(nsteps, 0, -1):
        x = traj[st - 1]
        hv = (force(x + eps_x * A) - force(x - eps_x * A)) / (2 * eps_x)
        Xt = X + hv - obj * force(x)
        for p in range(4):
            g[p] += np.sum(A * dtheta(force, x, p)) + obj * dtheta(energy, x, p)
        vh = V - gam * A + np.where(moving, dt * Xt, 0.0)
        X, V = Xt, vh
        A = np.where(moving, (2 * c if st > 1 else c) * vh, 0.0)
    for p in range(4):
        g[p] += np.sum(A * dtheta(force, x0, p))
    return g
