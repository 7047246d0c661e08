"""Replica sweeps: many uneven vesicles over a tension x temperature grid, stepped as one
batch (sm_create_batch).

The reference has no batch runner: a sweep there is one `MD` process per `.mpd` file,
each edited with the changeTension-style tools (analysis/changeTension.cpp:14-47).
`replica_sweep` builds the standard sweep in memory:

- lipid counts N_r = round(exp(U[ln lo, ln hi])), with U drawn from numpy's MT19937
  seeded with `seed` (one double per replica, numpy's 53-bit random());
- replica r is the reference liposome generator's vesicle with seed `seed_base + r` and
  N_r lipids at areal density 3.45 (box 400^3);
- tension and temperature on an 8 x 8 grid: tension in linspace(0, 1, 8), temperature in
  linspace(0.9, 1.1, 8) x the generator's temperature (3.0, so +-10 %);
- delta_lxy = 0.01 so the constant-tension box moves run (MD.cpp:589 skips them at 0).
"""
from __future__ import annotations

import numpy as np

from . import Mpd


def sweep_counts(n_rep: int = 64, seed: int = 20261202, lo: int = 2000, hi: int = 40000) -> np.ndarray:
    rng = np.random.Generator(np.random.MT19937(seed))
    u = rng.random(n_rep)
    return np.rint(np.exp(np.log(lo) + u * (np.log(hi) - np.log(lo)))).astype(int)


def replica_sweep(n_rep: int = 64, seed: int = 20261202, seed_base: int = 1000, lo: int = 2000,
                  hi: int = 40000, density: float = 3.45, delta_lxy: float = 0.01) -> list[dict]:
    counts = sweep_counts(n_rep, seed, lo, hi)
    side = int(np.ceil(np.sqrt(n_rep)))
    tensions = np.linspace(0.0, 1.0, side)
    systems = []
    for r, nl in enumerate(counts):
        d = Mpd.liposome(seed_base + r, int(nl), density).system()
        temps = np.linspace(0.9, 1.1, side) * d["temperature"]
        d["tension"] = float(tensions[r % side])
        d["temperature"] = float(temps[(r // side) % side])
        d["delta_lxy"] = delta_lxy
        systems.append(d)
    return systems
