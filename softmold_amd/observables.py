"""Area per lipid, and the leaflet split of the liposome generator.

- Bilayer: the box holds two leaflets, so the area per lipid is 2 Lx Ly / N_lipids, from
  the box lines of size_name.dat (`time Lx Ly Lz`, dataExtraction.h:992-1007) or from the
  solver's box.
- Vesicle: 4 pi R^2 / N_outer, with R the mean head distance from the heads' centre of
  mass (analysis/vesicleRadius.cpp:71-175: on the device, Solver.vesicle_radius; over a
  frames file, the vesicle_radius program) and N_outer the outer-leaflet lipid count of the
  liposome generator (liposome_leaflets).
"""
from __future__ import annotations

import math

import numpy as np


def read_size(name: str) -> np.ndarray:
    """Rows (time, Lx, Ly, Lz) of size_name.dat."""
    return np.atleast_2d(np.loadtxt(f"size_{name}.dat"))


def bilayer_area_per_lipid(box, n_lipids: int) -> np.ndarray:
    """2 Lx Ly / N_lipids for one box (3,) or rows of boxes (k, 3)."""
    b = np.asarray(box, dtype=float)
    return 2.0 * b[..., 0] * b[..., 1] / float(n_lipids)


def vesicle_area_per_lipid(radius, n_outer: int) -> np.ndarray:
    """4 pi R^2 / N_outer."""
    r = np.asarray(radius, dtype=float)
    return 4.0 * np.pi * r * r / float(n_outer)


def liposome_leaflets(n_lipids: int, density: float, lipid_length: int = 3, bond_length: float = 0.7):
    """(inner, outer) lipid counts of the liposome generator (models/lipidModels.h:1359-1365):
    R = sqrt((N / density) / (4 pi)), inner = int(N R^2 / (R^2 + (R + (2 L - 1) b)^2))."""
    r = math.sqrt((n_lipids / density) / (4.0 * math.pi))
    inner = int(n_lipids * math.pow(r, 2.0) /
                (math.pow(r, 2.0) + math.pow(r + float(lipid_length * 2 - 1) * bond_length, 2.0)))
    return inner, n_lipids - inner
