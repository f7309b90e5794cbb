"""Minimal PDB reader for test fixtures: ATOM/HETATM coordinates (Angstrom) and the
CRYST1 cell, converted to a unit-cell matrix whose columns are the lattice vectors
(the convention of Box, src/internals/Box.jl:191-203).

The cell follows the standard crystallographic setting used by PDB and GROMACS: a
along x, b in the xy plane, c completing the right-handed cell from the angles
alpha (b,c), beta (a,c), gamma (a,b)."""
from __future__ import annotations

import math
from typing import Optional, Tuple

import numpy as np

__all__ = ["read_pdb", "cryst1_cell"]


def cryst1_cell(a: float, b: float, c: float, alpha: float, beta: float, gamma: float) -> np.ndarray:
    ca, cb, cg = (math.cos(math.radians(v)) for v in (alpha, beta, gamma))
    sg = math.sin(math.radians(gamma))
    # exact zeros for right angles (cos(90 deg) is ~6e-17 in floating point)
    ca, cb, cg = (0.0 if abs(v) < 1e-12 else v for v in (ca, cb, cg))
    cx = c * cb
    cy = c * (ca - cb * cg) / sg
    cz = math.sqrt(max(c * c - cx * cx - cy * cy, 0.0))
    m = np.array([[a, b * cg, cx], [0.0, b * sg, cy], [0.0, 0.0, cz]])
    return m


def read_pdb(path: str) -> Tuple[np.ndarray, Optional[np.ndarray]]:
    """(coordinates (N, 3) float64, unit cell (3, 3) or None)."""
    xs, cell = [], None
    with open(path) as f:
        for line in f:
            if line.startswith("CRYST1"):
                v = [float(t) for t in line[6:].split()[:6]]
                cell = cryst1_cell(*v)
            elif line.startswith(("ATOM", "HETATM")):
                xs.append([float(line[30:38]), float(line[38:46]), float(line[46:54])])
    return np.array(xs, dtype=np.float64).reshape(-1, 3), cell
