"""GROMACS comparison of the reference (test/applications/gromacs/compare_with_gromacs.jl):
Lennard-Jones energies (c6/c12 form, cutoff 8 A, no modifier) of argon in a minimal
two-atom cube, a cube, a rhombic dodecahedron and a truncated octahedron.  GROMACS
printed "LJ (SR)" with 6 significant digits (its log excerpts, tests/golden/gromacs/*_energies.txt); the
reference compares at rtol 1e-5.  The coordinates are the PDB inputs of those GROMACS
runs (the reference reads the first frame of the .xtc, the same coordinates rounded to
the xtc precision of 0.01 A)."""
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
G = os.path.join(HERE, "golden", "gromacs")
C6, C12 = 0.00622127e6, 9.69576e6
# c12/d^12 - c6/d^6 == eps[(sig/d)^12 - 2 (sig/d)^6] with
SIG = (2.0 * C12 / C6) ** (1.0 / 6.0)
EPS = C6 * C6 / (4.0 * C12)
NAMES = ["minimal", "cubic", "dodecahedron", "octahedron"]


def lj_sr(name):
    lines = open(os.path.join(G, f"{name}_energies.txt")).read().splitlines()
    k = next(i for i, ln in enumerate(lines) if "LJ (SR)" in ln)
    return float(lines[k + 1].split()[0])


def system(clm, name):
    path = os.path.join(HERE, "golden", "argon_cubic.pdb") if name == "cubic" else os.path.join(G, f"{name}.pdb")
    return clm.pdb.read_pdb(path)


def test_cryst1_cells(clm):
    _, m = system(clm, "dodecahedron")
    assert np.allclose(m, [[24.364, 0, 12.182], [0, 24.364, 12.182], [0, 0, 17.22794962]], atol=1e-7)
    _, m = system(clm, "minimal")
    assert np.array_equal(m, np.diag([20.0, 20.0, 20.0]))


@pytest.mark.parametrize("lcell", [1, 2, 3])
@pytest.mark.parametrize("name", NAMES)
def test_oracle_lj_matches_gromacs(clm, O, name, lcell):
    x, m = system(clm, name)
    r = O.softcore(O.make_box(m, 8.0, lcell=lcell), x, EPS, SIG, forces=False, form=1)
    # the brute-force closure of the reference test, c12/d2^6 - c6/d2^3, on the same pairs
    i, j, d2, _ = O.pairs(O.make_box(m, 8.0, lcell=lcell), x, 0)
    direct = float(np.sum(C12 / d2 ** 6 - C6 / d2 ** 3))
    assert r["energy"] == pytest.approx(direct, rel=1e-12)
    assert r["energy"] == pytest.approx(lj_sr(name), rel=1e-5)


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("name", NAMES)
def test_gpu_lj_matches_gromacs(clm, O, name, dtype):
    print(f"[gromacs] {name} dtype={np.dtype(dtype).name}", flush=True)
    x, m = system(clm, name)
    x = x.astype(dtype)
    sys_ = clm.ParticleSystem(xpositions=x, unitcell=m, cutoff=8.0, output=clm.LennardJones(EPS, SIG))
    r = clm.pairwise_(None, sys_)
    assert r.energy == pytest.approx(lj_sr(name), rel=1e-5)
    o = O.softcore(O.make_box(m, 8.0), x.astype(np.float64), EPS, SIG, form=1)
    tol = 1e-10 if dtype == np.float64 else 1e-4
    assert abs(r.energy - o["energy"]) <= tol * abs(o["energy"])
    assert np.linalg.norm(r.forces - o["forces"]) <= tol * np.linalg.norm(o["forces"])
