"""NAMD comparison of the reference (test/applications/namd/ParticleSystem_vs_NAMD.jl):
Lennard-Jones energies of 9,999 neon atoms in orthorhombic, rotated and triclinic cells,
cutoff 10, read from the reference's DCD fixtures (tests/golden/namd/*.dcd)."""
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
EPS = 0.0441795
SIG = 2 * 1.64009
_T = np.array([[70.7107, 0.0, 0.0], [35.3553, 61.2372, 0.0], [35.3553, 20.4124, 57.735]]).T
# (file, unit cell with lattice vectors as columns, energy printed by the reference test)
CASES = [
    ("o1", np.array([50.0, 50.0, 50.0]), 32230.01699504111),
    ("o2", np.array([80.0, 70.0, 50.0]), 1093.7225407797744),
    ("o3", np.array([[50.0, 0.0, 50.0], [50.0, 50.0, 0.0], [0.0, 50.0, 50.0]]), 1724.3195067566828),
    ("o4", _T, 1754.0802503953591),
    ("o5", _T, 1765.1389457850137),
    ("o6", np.array([80.0, 80.0, 80.0]), -158.04751357760088),
    ("t1", np.array([[80.0, 0.0, 30.0], [30.0, 80.0, 0.0], [0.0, 40.0, 80.0]]), -116.53213607052128),
    ("t2", np.array([[50.0, 0.0, 0.0], [50.0, 50.0, 0.0], [0.0, 50.0, 50.0]]), 32096.48839031735),
]


def frame(clm, name):
    t = clm.dcd.read_dcd(os.path.join(HERE, "golden", "namd", f"{name}.dcd"))
    return t.frames[0]


def test_dcd_reader(clm):
    t = clm.dcd.read_dcd(os.path.join(HERE, "golden", "namd", "o1.dcd"))
    assert t.natoms == 9999 and len(t.frames) == 1
    assert t.frames[0].shape == (9999, 3) and t.frames[0].dtype == np.float32
    assert np.array_equal(t.cells[0], [50.0, 0.0, 50.0, 0.0, 0.0, 50.0])
    assert t.titles[0].startswith("REMARKS FILENAME=./o1.dcd")
    assert np.all(np.isfinite(t.frames[0]))


@pytest.mark.parametrize("lcell", [1, 3])
@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_oracle_lj_matches_namd(clm, O, case, lcell):
    name, uc, ref = case
    x = frame(clm, name).astype(np.float64)
    r = O.softcore(O.make_box(uc, 10.0, lcell=lcell), x, EPS, SIG, forces=False, form=1)
    assert r["energy"] == pytest.approx(ref, rel=1e-12)


def test_oracle_lj_closed_form(O):
    """Two atoms: u = eps[(sig/d)^12 - 2(sig/d)^6], F = -du/dd along the pair vector."""
    b = O.make_box([30.0, 30.0, 30.0], 10.0)
    for d in (2.5, 3.28018, 4.7, 9.99):
        x = np.array([[5.0, 5.0, 5.0], [5.0 + d, 5.0, 5.0]])
        r = O.softcore(b, x, EPS, SIG, form=1)
        q = SIG / d
        assert r["energy"] == pytest.approx(EPS * (q ** 12 - 2 * q ** 6), rel=1e-13)
        dudd = EPS * (-12 * q ** 12 + 12 * q ** 6) / d
        assert r["forces"][1, 0] == pytest.approx(-dudd, rel=1e-12, abs=1e-15)
        assert r["forces"][0, 0] == pytest.approx(dudd, rel=1e-12, abs=1e-15)


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_gpu_lj_matches_namd(clm, O, case, dtype):
    name, uc, ref = case
    print(f"[namd] {name} dtype={np.dtype(dtype).name}", flush=True)
    x = frame(clm, name).astype(dtype)
    sys_ = clm.ParticleSystem(xpositions=x, unitcell=uc, cutoff=10.0, output=clm.LennardJones(EPS, SIG))
    r = clm.pairwise_(None, sys_)
    tol = 1e-10 if dtype == np.float64 else 1e-4
    assert abs(r.energy - ref) <= tol * abs(ref)
    o = O.softcore(O.make_box(uc, 10.0), x.astype(np.float64), EPS, SIG, form=1)
    f_err = np.linalg.norm(r.forces - o["forces"]) / np.linalg.norm(o["forces"])
    print(f"     energy rel err {abs(r.energy - ref) / abs(ref):.3e}, forces L2 rel err {f_err:.3e}", flush=True)
    assert f_err < tol
