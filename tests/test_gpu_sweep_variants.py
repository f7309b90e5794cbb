"""The two sweep kernels (union rows / per-lane windows) against the oracle on the same
inputs: pair sets and distances bit-exact, energies and forces to the usual tolerance,
histogram counts and minimum distances exact.  Includes the pathological boundary
coordinates and the border-rule cells, where the windowed kernel's bounds are widened."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def canon(res):
    """(key, d) sorted by key, then d: a pair the reference's sweep reports twice (it does
    at float32 for points within rounding of the cell faces, see the faces case) then
    compares as a multiset."""
    a, b, d = res.canonical()
    k = a * (1 << 32) + b
    o = np.lexsort((d.astype(np.float64), k))
    return k[o], d[o]


def oracle_canon(O, box, x):
    i, j, d2, d = O.pairs(box, x, 1)
    k, _ = O.canonical_pairs(i, j)
    k = np.minimum(i, j).astype(np.int64) * (1 << 32) + np.maximum(i, j).astype(np.int64)
    o = np.lexsort((d, k))
    return k[o], d[o]


def boxes():
    rng = np.random.default_rng(4321)
    L = np.array([21.0, 24.5, 27.0])
    yield "ortho", L, rng.random((2500, 3)) * L
    M = np.array([[26.0, 6.0, 3.0], [0.0, 24.0, 5.0], [0.0, 0.0, 23.0]])
    yield "tric", M, (M @ rng.random((3, 2500))).T
    # points on the faces / corners of the unit cell (border rule, p == 1 wrap)
    side = 16.0
    f = rng.integers(0, 5, (1500, 3)) / 4.0
    yield "faces", np.array([side, side, side]), f * side + rng.normal(0, 1e-6, (1500, 3))
    M2 = np.array([[18.0, 4.0], [0.0, 17.0]])
    yield "tric2d", M2, (M2 @ rng.random((2, 900))).T
    yield "nonperiodic", None, rng.random((2000, 3)) * [30.0, 20.0, 25.0]


CASES = list(boxes())


@pytest.mark.parametrize("variant", [0, 1])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("lcell", [1, 2, 3])
@pytest.mark.parametrize("idx", range(len(CASES)))
def test_variant_neighborlist(clm, O, variant, dtype, lcell, idx):
    name, uc, x = CASES[idx]
    say(f"[variant {variant}] nl {name} {np.dtype(dtype).name} lcell={lcell}")
    x = x.astype(dtype)
    rc = 3.7
    sys_ = clm.ParticleSystem(xpositions=x, unitcell=uc, cutoff=rc, lcell=lcell, output=clm.NeighborListOp())
    sys_.set_sweep_variant(variant)
    kg, dg = canon(clm.pairwise_(None, sys_))
    box = O.make_box(uc, rc, lcell=lcell, dtype=dtype) if uc is not None else \
        O.make_nonperiodic_box(x, rc, lcell=lcell, dtype=dtype)
    ko, do = oracle_canon(O, box, x)
    assert np.array_equal(kg, ko)
    assert np.array_equal(dg.astype(np.float64), do)


@pytest.mark.parametrize("variant", [0, 1])
@pytest.mark.parametrize("idx", [0, 1, 2])
def test_variant_operators(clm, O, variant, idx):
    name, uc, x = CASES[idx]
    say(f"[variant {variant}] operators {name}")
    rc = 4.1
    sys_ = clm.ParticleSystem(xpositions=x, unitcell=uc, cutoff=rc, lcell=2)
    sys_.set_sweep_variant(variant)
    box = O.make_box(uc, rc)
    r = clm.pairwise_(clm.SoftCoreEnergyForces(0.2, 3.0), sys_)
    ref = O.softcore(box, x, 0.2, 3.0)
    assert abs(r.energy - ref["energy"]) <= 1e-10 * abs(ref["energy"])
    # The faces case holds pairs 1e-6 apart across the periodic boundary: their pair
    # vector, taken from a different (equally valid) image copy, carries an absolute
    # rounding of ulp(16) ~ 4e-15, i.e. 4e-9 of the vector, so forces are compared there
    # at 1e-8.
    ftol = 1e-8 if name == "faces" else 1e-10
    assert np.linalg.norm(r.forces - ref["forces"]) <= ftol * np.linalg.norm(ref["forces"])
    counts = clm.pairwise_(clm.DistanceHistogram(nbins=20, width=0.2), sys_)
    oc, _ = O.histogram(box, x, 20, 0.2, mode=1)
    assert np.array_equal(counts, oc)
    i, j, d = clm.pairwise_(clm.MinimumDistance(), sys_)
    assert (i, j, d) == O.mindist(box, x, 1)
    g = clm.energy_adjoint(sys_, 0.2, 3.0)
    ga = O.softcore(box, x, 0.2, 3.0, grads=True)
    assert abs(g.deps[0] - ga["deps"][0]) <= 1e-10 * abs(ga["deps"][0])


@pytest.mark.parametrize("variant", [0, 1])
@pytest.mark.parametrize("tric", [False, True])
def test_variant_cross(clm, O, variant, tric):
    say(f"[variant {variant}] cross tric={tric}")
    rng = np.random.default_rng(7 + tric)
    M = np.array([[22.0, 4.0, 2.0], [0.0, 25.0, 3.0], [0.0, 0.0, 21.0]]) if tric else np.array([22.0, 25.0, 21.0])
    Mm = M if tric else np.diag(M)
    x = (Mm @ rng.random((3, 1200))).T
    y = (Mm @ rng.random((3, 1700))).T
    sys_ = clm.ParticleSystem(xpositions=x, ypositions=y, unitcell=M, cutoff=3.9, lcell=2, output=clm.NeighborListOp())
    sys_.set_sweep_variant(variant)
    ig, jg, dg = clm.pairwise_(None, sys_).ordered()
    i, j, d2, d = O.pairs_cross(O.make_box(M, 3.9, lcell=2), x, y, 1)
    o = np.lexsort((j, i))
    assert np.array_equal(ig, i[o]) and np.array_equal(jg, j[o]) and np.array_equal(dg, d[o])
    r = clm.pairwise_(clm.SoftCoreEnergyForces(0.2, 3.0), sys_)
    b = O.make_box(M, 3.9)
    a = O.softcore(b, np.concatenate([x, y]), 0.2, 3.0)
    e = a["energy"] - O.softcore(b, x, 0.2, 3.0)["energy"] - O.softcore(b, y, 0.2, 3.0)["energy"]
    assert abs(r.energy - e) <= 1e-9 * abs(a["energy"])


@pytest.mark.parametrize("variant", [0, 1])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("lcell", [1, 2])
def test_triclinic_orientation(clm, variant, dtype, lcell):
    """Triclinic boxes: the reference skips index_i >= index_j (self.jl:174,
    vicinal_cells.jl:64), so every pair comes out as (lower index, higher index).  The
    windowed sweep finds a pair of two real particles from whichever comes first in the
    cell order and writes it in the same orientation."""
    for name, uc, x in CASES:
        if name not in ("tric", "tric2d"):
            continue
        say(f"[variant {variant}] orientation {name} {np.dtype(dtype).name} lcell={lcell}")
        sys_ = clm.ParticleSystem(xpositions=x.astype(dtype), unitcell=uc, cutoff=3.7, lcell=lcell,
                                  output=clm.NeighborListOp())
        sys_.set_sweep_variant(variant)
        r = clm.pairwise_(None, sys_)
        i = np.asarray(r.i)
        j = np.asarray(r.j)
        assert len(i) > 1000
        assert np.all(i < j)
