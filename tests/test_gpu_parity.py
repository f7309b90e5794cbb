"""GPU parity of the HIP library (through the C ABI) against the CPU oracle.

Bars: neighbour-list pair sets, distances, histogram counts and minimum distances
bit-exact in the working precision (float64 and float32 pipelines); energies,
forces and gradients within 1e-10 (float64) / 1e-4 (float32, vs the float64
oracle) relative, forces as a global L2 relative error.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

F64_TOL = 1e-10
F32_TOL = 1e-4


def say(*a):
    print(*a, flush=True)


def canon_gpu(res):
    a, b, d = res.canonical()
    return a * (1 << 32) + b, d


def canon_oracle(O, box, x, mode=1):
    i, j, d2, d = O.pairs(box, x, mode)
    k, order = O.canonical_pairs(i, j)
    return k, d[order]


def systems():
    rng = np.random.default_rng(1234)
    out = []
    L = np.array([23.0, 27.5, 31.0])
    x = rng.random((3000, 3)) * L
    x[::11] += 55.0
    out.append(("ortho", L, x))
    M = np.array([[30.0, 5.0, 2.0], [0.0, 28.0, 4.0], [0.0, 0.0, 26.0]])
    x = (M @ rng.random((3, 3000))).T
    x[::13] -= 40.0
    out.append(("tric", M, x))
    M2 = np.array([[100.0, 20.0, 10.0], [0.0, 100.0, 15.0], [0.0, 0.0, 100.0]]) * 0.35
    x = (M2 @ rng.random((3, 3500))).T
    out.append(("tric_cfg2shape", M2, x))
    return out


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("lcell", [1, 2, 3])
@pytest.mark.parametrize("idx", [0, 1, 2])
def test_neighborlist_bitexact(clm, O, dtype, lcell, idx):
    name, uc, x = systems()[idx]
    say(f"[nl] {name} dtype={np.dtype(dtype).name} lcell={lcell}")
    x = x.astype(dtype)
    rc = 5.0
    res = clm.neighborlist(xpositions=x, cutoff=rc, unitcell=uc, lcell=lcell)
    kg, dg = canon_gpu(res)
    box = O.make_box(uc, rc, lcell=lcell, dtype=dtype)
    ko, do = canon_oracle(O, box, x, mode=1)
    assert len(kg) == len(ko)
    assert np.array_equal(kg, ko)
    assert np.array_equal(dg.astype(np.float64), do)
    k0, d0 = canon_oracle(O, box, x, mode=0)   # reference sweep with projection filters
    assert np.array_equal(k0, ko)


def test_nonperiodic_and_argon(clm, O, argon):
    say("[nl] argon periodic / non-periodic")
    x32 = argon.astype(np.float32).astype(np.float64)
    res = clm.neighborlist(xpositions=x32, cutoff=8.0, unitcell=[21.0, 21.0, 21.0])
    assert len(res) == 1143
    rows = {(a, b): c for a, b, c in res}
    key = {(min(a, b), max(a, b)): c for (a, b), c in rows.items()}
    assert key[(0, 6)] == 3.3638756414119397
    assert key[(0, 19)] == 3.163779526466901
    assert key[(0, 46)] == 6.243868666689442
    assert key[(37, 67)] == 7.409287768713663
    assert key[(67, 89)] == 7.8758006026725464
    res = clm.neighborlist(xpositions=x32, cutoff=8.0)
    assert len(res) == 857
    kg, dg = canon_gpu(res)
    ko, do = canon_oracle(O, O.make_nonperiodic_box(x32, 8.0), x32, 1)
    assert np.array_equal(kg, ko) and np.array_equal(dg, do)
    sys_ = clm.ParticleSystem(xpositions=argon, unitcell=[21.0, 21.0, 21.0], cutoff=8.0,
                              output=clm.MinimumDistance())
    i, j, d = clm.pairwise_(None, sys_)
    assert d == 2.1991993997816563
    assert (i, j) == O.mindist(O.make_box([21.0] * 3, 8.0), argon, 1)[:2]


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("idx", [0, 1])
def test_energy_forces(clm, O, dtype, idx):
    name, uc, x = systems()[idx]
    say(f"[ef] {name} dtype={np.dtype(dtype).name}")
    x = x.astype(dtype)
    rc = 6.0
    sys_ = clm.ParticleSystem(xpositions=x, unitcell=uc, cutoff=rc, output=clm.SoftCoreEnergyForces(0.2, 3.0))
    r = clm.pairwise_(None, sys_)
    ref = O.softcore(O.make_box(uc, rc, dtype=np.float64), x.astype(np.float64), 0.2, 3.0)
    tol = F64_TOL if dtype == np.float64 else F32_TOL
    e_err = abs(r.energy - ref["energy"]) / abs(ref["energy"])
    f_err = np.linalg.norm(r.forces - ref["forces"]) / np.linalg.norm(ref["forces"])
    say(f"     energy rel err {e_err:.3e}, forces L2 rel err {f_err:.3e}")
    assert e_err < tol and f_err < tol


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_typed_adjoint(clm, O, dtype):
    say(f"[adjoint] dtype={np.dtype(dtype).name}")
    rng = np.random.default_rng(77)
    L = 40.0
    n = 6400
    x = (rng.random((n, 3)) * L).astype(dtype)
    types = rng.integers(0, 4, n).astype(np.int32)
    eps = np.array([0.1, 0.25, 0.4, 0.5])
    sig = np.array([2.5, 2.9, 3.2, 3.5])
    sys_ = clm.ParticleSystem(xpositions=x, unitcell=[L] * 3, cutoff=12.0)
    g = clm.energy_adjoint(sys_, eps, sig, types)
    ref = O.softcore(O.make_box([L] * 3, 12.0), x.astype(np.float64), eps, sig, types, grads=True)
    tol = F64_TOL if dtype == np.float64 else F32_TOL
    assert abs(g.energy - ref["energy"]) / abs(ref["energy"]) < tol
    assert np.linalg.norm(g.dedx + ref["forces"]) / np.linalg.norm(ref["forces"]) < tol
    assert np.allclose(g.deps, ref["deps"], rtol=tol)
    assert np.allclose(g.dsigma, ref["dsigma"], rtol=tol)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("idx", [0, 1])
def test_histogram_and_velocities(clm, O, dtype, idx):
    name, uc, x = systems()[idx]
    say(f"[hist] {name} dtype={np.dtype(dtype).name}")
    x = x.astype(dtype)
    rng = np.random.default_rng(5)
    v = rng.normal(size=x.shape).astype(dtype)
    rc = 6.0
    sys_ = clm.ParticleSystem(xpositions=x, unitcell=uc, cutoff=rc)
    counts = clm.pairwise_(clm.DistanceHistogram(30, 0.2), sys_)
    box = O.make_box(uc, rc, dtype=dtype)
    ref, _ = O.histogram(box, x, 30, 0.2, mode=1)
    assert np.array_equal(counts, ref)
    vh = clm.pairwise_(clm.PairwiseVelocities(30, 0.2, v), sys_)
    rc_, rv = O.histogram(box, x, 30, 0.2, v=v, mode=1)
    assert np.array_equal(vh.counts, rc_)
    tol = 1e-9 if dtype == np.float64 else 1e-3
    assert np.allclose(vh.vsum, rv, rtol=tol, atol=tol * np.abs(rv).max())


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_min_distance(clm, O, dtype):
    for name, uc, x in systems()[:2]:
        say(f"[mind] {name} dtype={np.dtype(dtype).name}")
        x = x.astype(dtype)
        sys_ = clm.ParticleSystem(xpositions=x, unitcell=uc, cutoff=5.0)
        got = clm.pairwise_(clm.MinimumDistance(), sys_)
        want = O.mindist(O.make_box(uc, 5.0, dtype=dtype), x, 1)
        assert got == want


def test_edge_cases(clm):
    say("[edge] empty, single, NaN, capacity")
    L = [10.0, 10.0, 10.0]
    assert len(clm.neighborlist(xpositions=np.zeros((0, 3)), cutoff=1.0, unitcell=L)) == 0
    assert len(clm.neighborlist(xpositions=np.array([[1.0, 1.0, 1.0]]), cutoff=1.0, unitcell=L)) == 0
    x = np.array([[0.5, 5.0, 5.0], [9.6, 5.0, 5.0]])
    r = clm.neighborlist(xpositions=x, cutoff=2.0, unitcell=L)
    assert len(r) == 1 and abs(list(r)[0][2] - 0.9) < 1e-12
    bad = np.array([[1.0, 2.0, np.nan], [0.0, 0.0, 0.0]])
    with pytest.raises(clm.ArgumentError):
        clm.neighborlist(xpositions=bad, cutoff=1.0, unitcell=L)
    with pytest.raises(clm.ArgumentError):
        clm.neighborlist(xpositions=x, cutoff=5.0, unitcell=L)
    rng = np.random.default_rng(0)
    y = rng.random((2000, 3)) * 20
    sys_ = clm.ParticleSystem(xpositions=y, unitcell=[20.0] * 3, cutoff=3.0)
    small = clm.pairwise_(clm.NeighborListOp(capacity=10), sys_)   # grows and retries
    full = clm.pairwise_(clm.NeighborListOp(), sys_)
    assert len(small) == len(full) > 10


def test_inplace_reuse_steps(clm, O):
    """InPlaceNeighborList across 10 update steps with Gaussian displacements (triclinic, f32)."""
    say("[inplace] 10 steps")
    rng = np.random.default_rng(2)
    M = np.array([[100.0, 20.0, 10.0], [0.0, 100.0, 15.0], [0.0, 0.0, 100.0]]) * (2e4 / 1e5) ** (1 / 3)
    x = (M @ rng.random((3, 20000))).T.astype(np.float32)
    nl = clm.InPlaceNeighborList(xpositions=x, cutoff=12.0, unitcell=M.astype(np.float32))
    box = O.make_box(M.astype(np.float32), 12.0, dtype=np.float32)
    for step in range(10):
        if step:
            x = (x + rng.normal(0, 0.05, x.shape)).astype(np.float32)
            clm.update_(nl, xpositions=x)
        res = clm.neighborlist_(nl)
        kg, dg = canon_gpu(res)
        ko, do = canon_oracle(O, box, x, 1)
        assert np.array_equal(kg, ko) and np.array_equal(dg.astype(np.float64), do), step


def test_cfg1_full_size(clm, O):
    """N=20,000, float64, density 0.1, cutoff 12, periodic cube: pair set and energy."""
    say("[cfg1] N=20000 f64")
    rng = np.random.default_rng(1)
    N = 20000
    L = (N / 0.1) ** (1 / 3)
    x = rng.random((N, 3)) * L
    res = clm.neighborlist(xpositions=x, cutoff=12.0, unitcell=[L] * 3)
    kg, dg = canon_gpu(res)
    box = O.make_box([L] * 3, 12.0)
    ko, do = canon_oracle(O, box, x, 1)
    assert np.array_equal(kg, ko) and np.array_equal(dg, do)
    assert abs(len(kg) - N * (N - 1) / 2 * 4 / 3 * np.pi * 12 ** 3 / L ** 3) < 0.01 * len(kg)
    sys_ = clm.ParticleSystem(xpositions=x, unitcell=[L] * 3, cutoff=12.0, output=clm.SoftCoreEnergy())
    r = clm.pairwise_(None, sys_)
    ref = O.softcore(box, x, forces=False)
    assert abs(r.energy - ref["energy"]) / abs(ref["energy"]) < F64_TOL
