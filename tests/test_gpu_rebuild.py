"""Cell-list rebuilds on an existing context: the first build bins in two passes (count,
scatter) and sizes the cell buckets; later builds bin in one pass into the buckets, and
a cell that outgrows its bucket makes the library repeat the pass with larger buckets.
Every build is checked against the oracle (pair sets and distances bit-exact)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def check(clm, O, sys_, box, x):
    a, b, d = clm.pairwise_(None, sys_).canonical()
    i, j, d2, do = O.pairs(box, x, 1)
    k, order = O.canonical_pairs(i, j)
    assert np.array_equal(a * (1 << 32) + b, k)
    assert np.array_equal(d.astype(np.float64), do[order])


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("tric", [False, True])
def test_rebuilds_match_oracle(clm, O, dtype, tric):
    say(f"[rebuild] dtype={np.dtype(dtype).name} tric={tric}")
    rng = np.random.default_rng(21 + tric)
    M = np.array([[24.0, 5.0, 3.0], [0.0, 22.0, 4.0], [0.0, 0.0, 23.0]]) if tric else np.diag([24.0, 22.0, 23.0])
    uc = M if tric else np.diag(M)
    x = (M @ rng.random((3, 3000))).T.astype(dtype)
    rc = 3.5
    sys_ = clm.ParticleSystem(xpositions=x, unitcell=uc, cutoff=rc, lcell=2, output=clm.NeighborListOp())
    box = O.make_box(uc, rc, lcell=2, dtype=dtype)
    check(clm, O, sys_, box, x)                       # two-pass build
    for step in range(3):                             # single-pass builds
        x = (x + rng.normal(0, 0.3, x.shape)).astype(dtype)
        clm.update_(sys_, xpositions=x)
        check(clm, O, sys_, box, x)
    # a dense clump: one cell far above the sized buckets -> the pass is repeated
    x2 = x.copy()
    x2[:400] = (x2[0] + rng.normal(0, 0.05, (400, 3))).astype(dtype)
    clm.update_(sys_, xpositions=x2)
    check(clm, O, sys_, box, x2)
    clm.update_(sys_, xpositions=x)
    check(clm, O, sys_, box, x)


def test_rebuild_energy_forces_batch(clm, O):
    """Batched systems rebuilt in one pass give the oracle's energies."""
    say("[rebuild] batch energies over updates")
    rng = np.random.default_rng(3)
    xs = [rng.random((n, 3)) * 20.0 for n in (800, 1500, 600)]
    b = clm.BatchedSystems(xs, [[20.0] * 3] * 3, [4.0, 5.0, 3.0], lcell=2)
    for step in range(3):
        e, f = b.energy_forces()
        for k in range(3):
            ref = O.softcore(O.make_box([20.0] * 3, [4.0, 5.0, 3.0][k]), xs[k], 0.2, 3.0)
            assert abs(e[k] - ref["energy"]) <= 1e-10 * abs(ref["energy"])
        xs = [x + rng.normal(0, 0.2, x.shape) for x in xs]
        b.set_positions(np.concatenate(xs))


@pytest.mark.parametrize("tric", [False, True])
def test_rebuild_past_entry_capacity(clm, O, tric):
    """Later builds bin in one pass without reading their status back; one that outgrows
    the entry capacity sized by the earlier builds (every particle on an edge of the box
    has three images) is found at the operator's read-back and rebuilt.  The first
    operator after each such update -- histogram, energy, neighbour list -- matches the
    oracle."""
    say(f"[rebuild] past the entry capacity, tric={tric}")
    dtype = np.float32
    rng = np.random.default_rng(5 + tric)
    M = np.array([[23.0, 4.0, 2.0], [0.0, 23.0, 3.0], [0.0, 0.0, 23.0]]) if tric else np.diag([23.0] * 3)
    uc = M if tric else np.diag(M)
    rc = 3.5
    box = O.make_box(uc, rc, lcell=2, dtype=dtype)
    x = (M @ rng.random((3, 2000))).T.astype(dtype)
    sys_ = clm.ParticleSystem(xpositions=x, unitcell=uc, cutoff=rc, lcell=2, output=clm.NeighborListOp())
    check(clm, O, sys_, box, x)                        # two-pass build
    x = (x + rng.normal(0, 0.2, x.shape)).astype(dtype)
    clm.update_(sys_, xpositions=x)
    check(clm, O, sys_, box, x)                        # one pass, status read at the end
    n1 = sys_.cell_list_info()["n_particles"]

    def edge():
        f = rng.random((2000, 3))
        f[:, 1] *= 0.15
        f[:, 2] *= 0.15
        return (M @ f.T).T.astype(dtype)

    x2 = edge()
    clm.update_(sys_, xpositions=x2)
    counts = clm.pairwise_(clm.DistanceHistogram(20, 0.2), sys_)
    ref, _ = O.histogram(box, x2, 20, 0.2, mode=1)
    assert np.array_equal(counts, ref)
    assert sys_.cell_list_info()["n_particles"] > 1.3 * n1  # the capacity was outgrown

    x3 = edge()
    clm.update_(sys_, xpositions=x3)
    e = clm.pairwise_(clm.SoftCoreEnergyForces(0.2, 3.0), sys_)
    ref = O.softcore(O.make_box(uc, rc, dtype=np.float64), x3.astype(np.float64), 0.2, 3.0)
    assert abs(e.energy - ref["energy"]) / abs(ref["energy"]) < 1e-4
    assert np.linalg.norm(e.forces - ref["forces"]) / np.linalg.norm(ref["forces"]) < 1e-4

    x4 = edge()
    clm.update_(sys_, xpositions=x4)
    check(clm, O, sys_, box, x4)
    clm.update_(sys_, xpositions=x)
    check(clm, O, sys_, box, x)


@pytest.mark.parametrize("tric", [False, True])
def test_later_build_with_nan(clm, O, tric):
    """A NaN coordinate in a later single-pass build (status read with the operator's
    results) raises the same ArgumentError as in a first build, and the system recovers
    on the next valid update."""
    say(f"[rebuild] NaN in a later build, tric={tric}")
    rng = np.random.default_rng(17 + tric)
    M = np.array([[20.0, 3.0, 2.0], [0.0, 21.0, 4.0], [0.0, 0.0, 22.0]]) if tric else np.diag([20.0, 21.0, 22.0])
    uc = M if tric else np.diag(M)
    rc = 3.0
    x = (M @ rng.random((3, 1500))).T
    sys_ = clm.ParticleSystem(xpositions=x, unitcell=uc, cutoff=rc, lcell=2, output=clm.NeighborListOp())
    box = O.make_box(uc, rc, lcell=2)
    check(clm, O, sys_, box, x)
    x = x + rng.normal(0, 0.1, x.shape)
    clm.update_(sys_, xpositions=x)
    check(clm, O, sys_, box, x)
    bad = x.copy()
    bad[123, 1] = np.nan
    clm.update_(sys_, xpositions=bad)
    with pytest.raises(clm.ArgumentError):
        clm.pairwise_(None, sys_)
    with pytest.raises(clm.ArgumentError):
        clm.pairwise_(clm.SoftCoreEnergyForces(0.2, 3.0), sys_)
    x = x + rng.normal(0, 0.1, x.shape)
    clm.update_(sys_, xpositions=x)
    check(clm, O, sys_, box, x)
    e = clm.pairwise_(clm.SoftCoreEnergyForces(0.2, 3.0), sys_)
    ref = O.softcore(box, x, 0.2, 3.0)
    assert abs(e.energy - ref["energy"]) <= 1e-10 * abs(ref["energy"])
