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
