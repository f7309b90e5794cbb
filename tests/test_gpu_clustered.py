"""Clustered cross-statistics (the 4M-point Gaussian-cluster workload, reduced to a size
the oracle finishes in seconds): dense cells of hundreds to thousands of entries, a
100-bin distance histogram (counts bit-exact) and the mean radial velocity per bin."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("lcell", [1, 2])
def test_clustered_histogram_velocity(clm, O, lcell):
    # 4M points / 2000 clusters in a 250 box, scaled to 30k points / 15 clusters at the
    # same cluster size and box volume per cluster
    n, k = 30_000, 15
    side = 250.0 * (k / 2000) ** (1.0 / 3.0)
    x, v, sides = clm.workloads.clustered(n=n, nclusters=k, side=side, seed=5)
    print(f"[clustered] n={n} side={side:.1f} lcell={lcell}", flush=True)
    rc, nb, width = 20.0, 100, 0.2
    sys_ = clm.ParticleSystem(xpositions=x, unitcell=sides, cutoff=rc, lcell=lcell)
    vh = clm.pairwise_(clm.PairwiseVelocities(nb, width, v), sys_)
    box = O.make_box(sides, rc, lcell=lcell, dtype=np.float32)
    counts, vsum = O.histogram(box, x, nb, width, v=v, mode=1)
    assert np.array_equal(vh.counts, counts)
    assert np.allclose(vh.vsum, vsum, rtol=1e-3, atol=1e-3 * np.abs(vsum).max())
    assert vh.counts.sum() > 1_000_000
