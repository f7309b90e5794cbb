"""The bench configuration at full size: N = 1,000,000 particles (float32, and float64) uniform in the
triclinic cell [[100,20,10],[0,100,15],[0,0,100]] * 10^(1/3) (lattice vectors as
columns), cutoff 12 -- about 3.6e8 pairs, more than the oracle's pair list can hold in
a test.  Checked through properties that hold at any size: the pair count against the
uniform-density expectation, each pair once and as (lower index, higher index), distances
in (0, cutoff], and the pair set and distances bit for bit between the two sweep kernels
(per-lane windows over the real / ghost split, and union rows over the reference's full
stencil with index_i < index_j)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_bench_config_full_size(clm, dtype):
    import torch
    say(f"[fullsize] N=1e6 {np.dtype(dtype).name} triclinic, two sweep kernels")
    N, rc = 1_000_000, 12.0
    rng = np.random.default_rng(2)
    M = np.array([[100.0, 20.0, 10.0], [0.0, 100.0, 15.0], [0.0, 0.0, 100.0]]) * 10.0 ** (1.0 / 3.0)
    x = np.ascontiguousarray((M @ rng.random((3, N))).T.astype(dtype))
    xt = torch.from_numpy(x).to("cuda:0")
    sorted_lists = {}
    for variant in (1, 0):
        nl = clm.InPlaceNeighborList(xpositions=xt, cutoff=rc, unitcell=M.astype(dtype), lcell=2,
                                     device_output=True)
        nl.sys.set_sweep_variant(variant)
        r = clm.neighborlist_(nl)
        i, j = r.i.long(), r.j.long()
        assert bool((i < j).all())  # the reference's orientation for triclinic cells
        key = i * (1 << 32) + j
        o = torch.argsort(key)
        sorted_lists[variant] = (key[o], r.d[o].clone())
        del nl, r, i, j, key, o
        torch.cuda.synchronize()
    k1, d1 = sorted_lists[1]
    k0, d0 = sorted_lists[0]
    P = k1.numel()
    expected = N * (N - 1) / 2 * 4.0 / 3.0 * np.pi * rc ** 3 / abs(np.linalg.det(M))
    say(f"     pairs {P}, expected {expected:.4g}")
    assert abs(P - expected) < 0.005 * P
    assert bool((k1[1:] > k1[:-1]).all())  # each pair once
    assert float(d1.min()) > 0.0 and float(d1.max()) <= rc
    assert torch.equal(k1, k0)
    assert torch.equal(d1, d0)


def test_bench_config_energy_forces_full_size(clm):
    """Soft-core energy and forces of the same full-size system at float32 and float64:
    within 1e-4 of each other (the float32 tolerance), and the float64 forces sum to zero
    (Newton's third law in a periodic box)."""
    import torch
    say("[fullsize] N=1e6 energy/forces, f32 against f64")
    N, rc = 1_000_000, 12.0
    rng = np.random.default_rng(2)
    M = np.array([[100.0, 20.0, 10.0], [0.0, 100.0, 15.0], [0.0, 0.0, 100.0]]) * 10.0 ** (1.0 / 3.0)
    x64 = np.ascontiguousarray((M @ rng.random((3, N))).T)
    out = {}
    for dtype in (np.float64, np.float32):
        xt = torch.from_numpy(x64.astype(dtype)).to("cuda:0")
        sys_ = clm.ParticleSystem(xpositions=xt, unitcell=M.astype(dtype), cutoff=rc, lcell=2)
        r = clm.pairwise_(clm.SoftCoreEnergyForces(0.2, 3.0), sys_, device_output=True)
        out[dtype] = (r.energy, r.forces.double())
        del sys_, xt
    e64, f64 = out[np.float64]
    e32, f32 = out[np.float32]
    say(f"     energy f64 {e64:.10g} f32 {e32:.10g}")
    assert abs(e32 - e64) <= 1e-4 * abs(e64)
    assert float(torch.linalg.norm(f32 - f64)) <= 1e-4 * float(torch.linalg.norm(f64))
    assert float(torch.linalg.norm(f64.sum(0))) <= 1e-8 * float(f64.abs().sum())
