"""GPU parity of the two-set (cross) path against the CPU oracle's cross sweep.

Bars: pair sets (i in x, j in y) and distances bit-exact in the working precision;
soft-core energies/forces within 1e-10 (float64) / 1e-4 (float32) relative, checked
against the decomposition E_xy = E(x u y) - E(x) - E(y) of the one-set oracle.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def oracle_ordered(O, box, x, y, mode=1):
    i, j, d2, d = O.pairs_cross(box, x, y, mode)
    i = i.astype(np.int64)
    j = j.astype(np.int64)
    o = np.lexsort((j, i))
    return i[o], j[o], d[o]


def cases():
    rng = np.random.default_rng(99)
    out = []
    L = np.array([23.0, 27.5, 31.0])
    x = rng.random((1800, 3)) * L
    y = rng.random((2500, 3)) * L
    x[::7] += 47.0
    y[::5] -= 29.0
    out.append(("ortho", L, x, y))
    M = np.array([[30.0, 5.0, 2.0], [0.0, 28.0, 4.0], [0.0, 0.0, 26.0]])
    x = (M @ rng.random((3, 2200))).T
    y = (M @ rng.random((3, 1500))).T
    x[::9] += 61.0
    out.append(("tric", M, x, y))
    L2 = np.array([25.0, 19.0])
    out.append(("ortho2d", L2, rng.random((900, 2)) * L2, rng.random((1300, 2)) * L2))
    M2 = np.array([[24.0, 6.0], [0.0, 21.0]])
    out.append(("tric2d", M2, (M2 @ rng.random((2, 1000))).T, (M2 @ rng.random((2, 800))).T))
    return out


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("lcell", [1, 2])
@pytest.mark.parametrize("idx", [0, 1, 2, 3])
def test_cross_neighborlist_bitexact(clm, O, dtype, lcell, idx):
    name, uc, x, y = cases()[idx]
    say(f"[cross-nl] {name} dtype={np.dtype(dtype).name} lcell={lcell}")
    x = x.astype(dtype)
    y = y.astype(dtype)
    rc = 4.5
    res = clm.neighborlist(xpositions=x, ypositions=y, cutoff=rc, unitcell=uc, lcell=lcell)
    ig, jg, dg = res.ordered()
    box = O.make_box(uc, rc, lcell=lcell, dtype=dtype)
    io, jo, do = oracle_ordered(O, box, x, y, 1)
    assert len(ig) == len(io) > 0
    assert np.array_equal(ig, io) and np.array_equal(jg, jo)
    assert np.array_equal(dg.astype(np.float64), do)
    i0, j0, _ = oracle_ordered(O, box, x, y, 0)  # reference sweep with projection filters
    assert np.array_equal(i0, io) and np.array_equal(j0, jo)


def test_cross_argon_goldens(clm, O, argon):
    """src/API/neighborlist.jl:292-312 (439 pairs, non-periodic) and
    src/API/ParticleSystem.jl:121-139 (sum of d^2 = 21886.196785000004, periodic)."""
    say("[cross-nl] argon non-periodic / periodic")
    x32 = argon.astype(np.float32).astype(np.float64)
    res = clm.neighborlist(xpositions=x32[:50], ypositions=x32[50:], cutoff=8.0)
    assert len(res) == 439
    rows = {(a, b): c for a, b, c in res}
    assert rows[(0, 10)] == 4.08865164675529
    assert rows[(0, 16)] == 5.939772435456664
    assert rows[(0, 29)] == 2.4572288423012236
    assert rows[(45, 47)] == 4.9269093987894745
    assert rows[(45, 0)] == 7.99947286297016
    ig, jg, dg = res.ordered()
    io, jo, do = oracle_ordered(O, O.make_nonperiodic_box_cross(x32[:50], x32[50:], 8.0), x32[:50], x32[50:], 1)
    assert np.array_equal(ig, io) and np.array_equal(jg, jo) and np.array_equal(dg, do)
    res = clm.neighborlist(xpositions=argon[:50], ypositions=argon[50:], cutoff=8.0, unitcell=[21.0, 21.0, 21.0])
    d = np.asarray(res.d)
    assert float(np.sum(d * d)) == pytest.approx(21886.196785000004, rel=1e-12)
    sys_ = clm.ParticleSystem(xpositions=argon[:50], ypositions=argon[50:], unitcell=[21.0, 21.0, 21.0], cutoff=8.0,
                              output=clm.MinimumDistance())
    i, j, dm = clm.pairwise_(None, sys_)
    oi, oj, od = O.mindist_cross(O.make_box([21.0] * 3, 8.0), argon[:50], argon[50:], 1)
    assert (i, j, dm) == (oi, oj, od)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("idx", [0, 1])
def test_cross_histogram_and_mindist(clm, O, dtype, idx):
    name, uc, x, y = cases()[idx]
    say(f"[cross-hist] {name} dtype={np.dtype(dtype).name}")
    x = x.astype(dtype)
    y = y.astype(dtype)
    rc = 5.0
    box = O.make_box(uc, rc, dtype=dtype)
    sys_ = clm.ParticleSystem(xpositions=x, ypositions=y, unitcell=uc, cutoff=rc,
                              output=clm.DistanceHistogram(nbins=25, width=0.2))
    counts = clm.pairwise_(None, sys_)
    _, _, do = oracle_ordered(O, box, x, y, 1)
    b = np.minimum(np.floor(do.astype(dtype) / dtype(0.2)).astype(np.int64), 24)
    assert np.array_equal(counts, np.bincount(b, minlength=25))
    i, j, d = clm.pairwise_(clm.MinimumDistance(), sys_)
    oi, oj, od = O.mindist_cross(box, x, y, 1)
    assert (i, j) == (oi, oj)
    assert d == od


def _softcore_cross_reference(O, uc, rc, x, y, eps, sig):
    b = O.make_box(uc, rc, dtype=np.float64)
    xy = np.concatenate([x, y]).astype(np.float64)
    a = O.softcore(b, xy, eps, sig)
    ax = O.softcore(b, x.astype(np.float64), eps, sig)
    ay = O.softcore(b, y.astype(np.float64), eps, sig)
    e = a["energy"] - ax["energy"] - ay["energy"]
    f = a["forces"].copy()
    f[: len(x)] -= ax["forces"]
    f[len(x):] -= ay["forces"]
    return e, f, np.linalg.norm(a["forces"])


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("idx", [0, 1, 2])
def test_cross_energy_forces(clm, O, dtype, idx):
    name, uc, x, y = cases()[idx]
    say(f"[cross-ef] {name} dtype={np.dtype(dtype).name}")
    x = x.astype(dtype)
    y = y.astype(dtype)
    rc = 5.0
    sys_ = clm.ParticleSystem(xpositions=x, ypositions=y, unitcell=uc, cutoff=rc,
                              output=clm.SoftCoreEnergyForces(0.2, 3.0))
    r = clm.pairwise_(None, sys_)
    e_ref, f_ref, fscale = _softcore_cross_reference(O, uc, rc, x, y, 0.2, 3.0)
    tol = 1e-10 if dtype == np.float64 else 1e-4
    assert r.forces.shape == (len(x) + len(y), len(uc))
    e_err = abs(r.energy - e_ref) / abs(e_ref)
    f_err = np.linalg.norm(r.forces - f_ref) / max(np.linalg.norm(f_ref), 1e-300)
    say(f"     energy rel err {e_err:.3e}, forces L2 rel err {f_err:.3e}")
    assert e_err < tol and f_err < tol
    # Newton's third law across the two sets
    assert np.linalg.norm(r.forces.astype(np.float64).sum(axis=0)) < tol * fscale


def test_cross_update_and_device_io(clm, O):
    import torch
    say("[cross] update!(sys; ypositions) and device tensors")
    _, uc, x, y = cases()[0]
    rc = 4.0
    sys_ = clm.ParticleSystem(xpositions=x, ypositions=y, unitcell=uc, cutoff=rc, output=clm.NeighborListOp())
    r1 = clm.pairwise_(None, sys_)
    io, jo, do = oracle_ordered(O, O.make_box(uc, rc), x, y, 1)
    assert np.array_equal(r1.ordered()[0], io)
    y2 = y[: len(y) // 2] + 0.37
    clm.update_(sys_, ypositions=y2)
    r2 = clm.pairwise_(None, sys_)
    io, jo, do = oracle_ordered(O, O.make_box(uc, rc), x, y2, 1)
    ig, jg, dg = r2.ordered()
    assert np.array_equal(ig, io) and np.array_equal(jg, jo) and np.array_equal(dg, do)
    xt = torch.tensor(x, device="cuda:0")
    yt = torch.tensor(y2, device="cuda:0")
    clm.update_(sys_, xpositions=xt, ypositions=yt)
    r3 = clm.pairwise_(None, sys_, device_output=True)
    assert r3.i.is_cuda
    ig, jg, dg = r3.ordered()
    assert np.array_equal(ig, io) and np.array_equal(jg, jo) and np.array_equal(dg, do)
    with pytest.raises(clm.ArgumentError):
        clm.update_(clm.ParticleSystem(xpositions=x, unitcell=uc, cutoff=rc, output=clm.NeighborListOp()),
                    ypositions=y)


def test_cross_nonperiodic_random(clm, O):
    """Non-periodic two-set lists: the box covers limits(x, y)."""
    rng = np.random.default_rng(5)
    for dtype in (np.float64, np.float32):
        x = (rng.random((1500, 3)) * [40.0, 25.0, 30.0]).astype(dtype)
        y = (rng.random((1100, 3)) * [30.0, 45.0, 20.0] - 7.0).astype(dtype)
        say(f"[cross-nl] non-periodic dtype={np.dtype(dtype).name}")
        res = clm.neighborlist(xpositions=x, ypositions=y, cutoff=3.5)
        ig, jg, dg = res.ordered()
        box = O.make_nonperiodic_box_cross(x, y, 3.5, dtype=dtype)
        io, jo, do = oracle_ordered(O, box, x, y, 1)
        assert np.array_equal(ig, io) and np.array_equal(jg, jo)
        assert np.array_equal(dg.astype(np.float64), do)
