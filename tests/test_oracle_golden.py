"""The oracle against the reference's own known answers and against brute force.

Known answers are the reference's doctests / tests (CellListMap.jl):
  src/API/ParticleSystem.jl:103-118   sum(d2), argon, periodic 21^3, rc 8 = 43774.54367600001
  src/API/neighborlist.jl:262-311     857 (non-periodic) / 1143 (periodic) pairs + shown rows
  src/API/parallel_custom.jl:169-193  minimum distance 2.1991993997816563
  docs/src/ParticleSystem/single_set_compound.md:93-110  sum(1/d) = 207.37593043370862, forces
  test/API/test_show.jl:13-29         Box([10,10,10],1): nc 13; 100 points in [0,1)^3 -> 800 entries, 1 cell with reals
  src/internals/Box.jl:163-187,320-326,354-370  nc / cell sizes of documented boxes
  test/API/ParticleSystem.jl:210-223,228-234    non-periodic sides, computing box
"""
import numpy as np
import pytest


def test_argon_sum_d2(O, argon):
    b = O.make_box([21.0, 21.0, 21.0], 8.0)
    for mode in (0, 1, 2):
        r = O.reduce(b, argon, mode)
        assert r["n"] == 1143
        assert r["sum_d2"] == pytest.approx(43774.54367600001, rel=1e-13)
        assert r["sum_invd"] == pytest.approx(207.37593043370862, rel=1e-13)


def test_argon_neighborlist_rows(O, argon):
    """The neighborlist doctest was produced from coordinates read as Float32 (and promoted
    to Float64 by the Float64 box): the printed rows then match bit for bit."""
    x32 = argon.astype(np.float32).astype(np.float64)
    b = O.make_box([21.0, 21.0, 21.0], 8.0)
    i, j, d2, d = O.pairs(b, x32, 0)
    rows = {(a, c): e for a, c, e in zip(i.tolist(), j.tolist(), d.tolist())}
    # first rows shown by the doctest (1-based there): (1,7), (1,20), (1,47); last (68,38), (68,90)
    assert rows[(0, 6)] == 3.3638756414119397
    assert rows[(0, 19)] == 3.163779526466901
    assert rows[(0, 46)] == 6.243868666689442
    key = {(min(a, c), max(a, c)): e for (a, c), e in rows.items()}
    assert key[(37, 67)] == 7.409287768713663
    assert key[(67, 89)] == 7.8758006026725464
    assert len(i) == 1143


def test_argon_nonperiodic(O, argon):
    b = O.make_nonperiodic_box(argon, 8.0)
    info = b.info()
    assert np.allclose(info["input"].diagonal(), [35.63, 35.76, 35.79], atol=1e-2)
    x32 = argon.astype(np.float32).astype(np.float64)
    i, j, d2, d = O.pairs(O.make_nonperiodic_box(x32, 8.0), x32, 0)
    assert len(i) == 857
    rows = {(a, c): e for a, c, e in zip(i.tolist(), j.tolist(), d.tolist())}
    assert rows[(0, 19)] == 3.163779526466901
    assert rows[(0, 60)] == 4.08865164675529
    assert rows[(0, 66)] == 5.939772435456664
    b10 = O.make_nonperiodic_box(argon, 10.0)
    inf = b10.info()
    assert np.allclose(inf["input"].diagonal(), [39.83, 39.96, 39.99], atol=1e-2)
    assert inf["nc"] == [6, 6, 6]
    assert np.allclose(inf["cell_size"], [13.28, 13.32, 13.33], atol=1e-2)


def test_argon_mindist(O, argon):
    b = O.make_box([21.0, 21.0, 21.0], 8.0)
    for mode in (0, 1):  # the cell-list arithmetic reproduces the printed value exactly
        i, j, d = O.mindist(b, argon, mode)
        assert d == 2.1991993997816563
    assert O.mindist(b, argon, 2)[2] == pytest.approx(2.1991993997816563, rel=1e-15)


def test_argon_forces_inverse_distance(O, argon):
    """forces of the 1/d example of single_set_compound.md:63-110, computed from oracle pairs."""
    b = O.make_box([21.0, 21.0, 21.0], 8.0)
    i, j, d2, d = O.pairs(b, argon, 2)
    L = 21.0
    f = np.zeros_like(argon)
    for a, c in zip(i, j):
        v = argon[c] - argon[a]
        v -= L * np.round(v / L)  # minimum image (y - x)
        dd = np.sqrt(v @ v)
        df = (1 / dd**2) * (1 / dd) * v
        f[a] += df
        f[c] -= df
    assert np.allclose(f[0], [0.02649383330735732, 0.18454277989323772, -0.012253902366284958], rtol=1e-10)
    assert np.allclose(f[1], [0.07782602581235692, 0.27910822337402613, 0.21926615329195248], rtol=1e-10)


@pytest.mark.parametrize("cell,ct,nc,cs", [
    ([10, 10, 10], None, [13, 13, 13], [1.0, 1.0, 1.0]),
    ([120, 150, 100], None, [15, 18, 13], [10.0, 10.0, 10.0]),
    (np.array([[100, 50, 0], [0, 120, 0], [0, 0, 130]]), None, [20, 13, 16], [10.0, 10.0, 10.0]),
    (np.array([[100, 0, 0], [0, 120, 0], [0, 0, 150]]), 0, [13, 15, 18], [10.0, 10.0, 10.0]),
])
def test_documented_boxes(O, cell, ct, nc, cs):
    rc = 1.0 if np.asarray(cell).ndim == 1 and cell[0] == 10 else 10.0
    b = O.make_box(cell, rc, celltype=ct)
    assert b.info()["nc"] == nc
    assert np.allclose(b.info()["cell_size"], cs)


def test_show_known_answer(O):
    rng = np.random.default_rng(1)
    x = rng.random((100, 3))
    st = O.cl_stats(O.make_box([10, 10, 10], 1), x)
    assert st["n_particles"] == 800
    assert st["n_cells_with_real_particles"] == 1
    assert st["number_of_cells"] == 13 ** 3


def test_computing_box(O):
    b = O.make_box([1, 1, 1], 0.1)
    assert np.allclose(b.info()["cbmin"], [-0.1] * 3) and np.allclose(b.info()["cbmax"], [1.1] * 3)


def test_errors(O):
    with pytest.raises(O.OracleError):
        O.make_box([10, 10, 10], 5.0)       # L <= 2 rc (Box.jl:597)
    with pytest.raises(O.OracleError):
        O.make_box([10, 10, 10], 1.0, lcell=0)
    x = np.zeros((3, 3)); x[1, 2] = np.nan
    with pytest.raises(O.OracleError):
        O.pairs(O.make_box([10, 10, 10], 1.0), x)


def test_pathological_neighborlists_2d(O):
    """test/API/neighborlists.jl:183-190 cases with a unit cell (2D)."""
    def nl(x, rc, uc):
        b = O.make_box(uc, rc)
        i, j, d2, d = O.pairs(b, np.array(x, float), 0)
        return list(zip(i.tolist(), j.tolist(), d.tolist()))
    assert nl([[0.0, 0.0]], 1.0, [2.0 + np.nextafter(1.0, 2), 2.0 + np.nextafter(1.0, 2)]) == []
    r = nl([[0.0, 0.0], [0.0, 1.0]], 1.0, [2.0 + np.nextafter(1.0, 2)] * 2)
    assert r in ([(0, 1, 1.0)], [(1, 0, 1.0)])
    assert nl([[0.0, 0.0], [0.0, 1.0]], np.nextafter(1.0, 0), [2.0, 2.0]) == []
    nf = np.nextafter(1.0, 2)
    r = nl([[0.0, 0.0], [0.0 + nf, 1.0 + nf]], np.nextafter(1.0, 0), [2.0, 2.0])
    assert [(min(a, c), max(a, c), e) for a, c, e in r] == [(0, 1, 0.9999999999999998)]
    for xx, expected in ((np.nextafter(0.1, 1), 0), (np.nextafter(0.9, 0), 0), (-0.1, 1), (0.1, 1), (0.9, 1)):
        assert len(nl([[0.0, 0.0], [xx, 0.0]], 0.1, [1.0, 1.0])) == expected


def test_pathological_nonperiodic(O):
    def nlnp(x, rc):
        x = np.array(x, float)
        b = O.make_nonperiodic_box(x, rc)
        return O.pairs(b, x, 0)[0].shape[0]
    assert nlnp([[0.0, 0.0, 1.0], [0.0, 0.0, 10.0], [0.0, 0.0, 7.0]], 2.0) == 0
    assert nlnp([[1, 2], [3, 4]], 3.0) == 1
    assert nlnp([[1.0, 1.0, 1.0], [10.0, 1.0, 1.0], [3.0, 1.0, 1.0]], 3.0) == 1


def _pairset(O, b, x, mode):
    i, j, d2, d = O.pairs(b, x, mode)
    k, order = O.canonical_pairs(i, j)
    return k, d2[order]


@pytest.mark.parametrize("lcell", [1, 2, 3])
@pytest.mark.parametrize("kind", ["ortho", "tric"])
def test_modes_agree_random(O, lcell, kind):
    """reference sweep with and without the projection filter, and brute force, give the
    same pair set at float64 (the three are independent code paths)."""
    rng = np.random.default_rng(100 + lcell)
    if kind == "ortho":
        uc = np.array([23.0, 27.5, 31.0])
        x = rng.random((1500, 3)) * uc
    else:
        uc = np.array([[30.0, 5.0, 2.0], [0.0, 28.0, 4.0], [0.0, 0.0, 26.0]])
        x = (uc @ rng.random((3, 1500))).T
    x[::7] -= 40.0  # particles outside the primary cell
    b = O.make_box(uc, 5.0, lcell=lcell)
    k0, d0 = _pairset(O, b, x, 0)
    k1, d1 = _pairset(O, b, x, 1)
    k2, d2 = _pairset(O, b, x, 2)
    assert np.array_equal(k0, k1) and np.array_equal(d0, d1)
    assert np.array_equal(k0, k2)
    assert np.allclose(d0, d2, rtol=1e-12)
    assert len(np.unique(k0)) == len(k0)  # each pair once


def test_exactly_once_scenarios(O):
    """Scenarios of test/internals/tests.jl:490-593 (pairs through periodic images,
    dense clusters at the boundary), counted against brute force."""
    x = np.array([[0.5, 5.0, 5.0], [9.6, 5.0, 5.0]])
    b = O.make_box([10.0, 10.0, 10.0], 2.0)
    assert O.reduce(b, x, 0)["n"] == 1 == O.reduce(b, x, 2)["n"]
    uc = np.array([[10.0, 3.0, 0.0], [0.0, 10.0, 2.0], [0.0, 0.0, 10.0]])
    x = np.array([[0.5, 0.5, 0.5], [2.0, 1.0, 1.0], [9.5, 9.5, 9.5], [8.5, 9.0, 9.0]])
    b = O.make_box(uc, 2.5)
    assert O.reduce(b, x, 0)["n"] == O.reduce(b, x, 2)["n"]
    pts = []
    for i in range(1, 6):
        for j in range(1, 6):
            for k in range(1, 6):
                pts.append([0.3 * i, 0.3 * j, 0.3 * k])
                pts.append([15 - 0.3 * i, 15 - 0.3 * j, 15 - 0.3 * k])
    x = np.array(pts)
    b = O.make_box([15.0, 15.0, 15.0], 2.5)
    assert O.reduce(b, x, 0)["n"] == O.reduce(b, x, 2)["n"]


def test_lcell_corner_margin(O):
    """test/internals/tests.jl:595-684: particles at cell corners, lcell 1,2,3,5."""
    rng = np.random.default_rng(5)
    for lcell in (1, 2, 3, 5):
        sides = np.array([20.0, 20.0, 20.0])
        x = list(rng.random((200, 3)) * sides)
        cs = sides / (2 * lcell + 1)
        for i in range(1, lcell + 2):
            for j in range(1, lcell + 2):
                for k in range(1, lcell + 2):
                    c = np.array([cs[0] * i + 0.1, cs[1] * j + 0.1, cs[2] * k + 0.1])
                    if np.all(c < sides):
                        x.append(c)
        x = np.array(x)
        b = O.make_box(sides, 2.5, lcell=lcell)
        assert O.reduce(b, x, 0)["sum_d"] == pytest.approx(O.reduce(b, x, 2)["sum_d"], rel=1e-10)


def test_random_cells_property(O):
    """check_random_cells (test/modules/Testing.jl:162-240) at reduced trial count."""
    rng = np.random.default_rng(11)
    trials = 0
    while trials < 150:
        dim = int(rng.integers(2, 4))
        tric = bool(rng.integers(0, 2))
        rc = 1 + rng.random()
        lcell = int(rng.integers(1, 4))
        if tric:
            m = -10 + 20 * rng.random((dim, dim))
        else:
            m = np.diag(-10 + 20 * rng.random(dim))
        try:
            b = O.make_box(m if tric else np.diag(m), rc, lcell=lcell)
        except O.OracleError:
            continue
        if np.prod(b.info()["nc"]) > 100000:
            continue
        x = 10 * rng.random((int(rng.integers(10, 21)), dim)) - 50
        r0 = O.reduce(b, x, 0)
        if r0["n"] == 0:
            continue
        r2 = O.reduce(b, x, 2)
        assert r0["n"] == r2["n"]
        assert r0["sum_d2"] == pytest.approx(r2["sum_d2"], rel=1e-10)
        trials += 1


def test_softcore_closed_forms(O):
    """Two particles: energy and force from the closed form; Newton's third law."""
    x = np.array([[1.0, 1.0, 1.0], [3.0, 2.0, 1.5]])
    b = O.make_box([20.0, 20.0, 20.0], 8.0)
    r = O.softcore(b, x, eps=0.2, sigma=3.0)
    d2 = 4.0 + 1.0 + 0.25
    s = 9.0 / (d2 + 1.0)
    assert r["energy"] == pytest.approx(0.2 * (s**6 - s**3), rel=1e-14)
    g = 2 * 0.2 * (6 * s**6 - 3 * s**3) / (d2 + 1.0)
    assert np.allclose(r["forces"][0], g * (x[0] - x[1]), rtol=1e-13)
    assert np.allclose(r["forces"].sum(axis=0), 0.0, atol=1e-12)


def test_softcore_gradient_finite_differences(O):
    """analytic forces and per-type parameter gradients vs central differences (h=1e-3)."""
    rng = np.random.default_rng(3)
    L = 24.0
    x = rng.random((120, 3)) * L
    types = rng.integers(0, 4, 120).astype(np.int32)
    eps = np.array([0.1, 0.25, 0.4, 0.5])
    sig = np.array([2.5, 2.9, 3.2, 3.5])
    b = O.make_box([L] * 3, 6.0)
    r = O.softcore(b, x, eps, sig, types, mode=2, grads=True)
    h = 1e-3
    for p in range(5):
        for k in range(3):
            xp = x.copy(); xp[p, k] += h
            xm = x.copy(); xm[p, k] -= h
            fd = (O.softcore(b, xp, eps, sig, types, mode=2, forces=False)["energy"]
                  - O.softcore(b, xm, eps, sig, types, mode=2, forces=False)["energy"]) / (2 * h)
            assert -r["forces"][p, k] == pytest.approx(fd, rel=1e-5, abs=1e-7)
    def fd_param(name, t, hh):
        arr = eps if name == "deps" else sig
        ap = arr.copy(); ap[t] += hh
        am = arr.copy(); am[t] -= hh
        ep = O.softcore(b, x, ap if name == "deps" else eps, ap if name == "dsigma" else sig, types, mode=2, forces=False)["energy"]
        em = O.softcore(b, x, am if name == "deps" else eps, am if name == "dsigma" else sig, types, mode=2, forces=False)["energy"]
        return (ep - em) / (2 * hh)

    for t in range(4):
        for name in ("deps", "dsigma"):
            # h = 1e-3 central difference carries an O(h^2) truncation error (s ~ sigma^2
            # enters as s^6), so it is checked at 1e-4 and its Richardson extrapolation tightly
            f1 = fd_param(name, t, h)
            f2 = fd_param(name, t, h / 2)
            assert r[name][t] == pytest.approx(f1, rel=1e-4, abs=1e-8)
            assert r[name][t] == pytest.approx((4 * f2 - f1) / 3, rel=1e-7, abs=1e-8)


def test_softcore_gradient_autograd(O):
    """an independent derivation: torch autograd through a brute-force float64 energy."""
    import torch
    rng = np.random.default_rng(9)
    L = 20.0
    n = 80
    x = rng.random((n, 3)) * L
    types = rng.integers(0, 4, n).astype(np.int32)
    eps = np.array([0.1, 0.25, 0.4, 0.5])
    sig = np.array([2.5, 2.9, 3.2, 3.5])
    b = O.make_box([L] * 3, 6.0)
    r = O.softcore(b, x, eps, sig, types, mode=0, grads=True)
    xt = torch.tensor(x, requires_grad=True)
    et = torch.tensor(eps, requires_grad=True)
    st = torch.tensor(sig, requires_grad=True)
    tt = torch.tensor(types, dtype=torch.long)
    dv = xt[:, None, :] - xt[None, :, :]
    dv = dv - L * torch.round(dv / L)
    d2 = (dv ** 2).sum(-1)
    iu = torch.triu_indices(n, n, 1)
    d2 = d2[iu[0], iu[1]]
    ti, tj = tt[iu[0]], tt[iu[1]]
    e = torch.sqrt(et[ti] * et[tj])
    s = ((st[ti] + st[tj]) / 2) ** 2 / (d2 + 1)
    mask = d2 <= 36.0
    E = (e * (s ** 6 - s ** 3))[mask].sum()
    E.backward()
    assert r["energy"] == pytest.approx(E.item(), rel=1e-12)
    assert np.allclose(-r["forces"], xt.grad.numpy(), rtol=1e-9, atol=1e-12)
    assert np.allclose(r["deps"], et.grad.numpy(), rtol=1e-9)
    assert np.allclose(r["dsigma"], st.grad.numpy(), rtol=1e-9)


def test_histogram_oracle_closed_form(O):
    """counts match a direct numpy histogram of brute-force distances; bins 0..nbins-1."""
    rng = np.random.default_rng(4)
    L = 30.0
    x = rng.random((600, 3)) * L
    b = O.make_box([L] * 3, 8.0)
    counts, _ = O.histogram(b, x, 16, 0.5, mode=0)
    i, j, d2, d = O.pairs(b, x, 2)
    ref = np.bincount(np.minimum(np.floor(d / 0.5).astype(int), 15), minlength=16)
    assert np.array_equal(counts, ref)
    v = rng.normal(size=(600, 3))
    c2, vs = O.histogram(b, x, 16, 0.5, v=v, mode=0)
    assert np.array_equal(c2, ref)
    vv = np.zeros(16)
    for a, c, dd in zip(i, j, d):
        r = x[a] - x[c]
        r -= L * np.round(r / L)
        vv[min(int(np.floor(dd / 0.5)), 15)] += np.dot(v[a] - v[c], r) / dd
    assert np.allclose(vs, vv, rtol=1e-9, atol=1e-9)


def test_argon_cross_golden(O, argon):
    """Two-set goldens: src/API/ParticleSystem.jl:121-139 (sum of d2, periodic) and
    src/API/neighborlist.jl:292-312 (non-periodic cross list, 439 pairs)."""
    b = O.make_box([21.0, 21.0, 21.0], 8.0)
    for mode in (0, 1, 2):
        r = O.reduce_cross(b, argon[:50], argon[50:], mode)
        assert r["sum_d2"] == pytest.approx(21886.196785000004, rel=1e-12)
    x32 = argon.astype(np.float32).astype(np.float64)
    bn = O.make_nonperiodic_box_cross(x32[:50], x32[50:], 8.0)
    i, j, d2, d = O.pairs_cross(bn, x32[:50], x32[50:], 0)
    assert len(i) == 439
    rows = {(a, c): e for a, c, e in zip(i.tolist(), j.tolist(), d.tolist())}
    assert len(rows) == 439
    assert rows[(0, 10)] == 4.08865164675529
    assert rows[(0, 16)] == 5.939772435456664
    assert rows[(0, 29)] == 2.4572288423012236
    assert rows[(45, 47)] == 4.9269093987894745
    assert rows[(45, 0)] == 7.99947286297016
    # the cross list of (x, y) is the x-y block of the single-set list of x ++ y
    ia, ja, _, da = O.pairs(O.make_nonperiodic_box(x32, 8.0), x32, 0)
    block = {}
    for a, c, e in zip(ia.tolist(), ja.tolist(), da.tolist()):
        if (a < 50) != (c < 50):
            xa, yc = (a, c - 50) if a < 50 else (c, a - 50)
            block[(xa, yc)] = e
    assert block == rows


def test_cross_random_cells_property(O):
    """Cross-set half of check_random_cells (test/modules/Testing.jl:162-240): the
    cell-list cross sweep against brute force, random boxes/cutoffs/lcell."""
    rng = np.random.default_rng(17)
    trials = 0
    while trials < 150:
        dim = int(rng.integers(2, 4))
        tric = bool(rng.integers(0, 2))
        rc = 1 + rng.random()
        lcell = int(rng.integers(1, 4))
        m = (-10 + 20 * rng.random((dim, dim))) if tric else np.diag(-10 + 20 * rng.random(dim))
        try:
            b = O.make_box(m if tric else np.diag(m), rc, lcell=lcell)
        except O.OracleError:
            continue
        if np.prod(b.info()["nc"]) > 100000:
            continue
        x = 10 * rng.random((int(rng.integers(10, 21)), dim)) - 50
        y = 10 * rng.random((int(rng.integers(10, 21)), dim)) - 50
        r0 = O.reduce_cross(b, x, y, 0)
        if r0["n"] == 0:
            continue
        r1 = O.reduce_cross(b, x, y, 1)
        r2 = O.reduce_cross(b, x, y, 2)
        assert r0["n"] == r1["n"] == r2["n"]
        assert r0["sum_d2"] == pytest.approx(r2["sum_d2"], rel=1e-10)
        i0, j0, _, _ = O.pairs_cross(b, x, y, 0)
        i2, j2, _, _ = O.pairs_cross(b, x, y, 2)
        assert sorted(zip(i0.tolist(), j0.tolist())) == sorted(zip(i2.tolist(), j2.tolist()))
        trials += 1


def test_cross_mindist(O, argon):
    """Cross minimum distance equals the brute-force minimum over x-y pairs."""
    b = O.make_box([21.0, 21.0, 21.0], 8.0)
    i, j, d = O.mindist_cross(b, argon[:50], argon[50:], 0)
    ii, jj, _, dd = O.pairs_cross(b, argon[:50], argon[50:], 0)
    k = int(np.argmin(dd))
    assert (i, j) == (int(ii[k]), int(jj[k]))
    assert d == dd[k]
    assert O.mindist_cross(b, argon[:50], argon[50:], 2)[2] == pytest.approx(d, rel=1e-15)
