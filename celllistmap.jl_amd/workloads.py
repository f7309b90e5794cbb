"""Synthetic workloads (numpy PCG64, fixed seeds) for tests and benchmarks.

uniform_cube        N uniform points in a periodic cube at a given density
triclinic_1m        N uniform points in the cell [[100,20,10],[0,100,15],[0,0,100]]*(N/1e5)^(1/3)
                    (columns are lattice vectors), density 0.1
batch_uneven        K orthorhombic cubes, N_k = floor(exp(U[ln 500, ln 50000])), density 0.1,
                    cutoffs U[8, 14]; a draw with L_k <= 2*rc_k is redrawn (the reference
                    rejects such a cell, src/internals/Box.jl:597)
typed               N points at density 0.1 with 4 types, eps_t ~ U[0.1,0.5], sigma_t ~ U[2.5,3.5]
clustered           Gaussian clusters in a periodic cube with Gaussian velocities
"""
from __future__ import annotations

import numpy as np

TRICLINIC_CELL = np.array([[100.0, 20.0, 10.0], [0.0, 100.0, 15.0], [0.0, 0.0, 100.0]])


def uniform_cube(n, density=0.1, seed=1, dtype=np.float64):
    rng = np.random.default_rng(seed)
    L = (n / density) ** (1.0 / 3.0)
    return (rng.random((n, 3)) * L).astype(dtype), np.array([L, L, L], dtype=np.float64)


def triclinic_1m(n=1_000_000, seed=2, dtype=np.float32):
    rng = np.random.default_rng(seed)
    M = TRICLINIC_CELL * (n / 1e5) ** (1.0 / 3.0)
    x = (M @ rng.random((n, 3)).T).T
    return x.astype(dtype), M.astype(dtype)


def batch_uneven(k=512, nmin=500, nmax=50_000, density=0.1, rc=(8.0, 14.0), seed=3, dtype=np.float32):
    rng = np.random.default_rng(seed)
    ns, cuts, sides, xs = [], [], [], []
    redraws = 0
    for _ in range(k):
        while True:
            n = int(np.floor(np.exp(rng.uniform(np.log(nmin), np.log(nmax)))))
            c = float(rng.uniform(rc[0], rc[1]))
            L = (n / density) ** (1.0 / 3.0)
            if L > 2 * c:
                break
            redraws += 1
        ns.append(n)
        cuts.append(c)
        sides.append(L)
    for n, L in zip(ns, sides):
        xs.append((rng.random((n, 3)) * L).astype(dtype))
    return dict(positions=xs, sides=[np.array([s, s, s]) for s in sides], cutoffs=np.array(cuts), counts=np.array(ns),
                redraws=redraws)


def typed(n=200_000, ntypes=4, density=0.1, seed=4, dtype=np.float32):
    rng = np.random.default_rng(seed)
    L = (n / density) ** (1.0 / 3.0)
    x = (rng.random((n, 3)) * L).astype(dtype)
    types = rng.integers(0, ntypes, n).astype(np.int32)
    eps = rng.uniform(0.1, 0.5, ntypes)
    sig = rng.uniform(2.5, 3.5, ntypes)
    return x, np.array([L, L, L]), types, eps, sig


def clustered(n=4_000_000, nclusters=2000, side=250.0, vsigma=300.0, seed=5, dtype=np.float32):
    rng = np.random.default_rng(seed)
    centres = rng.random((nclusters, 3)) * side
    sig = rng.uniform(0.5, 3.0, nclusters)
    sizes = rng.poisson(n / nclusters, nclusters)
    diff = n - int(sizes.sum())
    while diff != 0:
        k = int(rng.integers(0, nclusters))
        step = 1 if diff > 0 else -1
        if sizes[k] + step >= 0:
            sizes[k] += step
            diff -= step
    owner = np.repeat(np.arange(nclusters), sizes)
    x = centres[owner] + rng.normal(size=(n, 3)) * sig[owner][:, None]
    x = np.mod(x, side)
    v = rng.normal(scale=vsigma, size=(n, 3))
    return x.astype(dtype), v.astype(dtype), np.array([side, side, side])
