"""Autotuner for the sweep's cell size (lcell) and launch shape (waves per CU).

Candidates are timed on the actual workload with the library's per-kernel HIP-event
timers; the fastest configuration is cached on disk, keyed by the workload's shape
(dtype, cell kind, particle-count bucket and the number of neighbours per particle,
rho * rc^3), so a later call with a similar workload reuses the choice.  The
reference exposes the same knobs as keyword arguments (lcell, nbatches) chosen by
hand or by a fixed heuristic (src/internals/CellLists.jl:178-273).
"""
from __future__ import annotations

import json
import math
import os
from typing import Iterable, Optional

import numpy as np

CACHE = os.environ.get("CLM_AUTOTUNE_CACHE", os.path.join(os.path.expanduser("~"), ".cache", "celllistmap_amd",
                                                          "autotune.json"))
LCELLS = (1, 2, 3)
WAVES = (0, 8, 16)


def workload_key(kind: str, dtype, n: int, density: float, cutoff: float) -> str:
    nb = int(round(math.log2(max(n, 1))))
    nn = density * cutoff ** 3
    return f"{kind}|{np.dtype(dtype).name}|n2^{nb}|nn{int(round(math.log2(max(nn, 1e-9)) * 2))}"


def _load() -> dict:
    try:
        with open(CACHE) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store(d: dict):
    try:
        os.makedirs(os.path.dirname(CACHE), exist_ok=True)
        with open(CACHE, "w") as f:
            json.dump(d, f, indent=1, sort_keys=True)
    except OSError:
        pass


def _kernel_ms(ctx, name: str) -> float:
    import ctypes as C
    ms, cnt = C.c_double(0.0), C.c_int64(0)
    ctx.lib.clm_profile_get(ctx.ptr, name.encode(), C.byref(ms), C.byref(cnt))
    return ms.value / max(cnt.value, 1)


def tune_batch(make_batch, key: Optional[str] = None, lcells: Iterable[int] = LCELLS, waves: Iterable[int] = WAVES,
               reps: int = 3, use_cache: bool = True) -> dict:
    """Pick (lcell, waves_per_cu) for a batch.  ``make_batch(lcell)`` builds a
    BatchedSystems with that lcell for every system.  Returns the best configuration
    and the timing table (ms of sweep + binning per evaluation)."""
    cache = _load() if use_cache else {}
    if key and key in cache:
        return dict(cache[key], cached=True)
    table = []
    best = None
    for lc in lcells:
        try:
            b = make_batch(lc)
        except Exception as e:  # an lcell may be invalid for tiny boxes
            table.append({"lcell": lc, "error": str(e)})
            continue
        ctx = b._ctx
        for w in waves:
            b.set_tuning(w)
            b.energy_forces()  # warm-up (builds cell list and task tables)
            ctx.lib.clm_profile_reset(ctx.ptr)
            ctx.lib.clm_profile_enable(ctx.ptr, 1)
            for _ in range(reps):
                ctx.lib.clm_update_cell_list(ctx.ptr)
                b.energy_forces()
            ms = sum(_kernel_ms(ctx, k) for k in ("k_sweep_energy_forces", "k_bin_count", "k_bin_scatter"))
            ctx.lib.clm_profile_enable(ctx.ptr, 0)
            row = {"lcell": lc, "waves_per_cu": w, "ms": ms}
            table.append(row)
            if best is None or ms < best["ms"]:
                best = row
        del b
    out = {"lcell": best["lcell"], "waves_per_cu": best["waves_per_cu"], "ms": best["ms"], "table": table}
    if key and use_cache:
        cache[key] = out
        _store(cache)
    return dict(out, cached=False)
