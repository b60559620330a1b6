"""Autotuner for the zeta-solver layout of one basin.

The device solver has two tuning knobs with no counterpart in the
reference (its ILUT + BiCGStab is sequential):

  block_rows   rows per block-Jacobi ILU(0) block (one workgroup each)
  tri_sweeps   Jacobi sweeps per triangular factor (0 = exact substitution)

All settings solve the same system to the same tolerance; they trade
iterations against the cost of one iteration.  `tune` times a short run of
the real step for each candidate on the basin at hand, keeps the fastest,
and caches the choice per basin signature and precision.
"""
from __future__ import annotations

import hashlib
import itertools
import json
import os
import time

import numpy as np

CACHE_DIR = os.environ.get("SHYHIP_TUNE_CACHE", os.path.join(os.path.expanduser("~"), ".cache", "shyhip"))

DEFAULT_GRID = dict(block_rows=(128, 256, 512), tri_sweeps=(3, 4))


def basin_signature(basin) -> str:
    h = hashlib.sha1()
    for a in (basin.nen3v, basin.ilhv, basin.hm3v):
        h.update(np.ascontiguousarray(a).tobytes())
    return f"{basin.nkn}_{basin.nel}_{basin.nlv}_{h.hexdigest()[:16]}"


def _candidates(grid):
    keys = list(grid)
    for vals in itertools.product(*(grid[k] for k in keys)):
        yield dict(zip(keys, vals))


def tune(basin, precision: str, params: dict, forcing, steps: int = 30, warmup: int = 5, grid=None,
         cache: bool = True, verbose: bool = False, **model_kw) -> dict:
    """Pick block_rows / tri_sweeps for `basin`.

    forcing(model, n) applies the forcing of step n (boundary level, wind)
    and advances the model one step.  Returns the chosen settings and the
    measured ms/step of every candidate."""
    import shyhip
    grid = grid or DEFAULT_GRID
    key = f"{basin_signature(basin)}_{precision}"
    path = os.path.join(CACHE_DIR, "autotune.json")
    db = {}
    if cache and os.path.exists(path):
        try:
            db = json.load(open(path))
        except (OSError, ValueError):
            db = {}
        if key in db:
            return db[key]
    timings = []
    for cand in _candidates(grid):
        if cand["block_rows"] > basin.nkn:
            continue
        m = shyhip.Model(basin, precision, params=dict(params, **cand), **model_kw)
        try:
            for n in range(warmup):
                forcing(m, n)
            m.sync()
            t0 = time.perf_counter()
            for n in range(warmup, warmup + steps):
                forcing(m, n)
            m.sync()
            ms = 1e3 * (time.perf_counter() - t0) / steps
        except shyhip.ShyhipError:
            ms = float("inf")
        finally:
            m.close()
        timings.append(dict(cand, ms=ms))
        if verbose:
            print("autotune", cand, f"{ms:.3f} ms/step", flush=True)
    best = min(timings, key=lambda t: t["ms"])
    out = dict(choice={k: best[k] for k in grid}, ms=best["ms"], timings=timings)
    if cache:
        os.makedirs(CACHE_DIR, exist_ok=True)
        db[key] = out
        with open(path, "w") as f:
            json.dump(db, f, indent=1)
    return out
