"""Autotuner of the 2D deposit's work split (``s2g_tuning``), per shape class.

The work split decides which kernel deposits a particle (the small-footprint
scatter or the tile gather, by footprint box side) and how the work is cut
into blocks (particles per scatter item, pairs per gather item, image planes
per launch).  It does not change the arithmetic of a pixel weight, so every
setting gives the same maps up to rounding; only the time differs.  The
library's defaults suit large maps; a small map (few particles per CU) or a
map with many planes can run faster with another split.

``tune`` measures candidate settings on the caller's own call by coordinate
descent over ``SPACE`` (one field at a time, the others held at the best so
far, two rounds) and returns the fastest.  ``TuningCache`` keeps the results
in a JSON file keyed by ``shape_class`` (dimensions, image side, particle
count and median smoothing length in pixels on a log2 scale, quantities,
kernel, input type), and ``tuned`` looks a class up, tuning it on a miss.

There is no counterpart in the reference (its deposit is one loop per
worker process, cic_interpolation.jl:171-215).
"""
from __future__ import annotations

import json
import math
import os
import tempfile
import time
from typing import Callable

import numpy as np

from . import _lib

# candidate values per s2g_tuning field (the library accepts wider ranges)
SPACE = {
    "small_side": [18, 24, 28, 34],
    "small_item_min": [32, 64, 128, 256, 512],
    "small_items_per_cu": [4, 8, 16, 32],
    "gather_item_min": [16, 32, 64, 128],
    "gather_items_per_cu": [8, 16, 32, 64],
    "gather_planes": [2, 4],
    "small_planes": [2, 4],
    "gather_item_max": [0, 1024, 4096, 16384],
}


def shape_class(dims: int, npix: int, n: int, nq: int, kernel: str, dtype: str,
                hsml_px=None) -> str:
    """Key of a deposit's shape class.  ``hsml_px``: smoothing lengths in
    pixels (their median is bucketed) or None."""
    lg = lambda v: int(round(math.log2(v))) if v > 0 else 0
    h = ""
    if hsml_px is not None and np.size(hsml_px):
        h = f"_h2e{lg(float(np.median(hsml_px)))}"
    return f"{dims}d_npix{int(npix)}_n2e{lg(n)}_nq{int(nq)}_{kernel}_{dtype}{h}"


def batch_class(maps) -> str:
    """Key of a batch: the sorted shape classes of its maps, hashed."""
    import hashlib
    keys = sorted(shape_class(*m) for m in maps)
    return "batch_%d_%s" % (len(keys), hashlib.sha1("|".join(keys).encode()).hexdigest()[:16])


def _timed(run: Callable[[], object], reps: int) -> float:
    best = float("inf")
    for _ in range(reps):
        t0 = time.perf_counter()
        r = run()
        dt = (time.perf_counter() - t0) * 1e3
        best = min(best, float(r) if isinstance(r, (int, float)) and r > 0 else dt)
    return best


def tune(run: Callable[[], object], ctx, nplanes: int = 2, reps: int = 3, rounds: int = 2,
         space: dict | None = None, start: dict | None = None, log=None):
    """Coordinate descent over ``space`` on ``run``.

    ``run()`` performs one complete call on ``ctx`` (deposit, and reduce if
    the caller wants it timed) and returns when its results are ready; it may
    return its own time in ms, otherwise the wall time is used.  Each
    candidate is run once to warm up, then ``reps`` times; the minimum counts.
    Returns ``(best_tuning, best_ms, trials)`` and leaves ``ctx`` set to the
    best tuning."""
    space = dict(space or SPACE)
    if nplanes <= 2:  # one launch holds every plane whatever the split
        space.pop("gather_planes", None)
        space.pop("small_planes", None)
    best = dict(_lib.DEFAULT_TUNING, **(start or {}))
    trials = []
    seen = {}

    def measure(t):
        key = tuple(sorted(t.items()))
        if key in seen:  # measured in an earlier round
            return seen[key]
        ctx.set_tuning(t)
        run()
        ms = _timed(run, reps)
        trials.append((dict(t), ms))
        seen[key] = ms
        if log:
            log(f"tune {t} -> {ms:.3f} ms")
        return ms

    best_ms = measure(best)
    for _ in range(rounds):
        changed = False
        for field, values in space.items():
            for v in values:
                if v == best[field]:
                    continue
                cand = dict(best, **{field: v})
                ms = measure(cand)
                if ms < best_ms * 0.99:  # a 1 % margin against timing noise
                    best, best_ms, changed = cand, ms, True
        if not changed:
            break
    ctx.set_tuning(best)
    return best, best_ms, trials


class TuningCache:
    """JSON file of tuned settings: ``{class: {"tuning": {...}, "ms": t}}``.
    Path: the argument, else $SPH2GRID_TUNING_CACHE, else
    ~/.cache/sph2grid/tuning.json."""

    def __init__(self, path: str | None = None):
        self.path = path or os.environ.get("SPH2GRID_TUNING_CACHE") or os.path.join(
            os.path.expanduser("~"), ".cache", "sph2grid", "tuning.json")

    def load(self) -> dict:
        try:
            with open(self.path) as f:
                return json.load(f)
        except (FileNotFoundError, json.JSONDecodeError):
            return {}

    def get(self, key: str):
        e = self.load().get(key)
        return dict(e["tuning"]) if e else None

    def put(self, key: str, tuning: dict, ms: float):
        data = self.load()
        data[key] = {"tuning": dict(tuning), "ms": float(ms)}
        d = os.path.dirname(self.path) or "."
        os.makedirs(d, exist_ok=True)
        fd, tmp = tempfile.mkstemp(dir=d, prefix=".tuning.", suffix=".json")
        with os.fdopen(fd, "w") as f:
            json.dump(data, f, indent=1, sort_keys=True)
        os.replace(tmp, self.path)


def tuned(ctx, key: str, run: Callable[[], object], nplanes: int = 2,
          cache: TuningCache | None = None, retune: bool = False, log=None) -> dict:
    """Set ``ctx`` to the cached tuning of ``key``, tuning it with ``run`` on
    a miss (or when ``retune``)."""
    cache = cache or TuningCache()
    t = None if retune else cache.get(key)
    if t is None:
        t, ms, _ = tune(run, ctx, nplanes=nplanes, log=log)
        cache.put(key, t, ms)
    ctx.set_tuning(t)
    return t
