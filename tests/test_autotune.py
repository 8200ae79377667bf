"""Host logic of the work-split autotuner (CPU only): the search finds the
optimum of a known cost function, the cache round-trips, shape classes
bucket as documented, and the Python defaults and field order match the C
header and library source."""
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


class FakeCtx:
    def __init__(self):
        self.t = None
        self.calls = 0

    def set_tuning(self, t):
        self.t = dict(t)


def _cost(ctx):
    """A separable cost with its minimum at a non-default point."""
    t = ctx.t
    ctx.calls += 1
    return (10.0 + abs(t["small_side"] - 24) * 0.5 + abs(t["small_item_min"] - 64) / 64.0
            + (0.0 if t["gather_planes"] == 2 else 3.0) + (t["gather_item_min"] != 64) * 1.0)


def test_coordinate_descent_finds_the_optimum():
    from sph2grid import autotune, _lib
    ctx = FakeCtx()
    best, ms, trials = autotune.tune(lambda: _cost(ctx), ctx, nplanes=4, reps=1)
    assert best["small_side"] == 24 and best["small_item_min"] == 64
    assert best["gather_planes"] == 2 and best["gather_item_min"] == 64
    assert ms == pytest.approx(10.0)
    assert ctx.t == best  # the context is left at the best setting
    assert set(best) == set(_lib.DEFAULT_TUNING)
    # a candidate equal to the current best is never re-measured
    seen = [tuple(sorted(t.items())) for t, _ in trials]
    assert len(seen) == len(set(seen))


def test_planes_fields_skipped_for_two_planes():
    from sph2grid import autotune
    ctx = FakeCtx()
    _, _, trials = autotune.tune(lambda: _cost(ctx), ctx, nplanes=2, reps=1)
    assert all(t["gather_planes"] == 4 and t["small_planes"] == 4 for t, _ in trials)


def test_cache_round_trip_and_tuned(tmp_path):
    from sph2grid import autotune
    cache = autotune.TuningCache(str(tmp_path / "sub" / "tuning.json"))
    assert cache.get("k") is None
    ctx = FakeCtx()
    t = autotune.tuned(ctx, "k", lambda: _cost(ctx), nplanes=4, cache=cache)
    n1 = ctx.calls
    assert cache.get("k") == t and ctx.t == t
    # a hit does not run the call again
    ctx2 = FakeCtx()
    assert autotune.tuned(ctx2, "k", lambda: _cost(ctx2), cache=cache) == t
    assert ctx2.calls == 0 and ctx2.t == t and n1 > 10
    cache.put("other", {"small_side": 18}, 1.5)
    assert set(cache.load()) == {"k", "other"}


def test_shape_class_buckets():
    import numpy as np
    from sph2grid import autotune
    k = autotune.shape_class(2, 256, 65536, 1, "cubic", "f64", np.full(10, 12.0))
    assert k == "2d_npix256_n2e16_nq1_cubic_f64_h2e4"
    assert autotune.shape_class(2, 256, 70000, 1, "cubic", "f64") == "2d_npix256_n2e16_nq1_cubic_f64"
    b1 = autotune.batch_class([(2, 128, 1000, 3, "wendlandc6", "f32"),
                               (2, 256, 5000, 3, "wendlandc6", "f32")])
    b2 = autotune.batch_class([(2, 256, 5000, 3, "wendlandc6", "f32"),
                               (2, 128, 1000, 3, "wendlandc6", "f32")])
    assert b1 == b2 and b1.startswith("batch_2_")


def test_defaults_and_layout_match_the_library():
    from sph2grid import _lib
    hdr = open(os.path.join(ROOT, "include", "sph2grid.h")).read()
    body = re.search(r"typedef struct s2g_tuning \{(.*?)\} s2g_tuning;", hdr, re.S).group(1)
    fields = re.findall(r"int32_t (\w+);", body)
    assert fields == [f for f, _ in _lib.Tuning._fields_]
    src = open(os.path.join(ROOT, "sphtogrid.jl_amd", "csrc", "s2g_api.hip")).read()
    m = re.search(r"kDefaultTuning = \{kSmallSide2D, ([\d, ]+)\};", src)
    vals = [34] + [int(v) for v in m.group(1).split(",")]
    assert dict(zip(fields, vals)) == _lib.DEFAULT_TUNING
