"""The 2D deposit under non-default work splits (s2g_set_tuning): the split
moves particles between the scatter and the gather and recuts the blocks,
so every setting must agree with the FP64 oracle to the deposit's own
tolerances and with the default setting to rounding.  Also the argument
checks of s2g_set_tuning and a small autotuning run on the GPU."""
import numpy as np
import pytest

from conftest import max_rel, rel_l1

pytestmark = pytest.mark.gpu

SETTINGS = [
    dict(small_side=18),
    dict(small_side=5),  # the least: only footprints of at most 5 x 5 pixels scatter
    dict(small_item_min=16, small_items_per_cu=64, gather_item_min=8,
         gather_items_per_cu=256),
    dict(gather_planes=1, small_planes=1),
    dict(gather_planes=2, small_planes=3, small_side=24, small_item_min=4096),
    dict(gather_item_max=0),  # no cap: items from gather_items_per_cu alone
    dict(gather_item_min=8, gather_item_max=16, gather_items_per_cu=4),
]


def _particles(n, npix, nq, seed):
    r = np.random.default_rng(seed)
    x = r.uniform(-0.5, 0.5, (n, 3))
    h = np.exp(r.uniform(np.log(0.3 / npix), np.log(80.0 / npix), n))
    m, rho, w = (r.uniform(0.5, 2.0, n) for _ in range(3))
    q = r.uniform(0.1, 5.0, (n, nq))
    return x, h, m, rho, q, w


@pytest.fixture
def ctx(s2g):
    from sph2grid import _lib
    c = _lib.Context(0)
    yield c
    c.close()


@pytest.mark.parametrize("setting", range(len(SETTINGS)))
def test_settings_match_default_and_oracle(s2g, orc, ctx, setting):
    npix, n, nq = 200, 3000, 3
    x, h, m, rho, q, w = _particles(n, npix, nq, seed=11 + setting)
    kern = s2g.WendlandC6(2)
    base = s2g.deposit(x, h, m, rho, q, w, npix=npix, len2pix=float(npix), kernel=kern,
                       calc_mean=True, context=ctx)
    ctx.set_tuning(SETTINGS[setting])
    st = {}
    got = s2g.deposit(x, h, m, rho, q, w, npix=npix, len2pix=float(npix), kernel=kern,
                      calc_mean=True, context=ctx, stats=st)
    assert ctx.get_tuning() == dict(__import__("sph2grid")._lib.DEFAULT_TUNING,
                                    **SETTINGS[setting])
    ctx.set_tuning(None)
    ref = orc.cic_mapping_2D(x, h, m, rho, q, w, npix, float(npix), "wendlandc6", 2, True)
    print(SETTINGS[setting], {k: st[k] for k in ("n_small", "n_large", "n_pairs")}, flush=True)
    for k in range(nq + 1):
        assert rel_l1(got[:, k], ref[:, k]) <= 1e-5, k
        assert max_rel(got[:, k], ref[:, k]) <= 1e-4, k
        # same arithmetic per pixel weight; class moves change the
        # normalisation's summation only
        assert rel_l1(got[:, k], base[:, k]) <= 1e-7, k
    assert abs(got[:, -1].sum() / ref[:, -1].sum() - 1) <= 1e-6
    if SETTINGS[setting].get("small_side") == 5:
        assert st["n_small"] < 0.5 * n


def test_set_tuning_rejects_out_of_range(s2g, ctx):
    from sph2grid import _lib
    for bad in (dict(small_side=35), dict(small_side=4), dict(small_item_min=8),
                dict(gather_planes=5), dict(small_planes=0), dict(gather_items_per_cu=0),
                dict(gather_item_max=32), dict(gather_item_max=-1)):
        with pytest.raises(_lib.LibraryError, match="out of range"):
            ctx.set_tuning(bad)
    t = _lib.Tuning(**_lib.DEFAULT_TUNING)
    t.gather_item_max = (1 << 30) + 1
    rc = ctx.lib.s2g_set_tuning(ctx.handle, __import__("ctypes").byref(t))
    assert rc == _lib.S2G_EINVAL
    assert ctx.get_tuning() == _lib.DEFAULT_TUNING  # unchanged by the rejected calls


def test_autotune_small_map(s2g, orc, ctx, tmp_path):
    """A tuning run on a small map: the cached setting reproduces the map."""
    from sph2grid import autotune
    npix, n, nq = 128, 4000, 1
    x, h, m, rho, q, w = _particles(n, npix, nq, seed=5)
    kern = s2g.Cubic(2)
    run = lambda: s2g.deposit(x, h, m, rho, q[:, 0], w, npix=npix, len2pix=float(npix),
                              kernel=kern, calc_mean=False, context=ctx)
    cache = autotune.TuningCache(str(tmp_path / "t.json"))
    key = autotune.shape_class(2, npix, n, nq, "cubic", "f64", h * npix)
    t = autotune.tuned(ctx, key, run, nplanes=2, cache=cache)
    assert cache.get(key) == t and ctx.get_tuning() == t
    got = run()
    ctx.set_tuning(None)
    ref = orc.cic_mapping_2D(x, h, m, rho, q[:, 0], w, npix, float(npix), "cubic", 2, False)
    for k in range(2):
        assert rel_l1(got[:, k], ref[:, k]) <= 1e-5
