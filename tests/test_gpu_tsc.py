"""GPU TSC deposit against the oracle's textbook TSC (FP64 on both sides; a
cell sums its particles in brick-scan order on the GPU and in input order in
the oracle, hence rtol 1e-12)."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _tsc_gpu(s2g, pos, q, N, periodic, dtype):
    from sph2grid import _lib
    ctx = s2g.default_context()
    x = np.ascontiguousarray(pos, dtype=dtype)
    qq = np.ascontiguousarray(q, dtype=dtype)
    g = np.empty(2 * N ** 3)
    p = _lib.Params(npixels=N, len2pix=float(N))
    rc = ctx.lib.s2g_tsc3d(ctx.handle, _lib.S2G_F32 if dtype == np.float32 else _lib.S2G_F64,
                           x.ctypes.data_as(C.c_void_p), qq.ctypes.data_as(C.c_void_p),
                           x.shape[0], C.byref(p), _lib.S2G_FLAG_PERIODIC if periodic else 0,
                           g.ctypes.data_as(C.c_void_p), None)
    ctx.check(rc, "s2g_tsc3d")
    return g.reshape((N ** 3, 2), order="F")


@pytest.mark.parametrize("periodic", [False, True])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_tsc_matches_oracle(s2g, orc, periodic, dtype):
    N = 32
    r = np.random.default_rng(7)
    pos = r.uniform(-0.55, 0.55, (5000, 3)).astype(dtype)  # some outside: dropped / wrapped
    q = r.uniform(-1, 3, 5000).astype(dtype)
    g = _tsc_gpu(s2g, pos, q, N, periodic, dtype)
    ref = orc.tsc_3d(pos.astype(np.float64), q.astype(np.float64), N, float(N), periodic)
    np.testing.assert_allclose(g, ref, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("periodic", [False, True])
def test_tsc_is_reproducible(s2g, periodic):
    """The gather writes each cell once, summing in a fixed order: two calls
    give the same grid bit for bit (no atomics)."""
    r = np.random.default_rng(5)
    pos = r.normal(0.0, 0.2, (20000, 3))
    q = r.lognormal(0.0, 1.0, 20000)
    a = _tsc_gpu(s2g, pos, q, 48, periodic, np.float64)
    b = _tsc_gpu(s2g, pos, q, 48, periodic, np.float64)
    assert np.array_equal(a, b) and a[:, 1].sum() > 0


@pytest.mark.parametrize("periodic", [False, True])
def test_tsc_odd_grid_and_nonfinite(s2g, orc, periodic):
    # 20^3 is not a whole number of the 8^3 bricks the particles are ordered
    # by; particles with a NaN or infinite coordinate are dropped
    N = 20
    r = np.random.default_rng(11)
    pos = r.uniform(-0.7, 0.7, (4000, 3))
    q = r.uniform(-1, 3, 4000)
    bad = pos.copy()
    bad[:3] = [[np.nan, 0.1, 0.2], [0.1, np.inf, 0.0], [-np.inf, 0.0, 0.3]]
    g = _tsc_gpu(s2g, bad, q, N, periodic, np.float64)
    ref = orc.tsc_3d(np.ascontiguousarray(pos[3:]), q[3:], N, float(N), periodic)
    np.testing.assert_allclose(g, ref, rtol=1e-12, atol=1e-12)


def test_tsc_cell_centre(s2g):
    N = 8
    c = np.array([3, 4, 5])
    pos = ((c + 0.5) / N - 0.5)[None, :]
    g = _tsc_gpu(s2g, pos, np.array([1.0]), N, False, np.float64)[:, 1].reshape(N, N, N)
    assert abs(g[3, 4, 5] - 0.75 ** 3) < 1e-15
    assert abs(g[2, 4, 5] - 0.125 * 0.75 ** 2) < 1e-15
    assert abs(g.sum() - 1.0) < 1e-14


def test_tsc_mapping_average(s2g, orc):
    par = s2g.mappingParameters(x_lim=[0, 1], y_lim=[0, 1], z_lim=[0, 1], Npixels=24)
    r = np.random.default_rng(3)
    pos = r.random((3, 3000))
    q = r.uniform(1, 2, 3000)
    out = s2g.tsc_mapping(pos, q, param=par, dimensions=3, average=True)
    ref = orc.tsc_3d(np.ascontiguousarray((pos - 0.5).T), q, 24, 24.0)
    m = ref[:, 1] > 0
    exp = np.zeros(24 ** 3)
    exp[m] = ref[m, 0] / ref[m, 1]
    np.testing.assert_allclose(out.reshape(-1, order="F"), exp, rtol=1e-12, atol=1e-12)
    img2 = s2g.tsc_mapping(pos, q, param=par, dimensions=2, average=False)
    assert img2.shape == (24, 24)
    np.testing.assert_allclose(img2.sum(), ref[:, 0].sum(), rtol=1e-12)
