"""Bit-exact checks of the deposit's integer work (s2g_footprints) against
the oracle's restatement: per particle the footprint box of
pix_index_min_max (cic_shared.jl:46-52) after get_xyz's two roundings
(cic_shared.jl:85-100), the fallback flag of calculate_weights
(cic_2D.jl:49-69, cic_3D.jl:55-75, from the reference's own pixel loop), the
class of the work split (small / large closed-form / large explicit /
skipped), and the (tile, particle) pair lists and small-footprint buckets
restated from their definitions.  Image tolerances cannot see a one-column
footprint error at a tie (that column has zero overlap), so these compare
integers for equality."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

KIDS = {"cubic": 0, "quintic": 1, "wendlandc2": 2, "wendlandc4": 3, "wendlandc6": 4,
        "wendlandc8": 5}


def gpu_footprints(s2g, dims, pos, hsml, bq, npix, len2pix, kernel, dtype=np.float32,
                   calc_mean=True, exact_norm=False):
    from sph2grid import _lib
    ctx = s2g.default_context()
    n = hsml.shape[0]
    cv = lambda a: np.ascontiguousarray(a, dtype=dtype)
    x, h, q = cv(pos), cv(hsml), cv(bq)
    ones = cv(np.ones(n))
    nq = 1 if q.ndim == 1 else q.shape[1]
    box = np.zeros((n, 2 * dims), dtype=np.int32)
    cls = np.zeros(n, dtype=np.int32)
    fb = np.zeros(n, dtype=np.int32)
    T = 64 if dims == 2 else 16
    ntiles = ((npix + T - 1) // T) ** dims
    nst = ((npix + 31) // 32) ** 2 if dims == 2 else 1
    toff = np.zeros(ntiles + 1, dtype=np.int64)
    soff = np.zeros(nst + 1, dtype=np.int64)
    sord = np.zeros(max(n, 1), dtype=np.int32)
    counts = np.zeros(6, dtype=np.int64)
    par = _lib.Params(npixels=npix, len2pix=float(len2pix))
    flags = (_lib.S2G_FLAG_CALC_MEAN if calc_mean else 0) | \
        (_lib.S2G_FLAG_EXACT_NORM if exact_norm else 0)
    p = lambda a: a.ctypes.data_as(C.c_void_p)
    dt = _lib.S2G_F32 if dtype == np.float32 else _lib.S2G_F64

    def call(pairs, cap):
        rc = ctx.lib.s2g_footprints(ctx.handle, dims, dt, p(x), p(h), p(ones), p(ones), p(q), nq,
                                    p(ones), n, C.byref(par), KIDS[kernel], flags, p(box),
                                    p(cls), p(fb), p(toff), p(pairs), cap, p(soff), p(sord),
                                    p(counts))
        ctx.check(rc, "s2g_footprints")

    pairs = np.zeros(1, dtype=np.int32)
    call(pairs, 0)
    pairs = np.zeros(max(int(counts[3]), 1), dtype=np.int32)
    call(pairs, pairs.size)
    return dict(box=box, cls=cls, fb=fb, tile_off=toff, pairs=pairs[:counts[3]],
                small_off=soff, small_order=sord[:counts[0]], counts=counts.copy())


def compare(s2g, orc, dims, pos, hsml, bq, npix, len2pix, kernel, tag, dtype=np.float32,
            calc_mean=True, exact_norm=False):
    g = gpu_footprints(s2g, dims, pos, hsml, bq, npix, len2pix, kernel, dtype, calc_mean,
                       exact_norm)
    # the oracle sees the same values the GPU does (FP32 inputs promoted exactly)
    f = lambda a: np.asarray(a, dtype=dtype).astype(np.float64)
    box, cls, fb = orc.footprints(dims, f(pos), f(hsml), f(bq), npix, len2pix, kernel,
                                  calc_mean, exact_norm)
    print(f"[{tag}] n={hsml.shape[0]} classes {np.bincount(cls, minlength=4).tolist()} "
          f"fallback {int(fb.sum())} pairs {int(g['counts'][3])}", flush=True)
    for name, a, b in (("box", g["box"], box), ("cls", g["cls"], cls), ("fb", g["fb"], fb)):
        bad = np.flatnonzero(np.any((a != b).reshape(a.shape[0], -1), axis=1))
        assert bad.size == 0, f"{tag}: {name} differs for particles {bad[:10].tolist()}"
    off, pairs = orc.tile_lists(dims, box, cls, npix)
    np.testing.assert_array_equal(g["tile_off"], off, err_msg=tag)
    np.testing.assert_array_equal(g["pairs"], pairs, err_msg=tag)
    soff, sord = orc.small_buckets(dims, box, cls, npix)
    np.testing.assert_array_equal(g["small_off"], soff, err_msg=tag)
    np.testing.assert_array_equal(g["small_order"], sord, err_msg=tag)
    assert g["counts"][0] == (cls == 1).sum() and g["counts"][1] == (cls >= 2).sum()
    assert g["counts"][2] == (cls == 3).sum()
    return box, cls, fb


def test_cluster_2e17_512(s2g, orc):
    """The cluster particle set (2^17 particles, hsml from sub-pixel to ~50 px,
    many footprints clipped by the border) on 512^2, WendlandC6, FP32 inputs."""
    from sph2grid import workloads
    d = workloads.cluster(n=1 << 17, npix=512)
    par = s2g.mappingParameters(x_lim=d["lims"][0], y_lim=d["lims"][1], z_lim=d["lims"][2],
                                Npixels=512)
    pos, par2 = s2g.center_particles(d["Pos"].copy(), par)
    keep = s2g.filter_particles_in_image(pos, par2)
    _, cls, _ = compare(s2g, orc, 2, pos[:, keep].T, d["HSML"][keep], d["Bin_Quant"][keep], 512,
                        par2.len2pix, "wendlandc6", "cluster", calc_mean=False)
    assert np.all(np.bincount(cls, minlength=4)[1:] > 0)  # T > 0: nothing is skipped


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_lattice_ties(s2g, orc, dtype):
    """Centres on pixel edges and centres and hsml on half-pixel multiples, so
    x - h and x + h fall exactly on integers (floor ties), and offsets of one
    ulp either side."""
    npix, L = 64, 64.0
    r = np.random.default_rng(3)
    n = 4000
    cells = r.integers(2, npix - 2, (n, 3)) + r.choice([0.0, 0.5], (n, 3))
    eps = r.choice([0.0, 1.0, -1.0], (n, 3)) * np.spacing(cells.astype(dtype)).astype(np.float64)
    x = (cells + eps) / L - 0.5  # centred code units (len2pix = 64)
    h = r.choice([0.5, 1.0, 1.5, 2.5, 7.5, 16.0, 16.5, 17.0, 30.5], n) / L
    q = r.uniform(0.5, 2.0, n)
    compare(s2g, orc, 2, x, h, q, npix, L, "cubic", "ties2d", dtype=dtype)
    compare(s2g, orc, 3, x, h, q, npix, L, "wendlandc2", "ties3d", dtype=dtype)


@pytest.mark.parametrize("kernel", ["wendlandc4", "quintic"])
def test_clipped_and_degenerate(s2g, orc, kernel):
    """Footprints cut by every border and corner, particles outside the image
    whose footprints reach in or not, zero quantities without calc_mean,
    hsml <= 0, and the exact-normalisation flag."""
    npix, L = 200, 200.0
    r = np.random.default_rng(5)
    n = 3000
    x = r.uniform(-0.7, 0.7, (n, 3))
    h = np.exp(r.uniform(np.log(0.05 / L), np.log(150 / L), n))
    h[::97] = 0.0
    h[5::101] = -1.0 / L
    q = r.uniform(0.5, 2.0, n)
    q[::7] = 0.0
    for cm in (True, False):
        for ex in (False, True):
            compare(s2g, orc, 2, x, h, q, npix, L, kernel, f"clip cm={cm} exact={ex}",
                    calc_mean=cm, exact_norm=ex)
    compare(s2g, orc, 3, x, h * 0.3, q, 48, 48.0, kernel, "clip3d", calc_mean=False)


def test_cluster_3d(s2g, orc):
    """The cluster set on a 96^3 grid (small, large and clipped footprints in
    3D)."""
    from sph2grid import workloads
    d = workloads.cluster(n=1 << 15, npix=96)
    x = d["Pos"].T - 0.5
    compare(s2g, orc, 3, x, d["HSML"], d["Bin_Quant"], 96, 96.0, "wendlandc6", "cluster3d")


def test_multi_quantity_skip(s2g, orc):
    """nq = 3: a particle is skipped only if every quantity is zero."""
    npix, L = 96, 96.0
    r = np.random.default_rng(8)
    n = 2000
    x = r.uniform(-0.5, 0.5, (n, 3))
    h = np.exp(r.uniform(np.log(0.3 / L), np.log(40 / L), n))
    q = r.uniform(0.5, 2.0, (n, 3))
    q[::3, 0] = 0.0
    q[::5, :] = 0.0
    compare(s2g, orc, 2, x, h, q, npix, L, "wendlandc8", "nq3", calc_mean=False)
