"""Pins of the oracle's integer restatement (orc_footprints, tile_lists,
small_buckets) before it is used as the bit-exact checker of the GPU's
integer work.  CPU only.

* footprint box: brute force over every pixel of the axis -- pixel i is in
  the box iff [i, i + 1) meets [x - h, x + h], i.e. i <= x + h < ... and
  i + 1 > x - h (cic_shared.jl:46-52 with floor), clamped to the image;
* fallback flag: brute force -- no pixel centre (i + 1/2, j + 1/2) of the box
  lies strictly closer than h (the kernel weight is positive exactly there);
* classes: hand-picked particles on each side of every threshold;
* pair lists / small buckets: hand-made boxes with known tiles.
"""
import numpy as np
import pytest


def _brute_axis(c, h, npix):
    i = np.arange(npix)
    inside = (i <= c + h) & (i + 1 > c - h)
    idx = np.flatnonzero(inside)
    return (int(idx[0]), int(idx[-1])) if idx.size else None


@pytest.mark.parametrize("dims", [2, 3])
def test_footprint_box_brute_force(orc, dims):
    npix, L = 40, 40.0
    r = np.random.default_rng(17 + dims)
    n = 3000
    x = r.uniform(-0.7, 0.7, (n, 3))
    # a third of the particles on exact pixel edges / centres (floor ties)
    k = n // 3
    x[:k] = (r.integers(-24, 24, (k, 3)) + r.choice([0.0, 0.5], (k, 3))) / L
    h = np.exp(r.uniform(np.log(0.05 / L), np.log(30 / L), n))
    h[: k] = r.choice([0.5, 1.0, 2.0, 3.5], k) / L
    box, cls, _ = orc.footprints(dims, x, h, np.ones(n), npix, L, "cubic")
    for p in range(n):
        axes = []
        for d in range(dims):
            c = x[p, d] * L + 0.5 * npix  # get_xyz's two roundings
            axes.append(_brute_axis(c, h[p] * L, npix))
        if any(a is None for a in axes):
            assert cls[p] == orc.CLS_SKIP and np.all(box[p, 0::2] == 0) and \
                np.all(box[p, 1::2] == -1), p
        else:
            assert cls[p] != orc.CLS_SKIP, p
            assert [tuple(box[p, 2 * d:2 * d + 2]) for d in range(dims)] == axes, p


@pytest.mark.parametrize("L", [40.0, 4096.0 / 3.0])
def test_footprint_box_near_ties(orc, L):
    """Positions and hsml a few ulps either side of values that put x - h or
    x + h on a pixel edge: the box then depends on every rounding of
    get_xyz (x * len2pix, then + Npix / 2, then -/+ h, each in FP64).  The
    brute force evaluates the same expression in numpy's IEEE FP64, so a
    fused multiply-add or a reordered sum in the oracle changes boxes here."""
    npix = 48
    r = np.random.default_rng(int(L))
    n = 4000
    k = r.integers(-20, 20, (n, 3)) + r.choice([0.0, 0.5], (n, 3))
    x = k / L
    h = r.choice([0.5, 1.0, 1.5, 2.0, 3.0], n) / L
    for a, ulps in ((x, r.integers(-3, 4, (n, 3))), (h, r.integers(-3, 4, n))):
        step = np.where(ulps < 0, -np.inf, np.inf)
        for _ in range(3):
            move = np.abs(ulps) > 0
            a[move] = np.nextafter(a[move], step[move])
            ulps = ulps - np.sign(ulps)
    box, cls, _ = orc.footprints(2, x, h, np.ones(n), npix, L, "cubic")
    changed = 0
    for p in range(n):
        axes = []
        for d in range(2):
            c = x[p, d] * L + 0.5 * npix
            axes.append(_brute_axis(c, h[p] * L, npix))
            exact = (k[p, d] + 0.5 * npix)
            changed += np.floor(c - h[p] * L) != np.floor(exact - h[p] * L)
        if any(a is None for a in axes):
            assert cls[p] == orc.CLS_SKIP, p
        else:
            assert [tuple(box[p, 2 * d:2 * d + 2]) for d in range(2)] == axes, p
    assert changed > 100  # the nudges do move boxes across pixel edges


@pytest.mark.parametrize("dims", [2, 3])
def test_fallback_brute_force(orc, dims):
    npix, L = 32, 32.0
    r = np.random.default_rng(23 + dims)
    n = 4000
    x = r.uniform(-0.45, 0.45, (n, 3))
    h = np.exp(r.uniform(np.log(0.05 / L), np.log(3.0 / L), n))
    box, cls, fb = orc.footprints(dims, x, h, np.ones(n), npix, L, "wendlandc6")
    assert 0 < fb.sum() < n
    for p in np.flatnonzero(cls != orc.CLS_SKIP):
        c = x[p, :dims] * L + 0.5 * npix
        ax = [np.arange(box[p, 2 * d], box[p, 2 * d + 1] + 1) + 0.5 for d in range(dims)]
        grid = np.meshgrid(*ax, indexing="ij")
        d2 = sum((g - c[d]) ** 2 for d, g in enumerate(grid))
        assert fb[p] == int(not np.any(np.sqrt(d2) < h[p] * L)), p


def test_class_thresholds(orc):
    """Box side 34 px is small, 35 large (2D); 8 voxels small, 9 large (3D);
    large footprints take the closed form only if unclipped, not falling
    back, hsml >= the kernel's threshold (WendlandC6: 16 px in 2D, 14 voxels
    in 3D) and without the exact flag."""
    npix, L = 256, 256.0
    # centre on a pixel centre: x - h and x + h at +-h around it
    cx = 128.5 / L - 0.5

    def one(dims, hpix, shift=0.0, exact=False, kernel="wendlandc6", q=1.0, cm=True):
        x = np.array([[cx + shift, cx, cx]])
        b, c, f = orc.footprints(dims, x, np.array([hpix / L]), np.array([q]), npix, L, kernel,
                                 calc_mean=cm, exact_norm=exact)
        return int(c[0]), b[0]

    # 2D: side = floor(x + h) - floor(x - h) + 1 = 2*floor(h + 1/2) + 1 for x on a centre
    assert one(2, 16.4)[0] == orc.CLS_SMALL and one(2, 16.4)[1][1] - one(2, 16.4)[1][0] == 32
    c, b = one(2, 16.6)
    assert b[1] - b[0] + 1 == 35 and c == orc.CLS_LARGE_CLOSED
    assert one(2, 16.6, exact=True)[0] == orc.CLS_LARGE_EXPLICIT
    assert one(2, 16.6, kernel="wendlandc4")[0] == orc.CLS_LARGE_EXPLICIT  # below 28 px
    assert one(2, 28.0, kernel="wendlandc4")[0] == orc.CLS_LARGE_CLOSED
    assert one(2, 40.0, shift=-100.0 / L)[0] == orc.CLS_LARGE_EXPLICIT  # clipped at x = 0
    assert one(2, 40.0, shift=100.0 / L)[0] == orc.CLS_LARGE_EXPLICIT  # clipped at x = Npix
    assert one(2, 40.0, shift=80.0 / L)[0] == orc.CLS_LARGE_CLOSED  # 168.5 + 40 < 256
    assert one(2, 5.0, q=0.0, cm=False)[0] == orc.CLS_SKIP
    assert one(2, 5.0, q=0.0, cm=True)[0] == orc.CLS_SMALL
    # 3D: side 8 small, 9 large; threshold 14 voxels for WendlandC6
    assert one(3, 3.4)[0] == orc.CLS_SMALL
    assert one(3, 4.6)[0] == orc.CLS_LARGE_EXPLICIT
    assert one(3, 14.0)[0] == orc.CLS_LARGE_CLOSED


def test_tile_lists_and_buckets_by_hand(orc):
    npix = 130  # tiles of 64: 3 x 3 (the last partial); small tiles of 32: 5 x 5
    box = np.array([[0, 70, 0, 10],      # large: tiles (0,0), (1,0)
                    [60, 65, 120, 129],  # large: tiles (0,1),(0,2),(1,1),(1,2)
                    [33, 40, 100, 110],  # small: bucket (1, 3)
                    [0, -1, 0, -1],      # skipped
                    [129, 129, 0, 0],    # small: bucket (4, 0)
                    [1, 2, 99, 101]],    # small: bucket (0, 3)
                   dtype=np.int32)
    cls = np.array([2, 3, 1, 0, 1, 1], dtype=np.int32)
    off, pairs = orc.tile_lists(2, box, cls, npix)
    lists = [pairs[off[t]:off[t + 1]].tolist() for t in range(9)]
    assert lists == [[0], [1], [1], [0], [1], [1], [], [], []]
    soff, order = orc.small_buckets(2, box, cls, npix)
    assert soff.size == 26 and order.tolist() == [5, 2, 4]
    assert soff[3] == 0 and soff[4] == 1 and soff[8] == 1 and soff[9] == 2 and soff[20] == 2
    assert soff[21] == 3 and soff[25] == 3


def test_tile_list_order_by_hand(orc):
    """The tile lists take the L large particles e_0 .. e_{L-1} as
    e_{(j r) mod L}: L = 5 gives r = round(5 / phi) = 3 (coprime), so the
    order is e_0, e_3, e_1, e_4, e_2."""
    npix = 64
    box = np.array([[0, 40, 0, 40]] * 6, dtype=np.int32)
    cls = np.array([2, 3, 1, 2, 2, 3], dtype=np.int32)  # particle 2 is small
    off, pairs = orc.tile_lists(2, box, cls, npix)
    assert off.tolist() == [0, 5] and pairs.tolist() == [0, 4, 1, 5, 3]


def test_large_stride(orc):
    # nearest integer to L / phi, raised past common factors with L
    assert [orc.large_stride(L) for L in (0, 1, 2, 3, 5, 8, 10, 12)] == [1, 1, 1, 2, 3, 5, 7, 7]
    rng = np.random.default_rng(3)
    for L in list(rng.integers(3, 5000, 40)) + [2 ** 20, 3 * 5 * 7 * 11 * 13]:
        L = int(L)
        r = orc.large_stride(L)
        assert np.gcd(r, L) == 1 and abs(r - L / ((1 + 5 ** 0.5) / 2)) < 40
        assert np.unique(np.arange(L, dtype=np.int64) * r % L).size == L
