"""Pins of the oracle's HEALPix RING geometry and all-sky deposit against
closed forms (Gorski et al. 2005), brute force and invariants.  CPU only.

Pinned: ring structure and pixel counts; pix2ang -> ang2pix round trip over
every pixel; equal pixel areas (Monte Carlo); the disc query against brute
force over all pixels (pixel centres within the radius); healpix_map's
conservation (C7) and the reference's BoundsError behaviour.  Parity
unpinned: Healpix.jl's disc-edge membership at exactly the radius, where
the kernel is ~0.
"""
import numpy as np
import pytest


def _angdist(v, w):
    return np.arccos(np.clip(v @ w, -1.0, 1.0))


@pytest.mark.parametrize("nside", [1, 2, 4, 8, 16, 32])
def test_ring_structure_and_roundtrip(orc, nside):
    npix = orc.hp_npix(nside)
    assert npix == 12 * nside * nside
    zs = []
    for p in range(npix):
        th, ph = orc.hp_pix2ang_ring(nside, p)
        assert orc.hp_ang2pix_ring(nside, th, ph) == p
        zs.append(np.cos(th))
    zs = np.round(np.array(zs), 12)
    rings, counts = np.unique(-zs, return_counts=True)
    assert rings.size == 4 * nside - 1
    # 4i pixels in polar ring i, 4*nside on the equatorial rings
    expect = [4 * min(i, nside, 4 * nside - i) for i in range(1, 4 * nside)]
    assert list(counts) == expect


def test_equal_area_monte_carlo(orc):
    nside = 2
    npix = orc.hp_npix(nside)
    r = np.random.default_rng(1)
    n = 240000
    v = r.standard_normal((n, 3))
    v /= np.linalg.norm(v, axis=1)[:, None]
    hits = np.zeros(npix, dtype=int)
    for k in range(n):
        th, ph = orc.hp_vec2ang(v[k])
        hits[orc.hp_ang2pix_ring(nside, th, ph)] += 1
    mean = n / npix
    assert np.all(np.abs(hits - mean) < 5.5 * np.sqrt(mean)), hits


def test_pix2vec_unit_and_consistent(orc):
    nside = 8
    for p in range(0, orc.hp_npix(nside), 7):
        v = orc.hp_pix2vec_ring(nside, p)
        assert abs(np.linalg.norm(v) - 1) < 1e-14
        th, ph = orc.hp_vec2ang(v)
        assert orc.hp_ang2pix_ring(nside, th, ph) == p


@pytest.mark.parametrize("nside", [8, 16])
def test_query_disc_bruteforce(orc, nside):
    npix = orc.hp_npix(nside)
    centres = np.array([orc.hp_pix2vec_ring(nside, p) for p in range(npix)])
    r = np.random.default_rng(nside)
    cases = [(0.0, 0.3), (np.pi, 0.2), (0.05, 0.7), (np.pi - 0.01, 1.1), (1.2, 2.5)]
    for _ in range(40):
        cases.append((np.arccos(r.uniform(-1, 1)), r.uniform(0.01, 1.5)))
    for th, rad in cases:
        ph = r.uniform(0, 2 * np.pi)
        v = np.array([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)])
        got = set(orc.hp_query_disc_ring(nside, th, ph, rad).tolist())
        d = np.arccos(np.clip(centres @ v, -1, 1))
        inside = set(np.flatnonzero(d <= rad).tolist())
        edge = set(np.flatnonzero(np.abs(d - rad) < 1e-9).tolist())
        assert (got ^ inside) <= edge, (th, ph, rad)
        q = orc.hp_query_disc_ring(nside, th, ph, rad)
        assert len(q) == len(set(q.tolist()))  # no duplicates


def _shell(n, seed, hmin=1e-3, hmax=0.2):
    r = np.random.default_rng(seed)
    rad = np.sqrt(r.uniform(0.05 ** 2, 1.0, n))
    v = r.standard_normal((n, 3))
    v /= np.linalg.norm(v, axis=1)[:, None]
    pos = v * rad[:, None]
    h = np.exp(r.uniform(np.log(hmin), np.log(hmax), n))
    m, rho, w, q = (r.uniform(0.5, 2.0, n) for _ in range(4))
    return pos, h, m, rho, q, w


@pytest.mark.parametrize("kernel", ["cubic", "wendlandc4"])
def test_healpix_conservation_c7(orc, kernel):
    nside = 16
    pos, h, m, rho, q, w = _shell(400, 3)
    ang = np.sqrt(4 * np.pi / orc.hp_npix(nside))
    dx = np.linalg.norm(pos, axis=1)
    mp, wm = orc.healpix_map(pos.copy(), h, m, rho, q, w, Nside=nside, kernel=kernel)
    keep = dx >= h
    expect = (w * m / rho / (ang * dx) ** 2)[keep].sum()
    assert abs(wm.sum() / expect - 1) < 1e-12
    assert abs(mp.sum() / (q * w * m / rho / (ang * dx) ** 2)[keep].sum() - 1) < 1e-12


def test_healpix_bounds_error_semantics(orc):
    pos, h, m, rho, q, w = _shell(50, 4)
    # calc_mean=false works only if nothing is removed and every bin_q > 0
    orc.healpix_map(pos.copy(), h, m, rho, q, w, Nside=8, kernel="cubic", calc_mean=False)
    q2 = q.copy()
    q2[3] = 0.0
    with pytest.raises(orc.OracleBoundsError):
        orc.healpix_map(pos.copy(), h, m, rho, q2, w, Nside=8, kernel="cubic", calc_mean=False)
    with pytest.raises(orc.OracleBoundsError):
        orc.healpix_map(pos.copy(), h, m, rho, q, w, Nside=8, kernel="cubic", calc_mean=False,
                        radius_limits=(0.0, 0.5))


def test_mirror_filter_sort_semantics(s2g, orc):
    """The host mirror raises where the reference raises and applies the
    original-order shell mask to the radially sorted order."""
    pos = np.array([[0.9, 0.0, 0.0], [0.2, 0.0, 0.0], [0.5, 0.0, 0.0], [0.7, 0.0, 0.0]]).T
    one = np.ones(4)
    out = s2g.filter_sort_particles(pos.copy(), one * 0.01, one, one, np.arange(4.0), one,
                                    np.zeros(3), (0.0, 0.6), True)
    # dx = [.9, .2, .5, .7]; descending order = [0, 3, 2, 1]; mask (orig) = [F, T, T, F]
    np.testing.assert_array_equal(out[4], [3.0, 2.0])
    with pytest.raises(s2g.BoundsError):
        s2g.filter_sort_particles(pos.copy(), one, one, one, np.array([1.0, 0, 1, 1]), one,
                                  np.zeros(3), (0.0, np.inf), False)
    with pytest.raises(TypeError):
        s2g.healpix_map(pos.astype(np.float32), one, one, one, one, one, Nside=4,
                        kernel=s2g.Cubic(2))
