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


# ---------------------------------------------------------------------------
# Per-pixel area weights (pixel_weights.jl:6-8, 34-76, 120-137), pinned by
# the ring geometry of hp_geometry.py (built ring by ring from Gorski et al.
# 2005, not through the oracle's index inversion).
# ---------------------------------------------------------------------------
from hp_geometry import area_closed_form as _area_closed_form  # noqa: E402
from hp_geometry import ring_centres as _ring_centres  # noqa: E402


@pytest.mark.parametrize("nside", [4, 16])
def test_ring_centres_match_oracle(orc, nside):
    """The independent construction and the oracle agree on every centre."""
    v, _ = _ring_centres(nside)
    assert v.shape[0] == orc.hp_npix(nside)
    for p in range(0, v.shape[0], 3):
        np.testing.assert_allclose(orc.hp_pix2vec_ring(nside, p), v[p], atol=2e-15)


@pytest.mark.parametrize("kernel", ["wendlandc4", "cubic"])
def test_pole_centred_area_and_kernel_weights(orc, kernel):
    """A particle on the z axis: its disc is a set of whole polar rings.
    Every pixel's area weight and kernel weight follow from its ring's
    colatitude alone: A = contributing_area(d) / (ang_pix * Dx)^2 with the
    pixel-centre distance d, wk = shape(d / theta).  Pins the sign in
    proj_hsml - (r - ang_pix/2) and the /(ang_pix * Dx)^2 factor."""
    nside, Dx = 32, 0.8
    ang = np.sqrt(4 * np.pi / (12 * nside * nside))
    centres, ring = _ring_centres(nside)
    for theta_pix in (2.3, 4.7, 7.5):
        theta = theta_pix * ang
        h = Dx * np.sin(theta)
        pix, A, wk, info = orc.hp_particle_weights([0.0, 0.0, Dx], h, nside, kernel)
        assert not info["fallback"]
        assert abs(info["proj_hsml"] - theta) < 1e-15 and abs(info["ang_pix"] - ang) < 1e-16
        d = np.arccos(np.clip(centres[pix, 2], -1.0, 1.0))
        # the disc is every pixel of the rings whose colatitude is <= theta
        inside = np.flatnonzero(np.arccos(np.clip(centres[:, 2], -1, 1)) <= theta)
        assert sorted(pix.tolist()) == sorted(inside.tolist())
        s = 1.0 / (ang * Dx) ** 2
        np.testing.assert_allclose(A, s * _area_closed_form(d, theta, ang), rtol=1e-12)
        shape = np.array([orc.kernel_shape(kernel, x) for x in d / theta])
        norm = orc.kernel_norm(kernel, 2) / theta ** 2
        np.testing.assert_allclose(wk, norm * shape, rtol=1e-12)
        # equal weights within each ring
        for r in np.unique(ring[pix]):
            sel = ring[pix] == r
            assert np.ptp(A[sel]) <= 1e-12 * A[sel].max()
            assert np.ptp(wk[sel]) <= 1e-12 * max(wk[sel].max(), 1e-300)
        # the edge rings take the ramp: 1/2 <= A / s < 1 for theta - ang/2 < d <= theta
        edge = d > theta - 0.5 * ang
        assert np.all(A[~edge] == s) and np.all((A[edge] >= 0.5 * s) & (A[edge] < s))


def test_area_weight_values_at_the_ramp(orc):
    """The pixel holding a particle whose disc holds no pixel centre (theta =
    0.12 pixels): it is the particle's only pixel (constributing_pixels.jl:
    16-19), the particle falls back to wk = 1 (pixel_weights.jl:120-137) and
    A / s follows the ramp from 1 at d = theta - ang/2 to 0 at
    d = theta + ang/2, s = 1 / (ang_pix Dx)^2.  A disc whose edge passes
    exactly through a pixel centre gives that pixel A / s = 1/2."""
    nside, Dx = 16, 1.3
    ang = np.sqrt(4 * np.pi / (12 * nside * nside))
    centres, _ = _ring_centres(nside)
    c = centres[1000]
    # a unit vector perpendicular to c: move the particle by angle d
    t = np.cross(c, [0.0, 0.0, 1.0])
    t /= np.linalg.norm(t)
    s = 1.0 / (ang * Dx) ** 2
    theta = 0.12 * ang
    h = Dx * np.sin(theta)
    for d_pix, expect in ((0.2, 0.42), (0.35, 0.27), (0.6, 0.02)):
        d = d_pix * ang
        pix, A, wk, info = orc.hp_particle_weights(Dx * (np.cos(d) * c + np.sin(d) * t), h,
                                                   nside, "wendlandc4")
        assert pix.tolist() == [1000], (d_pix, pix)
        assert info["fallback"] and wk.tolist() == [1.0]
        assert abs(_area_closed_form(np.array([d]), theta, ang)[0] - expect) < 1e-12
        np.testing.assert_allclose(A[0], s * expect, rtol=1e-9)
    # a disc whose edge passes through the centre of pixel 1000
    d = 0.4 * ang
    pix, A, wk, info = orc.hp_particle_weights(Dx * (np.cos(d) * c + np.sin(d) * t),
                                               Dx * np.sin(d * (1 + 1e-12)), nside, "wendlandc4")
    k = pix.tolist().index(1000)
    np.testing.assert_allclose(A[k], 0.5 * s, rtol=1e-9)


def test_pole_centred_map_closed_form(orc):
    """The deposited weight map of one pole-centred particle equals
    V * wk_i * A_i / sum(wk * A) with V = w m / rho / (ang_pix Dx)^2
    (main.jl:25-45, 56-63, 193), everything from the geometry built here."""
    nside, Dx, kernel = 16, 0.6, "wendlandc4"
    ang = np.sqrt(4 * np.pi / (12 * nside * nside))
    centres, _ = _ring_centres(nside)
    theta = 5.2 * ang
    h = Dx * np.sin(theta)
    m, rho, w, q = 2.0, 3.0, 0.7, 5.0
    mp, wm = orc.healpix_map(np.array([[0.0, 0.0, Dx]]), [h], [m], [rho], [q], [w], Nside=nside,
                             kernel=kernel)
    d = np.arccos(np.clip(centres[:, 2], -1.0, 1.0))
    inside = d <= theta
    A = _area_closed_form(d, theta, ang) * inside
    wk = np.array([orc.kernel_shape(kernel, x) if x <= 1 else 0.0 for x in d / theta])
    V = w * m / rho / (ang * Dx) ** 2
    expect = V * wk * A / (wk * A).sum()
    np.testing.assert_allclose(wm, expect, rtol=1e-11, atol=1e-14 * expect.max())
    np.testing.assert_allclose(mp, q * expect, rtol=1e-11, atol=1e-14 * q * expect.max())
