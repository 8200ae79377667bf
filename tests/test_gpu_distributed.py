"""GPU tests of distributed_cic_map / distributed_allsky_map with the
device-resident sum.  A deposit is a sum over particles, so the map of a
particle set split into subfiles equals the map of the whole set (up to the
order of FP64 additions: <= 1e-12 relative here).  The reference's
NaN / Inf rule (cic.jl:63-73, healpix.jl:50-58) is checked against a host
restatement on maps that contain NaN and Inf."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _plain(n, seed, nq=1):
    r = np.random.default_rng(seed)
    pos = r.random((3, n))
    h = 2.0 * n ** (-1.0 / 3.0) * np.exp(0.3 * r.standard_normal(n))
    m = np.full(n, 1.0 / n)
    rho = 64.0 * m / (4.0 / 3.0 * np.pi * h ** 3)
    q = r.uniform(0.5, 2.0, (nq, n)) if nq > 1 else r.uniform(0.5, 2.0, n)
    return pos, h, m, rho, q


def _close(a, b, rtol=1e-12):
    scale = np.abs(b).max()
    assert np.abs(a - b).max() <= rtol * scale, np.abs(a - b).max() / scale


@pytest.mark.parametrize("nq", [1, 2])
def test_cic_particle_subfiles_equal_whole_set(s2g, tmp_path, nq):
    n, npix, nsub = 6000, 96, 5
    pos, h, m, rho, q = _plain(n, 11 + nq, nq)
    par = s2g.mappingParameters(x_lim=[0.0, 1.0], y_lim=[0.0, 1.0], z_lim=[0.0, 1.0],
                                Npixels=npix)
    edges = np.linspace(0, n, nsub + 1).astype(int)

    def sub(k):
        s = slice(edges[k], edges[k + 1])
        qq = q[..., s].copy()
        return s2g.Subfile(pos[:, s].copy(), h[s].copy(), m[s].copy(), rho[s].copy(), qq,
                           rho[s].copy(), kernel=s2g.WendlandC4(2), calc_mean=True)

    f = str(tmp_path / "map.fits")
    st = {}
    img = s2g.distributed_cic_map(f, nsub, sub, par, nq, prefetch=2, stats=st)
    print("stats", st, flush=True)
    ref = s2g.sphMapping(pos.copy(), h, m, rho, q, rho, param=par, kernel=s2g.WendlandC4(2),
                         calc_mean=True, show_progress=False)
    assert img.shape == ref.shape
    _close(img, ref)
    back, par2, snap, units = s2g.read_fits_image(f)
    np.testing.assert_array_equal(back, img[:, :, 0])
    assert st["subfiles"] == nsub and st["particles"] == n


def test_cic_map_subfiles_nan_inf_rule(s2g, tmp_path):
    n, npix, nsub = 3000, 64, 4
    pos, h, m, rho, q = _plain(n, 21)
    par = s2g.mappingParameters(x_lim=[0.0, 1.0], y_lim=[0.0, 1.0], z_lim=[0.0, 1.0],
                                Npixels=npix)
    edges = np.linspace(0, n, nsub + 1).astype(int)
    maps = []
    for k in range(nsub):
        s = slice(edges[k], edges[k + 1])
        flat = s2g.sphMapping(pos[:, s].copy(), h[s], m[s], rho[s], q[s], rho[s], param=par,
                              kernel=s2g.Cubic(2), return_both_maps=True, show_progress=False)
        mp, w = flat[:, :-1].copy(), flat[:, -1].copy()
        if k == 1:
            mp[5, 0], mp[77, 0], w[9] = np.nan, np.inf, -np.inf
        maps.append((mp, w))
    maps.append(None)  # a subfile without particles (cic.jl:62)
    got = s2g.distributed_cic_map(str(tmp_path / "m.fits"), nsub + 1,
                                  lambda k: maps[k], par, reduce_image=True)
    # host restatement of cic.jl:56-93
    sm, sw = np.zeros((npix * npix, 1)), np.zeros(npix * npix)
    for r in maps:
        if r is None:
            continue
        a, w = r
        sm += np.where(np.isfinite(a), a, 0.0)
        sw += np.where(np.isfinite(w), w, 0.0)
    expect = np.zeros((npix, npix, 1))
    img = sm[:, 0].reshape(npix, npix)  # [i, j] deposit order, i = x
    ww = sw.reshape(npix, npix)
    red = np.where(ww > 0, img / np.where(ww > 0, ww, 1.0), img)
    expect[:, :, 0] = red
    _close(got, expect)


def test_cic_3d_vtk(s2g, tmp_path):
    n, npix = 4000, 24
    pos, h, m, rho, q = _plain(n, 31)
    par = s2g.mappingParameters(x_lim=[0.0, 1.0], y_lim=[0.0, 1.0], z_lim=[0.0, 1.0],
                                Npixels=npix)
    half = n // 2
    sub = lambda k: s2g.Subfile(pos[:, k * half:(k + 1) * half].copy(),
                                h[k * half:(k + 1) * half], m[k * half:(k + 1) * half],
                                rho[k * half:(k + 1) * half], q[k * half:(k + 1) * half],
                                kernel=s2g.WendlandC6(3))
    img = s2g.distributed_cic_map(str(tmp_path / "cube"), 2, sub, par, Ndim=3, units="u")
    ref = s2g.sphMapping(pos.copy(), h, m, rho, q, param=par, kernel=s2g.WendlandC6(3),
                         dimensions=3, show_progress=False)
    _close(img, ref)
    back, _, units, _ = s2g.read_vtk_image(str(tmp_path / "cube.vtr"))
    np.testing.assert_array_equal(back, img)
    assert units == "u"


def _shell(n, seed):
    r = np.random.default_rng(seed)
    rad = np.sqrt(r.uniform(0.05 ** 2, 1.0, n))
    v = r.standard_normal((n, 3))
    v /= np.linalg.norm(v, axis=1)[:, None]
    pos = (v * rad[:, None]).T.copy()
    h = np.exp(r.uniform(np.log(2e-3), np.log(0.2), n))
    m, rho, w, q = (r.uniform(0.5, 2.0, n) for _ in range(4))
    return pos, h, m, rho, q, w


def test_allsky_particle_subfiles_equal_whole_set(s2g, tmp_path):
    n, nside, nsub = 1500, 32, 3
    pos, h, m, rho, q, w = _shell(n, 41)
    edges = np.linspace(0, n, nsub + 1).astype(int)

    def sub(k):
        s = slice(edges[k], edges[k + 1])
        return s2g.AllskySubfile(pos[:, s].copy(), h[s].copy(), m[s].copy(), rho[s].copy(),
                                 q[s].copy(), w[s].copy(), kernel=s2g.WendlandC4(2))

    f = str(tmp_path / "sky.fits")
    a, wt = s2g.distributed_allsky_map(f, nside, nsub, sub, reduce_image=False)
    ra, rw = s2g.healpix_map(pos.copy(), h, m, rho, q, w, Nside=nside, kernel=s2g.WendlandC4(2),
                             show_progress=False)
    _close(a, ra, 1e-11)
    _close(wt, rw, 1e-11)
    back, ns = s2g.read_healpix_fits(f)
    assert ns == nside
    np.testing.assert_array_equal(back, a)
    # reduced: healpix.jl:67-73
    a2, w2 = s2g.distributed_allsky_map(f, nside, nsub, sub, reduce_image=True)
    good = np.isfinite(rw) & (rw != 0)
    exp = ra.copy()
    exp[good] = ra[good] / rw[good]
    _close(a2, exp, 1e-11)


def test_allsky_maps_nan_rule(s2g, tmp_path):
    nside = 8
    npx = 12 * nside * nside
    r = np.random.default_rng(5)
    maps = [(r.random(npx), r.random(npx)) for _ in range(3)]
    maps[0][0][3] = np.nan
    maps[2][1][7] = np.inf
    maps[1][1][11] = 0.0
    a, w = s2g.distributed_allsky_map(str(tmp_path / "s.fits"), nside, 3, lambda k: maps[k],
                                      reduce_image=True)
    sa = sum(np.where(np.isfinite(x), x, 0.0) for x, _ in maps)
    sw = sum(np.where(np.isfinite(y), y, 0.0) for _, y in maps)
    good = np.isfinite(sw) & (sw != 0)
    exp = sa.copy()
    exp[good] = sa[good] / sw[good]
    _close(a, exp)
    _close(w, sw)
