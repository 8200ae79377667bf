"""GPU parity of the all-sky deposit (healpix_map's particle loop) against
the FP64 oracle.  Tolerances: rel-L1 <= 1e-5 per map, conservation of the
weight map <= 1e-6 (the oracle's own FP64 acos carries ~1.5e-8 absolute
error in pixel distances, the GPU uses the haversine form)."""
import ctypes as C

import numpy as np
import pytest

from conftest import rel_l1

pytestmark = pytest.mark.gpu

KNAMES = ["cubic", "quintic", "wendlandc2", "wendlandc4", "wendlandc6", "wendlandc8"]


def _kernel(s2g, name):
    return {"cubic": s2g.Cubic, "quintic": s2g.Quintic, "wendlandc2": s2g.WendlandC2,
            "wendlandc4": s2g.WendlandC4, "wendlandc6": s2g.WendlandC6,
            "wendlandc8": s2g.WendlandC8}[name](2)


def _shell(n, seed, hmin=2e-3, hmax=0.3, rmin=0.05):
    r = np.random.default_rng(seed)
    rad = np.sqrt(r.uniform(rmin ** 2, 1.0, n))
    v = r.standard_normal((n, 3))
    v /= np.linalg.norm(v, axis=1)[:, None]
    pos = v * rad[:, None]
    h = np.exp(r.uniform(np.log(hmin), np.log(hmax), n))
    m, rho, w, q = (r.uniform(0.5, 2.0, n) for _ in range(4))
    return pos, h, m, rho, q, w


def _compare(s2g, orc, pos, h, m, rho, q, w, nside, kname, calc_mean=True, tag=""):
    st = {}
    gm, gw = s2g.healpix_map(pos.T.copy(), h, m, rho, q, w, Nside=nside,
                             kernel=_kernel(s2g, kname), calc_mean=calc_mean,
                             show_progress=False, stats=st)
    om, ow = orc.healpix_map(pos.copy(), h, m, rho, q, w, Nside=nside, kernel=kname,
                             calc_mean=calc_mean)
    e1, e2 = rel_l1(gm, om), rel_l1(gw, ow)
    print(f"[{tag} {kname} nside={nside}] rel_l1 map {e1:.2e} weights {e2:.2e} "
          f"units {st.get('n_pairs')} dev {st.get('ms_device', 0):.2f} ms", flush=True)
    assert e1 <= 1e-5 and e2 <= 1e-5
    assert abs(gw.sum() / ow.sum() - 1) <= 1e-6
    return gm, gw


@pytest.mark.parametrize("kname", KNAMES)
def test_random_shell(s2g, orc, kname):
    pos, h, m, rho, q, w = _shell(600, 10 + KNAMES.index(kname))
    _compare(s2g, orc, pos, h, m, rho, q, w, 32, kname, tag="shell")


def test_nside_above_ring_lds(s2g, orc):
    """Nside 2048: rings of 8192 pixels exceed the ring-major deposit's LDS
    row, so the deposit takes the per-unit atomic path."""
    pos, h, m, rho, q, w = _shell(200, 77, hmin=2e-3, hmax=2e-2, rmin=0.3)
    _compare(s2g, orc, pos, h, m, rho, q, w, 2048, "wendlandc4", tag="atomic")


@pytest.mark.parametrize("nside", [1, 4, 64])
def test_nsides(s2g, orc, nside):
    pos, h, m, rho, q, w = _shell(300, 20 + nside)
    _compare(s2g, orc, pos, h, m, rho, q, w, nside, "wendlandc4", tag="nside")


def test_tiny_discs_fallback(s2g, orc):
    """Sub-pixel discs (no pixel centre inside the kernel): area weights."""
    pos, h, m, rho, q, w = _shell(400, 31, hmin=1e-4, hmax=2e-3)
    _compare(s2g, orc, pos, h, m, rho, q, w, 32, "wendlandc6", tag="tiny")


def test_poles_and_wrap(s2g, orc):
    """Discs containing a pole (full polar rings) and discs crossing phi = 0."""
    pts = []
    for th in (0.0, 1e-3, 0.05, np.pi - 0.02, np.pi):
        for ph in (0.0, 1.0, 6.2):
            pts.append([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)])
    for ph in (-0.01, 0.0, 0.01, 2 * np.pi - 1e-3):
        pts.append([np.cos(ph), np.sin(ph), 0.1])
    pos = np.array(pts) * 0.6
    n = pos.shape[0]
    r = np.random.default_rng(5)
    h = r.uniform(0.01, 0.5, n)
    m, rho, w, q = (r.uniform(0.5, 2.0, n) for _ in range(4))
    for nside in (8, 32):
        _compare(s2g, orc, pos, h, m, rho, q, w, nside, "cubic", tag="poles")


def test_calc_mean_false_and_skips(s2g, orc):
    pos, h, m, rho, q, w = _shell(200, 41)
    h[::9] = 2.0  # dx < hsml: skipped (main.jl:172-179)
    _compare(s2g, orc, pos, h, m, rho, q, w, 16, "wendlandc2", calc_mean=False, tag="cm0")
    with pytest.raises(s2g.BoundsError):
        q0 = q.copy()
        q0[0] = 0.0
        s2g.healpix_map(pos.T.copy(), h, m, rho, q0, w, Nside=16, kernel=s2g.Cubic(2),
                        calc_mean=False, show_progress=False)


def test_float32_inputs_through_abi(s2g, orc):
    """The C ABI also takes Float32 arrays (promoted exactly, as FP64 inputs)."""
    from sph2grid import _lib
    pos, h, m, rho, q, w = _shell(300, 51)
    arrs = [np.ascontiguousarray(a, dtype=np.float32) for a in (pos, h, m, rho, q, w)]
    nside = 16
    npx = 12 * nside * nside
    mp, wm = np.zeros(npx), np.zeros(npx)
    ctx = s2g.default_context()
    p = lambda a: a.ctypes.data_as(C.c_void_p)
    rc = ctx.lib.s2g_healpix_deposit(ctx.handle, _lib.S2G_F32, *[p(a) for a in arrs], 300,
                                     nside, 3, 2, _lib.S2G_FLAG_CALC_MEAN, p(mp), p(wm), None)
    ctx.check(rc, "s2g_healpix_deposit")
    om, ow = orc.healpix_map(arrs[0].astype(np.float64), *[a.astype(np.float64) for a in arrs[1:]],
                             Nside=nside, kernel="wendlandc4")
    assert rel_l1(mp, om) <= 1e-5 and rel_l1(wm, ow) <= 1e-5


def test_allsky_workload_scaled(s2g, orc):
    """Allsky particle set (shell 0.05-1, p(r) ~ r, hsml log-uniform
    [1e-3, 5e-2]) at 2^14 particles, Nside 128, WendlandC4."""
    from sph2grid import workloads
    d = workloads.allsky(n=1 << 14, nside=128)
    pos = d["Pos"].T.copy()
    _compare(s2g, orc, pos, d["HSML"], d["M"], d["Rho"], d["T"], d["Rho"], 128, "wendlandc4",
             tag="allsky")


@pytest.mark.parametrize("kname", ["wendlandc4", "cubic"])
def test_pole_centred_particle_closed_form(s2g, kname):
    """One particle on the z axis, no oracle involved: the weight map equals
    V * wk_i * A_i / sum(wk * A) (main.jl:25-45, 56-63, 193) with the ring
    geometry of hp_geometry.py, A the piecewise contributing_area
    (pixel_weights.jl:6-8) and wk the kernel shape at d / theta.  FP32 pixel
    weights on the GPU: bound 1e-5 relative per pixel (above 1e-3 of the
    maximum) and 1e-6 on the sum."""
    import oracle
    from hp_geometry import area_closed_form, ring_centres
    nside, Dx = 64, 0.45
    ang = np.sqrt(4 * np.pi / (12 * nside * nside))
    centres, _ = ring_centres(nside)
    d = np.arccos(np.clip(centres[:, 2], -1.0, 1.0))
    for theta_pix in (1.7, 6.3, 40.5):
        theta = theta_pix * ang
        h = Dx * np.sin(theta)
        m, rho, w, q = 1.5, 0.25, 2.0, 3.0
        pos = np.array([[0.0], [0.0], [Dx]])
        gm, gw = s2g.healpix_map(pos, np.array([h]), np.array([m]), np.array([rho]),
                                 np.array([q]), np.array([w]), Nside=nside,
                                 kernel=_kernel(s2g, kname), show_progress=False)
        A = area_closed_form(d, theta, ang) * (d <= theta)
        wk = np.array([oracle.kernel_shape(kname, x) if x <= 1 else 0.0 for x in d / theta])
        V = w * m / rho / (ang * Dx) ** 2
        expect = V * wk * A / (wk * A).sum()
        big = expect > 1e-3 * expect.max()
        err = np.abs(gw - expect)[big] / expect[big]
        print(f"[pole {kname} theta={theta_pix} px] max rel {err.max():.2e}", flush=True)
        assert err.max() <= 1e-5
        assert np.all(gw[d > theta + 1e-9] == 0.0)
        assert abs(gw.sum() / V - 1) <= 1e-6
        np.testing.assert_allclose(gm, q * gw, rtol=1e-12)


def test_fallback_particle_one_pixel(s2g):
    """A particle whose disc holds no pixel centre deposits its whole weight
    V into the pixel holding it (constributing_pixels.jl:16-19,
    pixel_weights.jl:120-137)."""
    from hp_geometry import ring_centres
    nside, Dx = 16, 1.3
    ang = np.sqrt(4 * np.pi / (12 * nside * nside))
    centres, _ = ring_centres(nside)
    c = centres[1000]
    t = np.cross(c, [0.0, 0.0, 1.0])
    t /= np.linalg.norm(t)
    for d_pix in (0.2, 0.35, 0.6):
        d = d_pix * ang
        pos = (Dx * (np.cos(d) * c + np.sin(d) * t)).reshape(3, 1)
        h = Dx * np.sin(0.12 * ang)
        gm, gw = s2g.healpix_map(pos.copy(), np.array([h]), np.array([2.0]), np.array([0.5]),
                                 np.array([7.0]), np.array([1.5]), Nside=nside,
                                 kernel=s2g.WendlandC4(2), show_progress=False)
        V = 1.5 * 2.0 / 0.5 / (ang * Dx) ** 2
        assert np.flatnonzero(gw).tolist() == [1000]
        assert abs(gw[1000] / V - 1) <= 1e-12 and abs(gm[1000] / (7.0 * V) - 1) <= 1e-12


def test_allsky_nside1024_ring_lds(s2g, orc):
    """The all-sky particle set at Nside 1024 (rings of up to 4096 pixels: the
    largest rows of the ring-major LDS deposit), 2^12 particles against the
    oracle."""
    from sph2grid import workloads
    d = workloads.allsky(n=1 << 12, nside=1024)
    pos = d["Pos"].T.copy()
    _compare(s2g, orc, pos, d["HSML"], d["M"], d["Rho"], d["T"], d["Rho"], 1024, "wendlandc4",
             tag="allsky1024")


@pytest.mark.parametrize("nside", [64, 2048])
def test_maps_of_one_particle_set(s2g, orc, nside):
    """healpix_maps: three maps of one particle set (rho with unit weights, T
    with rho weights, a third with zero quantities on some particles) in one
    deposit (shared disc queries, normalisation and ring lists; two-map and
    one-map groups; the ring-major path at Nside 64, the atomic one at 2048)
    equal the separate healpix_map calls (<= 1e-12 relative) and the oracle
    (rel-L1 <= 1e-5)."""
    pos, h, m, rho, q, w = _shell(500, 91, hmin=2e-3, hmax=0.25 if nside == 64 else 2e-2,
                                  rmin=0.05 if nside == 64 else 0.3)
    q3 = q.copy()
    q3[::4] = 0.0
    quant = [(rho, np.ones_like(rho)), (q, rho), (q3, w)]
    for cm in (True, False):
        if not cm:
            quant = [(rho, np.ones_like(rho)), (q, rho)]  # calc_mean=false needs Bin_q > 0
        got = s2g.healpix_maps(pos.T.copy(), h, m, rho, quant, Nside=nside,
                               kernel=s2g.WendlandC4(2), calc_mean=cm)
        for k, (bq, ww) in enumerate(quant):
            gm, gw = s2g.healpix_map(pos.T.copy(), h, m, rho, bq, ww, Nside=nside,
                                     kernel=s2g.WendlandC4(2), calc_mean=cm,
                                     show_progress=False)
            tol = lambda a: 1e-300 + 1e-13 * np.abs(a).max()
            np.testing.assert_allclose(got[k][0], gm, rtol=1e-12, atol=tol(gm))
            np.testing.assert_allclose(got[k][1], gw, rtol=1e-12, atol=tol(gw))
            om, ow = orc.healpix_map(pos.copy(), h, m, rho, bq, ww, Nside=nside,
                                     kernel="wendlandc4", calc_mean=cm)
            e1, e2 = rel_l1(got[k][0], om), rel_l1(got[k][1], ow)
            print(f"[maps nside={nside} cm={cm} map {k}] rel_l1 {e1:.2e} {e2:.2e}", flush=True)
            assert e1 <= 1e-5 and e2 <= 1e-5
