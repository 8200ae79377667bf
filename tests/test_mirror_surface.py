"""sphMapping's surface options that choose particles or code paths rather
than arithmetic (cic_interpolation.jl:35-273): `parallel`, `RM`, `stokes` and
`sort_z`.

The reference's effect, mirrored here:
  * `parallel = true` maps the same particles through Distributed workers
    and sums the images (:171-215); the GPU takes one call either way.
  * `RM` never changes a map: `stokes` is not passed on to cic_mapping_2D
    (:152-155), so faraday_rotate_pixel! computes the rotation and writes
    nothing (cic_shared.jl:129-159), and the parallel path drops RM.
  * `stokes = true` forces `sort_z` (and serial mode) (:74-81); `sort_z`
    selects reverse(sortperm(z))[mask] with the mask in the original order
    (filter_shift.jl:60-63).
CPU tests check the host selection; GPU tests check the maps against the
oracle on the same selection.
"""
import numpy as np
import pytest

from conftest import rel_l1


def _set(n=3000, seed=11):
    r = np.random.default_rng(seed)
    pos = r.uniform(-0.2, 1.2, (3, n))  # some particles outside the image
    h = np.exp(r.uniform(np.log(0.002), np.log(0.08), n))
    m, rho, w = (r.uniform(0.5, 2.0, n) for _ in range(3))
    q = r.uniform(0.1, 5.0, n)
    rm = r.uniform(-50.0, 50.0, n)
    return pos, h, m, rho, q, w, rm


def _par(s2g, npix=96):
    return s2g.mappingParameters(x_lim=[0.0, 1.0], y_lim=[0.0, 1.0], z_lim=[0.0, 1.0],
                                 Npixels=npix)


def _stokes_selection(pos_centred, par2):
    """reverse(sortperm(z))[mask] (filter_shift.jl:60-63) with the mask of
    filter_shift.jl:50-58, computed here from its definition."""
    lo, hi = par2.center - par2.halfsize, par2.center + par2.halfsize
    mask = np.all((lo[:, None] <= pos_centred) & (pos_centred <= hi[:, None]), axis=0)
    order = np.argsort(pos_centred[2], kind="stable")[::-1]
    return order[mask], mask


def test_stokes_selects_the_sort_z_order(s2g):
    """With stokes (or sort_z) the host mirror hands the deposit the particles
    reverse(sortperm(z))[mask], in that order; without, the masked particles
    in input order."""
    from sph2grid.mapping import _prepare
    pos, h, m, rho, q, w, _ = _set()
    par = _par(s2g)
    cen, par2 = s2g.center_particles(pos.copy(), par)
    sel, mask = _stokes_selection(cen, par2)
    assert sel.size == mask.sum() and not np.array_equal(np.sort(sel), np.flatnonzero(mask))
    for kw in ({"stokes": True}, {"sort_z": True}):
        pm = _prepare(pos.copy(), h, m, rho, q, w, param=par, kernel=s2g.Cubic(2),
                      show_progress=False, **kw)
        np.testing.assert_array_equal(pm.arrays[0], cen[:, sel].T)
        np.testing.assert_array_equal(pm.arrays[1], h[sel])
        np.testing.assert_array_equal(pm.arrays[4], q[sel])
    pm = _prepare(pos.copy(), h, m, rho, q, w, param=par, kernel=s2g.Cubic(2),
                  show_progress=False)
    np.testing.assert_array_equal(pm.arrays[0], cen[:, mask].T)


@pytest.mark.gpu
@pytest.mark.parametrize("dims", [2, 3])
def test_parallel_is_the_same_call(s2g, dims):
    pos, h, m, rho, q, w, _ = _set()
    par = _par(s2g, 96 if dims == 2 else 40)
    kw = dict(param=par, kernel=s2g.WendlandC6(dims), show_progress=False, dimensions=dims)
    a = s2g.sphMapping(pos.copy(), h, m, rho, q, w, parallel=False, **kw)
    b = s2g.sphMapping(pos.copy(), h, m, rho, q, w, parallel=True, **kw)
    np.testing.assert_allclose(b, a, rtol=1e-12, atol=1e-14 * np.abs(a).max())


@pytest.mark.gpu
def test_rm_has_no_effect(s2g):
    pos, h, m, rho, q, w, rm = _set()
    par = _par(s2g)
    kw = dict(param=par, kernel=s2g.Cubic(2), show_progress=False, calc_mean=True)
    a = s2g.sphMapping(pos.copy(), h, m, rho, q, w, **kw)
    b = s2g.sphMapping(pos.copy(), h, m, rho, q, w, rm, **kw)
    c = s2g.sphMapping(pos.copy(), h, m, rho, q, w, rm, parallel=True, **kw)
    for x in (b, c):
        np.testing.assert_allclose(x, a, rtol=1e-12, atol=1e-14 * np.abs(a).max())


@pytest.mark.gpu
def test_stokes_map_against_oracle(s2g, orc):
    """stokes = true (with RM): the map of the sort_z selection, against the
    oracle's deposit of the same particles; it differs from the plain map
    because the selection does (the reference's mask quirk)."""
    pos, h, m, rho, q, w, rm = _set()
    npix = 96
    par = _par(s2g, npix)
    kw = dict(param=par, kernel=s2g.WendlandC4(2), show_progress=False, calc_mean=True)
    got = s2g.sphMapping(pos.copy(), h, m, rho, q, w, rm, stokes=True, parallel=True, **kw)
    cen, par2 = s2g.center_particles(pos.copy(), par)
    sel, mask = _stokes_selection(cen, par2)
    flat = orc.cic_mapping_2D(np.ascontiguousarray(cen[:, sel].T), h[sel], m[sel], rho[sel],
                              q[sel], w[sel], npix, float(par2.len2pix), "wendlandc4", 2, True)
    ref = orc.reduce_image_2D(flat, npix, True)
    e = rel_l1(got, ref)
    print(f"stokes selection: rel_l1 {e:.2e}", flush=True)
    assert e <= 1e-5
    plain = s2g.sphMapping(pos.copy(), h, m, rho, q, w, **kw)
    assert rel_l1(plain, ref) > 1e-3
