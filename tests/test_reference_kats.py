"""Known-answer tests ported from the reference's own test suite
(test/runtests.jl), run against the oracle and against the host-side mirror
(sph2grid).  CPU only."""
import numpy as np
import pytest


# ---------------------------------------------------------------- runtests.jl:38-58
def test_mapping_parameters_errors_mirror(s2g):
    with pytest.raises(s2g.ErrorException) as e:
        s2g.mappingParameters()
    assert str(e.value) == "Giving a center position requires extent in x, y and z direction."
    with pytest.raises(s2g.ErrorException) as e:
        s2g.mappingParameters(center=[0.0, 0.0, 0.0], x_lim=[-1.0, 1.0], y_lim=[-1.0, 1.0],
                              z_lim=[-1.0, 1.0])
    assert str(e.value) == "Please specify pixelSideLength or number of pixels!"
    p = s2g.mappingParameters(center=[0.0, 0.0, 0.0], x_lim=[-1.0, 1.0], y_lim=[-1.0, 1.0],
                              z_lim=[-1.0, 1.0], Npixels=100)
    assert p.pixelSideLength == 0.02 and list(p.Npixels) == [100, 100, 100]
    p = s2g.mappingParameters(center=[0.0, 0.0, 0.0], x_lim=[-1.0, 1.0], y_lim=[-1.0, 1.0],
                              z_lim=[-1.0, 1.0], pixelSideLength=0.2)
    assert int(p.Npixels[0]) == 10


def test_mapping_parameters_errors_oracle(orc):
    with pytest.raises(orc.OracleParameterError, match="requires extent"):
        orc.mapping_parameters()
    with pytest.raises(orc.OracleParameterError, match="pixelSideLength or number"):
        orc.mapping_parameters(center=[0, 0, 0], x_lim=[-1, 1], y_lim=[-1, 1], z_lim=[-1, 1])
    p = orc.mapping_parameters(center=[0, 0, 0], x_lim=[-1, 1], y_lim=[-1, 1], z_lim=[-1, 1],
                               pixelSideLength=0.2)
    assert p.Npixels[0] == 10


def test_mapping_parameters_mirror_matches_oracle(s2g, orc):
    cases = [dict(center=[3.0, 3.0, 3.0], x_size=6.0, y_size=6.0, z_size=6.0, Npixels=500),
             dict(x_lim=[0.3, 5.7], y_lim=[0.1, 2.0], z_lim=[0.0, 1.0], Npixels=256,
                  boxsize=6.0),
             dict(x_lim=[-0.5, 0.5], y_lim=[-0.25, 0.75], z_lim=[0, 3], pixelSideLength=0.003)]
    for c in cases:
        a = s2g.mappingParameters(**c)
        b = orc.mapping_parameters(**c)
        assert a.len2pix == b.len2pix
        assert a.pixelSideLength == b.pixelSideLength
        assert list(a.Npixels) == list(b.Npixels)
        assert list(a.center) == list(b.center)
        assert list(a.halfsize) == list(b.halfsize)
        assert a.periodic == bool(b.periodic)


# ---------------------------------------------------------------- runtests.jl:60-77
def test_filter_particles(s2g, orc):
    par = s2g.mappingParameters(center=[3.0, 3.0, 3.0], x_size=6.0, y_size=6.0, z_size=6.0,
                                Npixels=500)
    x = np.zeros((3, 2))
    x[:, 0] = [1.0, 1.0, 1.0]
    x[:, 1] = [-5.0, -1.0, 1.0]
    keep = s2g.filter_particles_in_image(x, par)
    assert keep[0] and not keep[1]
    opar = orc.mapping_parameters(center=[3.0, 3.0, 3.0], x_size=6.0, y_size=6.0,
                                  z_size=6.0, Npixels=500)
    k2 = orc.filter_particles(np.ascontiguousarray(x.T), opar)
    assert list(k2) == [True, False]


# ---------------------------------------------------------------- runtests.jl:79-99
def test_shift_particles(s2g, orc):
    par = s2g.mappingParameters(center=[1.0, 1.0, 1.0], x_size=6.0, y_size=6.0, z_size=6.0,
                                Npixels=500)
    x = np.zeros((3, 2))
    x[:, 0] = [1.0, 1.0, 1.0]
    x[:, 1] = [-3.0, -1.0, 1.0]
    x, par2 = s2g.center_particles(x, par)
    np.testing.assert_allclose(x[:, 0], [0.0, 0.0, 0.0])
    np.testing.assert_allclose(x[:, 1], [-4.0, -2.0, 0.0])
    assert list(par.center) == [1.0, 1.0, 1.0]
    assert list(par2.center) == [0.0, 0.0, 0.0]
    opar = orc.mapping_parameters(center=[1.0, 1.0, 1.0], x_size=6.0, y_size=6.0,
                                  z_size=6.0, Npixels=500)
    xo = np.array([[1.0, 1.0, 1.0], [-3.0, -1.0, 1.0]])
    opar2 = orc.center_particles(xo, opar)
    np.testing.assert_array_equal(xo.T, x)
    assert list(opar2.center) == [0.0, 0.0, 0.0]
    assert opar2.len2pix == par2.len2pix


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("periodic", [False, True])
def test_center_particles_mirror_bitexact(s2g, orc, dtype, periodic):
    """The host mirror's in-place centring (FP64 arithmetic, stored in the
    array's type, boxsize/2 wrap) equals the oracle's bit for bit."""
    r = np.random.default_rng(7)
    n = 5000
    kw = dict(center=[0.37, 0.61, 0.5], x_size=0.8, y_size=0.6, z_size=0.7, Npixels=128)
    if periodic:
        kw["boxsize"] = 1.0
    x = (r.random((3, n)) * 1.4 - 0.2).astype(dtype)
    xo = np.ascontiguousarray(x.T).copy()
    par = s2g.mappingParameters(**kw)
    x2, par2 = s2g.center_particles(x, par)
    opar2 = orc.center_particles(xo, orc.mapping_parameters(**kw))
    np.testing.assert_array_equal(np.ascontiguousarray(x2.T), xo)
    keep = s2g.filter_particles_in_image(x2, par2)
    np.testing.assert_array_equal(keep, orc.filter_particles(xo, opar2))


def test_filter_sort_z_quirk(s2g):
    """filter_shift.jl:60-63 returns reverse(sortperm(z))[mask]: the mask in
    the original order applied to the sorted order.  Mirrored as is."""
    par = s2g.mappingParameters(center=[0.0, 0.0, 0.0], x_size=2.0, y_size=2.0, z_size=2.0,
                                Npixels=8)
    pos = np.array([[0.0, 0.0, 5.0, 0.0], [0.0, 0.0, 0.0, 0.0], [0.1, 0.3, 0.2, 0.9]])
    idx = s2g.filter_particles_in_image(pos, par, sort_z=True)
    # z order descending: [3, 1, 2, 0]; mask in original order: [T, T, F, T]
    assert list(idx) == [3, 1, 0]


# ---------------------------------------------------------------- runtests.jl:139-166
@pytest.mark.parametrize("dims", [2, 3])
def test_index_bijection_via_deposit(orc, dims):
    """calculate_index (indices.jl:6-17): a sub-pixel particle in pixel
    (i, j[, k]) lands at flat index i*N+j (2D) / i*N*N+j*N+k (3D), and the map
    from pixels to flat indices is a bijection over the grid."""
    N = 8
    seen = set()
    r = np.random.default_rng(3)
    for _ in range(40):
        ijk = r.integers(0, N, size=3)
        pos = ((ijk + 0.5) / N - 0.5)[None, :]  # centred coordinates, box [-0.5,0.5]
        h = np.array([0.1 / N])
        one = np.ones(1)
        if dims == 2:
            img = orc.cic_mapping_2D(pos, h, one, one, one, one, N, float(N), "cubic", 2)
            flat = int(ijk[0] * N + ijk[1])
        else:
            img = orc.cic_mapping_3D(pos, h, one, one, one, one, N, float(N), "cubic", 3)
            flat = int(ijk[0] * N * N + ijk[1] * N + ijk[2])
        nz = np.flatnonzero(img[:, -1])
        assert list(nz) == [flat]
        seen.add(flat)
    # bijection of the formula over the whole grid
    if dims == 2:
        idx = [i * N + j for i in range(N) for j in range(N)]
    else:
        idx = [i * N * N + j * N + k for i in range(N) for j in range(N) for k in range(N)]
    assert sorted(idx) == list(range(N ** dims))


# ---------------------------------------------------------------- runtests.jl:324-342
def test_mass_conservation_3d_oracle(orc):
    npix, r = 200, 64
    par = orc.mapping_parameters(x_lim=[-r, r], y_lim=[-r, r], z_lim=[-r, r], Npixels=npix)
    pos = np.array([[0.0101, -0.001, 0.001]])
    hs = np.array([5.0])
    mass = np.array([3.0])
    rho = np.ones(1)
    w = np.ones(1) * par.pixelSideLength * 1
    mp = orc.sph_mapping(pos, hs, mass, rho, rho, w, par, "cubic", kernel_dim=3,
                         dimensions=3, reduce_image=False)
    Vpix = (par.x_lim[1] - par.x_lim[0]) ** 3 / npix ** 3
    assert np.isclose(mass.sum(), Vpix * mp.sum(), rtol=1.5e-8)


# ---------------------------------------------------------------- runtests.jl:718-738
def test_particle_not_overlapping_centres_2d_oracle(orc):
    npix, r = 4, 10
    par = orc.mapping_parameters(x_lim=[-r, r], y_lim=[-r, r], z_lim=[-r, r], Npixels=npix)
    pos = np.array([[0.5, 0.0, 0.0]])
    one = np.ones(1)
    w = one * par.pixelSideLength
    mp = orc.sph_mapping(pos, one, one, one, one, w, par, "cubic", kernel_dim=3,
                         reduce_image=False)
    Apix = (par.x_lim[1] - par.x_lim[0]) ** 2 / npix ** 2
    assert np.isclose(1.0, Apix * mp.sum(), rtol=1e-10)
    # fallback: pure overlap-area weights (x = 2.1 px, y = 2.0 px, h = 0.2 px)
    expect = np.zeros((4, 4))
    expect[1, 1] = expect[1, 2] = 0.25 * 0.5 / Apix
    expect[2, 1] = expect[2, 2] = 0.75 * 0.5 / Apix
    np.testing.assert_allclose(mp[:, :, 0], expect, rtol=1e-12)


# ---------------------------------------------------------------- runtests.jl:488-514
def test_weight_functions(s2g):
    assert s2g.part_weight_one(1)[0] == 1.0
    par = s2g.mappingParameters(center=[3.0, 3.0, 3.0], x_size=6.0, y_size=6.0, z_size=6.0,
                                Npixels=200, boxsize=6.0)
    assert np.isclose(s2g.part_weight_physical(1, par)[0], par.pixelSideLength * 3.085678e21)
    assert np.isclose(s2g.part_weight_physical(1)[0], 3.085678e21)
    assert np.isclose(s2g.part_weight_emission([0.5, 0.5], [0.5, 0.5])[0], 0.1767766952966369)
    assert np.isclose(s2g.part_weight_spectroscopic([0.5, 0.5], [0.5, 0.5])[0],
                      0.4204482076268573)
    assert np.isclose(s2g.part_weight_XrayBand([0.5, 0.5], 0.5, 1.5)[0], 0.0)
