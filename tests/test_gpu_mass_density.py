"""GPU mass_density (effects/mass_density.jl:30-99) against the brute-force
FP64 oracle: hsml is a distance between two particles computed in FP64
without contraction on both sides, so it must agree to the last bit (the test
allows 1e-15 relative); rho sums the same terms in another order (1e-12)."""
import numpy as np
import pytest

from sph2grid import _lib

pytestmark = pytest.mark.gpu


def _clustered(n, seed, L=10.0):
    r = np.random.default_rng(seed)
    pos = r.random((3, n)) * L
    c = n // 2
    centres = r.random((3, 5)) * L
    pick = r.integers(0, 5, c)
    pos[:, :c] = centres[:, pick] + 0.05 * L * r.standard_normal((3, c))
    pos = np.mod(pos, L)
    return pos, r.uniform(0.5, 2.0, n)


def _check(s2g, orc, pos, m, kernel, kname, k, box, dtype=np.float64):
    st = {}
    rho, h = s2g.mass_density(pos.astype(dtype), m.astype(dtype), kernel=kernel, Nneighbors=k,
                              boxsize=box, stats=st)
    orho, oh = orc.mass_density(pos.astype(dtype).astype(np.float64),
                                m.astype(dtype).astype(np.float64), kname, kernel.dim, k,
                                box if box is not None else (0, 0, 0))
    eh = np.abs(h - oh).max() / oh.max()
    er = np.abs(rho - orho).max() / np.abs(orho).max()
    print(f"[{kname} k={k} box={box} {dtype.__name__}] hsml {eh:.1e} rho {er:.1e} "
          f"device {st.get('ms_device', 0):.2f} ms", flush=True)
    assert eh <= 1e-15 and er <= 1e-12


@pytest.mark.parametrize("kname", ["cubic", "wendlandc2", "wendlandc4", "wendlandc6"])
def test_periodic_uniform(s2g, orc, kname):
    r = np.random.default_rng(3)
    n = 3000
    pos, m = r.random((3, n)) * 6.0, r.uniform(0.5, 2.0, n)
    kern = {"cubic": s2g.Cubic, "wendlandc2": s2g.WendlandC2, "wendlandc4": s2g.WendlandC4,
            "wendlandc6": s2g.WendlandC6}[kname](3)
    _check(s2g, orc, pos, m, kern, kname, 32, (6.0, 6.0, 6.0))


@pytest.mark.parametrize("periodic", [False, True])
def test_clustered(s2g, orc, periodic):
    pos, m = _clustered(6000, 11 + periodic)
    _check(s2g, orc, pos, m, s2g.WendlandC4(3), "wendlandc4", 64,
           (10.0, 10.0, 10.0) if periodic else None)


def test_reference_neighbour_count_and_float32(s2g, orc):
    """Nneighbors = 195 as the reference's own test (runtests.jl:632-636),
    periodic box of 6, WendlandC4, Float32 inputs."""
    r = np.random.default_rng(5)
    n = 2500
    pos, m = r.random((3, n)) * 6.0, r.uniform(0.5, 2.0, n)
    _check(s2g, orc, pos, m, s2g.WendlandC4(3), "wendlandc4", 195, (6.0,) * 3, np.float32)


def test_lattice_and_kernel_dim_2(s2g, orc):
    L = 8
    g = np.arange(L) + 0.5
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    pos = np.stack([X.ravel(), Y.ravel(), Z.ravel()])
    m = np.full(pos.shape[1], 2.0)
    rho, h = s2g.mass_density(pos, m, kernel=s2g.Cubic(3), Nneighbors=26, boxsize=(8.0,) * 3)
    np.testing.assert_allclose(h, np.sqrt(3.0), rtol=1e-15)
    assert np.ptp(rho) <= 1e-13 * rho.max()
    # a 2D kernel normalisation in 3D space, as the reference evaluates it
    _check(s2g, orc, pos + 0.01 * np.random.default_rng(1).standard_normal(pos.shape),
           m, s2g.WendlandC6(2), "wendlandc6", 20, None)


@pytest.mark.parametrize("k", [64, 400])
def test_coincident_particles_and_buffer_cutbacks(s2g, orc, k):
    """40 copies of one particle and 500 pairs of coincident ones among 3000;
    400 neighbours make the candidate buffer overflow and be cut back while
    the cells nearest to a particle are read, before the outer ones."""
    r = np.random.default_rng(17)
    n = 3000
    pos, m = r.random((3, n)) * 6.0, r.uniform(0.5, 2.0, n)
    pos[:, 1:40] = pos[:, :1]
    pos[:, 2000:2500] = pos[:, 1000:1500]
    for box in ((6.0,) * 3, None):
        _check(s2g, orc, pos, m, s2g.WendlandC4(3), "wendlandc4", k, box)


@pytest.mark.parametrize("shape", ["plane", "line", "fewest"])
def test_degenerate_sets(s2g, orc, shape):
    """Particles in one plane (zero z extent) or on one line, and the fewest
    particles allowed (Nneighbors + 1, every particle a neighbour of every
    other), without a box."""
    r = np.random.default_rng(23)
    n, k = {"plane": (2000, 32), "line": (600, 20), "fewest": (33, 32)}[shape]
    pos, m = r.random((3, n)) * 5.0, r.uniform(0.5, 2.0, n)
    if shape == "plane":
        pos[2] = 1.5
    elif shape == "line":
        pos[1] = 2.0
        pos[2] = -1.0
    _check(s2g, orc, pos, m, s2g.WendlandC4(3), "wendlandc4", k, None)


def test_errors(s2g):
    r = np.random.default_rng(0)
    pos, m = r.random((3, 20)), np.ones(20)
    with pytest.raises(_lib.LibraryError) as e:
        s2g.mass_density(pos, m, Nneighbors=20)
    assert e.value.code == _lib.S2G_EINVAL
    with pytest.raises(_lib.LibraryError) as e:
        s2g.mass_density(r.random((3, 800)), np.ones(800), Nneighbors=700)
    assert e.value.code == _lib.S2G_EUNSUPPORTED
    with pytest.raises(_lib.LibraryError) as e:
        s2g.mass_density(pos, m, Nneighbors=4, boxsize=(1.0, 0.0, 1.0))
    assert e.value.code == _lib.S2G_EINVAL
    bad = pos.copy()
    bad[1, 3] = np.nan
    with pytest.raises(_lib.LibraryError) as e:
        s2g.mass_density(bad, m, Nneighbors=4)
    assert e.value.code == _lib.S2G_EINVAL
