"""Pins of the oracle's mass_density (effects/mass_density.jl:30-99).  CPU only.

* hsml against scipy's cKDTree (an independent k-nearest-neighbour
  implementation, periodic and not): the distance to the (N+1)-th nearest
  particle, the particle itself included.
* Density on a simple cubic lattice in a periodic box, where the 27 nearest
  particles are known in closed form (self, 6 at a, 12 at a*sqrt(2), 8 at
  a*sqrt(3)).
* The bias correction of Dehnen & Aly (2012, eqs. 18-19): eps100 and alpha of
  WendlandC2 / C4 / C6, no correction for the other kernels or when it would
  leave a non-positive density.
Parity unpinned: SPHKernels.jl's exact bias_correction code (not vendored);
the published formula is what is restated.
"""
import numpy as np
import pytest
from scipy.spatial import cKDTree


@pytest.mark.parametrize("periodic", [False, True])
def test_hsml_matches_ckdtree(orc, periodic):
    r = np.random.default_rng(7 + periodic)
    n, k = 2500, 40
    pos = r.random((3, n)) * 5.0
    pos[:, :800] = 2.5 + 0.2 * r.standard_normal((3, 800))  # a clump
    pos = np.clip(pos, 0.0, 5.0 - 1e-9)
    m = r.uniform(0.5, 1.5, n)
    box = (5.0, 5.0, 5.0) if periodic else (0.0, 0.0, 0.0)
    rho, h = orc.mass_density(pos, m, "cubic", 3, k, box)
    tree = cKDTree(pos.T, boxsize=5.0 if periodic else None)
    d, _ = tree.query(pos.T, k=k + 1)
    np.testing.assert_allclose(h, d[:, -1], rtol=1e-15)
    assert np.all(rho > 0)


@pytest.mark.parametrize("kernel", ["cubic", "wendlandc4"])
def test_lattice_closed_form(orc, kernel):
    L, a = 8, 1.0
    g = (np.arange(L) + 0.5) * a
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    pos = np.stack([X.ravel(), Y.ravel(), Z.ravel()])
    m = np.full(pos.shape[1], 2.0)
    rho, h = orc.mass_density(pos, m, kernel, 3, 26, (L * a,) * 3)
    hs = np.sqrt(3.0) * a
    np.testing.assert_allclose(h, hs, rtol=1e-15)
    norm = orc.kernel_norm(kernel, 3)
    s = lambda u: orc.kernel_shape(kernel, u)
    raw = 2.0 * norm / hs ** 3 * (s(0.0) + 6 * s(a / hs) + 12 * s(np.sqrt(2.0) * a / hs))
    if kernel == "wendlandc4":
        raw -= 0.01342 * (26 * 0.01) ** (-1.579) * 2.0 * norm / hs ** 3
    np.testing.assert_allclose(rho, raw, rtol=1e-13)


def test_bias_correction_constants(orc):
    for kern, e100, alpha in (("wendlandc2", 0.0294, 0.977), ("wendlandc4", 0.01342, 1.579),
                              ("wendlandc6", 0.0116, 2.236)):
        norm = orc.kernel_norm(kern, 3)
        for nn in (32, 100, 195):
            corr = e100 * (nn / 100.0) ** (-alpha) * 1.5 * norm * 2.0 ** 3
            got = orc.bias_correction(kern, 3, 1e3, 1.5, 2.0, nn)
            assert got == pytest.approx(1e3 - corr, rel=1e-14)
        # a correction that would leave a non-positive density is not applied
        assert orc.bias_correction(kern, 3, 1e-6, 1.5, 2.0, 32) == 1e-6
    for kern in ("cubic", "quintic", "wendlandc8"):
        assert orc.bias_correction(kern, 3, 7.0, 1.5, 2.0, 32) == 7.0
