"""Pins of the oracle's TSC deposit against closed forms (the reference's
TSC path is commented out, so TSC is defined by the textbook scheme).
CPU only."""
import numpy as np


def test_cell_centre_stencil(orc):
    N = 8
    c = np.array([3, 4, 5])
    pos = ((c + 0.5) / N - 0.5)[None, :]
    g = orc.tsc_3d(pos, np.array([2.0]), N, float(N))
    w = g[:, 1].reshape(N, N, N)
    axis = {-1: 0.125, 0: 0.75, 1: 0.125}
    for a in (-1, 0, 1):
        for b in (-1, 0, 1):
            for e in (-1, 0, 1):
                assert abs(w[c[0] + a, c[1] + b, c[2] + e] - axis[a] * axis[b] * axis[e]) < 1e-15
    assert abs(w.sum() - 1.0) < 1e-14
    np.testing.assert_allclose(g[:, 0], 2.0 * g[:, 1], rtol=0, atol=1e-15)


def test_partition_of_unity_and_linearity(orc):
    N = 16
    r = np.random.default_rng(2)
    pos = r.uniform(-0.4, 0.4, (200, 3))  # away from the border: nothing dropped
    q = r.uniform(-1, 3, 200)
    g = orc.tsc_3d(pos, q, N, float(N))
    assert abs(g[:, 1].sum() - 200) < 1e-10
    assert abs(g[:, 0].sum() - q.sum()) < 1e-10
    g2 = orc.tsc_3d(pos, 3 * q, N, float(N))
    np.testing.assert_allclose(g2[:, 0], 3 * g[:, 0], rtol=1e-13, atol=1e-14)


def test_1d_weights_closed_form(orc):
    """Per axis W(d) = 3/4 - d^2 (|d|<1/2), (3/2-|d|)^2/2 (1/2<=|d|<3/2)."""
    N = 8
    for off in (0.1, 0.37, 0.5, 0.83):
        pos = np.array([[(3 + off) / N - 0.5, 0.5 / N - 0.5 + 4.0 / N, 0.5 / N]])
        g = orc.tsc_3d(pos, np.ones(1), N, float(N))[:, 1].reshape(N, N, N)
        line = g[:, 4, 4] / (0.75 * 0.75)
        x = 3 + off
        for i in range(N):
            d = abs(x - (i + 0.5))
            exp = 0.75 - d * d if d < 0.5 else (0.5 * (1.5 - d) ** 2 if d < 1.5 else 0.0)
            assert abs(line[i] - exp) < 1e-14


def test_periodic_wrap(orc):
    N = 8
    pos = np.array([[-0.5 + 0.1 / N, 0.0, 0.0]])  # next to the x = 0 face
    g0 = orc.tsc_3d(pos, np.ones(1), N, float(N), periodic=False)
    g1 = orc.tsc_3d(pos, np.ones(1), N, float(N), periodic=True)
    assert g0[:, 1].sum() < 1.0 - 1e-3
    assert abs(g1[:, 1].sum() - 1.0) < 1e-14
    assert g1[:, 1].reshape(N, N, N)[N - 1].sum() > 0
