"""GPU parity of the 3D deposit (cic_mapping_3D + reduce_image_3D) against
the FP64 oracle, through the C ABI.  Same tolerances as the 2D tests:
rel-L1 <= 1e-5, max relative voxel error <= 1e-4 (voxels above 1e-3 of the
maximum), conservation <= 1e-6."""
import ctypes as C

import numpy as np
import pytest

from conftest import max_rel, rel_l1

pytestmark = pytest.mark.gpu

KNAMES = ["cubic", "quintic", "wendlandc2", "wendlandc4", "wendlandc6", "wendlandc8"]


def _kernel(s2g, name, dim=3):
    return {"cubic": s2g.Cubic, "quintic": s2g.Quintic, "wendlandc2": s2g.WendlandC2,
            "wendlandc4": s2g.WendlandC4, "wendlandc6": s2g.WendlandC6,
            "wendlandc8": s2g.WendlandC8}[name](dim)


def _check(gpu, ref, tag):
    print(f"[{tag}] rel_l1:", [f"{rel_l1(gpu[:, k], ref[:, k]):.2e}" for k in range(2)],
          flush=True)
    for k in range(2):
        assert rel_l1(gpu[:, k], ref[:, k]) <= 1e-5, (tag, k)
        assert max_rel(gpu[:, k], ref[:, k]) <= 1e-4, (tag, k)
    assert abs(gpu[:, 1].sum() / ref[:, 1].sum() - 1) <= 1e-6, tag


def _particles(n, npix, seed, hmin=0.05, hmax=14.0, dtype=np.float64):
    r = np.random.default_rng(seed)
    x = r.uniform(-0.5, 0.5, (n, 3))
    h = np.exp(r.uniform(np.log(hmin / npix), np.log(hmax / npix), n))
    m, rho, w, q = (r.uniform(0.5, 2.0, n) for _ in range(4))
    return [a.astype(dtype) for a in (x, h, m, rho, q, w)]


def test_mass_conservation_kat(s2g):
    """runtests.jl:324-342 through the GPU path."""
    npix, r = 200, 64
    par = s2g.mappingParameters(x_lim=[-r, r], y_lim=[-r, r], z_lim=[-r, r], Npixels=npix)
    pos = np.array([[0.0101], [-0.001], [0.001]])
    hs = np.array([5.0])
    mass = np.array([3.0])
    rho = np.ones(1)
    w = s2g.part_weight_physical(1, par, 1)
    mp = s2g.sphMapping(pos, hs, mass, rho, rho, w, param=par, dimensions=3,
                        kernel=s2g.Cubic(), reduce_image=False, show_progress=False)
    Vpix = (par.x_lim[1] - par.x_lim[0]) ** 3 / npix ** 3
    assert mp.shape == (200, 200, 200)
    assert np.isclose(mass.sum(), Vpix * mp.sum(), rtol=1.5e-8)


@pytest.mark.parametrize("kernel", KNAMES)
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_flat_image_kernels_dtypes(s2g, orc, kernel, dtype):
    npix, n = 40, 700
    x, h, m, rho, q, w = _particles(n, npix, 300 + KNAMES.index(kernel), dtype=dtype)
    st = {}
    gpu = s2g.deposit(x, h, m, rho, q, w, npix=npix, len2pix=float(npix), dimensions=3,
                      kernel=_kernel(s2g, kernel), calc_mean=False, stats=st)
    ref = orc.cic_mapping_3D(x, h, m, rho, q, w, npix, float(npix), kernel, 3, False)
    print(kernel, dtype.__name__, st, flush=True)
    assert st["n_small"] > 0 and st["n_large"] > 0
    _check(gpu, ref, f"{kernel}/{dtype.__name__}")


def test_zero_quantity_skip_and_calc_mean(s2g, orc):
    npix, n = 32, 500
    x, h, m, rho, q, w = _particles(n, npix, 17)
    q[::3] = 0.0
    for cm in (False, True):
        gpu = s2g.deposit(x, h, m, rho, q, w, npix=npix, len2pix=float(npix), dimensions=3,
                          kernel=s2g.WendlandC6(3), calc_mean=cm)
        ref = orc.cic_mapping_3D(x, h, m, rho, q, w, npix, float(npix), "wendlandc6", 3, cm)
        _check(gpu, ref, f"calc_mean={cm}")


def test_exact_norm_matches_closed_form(s2g):
    npix, n = 96, 200
    x, h, m, rho, q, w = _particles(n, npix, 5, hmin=15.0, hmax=30.0)
    x *= 0.3
    sa, sb = {}, {}
    a = s2g.deposit(x, h, m, rho, q, w, npix=npix, len2pix=float(npix), dimensions=3,
                    kernel=s2g.WendlandC6(3), stats=sa)
    b = s2g.deposit(x, h, m, rho, q, w, npix=npix, len2pix=float(npix), dimensions=3,
                    kernel=s2g.WendlandC6(3), exact_norm=True, stats=sb)
    assert sa["n_large_exact"] < sa["n_large"] == sb["n_large_exact"]
    assert rel_l1(a[:, 1], b[:, 1]) < 1e-7


@pytest.mark.parametrize("kernel", KNAMES)
def test_clipped_footprints_border_sum(s2g, orc, kernel):
    """Large footprints cut by the grid border: closed form minus the cut-off
    part where that part is the smaller one, else the direct sum; both equal
    the exact sum (S2G_FLAG_EXACT_NORM) and the oracle."""
    npix, n = 128, 60
    r = np.random.default_rng(500 + KNAMES.index(kernel))
    x = r.uniform(-0.25, 0.25, (n, 3))
    side = r.integers(0, 3, n)
    x[np.arange(n), side] = np.where(r.random(n) < 0.5, 1, -1) * r.uniform(0.36, 0.49, n)
    h = r.uniform(41.0, 50.0, n) / npix  # >= every kernel's 3D closed-form threshold
    m, rho, w, q = (r.uniform(0.5, 2.0, n) for _ in range(4))
    sa, sb = {}, {}
    a = s2g.deposit(x, h, m, rho, q, w, npix=npix, len2pix=float(npix), dimensions=3,
                    kernel=_kernel(s2g, kernel), calc_mean=False, stats=sa)
    b = s2g.deposit(x, h, m, rho, q, w, npix=npix, len2pix=float(npix), dimensions=3,
                    kernel=_kernel(s2g, kernel), calc_mean=False, exact_norm=True, stats=sb)
    assert sa["n_large_exact"] == n, sa
    for k in range(2):
        e = rel_l1(a[:, k], b[:, k])
        print(kernel, "plane", k, "border sum vs exact sum rel_l1", e, flush=True)
        # the closed form's own error (< 2e-9) and the FP32 weights' bias,
        # relative to the kept part: the bound of test_exact_norm_matches_closed_form
        assert e <= 1e-7, (kernel, k, e)
    ref = orc.cic_mapping_3D(x, h, m, rho, q, w, npix, float(npix), kernel, 3, False)
    _check(a, ref, f"border3d/{kernel}")


def test_large_only_is_deterministic(s2g):
    npix, n = 48, 150
    x, h, m, rho, q, w = _particles(n, npix, 8, hmin=4.0, hmax=12.0)
    a = s2g.deposit(x, h, m, rho, q, w, npix=npix, len2pix=float(npix), dimensions=3,
                    kernel=s2g.WendlandC4(3))
    b = s2g.deposit(x, h, m, rho, q, w, npix=npix, len2pix=float(npix), dimensions=3,
                    kernel=s2g.WendlandC4(3))
    np.testing.assert_array_equal(a, b)


def test_reduce_3d_bitexact(s2g, orc):
    from sph2grid import _lib
    ctx = s2g.default_context()
    r = np.random.default_rng(4)
    for npix in (1, 7, 33):
        img = np.asfortranarray(r.uniform(-1, 2, (npix ** 3, 2)))
        img[::4, 1] = 0.0
        for reduce in (True, False):
            out = np.empty(npix ** 3)
            rc = ctx.lib.s2g_reduce(ctx.handle, 3, npix, 1, _lib.S2G_FLAG_REDUCE if reduce else 0,
                                    img.ctypes.data_as(C.c_void_p), out.ctypes.data_as(C.c_void_p))
            ctx.check(rc, "s2g_reduce")
            ref_img = img.copy(order="F")
            if not reduce:
                ref_img[:, 1] = 1.0
            np.testing.assert_array_equal(out.reshape((npix,) * 3, order="F"),
                                          orc.reduce_image_3D(ref_img, npix))


def test_sphmapping_3d_cluster_scaled(s2g, orc):
    """Cluster-like set at 2^15 particles on a 64^3 grid through sphMapping."""
    from sph2grid import workloads
    d = workloads.cluster(n=1 << 15, npix=64)
    par = s2g.mappingParameters(x_lim=d["lims"][0], y_lim=d["lims"][1], z_lim=d["lims"][2],
                                Npixels=64)
    st = {}
    gpu = s2g.sphMapping(d["Pos"].copy(), d["HSML"], d["M"], d["Rho"], d["Bin_Quant"],
                         d["Weights"], param=par, kernel=s2g.WendlandC6(3), dimensions=3,
                         show_progress=False, stats=st)
    opar = orc.mapping_parameters(x_lim=d["lims"][0], y_lim=d["lims"][1], z_lim=d["lims"][2],
                                  Npixels=64)
    ref = orc.sph_mapping(np.ascontiguousarray(d["Pos"].T), d["HSML"], d["M"], d["Rho"],
                          d["Bin_Quant"], d["Weights"], opar, "wendlandc6", kernel_dim=3,
                          dimensions=3, nworkers=8)
    print("cluster 3D:", st, "rel_l1", rel_l1(gpu, ref), flush=True)
    assert gpu.shape == ref.shape == (64, 64, 64)
    assert rel_l1(gpu, ref) <= 1e-5
    assert max_rel(gpu, ref) <= 1e-4
