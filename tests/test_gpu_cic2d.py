"""GPU parity of the 2D deposit (cic_mapping_2D + reduce_image_2D) against
the FP64 oracle, through the C ABI.  Tolerances (FP32 per-pixel kernel
weights, FP64 accumulation):
  rel-L1 <= 1e-5 (the target of the workload description),
  max relative pixel error <= 1e-4 over pixels above 1e-3 of the maximum,
  conservation (sum of the weight plane) <= 1e-6 relative.
"""
import numpy as np
import pytest

from conftest import max_rel, rel_l1

pytestmark = pytest.mark.gpu

KNAMES = ["cubic", "quintic", "wendlandc2", "wendlandc4", "wendlandc6", "wendlandc8"]
RTOL_L1 = 1e-5
RTOL_MAX = 1e-4
RTOL_CONS = 1e-6


def _kernel(s2g, name, dim=2):
    return {"cubic": s2g.Cubic, "quintic": s2g.Quintic, "wendlandc2": s2g.WendlandC2,
            "wendlandc4": s2g.WendlandC4, "wendlandc6": s2g.WendlandC6,
            "wendlandc8": s2g.WendlandC8}[name](dim)


def _check_flat(gpu, ref, tag=""):
    print(f"[{tag}] rel_l1 per plane:",
          [f"{rel_l1(gpu[:, k], ref[:, k]):.2e}" for k in range(ref.shape[1])], flush=True)
    for k in range(ref.shape[1]):
        assert rel_l1(gpu[:, k], ref[:, k]) <= RTOL_L1, (tag, k)
        assert max_rel(gpu[:, k], ref[:, k]) <= RTOL_MAX, (tag, k)
    w_g, w_r = gpu[:, -1].sum(), ref[:, -1].sum()
    if w_r:
        assert abs(w_g / w_r - 1) <= RTOL_CONS, tag


def _random_particles(n, npix, seed, hmin_px=0.05, hmax_px=60.0, dtype=np.float64):
    r = np.random.default_rng(seed)
    x = r.uniform(-0.5, 0.5, (n, 3))
    h = np.exp(r.uniform(np.log(hmin_px / npix), np.log(hmax_px / npix), n))
    m, rho, w = (r.uniform(0.5, 2.0, n) for _ in range(3))
    return [a.astype(dtype) for a in (x, h, m, rho, w)]


def test_fallback_conservation_kat(s2g):
    """runtests.jl:718-738 through the GPU path."""
    npix, r = 4, 10
    par = s2g.mappingParameters(x_lim=[-r, r], y_lim=[-r, r], z_lim=[-r, r], Npixels=npix)
    pos = np.array([[0.5], [0.0], [0.0]])
    one = np.ones(1)
    w = s2g.part_weight_physical(1, par, 1)
    mp = s2g.sphMapping(pos, one, one, one, one, w, param=par, kernel=s2g.Cubic(),
                        reduce_image=False, show_progress=False)
    Apix = (par.x_lim[1] - par.x_lim[0]) ** 2 / npix ** 2
    assert np.isclose(1.0, Apix * mp.sum(), rtol=1e-10)


@pytest.mark.parametrize("kernel", KNAMES)
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_flat_image_kernels_dtypes(s2g, orc, kernel, dtype):
    npix, n = 96, 1500
    x, h, m, rho, w = _random_particles(n, npix, seed=100 + KNAMES.index(kernel), dtype=dtype)
    q = np.random.default_rng(1).uniform(0.1, 5.0, n).astype(dtype)
    st = {}
    gpu = s2g.deposit(x, h, m, rho, q, w, npix=npix, len2pix=float(npix),
                      kernel=_kernel(s2g, kernel), calc_mean=True, stats=st)
    ref = orc.cic_mapping_2D(x, h, m, rho, q, w, npix, float(npix), kernel, 2, True)
    print(kernel, dtype.__name__, st, flush=True)
    assert st["n_small"] > 0 and st["n_large"] > 0
    _check_flat(gpu, ref, f"{kernel}/{dtype.__name__}")


@pytest.mark.parametrize("kernel", KNAMES)
def test_clipped_footprints_border_sum(s2g, orc, kernel):
    """Large footprints cut by the image border: the normalisation is the
    closed form minus the sum over the cut-off part where that part is the
    smaller one, else the direct sum.  Both must equal the exact sum
    (S2G_FLAG_EXACT_NORM) and the oracle."""
    npix, n = 160, 400
    r = np.random.default_rng(300 + KNAMES.index(kernel))
    x = r.uniform(-0.5, 0.5, (n, 3))
    side = r.integers(0, 2, n)
    x[np.arange(n), side] = np.where(r.random(n) < 0.5, 1, -1) * r.uniform(0.38, 0.499, n)
    h = r.uniform(70.0, 160.0, n) / npix  # >= every kernel's closed-form threshold
    m, rho, w = (r.uniform(0.5, 2.0, n) for _ in range(3))
    q = r.uniform(0.1, 5.0, n)
    st_a, st_b = {}, {}
    a = s2g.deposit(x, h, m, rho, q, w, npix=npix, len2pix=float(npix),
                    kernel=_kernel(s2g, kernel), calc_mean=True, stats=st_a)
    b = s2g.deposit(x, h, m, rho, q, w, npix=npix, len2pix=float(npix),
                    kernel=_kernel(s2g, kernel), calc_mean=True, exact_norm=True, stats=st_b)
    assert st_a["n_large_exact"] > n // 2, st_a
    for k in range(2):
        e = rel_l1(a[:, k], b[:, k])
        print(kernel, "plane", k, "border sum vs exact sum rel_l1", e, flush=True)
        assert e <= 1e-8, (kernel, k, e)
    ref = orc.cic_mapping_2D(x, h, m, rho, q, w, npix, float(npix), kernel, 2, True)
    _check_flat(a, ref, f"border/{kernel}")


@pytest.mark.parametrize("kernel", KNAMES)
def test_small_footprints_conserve(s2g, orc, kernel):
    """Particles of 0.3-3 px, none clipped: each deposits exactly its weight
    (the sum over the weight plane equals the oracle's to 1e-12, including
    particles whose only pixel centres inside the kernel sit near its edge),
    and the pixels match the oracle."""
    npix, n = 64, 3000
    r = np.random.default_rng(400 + KNAMES.index(kernel))
    x = r.uniform(-0.4, 0.4, (n, 3))
    h = np.exp(r.uniform(np.log(0.3), np.log(3.0), n)) / npix
    m, rho, w = (r.uniform(0.5, 2.0, n) for _ in range(3))
    q = r.uniform(0.1, 5.0, n)
    gpu = s2g.deposit(x, h, m, rho, q, w, npix=npix, len2pix=float(npix),
                      kernel=_kernel(s2g, kernel), calc_mean=True)
    ref = orc.cic_mapping_2D(x, h, m, rho, q, w, npix, float(npix), kernel, 2, True)
    cons = abs(gpu[:, 1].sum() / ref[:, 1].sum() - 1)
    print(kernel, "weight-plane sum vs oracle", cons, "rel_l1", rel_l1(gpu[:, 1], ref[:, 1]),
          flush=True)
    assert cons <= 1e-12, (kernel, cons)
    _check_flat(gpu, ref, f"small/{kernel}")
    for k in range(2):  # FP32 weights: ~1e-6 at 1e-3 of the peak for the C2-like kernels
        assert max_rel(gpu[:, k], ref[:, k]) <= 1e-5, (kernel, k)


@pytest.mark.parametrize("nq", [1, 2, 3, 4, 6])
def test_multi_quantity_planes(s2g, orc, nq):
    npix, n = 80, 1200
    x, h, m, rho, w = _random_particles(n, npix, seed=200 + nq)
    q = np.random.default_rng(nq).uniform(-1.0, 4.0, (n, nq))
    q[::7] = 0.0  # all-zero rows: skipped without calc_mean
    for calc_mean in (False, True):
        gpu = s2g.deposit(x, h, m, rho, q, w, npix=npix, len2pix=float(npix),
                          kernel=s2g.WendlandC6(2), calc_mean=calc_mean)
        ref = orc.cic_mapping_2D(x, h, m, rho, q, w, npix, float(npix), "wendlandc6", 2,
                                 calc_mean)
        _check_flat(gpu, ref, f"nq={nq} calc_mean={calc_mean}")


def test_exact_norm_matches_closed_form(s2g):
    npix, n = 256, 800
    x, h, m, rho, w = _random_particles(n, npix, seed=7, hmin_px=20.0, hmax_px=90.0)
    x[:, :2] *= 0.4  # mostly unclipped
    q = np.ones(n)
    st_a, st_b = {}, {}
    a = s2g.deposit(x, h, m, rho, q, w, npix=npix, len2pix=float(npix),
                    kernel=s2g.WendlandC6(2), stats=st_a)
    b = s2g.deposit(x, h, m, rho, q, w, npix=npix, len2pix=float(npix),
                    kernel=s2g.WendlandC6(2), exact_norm=True, stats=st_b)
    assert st_a["n_large_exact"] < st_a["n_large"]
    assert st_b["n_large_exact"] == st_b["n_large"]
    assert rel_l1(a[:, 1], b[:, 1]) < 1e-7


def test_plain_workload_full_size(s2g, orc):
    """Plain workload: 65,536 particles, 256^2, Cubic, Float64, density map."""
    from sph2grid import workloads
    d = workloads.plain()
    par = s2g.mappingParameters(x_lim=d["lims"][0], y_lim=d["lims"][1], z_lim=d["lims"][2],
                                Npixels=d["Npixels"])
    opar = orc.mapping_parameters(x_lim=d["lims"][0], y_lim=d["lims"][1],
                                  z_lim=d["lims"][2], Npixels=d["Npixels"])
    pos_o = np.ascontiguousarray(d["Pos"].T)
    ref = orc.sph_mapping(pos_o, d["HSML"], d["M"], d["Rho"], d["Bin_Quant"], d["Weights"],
                          opar, "cubic", kernel_dim=2, reduce_image=True)
    st = {}
    gpu = s2g.sphMapping(d["Pos"].copy(), d["HSML"], d["M"], d["Rho"], d["Bin_Quant"],
                         d["Weights"], param=par, kernel=s2g.Cubic(2), show_progress=False,
                         stats=st)
    print("plain:", st, "rel_l1", rel_l1(gpu, ref), flush=True)
    assert gpu.shape == ref.shape == (256, 256, 1)
    assert rel_l1(gpu, ref) <= RTOL_L1
    assert max_rel(gpu, ref) <= RTOL_MAX


def test_return_both_maps_and_reduce(s2g, orc):
    from sph2grid import workloads
    d = workloads.sedov(n=40000, npix=128, nq=2, dtype=np.float32)
    par = s2g.mappingParameters(x_lim=d["lims"][0], y_lim=d["lims"][1], z_lim=d["lims"][2],
                                Npixels=d["Npixels"])
    flat = s2g.sphMapping(d["Pos"].copy(), d["HSML"], d["M"], d["Rho"], d["Bin_Quant"],
                          d["Weights"], param=par, kernel=s2g.WendlandC6(2),
                          show_progress=False, return_both_maps=True, calc_mean=True)
    red = s2g.sphMapping(d["Pos"].copy(), d["HSML"], d["M"], d["Rho"], d["Bin_Quant"],
                         d["Weights"], param=par, kernel=s2g.WendlandC6(2),
                         show_progress=False, calc_mean=True)
    opar = orc.mapping_parameters(x_lim=d["lims"][0], y_lim=d["lims"][1],
                                  z_lim=d["lims"][2], Npixels=d["Npixels"])
    pos_o = np.ascontiguousarray(d["Pos"].T)
    ref_flat = orc.sph_mapping(pos_o, d["HSML"], d["M"], d["Rho"], d["Bin_Quant"].T,
                               d["Weights"], opar, "wendlandc6", return_both_maps=True,
                               calc_mean=True)
    _check_flat(flat, ref_flat, "sedov flat")
    # reducing the GPU flat image on the CPU oracle gives the GPU reduced map
    # (two separate deposits: the atomic scatter order differs at the ulp level)
    np.testing.assert_allclose(orc.reduce_image_2D(flat, 128, True), red, rtol=1e-12)


def test_reduce_entry_point_bitexact(s2g, orc):
    """s2g_reduce (reduce_image.jl:8-31) is an exact transpose + IEEE division."""
    import ctypes as C
    from sph2grid import _lib
    ctx = s2g.default_context()
    r = np.random.default_rng(3)
    for npix, nq in ((1, 1), (70, 2), (129, 3)):
        img = r.uniform(-1, 2, (npix * npix, nq + 1))
        img[::5, -1] = 0.0
        img = np.asfortranarray(img)
        for reduce in (True, False):
            out = np.empty(npix * npix * nq)
            rc = ctx.lib.s2g_reduce(ctx.handle, 2, npix, nq,
                                    _lib.S2G_FLAG_REDUCE if reduce else 0,
                                    img.ctypes.data_as(C.c_void_p),
                                    out.ctypes.data_as(C.c_void_p))
            ctx.check(rc, "s2g_reduce")
            np.testing.assert_array_equal(out.reshape((npix, npix, nq), order="F"),
                                          orc.reduce_image_2D(img, npix, reduce))


def test_empty_and_all_filtered(s2g):
    par = s2g.mappingParameters(x_lim=[0, 1], y_lim=[0, 1], z_lim=[0, 1], Npixels=32)
    z = np.zeros(0)
    out = s2g.sphMapping(np.zeros((3, 0)), z, z, z, z, z, param=par, kernel=s2g.Cubic(),
                         show_progress=False)
    assert out.shape == (32, 32, 1) and not out.any()
    pos = np.full((3, 5), 3.0)  # all outside
    one = np.ones(5)
    out = s2g.sphMapping(pos, one, one, one, one, one, param=par, kernel=s2g.Cubic(),
                         show_progress=False)
    assert not out.any()


def test_large_only_is_deterministic(s2g):
    npix, n = 192, 600
    x, h, m, rho, w = _random_particles(n, npix, seed=9, hmin_px=17.0, hmax_px=70.0)
    q = np.random.default_rng(2).uniform(0, 1, n)
    a = s2g.deposit(x, h, m, rho, q, w, npix=npix, len2pix=float(npix), kernel=s2g.WendlandC4())
    b = s2g.deposit(x, h, m, rho, q, w, npix=npix, len2pix=float(npix), kernel=s2g.WendlandC4())
    np.testing.assert_array_equal(a, b)


def test_cluster_scaled(s2g, orc):
    """Cluster-like particle set (halo + background, h log-uniform over three
    decades) at 2^17 particles on 512^2."""
    from sph2grid import workloads
    d = workloads.cluster(n=1 << 17, npix=512)
    par = s2g.mappingParameters(x_lim=d["lims"][0], y_lim=d["lims"][1], z_lim=d["lims"][2],
                                Npixels=d["Npixels"])
    st = {}
    gpu = s2g.sphMapping(d["Pos"].copy(), d["HSML"], d["M"], d["Rho"], d["Bin_Quant"],
                         d["Weights"], param=par, kernel=s2g.WendlandC6(2), show_progress=False,
                         return_both_maps=True, stats=st)
    opar = orc.mapping_parameters(x_lim=d["lims"][0], y_lim=d["lims"][1],
                                  z_lim=d["lims"][2], Npixels=d["Npixels"])
    ref = orc.sph_mapping(np.ascontiguousarray(d["Pos"].T), d["HSML"], d["M"], d["Rho"],
                          d["Bin_Quant"], d["Weights"], opar, "wendlandc6",
                          return_both_maps=True, nworkers=8)
    print("cluster scaled:", st, flush=True)
    _check_flat(gpu, ref, "cluster 2^17")


@pytest.mark.parametrize("dims", [2, 3])
def test_degenerate_particles_deposit_nothing(s2g, orc, dims):
    """h = 0, h < 0 and NaN h deposit nothing; the other particles' image is
    unchanged.  In 3D the reference does the same for h = 0 (the footprint's
    overlap volume is 0, so every pix_weight is 0, cic_3D.jl:168-188).  In 2D
    it differs on purpose: area = (2h)^2 = 0 and dz = Inf make area_norm NaN,
    which poisons the particle's pixel (cic_2D.jl:187-222, reproduced by the
    oracle).  A NaN h throws InexactError in floor(Integer, x - hsml)
    (cic_shared.jl:46-52)."""
    npix, n = 48 if dims == 2 else 24, 400
    x, h, m, rho, w = _random_particles(n, npix, seed=31 + dims, hmin_px=0.3, hmax_px=12.0)
    q = np.random.default_rng(5).uniform(0.5, 2.0, n)
    bad = np.array([3, 50, 51, 200])
    h_bad = h.copy()
    h_bad[bad] = [0.0, -1.0 / npix, np.nan, 0.0]
    good = np.setdiff1d(np.arange(n), bad)
    kw = dict(npix=npix, len2pix=float(npix), kernel=s2g.WendlandC4(dims), dimensions=dims)
    gpu = s2g.deposit(x, h_bad, m, rho, q, w, **kw)
    gpu_good = s2g.deposit(x[good], h[good], m[good], rho[good], q[good], w[good], **kw)
    assert np.isfinite(gpu).all()
    assert rel_l1(gpu, gpu_good) <= 1e-14
    f = orc.cic_mapping_2D if dims == 2 else orc.cic_mapping_3D
    ref = f(x[good], h[good], m[good], rho[good], q[good], w[good], npix, float(npix),
            "wendlandc4", dims, True)
    _check_flat(gpu, ref, f"degenerate {dims}D")
    # the oracle follows the reference for h = 0 inside the image: NaN in
    # every plane in 2D, nothing in 3D
    ref_h0 = f(x[:4], np.zeros(4), m[:4], rho[:4], q[:4], w[:4], npix, float(npix),
               "wendlandc4", dims, True)
    if dims == 2:
        assert np.isnan(ref_h0).any(axis=0).all()
    else:
        assert not ref_h0.any()


@pytest.mark.parametrize("dims", [2, 3])
def test_lattice_aligned_ties(s2g, orc, dims):
    """Particles on pixel centres, edges and corners with h a whole or half
    number of pixels: footprint bounds floor(x -/+ h) land exactly on
    integers, neighbouring centres sit at u = 1 exactly, and h = 0.5 on a
    corner has every centre at u = 1 (the fallback).  Float64 inputs, so
    every coordinate is exact."""
    npix = 40 if dims == 2 else 20
    offs = [0.0, 0.5, 0.25]  # pixel corner, pixel centre, quarter
    hs = [0.5, 1.0, 1.5, 2.0, 4.0, 7.5] + ([16.0] if dims == 2 else [])
    rows = []
    r = np.random.default_rng(77 + dims)
    for hp in hs:
        for o in offs:
            c = r.integers(2, npix - 2, size=3) + o  # pixel coordinates
            rows.append((c, hp))
    n = len(rows)
    x = np.array([(c / npix) - 0.5 for c, _ in rows])
    h = np.array([hp / npix for _, hp in rows])
    m, rho, w = (r.uniform(0.5, 2.0, n) for _ in range(3))
    q = r.uniform(0.5, 2.0, n)
    for kname in ("cubic", "wendlandc6"):
        gpu = s2g.deposit(x, h, m, rho, q, w, npix=npix, len2pix=float(npix),
                          kernel=_kernel(s2g, kname, dims), dimensions=dims)
        f = orc.cic_mapping_2D if dims == 2 else orc.cic_mapping_3D
        ref = f(x, h, m, rho, q, w, npix, float(npix), kname, dims, True)
        _check_flat(gpu, ref, f"ties {dims}D {kname}")
        # with every footprint summed explicitly, the deposit conserves to
        # FP64 rounding; the default closed-form sum for large unclipped 3D
        # footprints is within RTOL_CONS (measured 2.6e-10 here)
        ex = s2g.deposit(x, h, m, rho, q, w, npix=npix, len2pix=float(npix),
                         kernel=_kernel(s2g, kname, dims), dimensions=dims, exact_norm=True)
        _check_flat(ex, ref, f"ties {dims}D {kname} exact")
        assert abs(ex[:, -1].sum() / ref[:, -1].sum() - 1) <= 1e-12


@pytest.mark.parametrize("dims", [2, 3])
def test_explicit_norm_matches_deposit(s2g, dims):
    """With every footprint summed explicitly (exact_norm), the normalisation
    pass forms each weight with the gather's FP32 operations in the gather's
    order, so every large particle deposits exactly w m / (rho / len2pix^3)
    (x len2pix in 3D) up to FP64 rounding; clipped and unclipped footprints,
    offsets / h not representable in FP32.  A different rounding of one FP32
    term (an FMA in one pass and not the other) shows as ~1e-9."""
    npix = 96 if dims == 2 else 40
    n = 300
    r = np.random.default_rng(900 + dims)
    x = r.uniform(-0.5, 0.5, (n, 3))
    h = (r.uniform(18.0, 45.0, n) if dims == 2 else r.uniform(4.6, 16.0, n)) / npix
    m, rho, w = (r.uniform(0.5, 2.0, n) for _ in range(3))
    q = r.uniform(0.5, 2.0, n)
    st = {}
    gpu = s2g.deposit(x, h, m, rho, q, w, npix=npix, len2pix=float(npix),
                      kernel=s2g.WendlandC6(dims), dimensions=dims, exact_norm=True, stats=st)
    assert st["n_large"] >= 0.9 * n  # a few clipped boxes fall to the small path
    exp = (w * m / rho).sum() * float(npix) ** (3 if dims == 2 else 4)
    cons = abs(gpu[:, -1].sum() / exp - 1)
    print(f"{dims}D explicit-norm conservation {cons:.2e}", flush=True)
    assert cons <= 1e-12


def test_deposit_argument_errors(s2g):
    """s2g_deposit / s2g_reduce reject bad arguments with the documented
    codes and a message in s2g_last_error, and the context stays usable."""
    import ctypes as C
    from sph2grid import _lib
    ctx = s2g.default_context()
    L = ctx.lib
    n, npix = 8, 16
    r = np.random.default_rng(3)
    x = np.ascontiguousarray(r.uniform(-0.4, 0.4, (n, 3)))
    h = np.full(n, 2.0 / npix)
    one = np.ones(n)
    img = np.zeros(2 * npix ** 3)
    P = lambda a: a.ctypes.data_as(C.c_void_p)

    def call(dims=2, dtype=_lib.S2G_F64, pos=True, nq=1, nn=n, par="ok", kernel=0,
             kdim=2, npx=npix, l2p=float(npix)):
        prm = _lib.Params(npixels=npx, len2pix=l2p) if par == "ok" else None
        return L.s2g_deposit(ctx.handle, dims, dtype, P(x) if pos else None, P(h), P(one),
                             P(one), P(one), nq, P(one), nn,
                             C.byref(prm) if prm is not None else None, kernel, kdim, 0,
                             P(img), None)

    cases = [
        (dict(dims=4), _lib.S2G_EINVAL, "dims"),
        (dict(nq=0), _lib.S2G_EINVAL, "nq"),
        (dict(nq=65), _lib.S2G_EINVAL, "nq"),
        (dict(dims=3, kdim=3, nq=2), _lib.S2G_EINVAL, "scalar"),
        (dict(dtype=7), _lib.S2G_EINVAL, "dtype"),
        (dict(kernel=99), _lib.S2G_EUNSUPPORTED, "kernel"),
        (dict(kdim=1), _lib.S2G_EINVAL, "kernel_dim"),
        (dict(nn=-1), _lib.S2G_EINVAL, "particle count"),
        (dict(par=None), _lib.S2G_EINVAL, "params"),
        (dict(npx=0), _lib.S2G_EINVAL, "npixels"),
        (dict(l2p=float("nan")), _lib.S2G_EINVAL, "len2pix"),
        (dict(pos=False), _lib.S2G_EINVAL, "NULL"),
    ]
    for kw, code, word in cases:
        rc = call(**kw)
        msg = L.s2g_last_error(ctx.handle).decode()
        print(kw, rc, msg, flush=True)
        assert rc == code, (kw, rc, msg)
        assert word in msg, (kw, msg)
    # the reduce entry point: bad dims and NULL image
    out = np.zeros(npix * npix)
    assert L.s2g_reduce(ctx.handle, 5, npix, 1, 0, P(img), P(out)) == _lib.S2G_EINVAL
    assert L.s2g_reduce(ctx.handle, 2, npix, 1, 0, None, P(out)) == _lib.S2G_EINVAL
    # still usable: a valid call after the errors matches the Python surface
    assert call() == _lib.S2G_OK
    ref = s2g.deposit(x, h, one, one, one, one, npix=npix, len2pix=float(npix),
                      kernel=s2g.Cubic(2))
    np.testing.assert_allclose(img[:2 * npix * npix], ref.ravel(order="F"), rtol=1e-12,
                               atol=0)


@pytest.mark.parametrize("nq", [1, 3])
def test_fp32_accumulation(s2g, orc, nq):
    """S2G_FLAG_FP32_ACCUM: the gather's per-item pixel sums in FP32 stay
    within the deposit's tolerances of the FP64 oracle."""
    npix, n = 256, 6000
    x, h, m, rho, w = _random_particles(n, npix, seed=77 + nq, hmin_px=5.0, hmax_px=120.0)
    q = np.random.default_rng(3).lognormal(0.0, 1.0, (n, nq))
    kern = _kernel(s2g, "wendlandc6")
    st = {}
    got = s2g.deposit(x, h, m, rho, q, w, npix=npix, len2pix=float(npix), kernel=kern,
                      calc_mean=True, fp32_accum=True, stats=st)
    f64 = s2g.deposit(x, h, m, rho, q, w, npix=npix, len2pix=float(npix), kernel=kern,
                      calc_mean=True)
    ref = orc.cic_mapping_2D(x, h, m, rho, q, w, npix, float(npix), "wendlandc6", 2, True)
    assert st["n_large"] > n // 2
    print("fp32 accumulation rel_l1 vs oracle:",
          [f"{rel_l1(got[:, k], ref[:, k]):.2e}" for k in range(nq + 1)],
          "fp64:", [f"{rel_l1(f64[:, k], ref[:, k]):.2e}" for k in range(nq + 1)], flush=True)
    _check_flat(got, ref, f"fp32acc/nq{nq}")
    for k in range(nq + 1):
        assert rel_l1(got[:, k], f64[:, k]) <= 1e-5


@pytest.mark.parametrize("dims", [2, 3])
def test_particle_order_does_not_change_the_map(s2g, dims):
    """A deposit is a sum over particles: the cluster-like set in its
    generated order, sorted in space (by the 64-pixel tile, or 16-voxel
    brick, of the centre: the order of a Peano-Hilbert-sorted snapshot, which
    the tile lists' stride order exists for) and randomly permuted must give
    the same image up to the order of the FP64 sums (1e-12 relative)."""
    from sph2grid import workloads
    npix = 512 if dims == 2 else 96
    d = workloads.cluster(n=1 << 16, npix=npix)
    par = s2g.mappingParameters(x_lim=d["lims"][0], y_lim=d["lims"][1], z_lim=d["lims"][2],
                                Npixels=npix)
    pos = d["Pos"]
    n = pos.shape[1]
    T = 64 if dims == 2 else 16
    t = np.clip(((pos - np.array(d["lims"])[:, :1]) /
                 (np.array(d["lims"])[:, 1:] - np.array(d["lims"])[:, :1]) * npix // T),
                0, npix // T - 1).astype(np.int64)
    key = (t[0] * (npix // T) + t[1]) * (npix // T) + t[2]
    orders = {"generated": np.arange(n), "sorted": np.argsort(key, kind="stable"),
              "permuted": np.random.default_rng(3).permutation(n)}
    q = d["Bin_Quant"]
    imgs = {}
    for name, o in orders.items():
        imgs[name] = s2g.sphMapping(pos[:, o].copy(), d["HSML"][o], d["M"][o], d["Rho"][o],
                                    q[o], d["Weights"][o], param=par,
                                    kernel=s2g.WendlandC6(dims), show_progress=False,
                                    return_both_maps=True, dimensions=dims)
    ref = imgs["generated"]
    for name in ("sorted", "permuted"):
        e = max_rel(imgs[name], ref)
        print(f"{dims}D {name} vs generated: max rel diff {e:.2e}", flush=True)
        assert e <= 1e-12, name
