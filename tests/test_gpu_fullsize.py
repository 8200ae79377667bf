"""The headline cluster workload at full size (2^24 particles, WendlandC6,
Float32 inputs resident on the device), through s2g_deposit_device on a
4096^2 map and a 256^3 grid.  The oracle cannot finish these sizes in test
time, so the checks are size-independent properties of the deposit:

* conservation -- every particle deposits exactly w m / (rho / len2pix^3)
  into the weight plane (2D, cic_2D.jl:187-222: area_norm * sum(wk A) =
  area * w * dz) and w m / (rho / len2pix^3) * len2pix in 3D
  (cic_3D.jl:199-236), clipped or not, and q times that into the value
  plane.  Tolerance 1e-6 relative: the closed-form kernel sum used for
  unclipped large footprints agrees with the discrete sum to about 1e-7.
* linearity -- the deposit of the whole set equals the sum of the deposits
  of two interleaved halves (every per-particle decision depends on that
  particle only; the halves change every chunk, pair list and tile
  schedule).  FP64 accumulation of positive terms in another order:
  rel-L1 <= 1e-12, max relative pixel error <= 1e-10 over pixels above
  1e-6 of the maximum.  A lost update or a dropped chunk fails this.
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

RTOL_CONS = 1e-6
RTOL_LIN_L1 = 1e-12
RTOL_LIN_MAX = 1e-10


@pytest.fixture(scope="module")
def cluster(s2g):
    import torch
    from sph2grid import workloads
    import bench
    p = bench.prepare(workloads.cluster(), s2g)
    assert p["x"].dtype == np.float32 and p["nq"] == 1
    assert p["kernel"] == "wendlandc6" and not p["calc_mean"]
    dev = torch.device("cuda:0")
    arrs = {k: torch.from_numpy(np.ascontiguousarray(p[k])).to(dev)
            for k in ("x", "hsml", "m", "rho", "w", "bq")}
    x64 = {k: p[k].astype(np.float64) for k in ("m", "rho", "w", "bq")}
    live = x64["bq"] != 0.0  # calc_mean is off: zero-quantity particles are skipped
    per = np.where(live, x64["w"] * x64["m"] / x64["rho"], 0.0)
    return dict(arrs=arrs, n=p["x"].shape[0], w_sum=float(per.sum()),
                q_sum=float((per * x64["bq"]).sum()), torch=torch)


def _deposit(s2g, c, dims, npix, idx=None):
    import bench
    from sph2grid import _lib
    torch = c["torch"]
    a = c["arrs"]
    if idx is not None:
        a = {k: v.index_select(0, idx).contiguous() for k, v in a.items()}
    n = a["hsml"].shape[0]
    img = torch.empty(2 * npix ** dims, dtype=torch.float64, device=a["hsml"].device)
    ctx = s2g.default_context()
    par = _lib.Params(npixels=npix, len2pix=float(npix))
    vp = lambda t: C.c_void_p(t.data_ptr())
    st = _lib.Stats()
    ctx.check(ctx.lib.s2g_deposit_device(ctx.handle, dims, _lib.S2G_F32, vp(a["x"]),
                                         vp(a["hsml"]), vp(a["m"]), vp(a["rho"]),
                                         vp(a["bq"]), 1, vp(a["w"]), n, C.byref(par),
                                         bench.KIDS["wendlandc6"], dims, 0, vp(img), None,
                                         C.byref(st)), f"{dims}D deposit")
    return img.view(2, -1)


def _check(s2g, c, dims, npix):
    torch = c["torch"]
    scale = float(npix) ** (3 if dims == 2 else 4)
    print(f"[cluster {c['n']} particles, {dims}D, {npix}^{dims}] whole set", flush=True)
    whole = _deposit(s2g, c, dims, npix)
    cons_w = abs(whole[1].sum().item() / (c["w_sum"] * scale) - 1)
    cons_q = abs(whole[0].sum().item() / (c["q_sum"] * scale) - 1)
    print(f"  conservation: weight {cons_w:.2e} value {cons_q:.2e}", flush=True)
    assert cons_w <= RTOL_CONS and cons_q <= RTOL_CONS
    dev = whole.device
    even = torch.arange(0, c["n"], 2, device=dev)
    odd = torch.arange(1, c["n"], 2, device=dev)
    print("  even/odd halves", flush=True)
    parts = _deposit(s2g, c, dims, npix, even) + _deposit(s2g, c, dims, npix, odd)
    for k in range(2):
        diff = (parts[k] - whole[k]).abs()
        l1 = (diff.sum() / whole[k].abs().sum()).item()
        big = whole[k].abs() > 1e-6 * whole[k].abs().max()
        mx = (diff[big] / whole[k].abs()[big]).max().item()
        print(f"  linearity plane {k}: rel-L1 {l1:.2e} max-rel {mx:.2e}", flush=True)
        assert l1 <= RTOL_LIN_L1 and mx <= RTOL_LIN_MAX, (dims, k, l1, mx)


def test_cluster_2d_4096(s2g, cluster):
    _check(s2g, cluster, 2, 4096)


def test_cluster_3d_256(s2g, cluster):
    _check(s2g, cluster, 3, 256)
