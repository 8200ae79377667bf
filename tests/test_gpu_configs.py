"""GPU tests at the full size of the batched and the all-sky workloads, where
the oracle cannot run over everything in test time:

* the batched set (64 sedov-like maps, Nq = 3, sides 128-2048, about 5e7
  particles) through one s2g_deposit_batch call: the five maps with the
  fewest footprint pixels against the oracle (rel-L1 <= 1e-5 per plane), and
  conservation of every map's weight plane against
  sum(w m / rho) * len2pix^3 (<= 1e-6);
* TSC on a 512^3 grid of the all-sky particle set (2^22 particles): the sums
  of both planes against their closed forms (each particle deposits
  prod_d S_d, S_d the sum of its axis weights on cells inside the grid, which
  is 1 away from the border), and a 2^14-particle subsample against the
  oracle cell by cell.
"""
import ctypes as C

import numpy as np
import pytest

from conftest import max_rel, rel_l1

pytestmark = pytest.mark.gpu


def test_batched_64_maps(s2g, orc):
    import torch
    from sph2grid import _lib, workloads
    from sph2grid.mapping import _prepare
    dev = torch.device("cuda:0")
    maps = workloads.batched(nmaps=64)
    descs, keep, prep = [], [], []
    for d in maps:
        par = s2g.mappingParameters(x_lim=d["lims"][0], y_lim=d["lims"][1], z_lim=d["lims"][2],
                                    Npixels=d["Npixels"])
        pm = _prepare(d["Pos"].copy(), d["HSML"], d["M"], d["Rho"], d["Bin_Quant"],
                      d["Weights"], param=par, kernel=s2g.WendlandC6(2), show_progress=False,
                      calc_mean=True)
        t = [torch.from_numpy(a).to(dev) for a in pm.arrays]
        out = torch.empty(pm.npix ** 2 * (pm.nq + 1), dtype=torch.float64, device=dev)
        keep.append((t, out))
        prep.append(pm)
        descs.append(dict(pos=t[0], hsml=t[1], m=t[2], rho=t[3], bin_q=t[4], weights=t[5],
                          out=out, n=pm.n, nq=pm.nq, npix=pm.npix, len2pix=pm.params.len2pix,
                          kernel=pm.kid, calc_mean=True))
    st = {}
    s2g.deposit_batch_device(descs, dtype="f32", stats=st)
    torch.cuda.synchronize()
    ntot = sum(pm.n for pm in prep)
    print(f"64 maps, {ntot} particles: {st['ms_device']:.1f} ms, small {st['n_small']} "
          f"large {st['n_large']} pairs {st['n_pairs']}", flush=True)
    assert st["n_particles"] == ntot and ntot > 10_000_000
    # conservation of every map's weight plane
    worst = 0.0
    fps = []
    for pm, (t, out) in zip(prep, keep):
        x, h, m, rho, bq, w = (a.astype(np.float64) for a in pm.arrays)
        L = pm.params.len2pix
        expect = float((w * m / rho).sum() * L ** 3)
        got = float(out.view(pm.nq + 1, -1)[pm.nq].sum().item())
        worst = max(worst, abs(got / expect - 1))
        hp = h * L
        xs, ys = x[:, 0] * L + pm.npix / 2, x[:, 1] * L + pm.npix / 2
        wi = np.minimum(np.floor(xs + hp), pm.npix - 1) - np.maximum(np.floor(xs - hp), 0) + 1
        wj = np.minimum(np.floor(ys + hp), pm.npix - 1) - np.maximum(np.floor(ys - hp), 0) + 1
        fps.append(float((np.maximum(wi, 0) * np.maximum(wj, 0)).sum()))
    print(f"worst weight conservation error over 64 maps: {worst:.2e}", flush=True)
    assert worst <= 1e-6
    # the five cheapest maps against the oracle, plane by plane
    for k in np.argsort(fps)[:5]:
        pm = prep[int(k)]
        x, h, m, rho, bq, w = (a.astype(np.float64) for a in pm.arrays)
        ref = orc.cic_mapping_2D(x, h, m, rho, bq, w, pm.npix, pm.params.len2pix, "wendlandc6",
                                 2, True, nworkers=8)
        got = keep[int(k)][1].view(pm.nq + 1, -1).T.cpu().numpy()
        for p in range(pm.nq + 1):
            e1, em = rel_l1(got[:, p], ref[:, p]), max_rel(got[:, p], ref[:, p])
            print(f"map {int(k)} (n={pm.n}, {pm.npix}^2) plane {p}: rel_l1 {e1:.2e} "
                  f"max_rel {em:.2e}", flush=True)
            assert e1 <= 1e-5 and em <= 1e-4


def _axis_sums(g, npix):
    """Per particle and axis: the sum of the three TSC weights whose cells
    lie in [0, npix) (1 when all three do)."""
    c = np.floor(g)
    s = np.zeros_like(g)
    for o in (-1, 0, 1):
        d = np.abs(g - (c + o + 0.5))
        w = np.where(d < 0.5, 0.75 - d * d, np.where(d < 1.5, 0.5 * (1.5 - d) ** 2, 0.0))
        cell = c + o
        s += np.where((cell >= 0) & (cell < npix), w, 0.0)
    return s


def test_tsc_allsky_512(s2g, orc):
    import torch
    from sph2grid import _lib, workloads
    d = workloads.allsky()
    n, N = d["HSML"].shape[0], 512
    L = N / 2.0  # grid over [-1, 1]^3
    pos = np.ascontiguousarray(d["Pos"].T)
    q = d["T"]
    dev = torch.device("cuda:0")
    dpos, dq = torch.from_numpy(pos).to(dev), torch.from_numpy(q).to(dev)
    grid = torch.empty(2 * N ** 3, dtype=torch.float64, device=dev)
    ctx = s2g.default_context()
    par = _lib.Params(npixels=N, len2pix=L)
    vp = lambda t: C.c_void_p(t.data_ptr())
    ctx.check(ctx.lib.s2g_tsc3d_device(ctx.handle, _lib.S2G_F64, vp(dpos), vp(dq), n,
                                       C.byref(par), 0, vp(grid), None), "s2g_tsc3d_device")
    g = pos * L + 0.5 * N
    S = _axis_sums(g, N).prod(axis=1)
    gw = grid.view(2, -1)
    sw, sq = float(gw[1].sum().item()), float(gw[0].sum().item())
    ew, eq = float(S.sum()), float((q * S).sum())
    print(f"TSC 512^3, 2^22: sum W {sw:.10e} (closed form {ew:.10e}), sum qW rel err "
          f"{abs(sq / eq - 1):.2e}", flush=True)
    assert abs(sw / ew - 1) <= 1e-12 and abs(sq / eq - 1) <= 1e-12
    assert float(gw[1].min().item()) >= 0.0
    # a 2^14 subsample cell by cell against the oracle
    sub = np.sort(np.random.default_rng(9).choice(n, 1 << 14, replace=False))
    # named tensors: a temporary's memory could be reused before the kernel reads it
    spos, sqq = dpos[sub].contiguous(), dq[sub].contiguous()
    ctx.check(ctx.lib.s2g_tsc3d_device(ctx.handle, _lib.S2G_F64, vp(spos), vp(sqq), sub.size,
                                       C.byref(par), 0, vp(grid), None), "s2g_tsc3d_device")
    ref = orc.tsc_3d(pos[sub], q[sub], N, L, False)
    nzr = np.flatnonzero(ref[:, 1])
    got = grid.view(2, -1).T
    nzg = torch.nonzero(got[:, 1]).flatten().cpu().numpy()
    np.testing.assert_array_equal(np.sort(nzg), nzr)
    gv = got[torch.from_numpy(nzr).to(dev)].cpu().numpy()
    np.testing.assert_allclose(gv, ref[nzr], rtol=1e-12, atol=0)
