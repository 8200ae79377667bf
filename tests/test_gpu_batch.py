"""GPU tests of the batched 2D deposit (s2g_deposit_batch): a batch must give
each map what a separate deposit of that map gives (same kernels, different
work split: <= 1e-12 relative), and the oracle's result for the map
(rel-L1 <= 1e-5, as the single-map tests).  Covers mixed kernels and nq
(separate passes), sides that are not multiples of the 64-pixel tile, empty
maps, flat and reduced outputs, Float32 and Float64, and the argument checks."""
import ctypes as C

import numpy as np
import pytest

from conftest import max_rel, rel_l1

pytestmark = pytest.mark.gpu


def _particles(n, seed, hmin_px, hmax_px, npix, nq, dtype):
    r = np.random.default_rng(seed)
    pos = r.uniform(0.0, 1.0, (3, n))
    h = np.exp(r.uniform(np.log(hmin_px / npix), np.log(hmax_px / npix), n))
    m, rho, w = (r.uniform(0.5, 2.0, n) for _ in range(3))
    q = r.uniform(0.1, 5.0, (nq, n)) if nq > 1 else r.uniform(0.1, 5.0, n)
    c = lambda a: np.asarray(a, dtype=dtype)
    return dict(Pos=c(pos), HSML=c(h), M=c(m), Rho=c(rho), Bin_Quant=c(q), Weights=c(w))


def _calls(s2g, dtype=np.float64):
    spec = [  # (n, npix, nq, kernel, reduce, calc_mean, hmin_px, hmax_px)
        (900, 96, 1, s2g.WendlandC6(2), True, False, 0.1, 40.0),
        (0, 64, 1, s2g.WendlandC6(2), True, False, 0.1, 10.0),
        (1500, 130, 2, s2g.Cubic(2), False, True, 0.05, 70.0),
        (400, 17, 1, s2g.WendlandC6(2), False, True, 0.5, 30.0),
        (2000, 200, 3, s2g.WendlandC6(2), True, True, 0.2, 90.0),
        (700, 64, 2, s2g.Cubic(2), True, False, 0.1, 20.0),
        (1200, 150, 1, s2g.WendlandC6(2), True, True, 0.05, 120.0),
    ]
    calls = []
    for k, (n, npix, nq, ker, red, cm, hmin, hmax) in enumerate(spec):
        d = _particles(n, 700 + k, hmin, hmax, npix, nq, dtype)
        par = s2g.mappingParameters(x_lim=[0.0, 1.0], y_lim=[0.0, 1.0], z_lim=[0.0, 1.0],
                                    Npixels=npix)
        calls.append(dict(d, param=par, kernel=ker, reduce_image=red,
                          return_both_maps=not red, calc_mean=cm))
    return calls


def _copy_call(c):
    return {k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in c.items()}


def _oracle_result(orc, c):
    """The oracle's result for one call, laid out as sphMapping returns it."""
    npix = int(c["param"].Npixels[0])
    x = (c["Pos"].astype(np.float64) - 0.5).T.copy()
    q = c["Bin_Quant"].T.copy() if c["Bin_Quant"].ndim > 1 else c["Bin_Quant"]
    kname = type(c["kernel"]).__name__.lower()
    flat = orc.cic_mapping_2D(x, c["HSML"], c["M"], c["Rho"], q, c["Weights"], npix,
                              float(npix), kname, 2, c["calc_mean"])
    if c["return_both_maps"]:
        return flat
    return orc.reduce_image_2D(flat, npix, c["reduce_image"])


def _where(a, b):
    """Planes and pixel box where a and b differ by more than 1e-6 relative."""
    bad = np.abs(a - b) > 1e-6 * np.abs(b).max()
    if not bad.any():
        return "none"
    idx = np.argwhere(bad)
    return (f"{bad.sum()} values, index ranges "
            f"{[(int(idx[:, d].min()), int(idx[:, d].max())) for d in range(idx.shape[1])]}")


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_batch_matches_single_maps(s2g, orc, dtype):
    calls = _calls(s2g, dtype)
    st = {}
    got = s2g.map_batch([_copy_call(c) for c in calls], stats=st)
    print("batch stats", st, flush=True)
    assert st["n_particles"] == sum(c["M"].shape[0] for c in calls)
    for k, c in enumerate(calls):
        kw = _copy_call(c)
        args = [kw.pop(x) for x in ("Pos", "HSML", "M", "Rho", "Bin_Quant", "Weights")]
        ref = s2g.sphMapping(*args, show_progress=False, **kw)
        assert got[k].shape == ref.shape, k
        tol = 1e-300 + 1e-13 * np.abs(ref).max()
        if not np.allclose(got[k], ref, rtol=1e-12, atol=tol):
            # name the side that departs from the FP64 oracle, and where
            o = _oracle_result(orc, c)
            print(f"map {k}: batch vs oracle: {_where(got[k], o)}; "
                  f"single vs oracle: {_where(ref, o)}", flush=True)
        np.testing.assert_allclose(got[k], ref, rtol=1e-12, atol=tol, err_msg=f"map {k}")


def test_batch_vs_oracle_flat(s2g, orc):
    calls = _calls(s2g, np.float64)
    got = s2g.map_batch([dict(_copy_call(c), reduce_image=False, return_both_maps=True)
                         for c in calls])
    for k, c in enumerate(calls):
        n = c["M"].shape[0]
        if n == 0:
            assert not np.any(got[k]), k
            continue
        npix = int(c["param"].Npixels[0])
        x = (c["Pos"] - 0.5).T.copy()  # centred on (0.5, 0.5, 0.5); every particle is inside
        q = c["Bin_Quant"].T.copy() if c["Bin_Quant"].ndim > 1 else c["Bin_Quant"]
        kname = type(c["kernel"]).__name__.lower()
        ref = orc.cic_mapping_2D(x, c["HSML"], c["M"], c["Rho"], q, c["Weights"], npix,
                                 float(npix), kname, 2, c["calc_mean"])
        for p in range(ref.shape[1]):
            e1, em = rel_l1(got[k][:, p], ref[:, p]), max_rel(got[k][:, p], ref[:, p])
            print(f"map {k} plane {p}: rel_l1 {e1:.2e} max_rel {em:.2e}", flush=True)
            assert e1 <= 1e-5 and em <= 1e-4, (k, p)


def test_batch_workload_scaled(s2g):
    """The batched workload (sedov-like maps, Nq = 3, Weights = Rho, sides
    128-2048) at 1/200 of the particles, 12 maps, against per-map sphMapping."""
    from sph2grid import workloads
    maps = workloads.batched(nmaps=12, scale=1.0 / 200)
    calls = []
    for d in maps:
        par = s2g.mappingParameters(x_lim=d["lims"][0], y_lim=d["lims"][1], z_lim=d["lims"][2],
                                    Npixels=d["Npixels"])
        calls.append(dict(Pos=d["Pos"], HSML=d["HSML"], M=d["M"], Rho=d["Rho"],
                          Bin_Quant=d["Bin_Quant"], Weights=d["Weights"], param=par,
                          kernel=s2g.WendlandC6(2), reduce_image=True, calc_mean=True))
    st = {}
    got = s2g.map_batch([_copy_call(c) for c in calls], stats=st)
    print("batched workload stats", st, flush=True)
    for k, c in enumerate(calls):
        kw = _copy_call(c)
        args = [kw.pop(x) for x in ("Pos", "HSML", "M", "Rho", "Bin_Quant", "Weights")]
        ref = s2g.sphMapping(*args, show_progress=False, **kw)
        err = np.abs(got[k] - ref).max() / max(np.abs(ref).max(), 1e-300)
        assert err <= 1e-12, (k, err)


def test_batch_repeat_and_empty(s2g):
    """Repeated batches reuse the context's buffers; an empty batch is a no-op."""
    calls = _calls(s2g, np.float32)
    a = s2g.map_batch([_copy_call(c) for c in calls])
    b = s2g.map_batch([_copy_call(c) for c in calls[2:4]])
    c2 = s2g.map_batch([_copy_call(c) for c in calls])
    for k in range(len(calls)):
        np.testing.assert_allclose(a[k], c2[k], rtol=1e-12, atol=1e-13 * np.abs(a[k]).max())
    assert len(b) == 2
    assert s2g.map_batch([]) == []
    s2g.deposit_batch_device([])


def test_batch_argument_errors(s2g):
    import torch
    from sph2grid import _lib
    dev = torch.device("cuda:0")
    n, npix = 10, 32
    t = lambda k: torch.zeros(k, dtype=torch.float32, device=dev)
    good = dict(pos=t(3 * n), hsml=t(n), m=t(n), rho=t(n), bin_q=t(n), weights=t(n),
                out=torch.zeros(2 * npix * npix, dtype=torch.float64, device=dev), n=n, nq=1,
                npix=npix, len2pix=float(npix), kernel=s2g.Cubic(2))
    s2g.deposit_batch_device([good])
    ctx = s2g.default_context()
    with pytest.raises(_lib.LibraryError) as e:
        s2g.deposit_batch_device([good, dict(good, out=0)])
    assert e.value.code == _lib.S2G_EINVAL and "map 1" in str(e.value)
    with pytest.raises(_lib.LibraryError) as e:
        s2g.deposit_batch_device([dict(good, kernel=17)])
    assert e.value.code == _lib.S2G_EUNSUPPORTED
    with pytest.raises(_lib.LibraryError) as e:
        s2g.deposit_batch_device([dict(good, npix=0)])
    assert e.value.code == _lib.S2G_EINVAL
    d = (_lib.MapDesc * 1)()
    d[0] = _lib.MapDesc(dims=3, nq=1, n=0, params=_lib.Params(npixels=8, len2pix=8.0),
                        kernel=0, flags=0, d_out=good["out"].data_ptr())
    rc = ctx.lib.s2g_deposit_batch(ctx.handle, _lib.S2G_F32, d, 1, None, None)
    assert rc == _lib.S2G_EUNSUPPORTED
    rc = ctx.lib.s2g_deposit_batch(ctx.handle, 7, d, 1, None, None)
    assert rc == _lib.S2G_EINVAL
    rc = ctx.lib.s2g_deposit_batch(ctx.handle, _lib.S2G_F32, None, 2, None, None)
    assert rc == _lib.S2G_EINVAL


def test_inputs_written_on_the_null_stream_just_before(s2g, orc):
    """Ordering with the caller's stream: torch queues a long run of work and
    then the writes of the inputs (a device copy from scratch values) on its
    current (null) stream, and immediately calls s2g_deposit_batch and
    s2g_deposit_device with stream NULL (the context's stream, which is
    blocking, so it waits for that work).  The map is the nq = 3, 200^2,
    reduced, calc_mean map of the batch tests; the result must be the
    oracle's for the written inputs, on both entry points and on repeats."""
    import torch
    from sph2grid import _lib
    dev = torch.device("cuda:0")
    c = _calls(s2g, np.float32)[4]
    npix, nq = 200, 3
    pm_args = [c[k] for k in ("Pos", "HSML", "M", "Rho", "Bin_Quant", "Weights")]
    from sph2grid.mapping import _prepare
    pm = _prepare(*[a.copy() for a in pm_args], param=c["param"], kernel=c["kernel"],
                  show_progress=False, reduce_image=True, calc_mean=True)
    src = [torch.from_numpy(a).to(dev) for a in pm.arrays]
    want = _oracle_result(orc, c)
    ctx = s2g.default_context()
    big = torch.randn(3072, 3072, device=dev)
    for rep in range(3):
        # wrong values first (zero hsml: every particle would be skipped)
        dst = [torch.zeros_like(a) for a in src]
        torch.cuda.synchronize()
        for _ in range(6):
            big = (big @ big).mul_(1e-3)  # keeps the null stream busy
        for d, a in zip(dst, src):
            d.copy_(a)  # queued behind the matmuls, on the null stream
        out = torch.empty(npix * npix * nq, dtype=torch.float64, device=dev)
        s2g.deposit_batch_device([dict(pos=dst[0], hsml=dst[1], m=dst[2], rho=dst[3],
                                       bin_q=dst[4], weights=dst[5], out=out, n=pm.n, nq=nq,
                                       npix=npix, len2pix=pm.params.len2pix, kernel=pm.kid,
                                       calc_mean=True, reduce_image=True)], dtype="f32")
        got = out.cpu().numpy().reshape((npix, npix, nq), order="F")
        e = rel_l1(got, want)
        print(f"rep {rep}: batch rel_l1 vs oracle {e:.2e}", flush=True)
        assert e <= 1e-5
        # the single-map device entry point, same ordering
        dst = [torch.zeros_like(a) for a in src]
        torch.cuda.synchronize()
        for _ in range(6):
            big = (big @ big).mul_(1e-3)
        for d, a in zip(dst, src):
            d.copy_(a)
        img = torch.empty(npix * npix * (nq + 1), dtype=torch.float64, device=dev)
        vp = lambda t: C.c_void_p(t.data_ptr())
        ctx.check(ctx.lib.s2g_deposit_device(ctx.handle, 2, _lib.S2G_F32, vp(dst[0]), vp(dst[1]),
                                             vp(dst[2]), vp(dst[3]), vp(dst[4]), nq, vp(dst[5]),
                                             pm.n, C.byref(pm.params), pm.kid, 2,
                                             _lib.S2G_FLAG_CALC_MEAN, vp(img), None, None),
                  "s2g_deposit_device")
        red = torch.empty(npix * npix * nq, dtype=torch.float64, device=dev)
        ctx.check(ctx.lib.s2g_reduce_device(ctx.handle, 2, npix, nq, _lib.S2G_FLAG_REDUCE,
                                            vp(img), vp(red), None), "s2g_reduce_device")
        got = red.cpu().numpy().reshape((npix, npix, nq), order="F")
        e = rel_l1(got, want)
        print(f"rep {rep}: device entry point rel_l1 vs oracle {e:.2e}", flush=True)
        assert e <= 1e-5


def test_inputs_written_on_a_caller_stream_just_before(s2g, orc):
    """Ordering with a caller's own (non-blocking) stream: torch queues a long
    run of work and then the writes of the inputs on a torch.cuda.Stream, and
    the deposits are enqueued on that same stream with no host
    synchronisation in between.  The deposits fork their side stream (the
    normalisation and the small-footprint scatter) from the caller's stream
    and join it back, so every stage must see the written inputs: the 2D
    deposit + reduce and the 3D deposit must give the oracle's result on
    three repeats (the 3D one against the inputs' 2D oracle's total weight:
    both conserve it)."""
    import torch
    from sph2grid import _lib
    dev = torch.device("cuda:0")
    c = _calls(s2g, np.float32)[4]
    npix, nq = 200, 3
    pm_args = [c[k] for k in ("Pos", "HSML", "M", "Rho", "Bin_Quant", "Weights")]
    from sph2grid.mapping import _prepare
    pm = _prepare(*[a.copy() for a in pm_args], param=c["param"], kernel=c["kernel"],
                  show_progress=False, reduce_image=True, calc_mean=True)
    src = [torch.from_numpy(a).to(dev) for a in pm.arrays]
    want = _oracle_result(orc, c)
    ctx = s2g.default_context()
    vp = lambda t: C.c_void_p(t.data_ptr())
    user = torch.cuda.Stream(device=dev)  # default priority, non-blocking
    big = torch.randn(3072, 3072, device=dev)
    npix3 = 48
    par3 = _lib.Params(npixels=npix3, len2pix=float(npix3))
    for rep in range(3):
        dst = [torch.zeros_like(a) for a in src]  # wrong values first
        img = torch.empty(npix * npix * (nq + 1), dtype=torch.float64, device=dev)
        red = torch.empty(npix * npix * nq, dtype=torch.float64, device=dev)
        img3 = torch.empty(2 * npix3 ** 3, dtype=torch.float64, device=dev)
        torch.cuda.synchronize()
        with torch.cuda.stream(user):
            for _ in range(6):
                big = (big @ big).mul_(1e-3)  # keeps the caller's stream busy
            for d, a in zip(dst, src):
                d.copy_(a)  # queued behind the matmuls, on the caller's stream
            sp = C.c_void_p(user.cuda_stream)
            ctx.check(ctx.lib.s2g_deposit_device(
                ctx.handle, 2, _lib.S2G_F32, vp(dst[0]), vp(dst[1]), vp(dst[2]), vp(dst[3]),
                vp(dst[4]), nq, vp(dst[5]), pm.n, C.byref(pm.params), pm.kid, 2,
                _lib.S2G_FLAG_CALC_MEAN, vp(img), sp, None), "s2g_deposit_device")
            ctx.check(ctx.lib.s2g_reduce_device(ctx.handle, 2, npix, nq, _lib.S2G_FLAG_REDUCE,
                                                vp(img), vp(red), sp), "s2g_reduce_device")
            # 3D deposit of the same particles (first quantity), same stream
            q0 = dst[4].view(-1, nq)[:, 0].contiguous()
            ctx.check(ctx.lib.s2g_deposit_device(
                ctx.handle, 3, _lib.S2G_F32, vp(dst[0]), vp(dst[1]), vp(dst[2]), vp(dst[3]),
                vp(q0), 1, vp(dst[5]), pm.n, C.byref(par3), pm.kid, 3, 0, vp(img3), sp, None),
                "s2g_deposit_device 3D")
        user.synchronize()
        got = red.cpu().numpy().reshape((npix, npix, nq), order="F")
        e = rel_l1(got, want)
        # 3D: the weight plane's total is sum(w m / rho) len2pix^3 * len2pix over the
        # particles whose footprint meets the grid (all of them here: centres inside)
        x = dst[0].double().view(-1, 3).cpu().numpy()
        w, m, rho = (t.double().cpu().numpy() for t in (dst[5], dst[2], dst[3]))
        want3 = float((w * m / rho).sum()) * npix3 ** 4
        w3 = float(img3.view(2, -1)[1].sum().item())
        print(f"rep {rep}: 2D rel_l1 vs oracle {e:.2e}; 3D weight total rel err "
              f"{abs(w3 / want3 - 1):.2e} ({x.shape[0]} particles)", flush=True)
        assert e <= 1e-5
        assert abs(w3 / want3 - 1) <= 1e-6
