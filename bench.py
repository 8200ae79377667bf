#!/usr/bin/env python3
"""Benchmark of the deposition hot path on one MI355X.

Default workload: the cluster-like particle set (2^24 particles, 80% in an
r^-2 halo + 20% uniform background, hsml log-uniform over [1e-4, 1e-1] box
lengths), WendlandC6, 2D 4096^2 map of T weighted by Rho with reduce_image.
One step = one deposit (cic_mapping_2D equivalent) + reduce_image_2D over the
whole particle set, inputs already resident in HBM.

Prints ONE JSON line (stdout).  Everything else goes to stderr.
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (ROOT, os.path.join(ROOT, "sphtogrid.jl_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

import numpy as np  # noqa: E402

PEAK_HBM_GBPS = 8000.0
PEAK_FP32_TFLOPS = 157.3
# the FP32 vector rate without packed (two-wide) FP32 instructions, which this
# build does not emit: one FP32 FMA per lane and cycle
PEAK_FP32_NONPACKED_TFLOPS = 78.6
# flops per pixel-update of the 2D deposit (roofline convention): 10 + 2*Nq + K
K_FLOPS = {"cubic": 6, "quintic": 16, "wendlandc2": 6, "wendlandc4": 9, "wendlandc6": 11,
           "wendlandc8": 14}
KIDS = {"cubic": 0, "quintic": 1, "wendlandc2": 2, "wendlandc4": 3, "wendlandc6": 4,
        "wendlandc8": 5}


# workloads that --only can select
ONLY = ("cluster3d", "cluster2d_oracle", "plain", "sedov", "allsky", "tsc", "batched", "density",
        "streaming")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def footprints(x, h, npix, len2pix):
    """Integer footprint sizes per particle (cic_shared.jl:46-52, FP64)."""
    xs = x[:, 0].astype(np.float64) * len2pix + 0.5 * npix
    ys = x[:, 1].astype(np.float64) * len2pix + 0.5 * npix
    hp = h.astype(np.float64) * len2pix
    i0 = np.maximum(np.floor(xs - hp), 0)
    i1 = np.minimum(np.floor(xs + hp), npix - 1)
    j0 = np.maximum(np.floor(ys - hp), 0)
    j1 = np.minimum(np.floor(ys + hp), npix - 1)
    wi = np.maximum(i1 - i0 + 1, 0)
    wj = np.maximum(j1 - j0 + 1, 0)
    return wi, wj


def prepare(d, s2g):
    """Host-side part of sphMapping: centre + filter (cic_interpolation.jl:59-129)."""
    par = s2g.mappingParameters(x_lim=d["lims"][0], y_lim=d["lims"][1], z_lim=d["lims"][2],
                                Npixels=d["Npixels"])
    pos = d["Pos"].copy()
    pos, par2 = s2g.center_particles(pos, par)
    keep = s2g.filter_particles_in_image(pos, par2)
    bq = d["Bin_Quant"]
    out = dict(x=np.ascontiguousarray(pos[:, keep].T), hsml=d["HSML"][keep], m=d["M"][keep],
               rho=d["Rho"][keep], w=d["Weights"][keep],
               bq=(bq[keep] if bq.ndim == 1 else np.ascontiguousarray(bq[:, keep].T)),
               npix=int(par2.Npixels[0]), len2pix=float(par2.len2pix),
               kernel=d["kernel"], calc_mean=d["calc_mean"], reduce=d["reduce_image"])
    out["nq"] = 1 if out["bq"].ndim == 1 else out["bq"].shape[1]
    return out


def cpu_baseline(p, threads, frac=1.0 / 16):
    """Oracle (FP64 C restatement, reference parallel=true layout: contiguous
    batches, private images, summed in worker order) on a seeded uniform
    particle subsample of `frac` of the set; the full-set time is a + b *
    (footprint pixels), with a (the per-call cost of the full-size images of
    the reference layout) from a second, tiny sample."""
    import oracle
    n = p["x"].shape[0]
    wi, wj = footprints(p["x"], p["hsml"], p["npix"], p["len2pix"])
    fp = wi * wj
    fp_full = float(fp.sum())
    rng = np.random.default_rng(12345)

    def timed(m):
        idx = np.sort(rng.choice(n, size=max(1, int(m)), replace=False))
        sub = {k: p[k][idx] for k in ("x", "hsml", "m", "rho", "w", "bq")}
        t0 = time.perf_counter()
        img = oracle.cic_mapping_2D(sub["x"], sub["hsml"], sub["m"], sub["rho"], sub["bq"],
                                    sub["w"], p["npix"], p["len2pix"], p["kernel"], 2,
                                    p["calc_mean"], nworkers=threads)
        oracle.reduce_image_2D(img, p["npix"], p["reduce"])
        dt = time.perf_counter() - t0
        fs = float(fp[idx].sum())
        log(f"cpu baseline sample: {idx.size} particles, {fs:.3g} footprint pixels, "
            f"{dt:.2f} s on {threads} threads")
        return idx.size, fs, dt

    _, f1, t1 = timed(max(64, n * 3e7 / max(fp_full, 1.0)))
    n2, f2, t2 = timed(n * frac)
    b = max((t2 - t1) / max(f2 - f1, 1.0), 1e-15)
    a = max(t1 - b * f1, 0.0)
    t_full = a + b * fp_full
    sample = (f"{n2} of {n} particles (seeded uniform 1/{round(1 / frac)} subsample, {f2:.3g} "
              f"of {fp_full:.3g} footprint pixels) in {t2:.1f} s on {threads} threads; full "
              f"set = a + b*footprint with a = {a:.2f} s per call (from a {t1:.2f} s sample of "
              f"{f1:.3g} pixels) and b = {b:.3g} s per pixel-update: {t_full:.1f} s")
    return {"value": n / t_full, "unit": "particles/s", "cores": threads, "kind": "port",
            "sample": sample}


def _time_steps(fn, torch, steps=3, warmup=1):
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) * 1e3 / steps


def _errs_max(a, b):
    """max |a - b| / max |b| of two device tensors."""
    return float((a - b).abs().max().item()) / max(float(b.abs().max().item()), 1e-300)


def _errs(gpu, ref):
    gpu = np.asarray(gpu, dtype=np.float64).ravel(order="F")
    ref = np.asarray(ref, dtype=np.float64).ravel(order="F")
    den = np.abs(ref).sum()
    m = np.abs(ref) > 1e-3 * np.abs(ref).max() if ref.size else np.zeros(0, bool)
    return {"rel_l1": float(np.abs(gpu - ref).sum() / den) if den else 0.0,
            "max_pixel_rel": float((np.abs(gpu - ref)[m] / np.abs(ref)[m]).max()) if m.any() else 0.0}


def extras(args, p, dev_arrays, torch, s2g, _lib):
    """Other workloads and oracle comparisons (reported beside the headline);
    with --only, the selected ones."""
    out = {}
    want = lambda name: not args.only or name in args.only
    ctx = s2g.default_context()
    stream = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    st = _lib.Stats()

    # --- cluster 3D: the same resident particles on a 256^3 grid (WendlandC6)
    if want("cluster3d"):
        out.update(cluster3d_extra(ctx, p, dev_arrays, torch, _lib, st, stream))
    if want("cluster2d_oracle") and not args.no_oracle:
        out.update(cluster2d_oracle_extra(s2g))
    if want("plain"):
        out.update(plain_extra(args, ctx, torch, s2g, _lib, st, stream))
    if want("sedov"):
        out.update(sedov_extra(args, ctx, torch, s2g, _lib, st, stream))
    if want("allsky") or want("tsc"):
        out.update(allsky_extra(args, ctx, torch, s2g, _lib, st, stream, want))
    if want("batched"):
        out["batched_64maps"] = batched_extra(args, torch, s2g, _lib)
    if want("density"):
        out["mass_density_2e20_k64"] = density_extra(args, torch, s2g, _lib)
    return out


def cluster3d_extra(ctx, p, dev_arrays, torch, _lib, st, stream):
    """The headline's resident particles deposited on a 256^3 grid."""
    vp = lambda t: C.c_void_p(t.data_ptr())
    dev = torch.device("cuda:0")
    out = {}
    dx, dh, dm, drho, dw, dbq = dev_arrays
    n = p["x"].shape[0]
    N3 = 256
    img3 = torch.empty(2 * N3 ** 3, dtype=torch.float64, device=dev)
    map3 = torch.empty(N3 ** 3, dtype=torch.float64, device=dev)
    par3 = _lib.Params(npixels=N3, len2pix=float(N3))
    dt = _lib.S2G_F32 if p["x"].dtype == np.float32 else _lib.S2G_F64

    def step3():
        ctx.check(ctx.lib.s2g_deposit_device(ctx.handle, 3, dt, vp(dx), vp(dh), vp(dm), vp(drho),
                                             vp(dbq), 1, vp(dw), n, C.byref(par3),
                                             KIDS[p["kernel"]], 3, 0, vp(img3), stream,
                                             C.byref(st)), "3D deposit")
        ctx.check(ctx.lib.s2g_reduce_device(ctx.handle, 3, N3, 1, _lib.S2G_FLAG_REDUCE, vp(img3),
                                            vp(map3), stream), "3D reduce")
    ms3 = _time_steps(step3, torch)
    w3 = float(img3.view(2, -1)[1].sum().item())
    x64 = {k: p[k].astype(np.float64) for k in ("hsml", "m", "rho", "w")}
    w3_exp = float((x64["w"] * x64["m"] / x64["rho"]).sum() * N3 ** 4)
    out["cluster_3d_256"] = {"ms_per_step": ms3, "particles_per_s": n / (ms3 / 1e3),
                             "stage_ms": {"gather": st.ms_gather, "norm": st.ms_norm,
                                          "small": st.ms_small, "sort": st.ms_sort},
                             "counts": {"small": st.n_small, "large": st.n_large,
                                        "exact": st.n_large_exact, "pairs": st.n_pairs},
                             "conservation_rel_err": abs(w3 / w3_exp - 1)}
    log(f"extra cluster 3D: {ms3:.1f} ms")
    return out


def cluster2d_oracle_extra(s2g):
    """Cluster 2D subsample against the FP64 oracle (same distribution, 2^18)."""
    import oracle
    from sph2grid import workloads
    out = {}
    if True:
        d = workloads.cluster(n=1 << 18, npix=4096)
        ps = prepare(d, s2g)
        gpu = s2g.deposit(ps["x"], ps["hsml"], ps["m"], ps["rho"], ps["bq"], ps["w"], npix=4096,
                          len2pix=ps["len2pix"], kernel=s2g.WendlandC6(2))
        ref = oracle.cic_mapping_2D(ps["x"], ps["hsml"], ps["m"], ps["rho"], ps["bq"], ps["w"],
                                    4096, ps["len2pix"], "wendlandc6", 2, False,
                                    nworkers=min(16, os.cpu_count() or 1))
        e = _errs(oracle.reduce_image_2D(gpu, 4096, True), oracle.reduce_image_2D(ref, 4096, True))
        e["conservation_rel_err"] = abs(gpu[:, 1].sum() / ref[:, 1].sum() - 1)
        out["cluster_2d_4096_subsample_2e18_vs_oracle"] = e
        log(f"extra cluster 2D oracle check: {e}")
    return out


def plain_extra(args, ctx, torch, s2g, _lib, st, stream):
    """The plain workload (65,536 particles, 256^2, Cubic, Float64)."""
    import oracle
    from sph2grid import workloads
    vp = lambda t: C.c_void_p(t.data_ptr())
    dev = torch.device("cuda:0")
    out = {}
    d = workloads.plain()
    pp = prepare(d, s2g)
    T = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    arr = [T(pp[k]) for k in ("x", "hsml", "m", "rho", "w", "bq")]
    img = torch.empty(256 * 256 * 2, dtype=torch.float64, device=dev)
    mp = torch.empty(256 * 256, dtype=torch.float64, device=dev)
    parp = _lib.Params(npixels=256, len2pix=pp["len2pix"])

    def stepp():
        ctx.check(ctx.lib.s2g_deposit_device(ctx.handle, 2, _lib.S2G_F64, vp(arr[0]), vp(arr[1]),
                                             vp(arr[2]), vp(arr[3]), vp(arr[5]), 1, vp(arr[4]),
                                             pp["x"].shape[0], C.byref(parp), 0, 2, 0, vp(img),
                                             stream, C.byref(st)), "plain")
        ctx.check(ctx.lib.s2g_reduce_device(ctx.handle, 2, 256, 1, _lib.S2G_FLAG_REDUCE,
                                            vp(img), vp(mp), stream), "plain reduce")
    msp = _time_steps(stepp, torch, steps=10, warmup=2)
    rec = {"ms_per_step": msp, "particles_per_s": pp["x"].shape[0] / (msp / 1e3)}
    if not args.no_oracle:
        t0 = time.perf_counter()
        ref = oracle.cic_mapping_2D(pp["x"], pp["hsml"], pp["m"], pp["rho"], pp["bq"], pp["w"],
                                    256, pp["len2pix"], "cubic", 2, False)
        rec["oracle_serial_s"] = time.perf_counter() - t0
        g = img.view(2, -1).T.cpu().numpy()
        rec.update(_errs(oracle.reduce_image_2D(np.asfortranarray(g), 256, True),
                         oracle.reduce_image_2D(ref, 256, True)))
        rec["conservation_rel_err"] = abs(g[:, 1].sum() / ref[:, 1].sum() - 1)
    # the same call as a replayed plan (s2g_plan_*: one hipGraph, no host
    # read-back inside); its output against the eager call's
    ref_map = mp.clone()
    mp_plan = torch.empty_like(mp)
    pdesc = (_lib.MapDesc * 1)()
    pdesc[0] = _lib.MapDesc(dims=2, nq=1, n=pp["x"].shape[0], params=parp, kernel=0,
                            flags=_lib.S2G_FLAG_REDUCE, d_pos=arr[0].data_ptr(),
                            d_hsml=arr[1].data_ptr(), d_m=arr[2].data_ptr(),
                            d_rho=arr[3].data_ptr(), d_bin_q=arr[5].data_ptr(),
                            d_weights=arr[4].data_ptr(), d_out=mp_plan.data_ptr())

    def time_plan():
        ph = C.c_void_p()
        ctx.check(ctx.lib.s2g_plan_create(ctx.handle, _lib.S2G_F64, pdesc, 1, 1.25,
                                          C.byref(ph)), "s2g_plan_create")
        try:
            def stepplan():
                ctx.check(ctx.lib.s2g_plan_run(ph, stream, None), "s2g_plan_run")
            msplan = _time_steps(stepplan, torch, steps=50, warmup=3)
        finally:
            ctx.lib.s2g_plan_destroy(ph)
        return {"ms_per_step": msplan, "particles_per_s": pp["x"].shape[0] / (msplan / 1e3),
                "max_rel_diff_vs_eager": float((mp_plan - ref_map).abs().max().item() /
                                               ref_map.abs().max().item())}
    rec["plan"] = time_plan()
    if not args.no_tune:
        # work split tuned for this call (sph2grid.autotune; tuning not timed)
        from sph2grid import autotune

        def run():
            stepp()
            torch.cuda.synchronize()
        t0 = time.perf_counter()
        tp, _, _ = autotune.tune(run, ctx, nplanes=2, reps=5)
        tune_s = time.perf_counter() - t0
        mst = _time_steps(stepp, torch, steps=20, warmup=2)
        plan_tuned = time_plan()  # a plan takes the context's tuning at creation
        ctx.set_tuning(None)
        rec["tuned"] = {"tuning": tp, "ms_per_step": mst, "plan": plan_tuned,
                        "particles_per_s": pp["x"].shape[0] / (mst / 1e3),
                        "max_rel_diff_vs_default": float((mp - ref_map).abs().max().item() /
                                                         ref_map.abs().max().item()),
                        "tuning_s": tune_s}
    out["plain_65536_256"] = rec
    log(f"extra plain: {rec}")
    return out


def sedov_extra(args, ctx, torch, s2g, _lib, st, stream):
    """Sedov-like blast (2^20 particles, 1024^2, WendlandC6, [rho; T], calc_mean)."""
    import oracle
    from sph2grid import workloads
    vp = lambda t: C.c_void_p(t.data_ptr())
    dev = torch.device("cuda:0")
    T = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    out = {}
    d = workloads.sedov()
    ps = prepare(d, s2g)
    ns = ps["x"].shape[0]
    arr = [T(ps[k]) for k in ("x", "hsml", "m", "rho", "w", "bq")]
    img = torch.empty(1024 * 1024 * 3, dtype=torch.float64, device=dev)
    mp = torch.empty(1024 * 1024 * 2, dtype=torch.float64, device=dev)
    pars = _lib.Params(npixels=1024, len2pix=ps["len2pix"])

    def steps():
        ctx.check(ctx.lib.s2g_deposit_device(ctx.handle, 2, _lib.S2G_F32, vp(arr[0]), vp(arr[1]),
                                             vp(arr[2]), vp(arr[3]), vp(arr[5]), 2, vp(arr[4]),
                                             ns, C.byref(pars), 4, 2, _lib.S2G_FLAG_CALC_MEAN,
                                             vp(img), stream, C.byref(st)), "sedov")
        ctx.check(ctx.lib.s2g_reduce_device(ctx.handle, 2, 1024, 2, _lib.S2G_FLAG_REDUCE,
                                            vp(img), vp(mp), stream), "sedov reduce")
    mss = _time_steps(steps, torch)
    rec = {"ms_per_step": mss, "particles_per_s": ns / (mss / 1e3),
           "stage_ms": {"gather": st.ms_gather, "small": st.ms_small, "sort": st.ms_sort}}
    # the same with S2G_FLAG_FP32_ACCUM (three planes: the FP32-accumulator gather)
    mp64 = mp.clone()

    def steps32():
        ctx.check(ctx.lib.s2g_deposit_device(ctx.handle, 2, _lib.S2G_F32, vp(arr[0]), vp(arr[1]),
                                             vp(arr[2]), vp(arr[3]), vp(arr[5]), 2, vp(arr[4]),
                                             ns, C.byref(pars), 4, 2,
                                             _lib.S2G_FLAG_CALC_MEAN | _lib.S2G_FLAG_FP32_ACCUM,
                                             vp(img), stream, C.byref(st)), "sedov fp32")
        ctx.check(ctx.lib.s2g_reduce_device(ctx.handle, 2, 1024, 2, _lib.S2G_FLAG_REDUCE,
                                            vp(img), vp(mp), stream), "sedov reduce")
    mss32 = _time_steps(steps32, torch)
    rec["fp32_accum"] = {"ms_per_step": mss32, "gather_ms": st.ms_gather,
                         "rel_l1_vs_fp64": float((mp - mp64).abs().sum().item() /
                                                 mp64.abs().sum().item())}
    steps()  # leave the FP64 result in img for the oracle check below
    if not args.no_oracle:
        ref = oracle.cic_mapping_2D(ps["x"], ps["hsml"], ps["m"], ps["rho"], ps["bq"], ps["w"],
                                    1024, ps["len2pix"], "wendlandc6", 2, True,
                                    nworkers=min(16, os.cpu_count() or 1))
        g = img.view(3, -1).T.cpu().numpy()
        rec.update(_errs(oracle.reduce_image_2D(np.asfortranarray(g), 1024, True),
                         oracle.reduce_image_2D(ref, 1024, True)))
        rec["conservation_rel_err"] = abs(g[:, 2].sum() / ref[:, 2].sum() - 1)
    out["sedov_2e20_1024"] = rec
    log(f"extra sedov: {rec}")
    return out


def allsky_extra(args, ctx, torch, s2g, _lib, st, stream, want):
    """All-sky: 2^22 particles, Nside 1024, WendlandC4: rho (Weights = 1) and
    T (Weights = Rho), plus TSC 512^3 of the same particles."""
    import oracle
    from sph2grid import workloads
    vp = lambda t: C.c_void_p(t.data_ptr())
    dev = torch.device("cuda:0")
    T = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    out = {}
    d = workloads.allsky()
    na = d["HSML"].shape[0]
    nside = d["Nside"]
    npx = 12 * nside * nside
    apos = T(np.ascontiguousarray(d["Pos"].T))
    ah, am, arho, aT = (T(d[k]) for k in ("HSML", "M", "Rho", "T"))
    ones = torch.ones(na, dtype=torch.float64, device=dev)
    maps = torch.empty(4 * npx, dtype=torch.float64, device=dev)
    st2 = _lib.Stats()

    def stepa_separate():
        ctx.check(ctx.lib.s2g_healpix_deposit_device(
            ctx.handle, _lib.S2G_F64, vp(apos), vp(ah), vp(am), vp(arho), vp(arho), vp(ones), na,
            nside, 3, 2, _lib.S2G_FLAG_CALC_MEAN, C.c_void_p(maps.data_ptr()),
            C.c_void_p(maps.data_ptr() + 8 * npx), stream, C.byref(st)), "allsky rho")
        ctx.check(ctx.lib.s2g_healpix_deposit_device(
            ctx.handle, _lib.S2G_F64, vp(apos), vp(ah), vp(am), vp(arho), vp(aT), vp(arho), na,
            nside, 3, 2, _lib.S2G_FLAG_CALC_MEAN, C.c_void_p(maps.data_ptr() + 16 * npx),
            C.c_void_p(maps.data_ptr() + 24 * npx), stream, C.byref(st2)), "allsky T")

    if want("allsky"):
        out.update(_allsky_maps(args, ctx, s2g, _lib, st, stream, d, apos, ah, am, arho, aT,
                                stepa_separate, ones, maps, torch))
    if want("tsc"):
        out.update(_tsc512(ctx, _lib, stream, apos, aT, na, torch))
    return out


def _allsky_maps(args, ctx, s2g, _lib, st, stream, d, apos, ah, am, arho, aT, stepa_separate,
                 ones, maps, torch):
    import oracle
    vp = lambda t: C.c_void_p(t.data_ptr())
    out = {}
    na = d["HSML"].shape[0]
    nside = d["Nside"]
    npx = 12 * nside * nside
    st2 = _lib.Stats()
    # both maps in one s2g_healpix_deposit_maps_device call: one disc query,
    # normalisation and ring-list pass for the particle set
    pa = lambda *xs: (C.c_void_p * len(xs))(*[C.c_void_p(x) for x in xs])
    qs, ws = pa(arho.data_ptr(), aT.data_ptr()), pa(ones.data_ptr(), arho.data_ptr())
    mo = pa(maps.data_ptr(), maps.data_ptr() + 16 * npx)
    wo = pa(maps.data_ptr() + 8 * npx, maps.data_ptr() + 24 * npx)

    def stepa():
        ctx.check(ctx.lib.s2g_healpix_deposit_maps_device(
            ctx.handle, _lib.S2G_F64, vp(apos), vp(ah), vp(am), vp(arho), 2, qs, ws, na, nside,
            3, 2, _lib.S2G_FLAG_CALC_MEAN, mo, wo, stream, C.byref(st2)), "allsky maps")
    msa_sep = _time_steps(stepa_separate, torch)
    msa = _time_steps(stepa, torch)
    mp = maps.view(4, -1)
    pos64 = d["Pos"].T
    dxa = np.linalg.norm(pos64, axis=1)
    ang = np.sqrt(4 * np.pi / npx)
    keep = dxa >= d["HSML"]
    wexp = float((d["M"] / d["Rho"] / (ang * dxa) ** 2)[keep].sum())
    rec = {"ms_per_step_two_maps": msa, "particles_per_s": 2 * na / (msa / 1e3),
           "ms_two_separate_calls": msa_sep,
           "stage_ms": {"norm": st2.ms_norm, "lists_and_deposit": st2.ms_small,
                        "setup": st2.ms_setup},
           "units": st2.n_pairs,
           "conservation_rel_err_rho_weights": abs(float(mp[1].sum().item()) / wexp - 1)}
    if not args.no_oracle:
        sub = np.random.default_rng(7).choice(na, 1 << 16, replace=False)
        gm, gw = s2g.healpix_map(d["Pos"][:, sub].copy(), d["HSML"][sub], d["M"][sub],
                                 d["Rho"][sub], d["T"][sub], d["Rho"][sub], Nside=nside,
                                 kernel=s2g.WendlandC4(2), show_progress=False)
        om, ow = oracle.healpix_deposit_parallel(pos64[sub], d["HSML"][sub], d["M"][sub],
                                                 d["Rho"][sub], d["T"][sub], d["Rho"][sub], nside,
                                                 "wendlandc4", 2, True,
                                                 nworkers=min(16, os.cpu_count() or 1))
        e = _errs(gm, om)
        e["weights_rel_l1"] = _errs(gw, ow)["rel_l1"]
        e["conservation_rel_err"] = abs(gw.sum() / ow.sum() - 1)
        rec["subsample_2e16_vs_oracle"] = e
    out["allsky_2e22_nside1024"] = rec
    log(f"extra allsky: {rec}")
    return out


def _tsc512(ctx, _lib, stream, apos, aT, na, torch):
    vp = lambda t: C.c_void_p(t.data_ptr())
    dev = torch.device("cuda:0")
    out = {}
    N5 = 512
    grid = torch.empty(2 * N5 ** 3, dtype=torch.float64, device=dev)
    part = _lib.Params(npixels=N5, len2pix=N5 / 2.0)

    def stept():
        ctx.check(ctx.lib.s2g_tsc3d_device(ctx.handle, _lib.S2G_F64, vp(apos), vp(aT), na,
                                           C.byref(part), 0, vp(grid), stream), "tsc")
    mst = _time_steps(stept, torch)
    wsum = float(grid.view(2, -1)[1].sum().item())
    out["tsc_2e22_512"] = {"ms_per_step": mst, "particles_per_s": na / (mst / 1e3),
                           "weight_sum_over_n": wsum / na}
    log(f"extra tsc: {out['tsc_2e22_512']}")
    return out


def batched_extra(args, torch, s2g, _lib):
    """The batched workload: 64 sedov-like maps (WendlandC6, Nq = 3 with
    Weights = Rho, calc_mean, reduce_image; N log-uniform [1e4, 4e6], sides
    128-2048).  One step maps all 64: either one s2g_deposit_batch call, or
    64 x (s2g_deposit_device + s2g_reduce_device), the reference's
    one-sphMapping-per-map loop on device-resident inputs."""
    import oracle
    from sph2grid import workloads
    ctx = s2g.default_context()
    dev = torch.device("cuda:0")
    vp = lambda t: C.c_void_p(t.data_ptr())
    stream = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    T = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    t0 = time.perf_counter()
    maps = workloads.batched(nmaps=64)
    flags = _lib.S2G_FLAG_CALC_MEAN | _lib.S2G_FLAG_REDUCE
    kid = KIDS["wendlandc6"]
    res, fps = [], []
    for d in maps:
        pp = prepare(d, s2g)
        npix, nq = pp["npix"], pp["nq"]
        nd = npix * npix
        wi, wj = footprints(pp["x"], pp["hsml"], npix, pp["len2pix"])
        fps.append(float((wi * wj).sum()))
        res.append(dict(pp=pp, t=[T(pp[k]) for k in ("x", "hsml", "m", "rho", "bq", "w")],
                        n=pp["x"].shape[0], npix=npix, nq=nq,
                        par=_lib.Params(npixels=npix, len2pix=pp["len2pix"]),
                        img=torch.empty(nd * (nq + 1), dtype=torch.float64, device=dev),
                        o_plain=torch.empty(nd * nq, dtype=torch.float64, device=dev),
                        o_batch=torch.empty(nd * nq, dtype=torch.float64, device=dev)))
    log(f"batched workload ready in {time.perf_counter() - t0:.1f} s")
    ntot = sum(r["n"] for r in res)

    def step_plain():
        for r in res:
            x, h, m, rho, bq, w = r["t"]
            ctx.check(ctx.lib.s2g_deposit_device(ctx.handle, 2, _lib.S2G_F32, vp(x), vp(h), vp(m),
                                                 vp(rho), vp(bq), r["nq"], vp(w), r["n"],
                                                 C.byref(r["par"]), kid, 2, flags, vp(r["img"]),
                                                 stream, None), "deposit")
            ctx.check(ctx.lib.s2g_reduce_device(ctx.handle, 2, r["npix"], r["nq"], flags,
                                                vp(r["img"]), vp(r["o_plain"]), stream), "reduce")

    descs = (_lib.MapDesc * len(res))()
    for k, r in enumerate(res):
        x, h, m, rho, bq, w = r["t"]
        descs[k] = _lib.MapDesc(dims=2, nq=r["nq"], n=r["n"], params=r["par"], kernel=kid,
                                flags=flags, d_pos=x.data_ptr(), d_hsml=h.data_ptr(),
                                d_m=m.data_ptr(), d_rho=rho.data_ptr(), d_bin_q=bq.data_ptr(),
                                d_weights=w.data_ptr(), d_out=r["o_batch"].data_ptr())
    st = _lib.Stats()

    def step_batch():
        ctx.check(ctx.lib.s2g_deposit_batch(ctx.handle, _lib.S2G_F32, descs, len(res), stream,
                                            C.byref(st)), "batch")

    ms_plain = _time_steps(step_plain, torch, steps=3, warmup=1)
    ms_batch = _time_steps(step_batch, torch, steps=3, warmup=1)

    def max_diff():
        diff = 0.0
        for r in res:
            a, b = r["o_batch"], r["o_plain"]
            diff = max(diff, float((a - b).abs().max().item()) /
                       max(float(b.abs().max().item()), 1e-300))
        return diff
    diff = max_diff()
    rec = {"maps": len(res), "particles": ntot, "sum_footprint": sum(fps),
           "plain_ms_per_step": ms_plain, "batched_ms_per_step": ms_batch,
           "plain_maps_per_s": len(res) / (ms_plain / 1e3),
           "batched_maps_per_s": len(res) / (ms_batch / 1e3),
           "batched_particles_per_s": ntot / (ms_batch / 1e3),
           "speedup_batched_vs_plain": ms_plain / ms_batch,
           "batched_stage_ms": {"setup": st.ms_setup, "norm": st.ms_norm, "sort": st.ms_sort,
                                "gather": st.ms_gather, "finalize": st.ms_finalize,
                                "small": st.ms_small, "device": st.ms_device},
           "counts": {"small": st.n_small, "large": st.n_large, "pairs": st.n_pairs},
           "max_rel_diff_batched_vs_plain": diff}
    if not args.no_tune:
        # work split tuned (sph2grid.autotune) for the batch pass and, for the
        # plain loop, for each map's own call; tuning runs are not timed
        from sph2grid import autotune

        def sync(fn):
            def run():
                fn()
                torch.cuda.synchronize()
            return run
        t0 = time.perf_counter()
        tb, _, _ = autotune.tune(sync(step_batch), ctx, nplanes=4, reps=2)
        tplain = []
        for r in res:
            def one(r=r):
                x, h, m, rho, bq, w = r["t"]
                ctx.check(ctx.lib.s2g_deposit_device(
                    ctx.handle, 2, _lib.S2G_F32, vp(x), vp(h), vp(m), vp(rho), vp(bq), r["nq"],
                    vp(w), r["n"], C.byref(r["par"]), kid, 2, flags, vp(r["img"]), stream,
                    None), "deposit")
                ctx.check(ctx.lib.s2g_reduce_device(ctx.handle, 2, r["npix"], r["nq"], flags,
                                                    vp(r["img"]), vp(r["o_plain"]), stream),
                          "reduce")
            t, _, _ = autotune.tune(sync(one), ctx, nplanes=4, reps=2)
            tplain.append(t)
        tune_s = time.perf_counter() - t0

        def step_plain_tuned():
            for r, t in zip(res, tplain):
                ctx.set_tuning(t)
                x, h, m, rho, bq, w = r["t"]
                ctx.check(ctx.lib.s2g_deposit_device(
                    ctx.handle, 2, _lib.S2G_F32, vp(x), vp(h), vp(m), vp(rho), vp(bq), r["nq"],
                    vp(w), r["n"], C.byref(r["par"]), kid, 2, flags, vp(r["img"]), stream,
                    None), "deposit")
                ctx.check(ctx.lib.s2g_reduce_device(ctx.handle, 2, r["npix"], r["nq"], flags,
                                                    vp(r["img"]), vp(r["o_plain"]), stream),
                          "reduce")
        ms_plain_t = _time_steps(step_plain_tuned, torch, steps=3, warmup=1)
        ctx.set_tuning(tb)
        ms_batch_t = _time_steps(step_batch, torch, steps=3, warmup=1)
        ctx.set_tuning(None)
        rec["tuned"] = {"batch_tuning": tb, "plain_ms_per_step": ms_plain_t,
                        "batched_ms_per_step": ms_batch_t,
                        "speedup_batched_vs_plain": ms_plain_t / ms_batch_t,
                        "speedup_batched_tuned_vs_plain_default": ms_plain / ms_batch_t,
                        "batched_stage_ms": {"gather": st.ms_gather, "small": st.ms_small,
                                             "device": st.ms_device},
                        "max_rel_diff_batched_vs_plain": max_diff(),
                        "tuning_s": tune_s}
    # the batch as a replayed plan (s2g_plan_*: one hipGraph, sized at creation)
    pdescs = (_lib.MapDesc * len(res))()
    for k, r in enumerate(res):
        r["o_plan"] = torch.empty_like(r["o_batch"])
        pdescs[k] = descs[k]
        pdescs[k].d_out = r["o_plan"].data_ptr()
    ph = C.c_void_p()
    t0 = time.perf_counter()
    ctx.check(ctx.lib.s2g_plan_create(ctx.handle, _lib.S2G_F32, pdescs, len(res), 1.25,
                                      C.byref(ph)), "s2g_plan_create")
    create_s = time.perf_counter() - t0
    stp = _lib.Stats()
    try:
        def step_plan():
            ctx.check(ctx.lib.s2g_plan_run(ph, stream, C.byref(stp)), "s2g_plan_run")
        ms_plan = _time_steps(step_plan, torch, steps=3, warmup=1)
        step_batch()
        torch.cuda.synchronize()
        pdiff = max(_errs_max(r["o_plan"], r["o_batch"]) for r in res)
    finally:
        ctx.lib.s2g_plan_destroy(ph)
    rec["plan"] = {"ms_per_step": ms_plan, "create_s": create_s,
                   "speedup_vs_plain": ms_plain / ms_plan, "device_ms": stp.ms_device,
                   "max_rel_diff_vs_batch": pdiff}
    for r in res:
        del r["o_plan"]
    # S2G_FLAG_FP32_ACCUM on every map (FP32 per-item pixel sums in the gather),
    # both sides; errors against the FP64 default's maps
    ref64 = [r["o_batch"].clone() for r in res]
    f32 = flags | _lib.S2G_FLAG_FP32_ACCUM
    fdescs = (_lib.MapDesc * len(res))()
    for k in range(len(res)):
        fdescs[k] = descs[k]
        fdescs[k].flags = f32
    st32 = _lib.Stats()

    def step_batch32():
        ctx.check(ctx.lib.s2g_deposit_batch(ctx.handle, _lib.S2G_F32, fdescs, len(res), stream,
                                            C.byref(st32)), "batch fp32 accumulation")

    def step_plain32():
        for r in res:
            x, h, m, rho, bq, w = r["t"]
            ctx.check(ctx.lib.s2g_deposit_device(ctx.handle, 2, _lib.S2G_F32, vp(x), vp(h), vp(m),
                                                 vp(rho), vp(bq), r["nq"], vp(w), r["n"],
                                                 C.byref(r["par"]), kid, 2, f32, vp(r["img"]),
                                                 stream, None), "deposit")
            ctx.check(ctx.lib.s2g_reduce_device(ctx.handle, 2, r["npix"], r["nq"], flags,
                                                vp(r["img"]), vp(r["o_plain"]), stream), "reduce")
    ms_plain32 = _time_steps(step_plain32, torch, steps=3, warmup=1)
    ms_batch32 = _time_steps(step_batch32, torch, steps=3, warmup=1)
    num = den = 0.0
    worst = 0.0
    for r, a in zip(res, ref64):
        d = (r["o_batch"] - a).abs()
        num += float(d.sum().item())
        den += float(a.abs().sum().item())
        m = a.abs() > 1e-3 * float(a.abs().max().item())
        if bool(m.any()):
            worst = max(worst, float((d[m] / a.abs()[m]).max().item()))
    rec["fp32_accum"] = {"plain_ms_per_step": ms_plain32, "batched_ms_per_step": ms_batch32,
                         "speedup_batched_vs_plain": ms_plain32 / ms_batch32,
                         "speedup_vs_fp64_plain": ms_plain / ms_batch32,
                         "batched_stage_ms": {"gather": st32.ms_gather, "small": st32.ms_small,
                                              "device": st32.ms_device},
                         "rel_l1_vs_fp64_all_maps": num / den if den else 0.0,
                         "max_pixel_rel_vs_fp64": worst,
                         "max_rel_diff_batched_vs_plain": max_diff()}
    if not args.no_oracle:
        # the three maps with the fewest footprint pixels against the oracle
        errs = {}
        for k in np.argsort(fps)[:3]:
            r = res[int(k)]
            pp = r["pp"]
            ref = oracle.cic_mapping_2D(pp["x"], pp["hsml"], pp["m"], pp["rho"], pp["bq"], pp["w"],
                                        r["npix"], pp["len2pix"], "wendlandc6", 2, True,
                                        nworkers=min(16, os.cpu_count() or 1))
            ref = oracle.reduce_image_2D(ref, r["npix"], True)
            got = r["o_batch"].cpu().numpy().reshape((r["npix"], r["npix"], r["nq"]), order="F")
            errs[f"map{int(k)}_n{r['n']}_npix{r['npix']}"] = _errs(got, ref)
        rec["oracle_checks"] = errs
    log(f"extra batched: {rec}")
    del res
    torch.cuda.empty_cache()
    return rec


def streaming_extra(args, d, torch, s2g, _lib):
    """distributed_cic_map over 16 subfiles of the cluster set (2^20
    particles each, 4096^2, WendlandC6, T weighted by Rho): the device-resident
    sum with prefetched, overlapped uploads, against the reference's flow on
    the same GPU deposit (each subfile's flat map returned to the host and
    summed there with the NaN / Inf rule, cic.jl:56-93)."""
    import contextlib
    import tempfile
    nsub = 16
    n = d["M"].shape[0]
    edges = np.linspace(0, n, nsub + 1).astype(int)
    lims = d["lims"]
    par = s2g.mappingParameters(x_lim=lims[0], y_lim=lims[1], z_lim=lims[2],
                                Npixels=d["Npixels"])
    kernel = s2g.WendlandC6(2)

    def sub(k):
        s = slice(edges[k], edges[k + 1])
        return s2g.Subfile(d["Pos"][:, s].copy(), d["HSML"][s], d["M"][s], d["Rho"][s],
                           d["Bin_Quant"][s], d["Weights"][s], kernel=kernel)

    out = tempfile.mkdtemp()
    fits = os.path.join(out, "stream.fits")
    with contextlib.redirect_stdout(sys.stderr):
        s2g.distributed_cic_map(fits, 2, sub, par)  # warm-up (allocations)
        st = {}
        t0 = time.perf_counter()
        img = s2g.distributed_cic_map(fits, nsub, sub, par, stats=st)
        t_stream = time.perf_counter() - t0

        npix = int(d["Npixels"])
        t0 = time.perf_counter()
        sm = np.zeros(npix * npix)
        sw = np.zeros(npix * npix)
        for k in range(nsub):
            s = sub(k)
            flat = s2g.sphMapping(s["Pos"], s["HSML"], s["M"], s["Rho"], s["Bin_Quant"],
                                  s["Weights"], param=par, kernel=kernel, return_both_maps=True,
                                  show_progress=False)
            a, w = flat[:, 0], flat[:, 1]
            sm += np.where(np.isfinite(a), a, 0.0)
            sw += np.where(np.isfinite(w), w, 0.0)
        red = np.where(sw > 0, sm / np.where(sw > 0, sw, 1.0), sm)
        ref = red.reshape(npix, npix)[:, :, None]  # [x, y, q]
        t_host = time.perf_counter() - t0
    diff = float(np.abs(img - ref).max() / np.abs(ref).max())
    t_loop = st["wall_s"]  # subfiles + reduce + copy-back, without the FITS write
    rec = {"subfiles": nsub, "particles": n, "npix": int(d["Npixels"]),
           "device_resident_s": t_loop, "device_resident_subfiles_per_s": nsub / t_loop,
           "device_resident_particles_per_s": n / t_loop,
           "device_resident_with_fits_write_s": t_stream,
           "device_ms_in_deposits": st.get("ms_device"), "wait_for_producer_s": st.get("wait_s"),
           "host_sum_flow_s": t_host, "host_sum_flow_subfiles_per_s": nsub / t_host,
           "speedup_resident_vs_host_sum": t_host / t_loop,
           "max_rel_diff_vs_host_sum": diff}
    log(f"extra streaming: {rec}")
    return rec


def density_extra(args, torch, s2g, _lib):
    """mass_density (effects/mass_density.jl:30-99) of 2^20 clustered
    particles (half in 64 Gaussian clumps, half uniform) in a periodic box,
    WendlandC4, 64 neighbours: device time, and 512 particles against the
    brute-force oracle (all 2^20 candidates each)."""
    import oracle
    r = np.random.default_rng(0x5348544F47520006)
    n, L = 1 << 20, 10.0
    pos = r.random((3, n)) * L
    c = n // 2
    centres = r.random((3, 64)) * L
    pos[:, :c] = np.mod(centres[:, r.integers(0, 64, c)] + 0.02 * L * r.standard_normal((3, c)),
                        L)
    m = r.uniform(0.5, 2.0, n)
    ctx = s2g.default_context()
    dev = torch.device("cuda:0")
    T = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    dpos, dm = T(pos.T.copy()), T(m)
    rho = torch.empty(n, dtype=torch.float64, device=dev)
    hs = torch.empty(n, dtype=torch.float64, device=dev)
    box = (C.c_double * 3)(L, L, L)
    st = _lib.Stats()
    vp = lambda t: C.c_void_p(t.data_ptr())

    def step():
        ctx.check(ctx.lib.s2g_mass_density_device(ctx.handle, _lib.S2G_F64, vp(dpos), vp(dm), n,
                                                  3, 3, 64, box, vp(rho), vp(hs), None,
                                                  C.byref(st)), "mass_density")
    ms = _time_steps(step, torch, steps=3, warmup=1)
    rec = {"particles": n, "nneighbors": 64, "kernel": "wendlandc4", "ms_per_call": ms,
           "particles_per_s": n / (ms / 1e3), "ms_sort": st.ms_sort, "ms_search": st.ms_gather}
    if not args.no_oracle:
        idx = r.choice(n, 512, replace=False)
        orho, oh = oracle.mass_density_subset(pos, m, idx, "wendlandc4", 3, 64, (L, L, L))
        g_rho, g_h = rho.cpu().numpy()[idx], hs.cpu().numpy()[idx]
        rec["sample_512_vs_oracle"] = {
            "hsml_max_rel": float(np.abs(g_h - oh).max() / oh.max()),
            "rho_max_rel": float((np.abs(g_rho - orho) / np.abs(orho)).max())}
    log(f"extra mass_density: {rec}")
    return rec


def measured_traffic(args, n, npix):
    """HBM bytes per launch of k2d_gather for the default cluster workload, as
    measured by the two rocprofv3 --pmc passes summarised under profiles/
    (FETCH_SIZE and WRITE_SIZE, scaled by a 1 GiB copy).  None otherwise."""
    if args.workload != "cluster" or n != 1 << 24 or npix != 4096 or args.exact_norm:
        return None, None
    path = os.path.join(ROOT, "profiles", "k2d_gather.cluster_2e24_4096.traffic.md")
    try:
        for line in open(path):
            if line.startswith("traffic_bytes_per_launch k2d_gather:"):
                return float(line.split(":")[1]), os.path.relpath(path, ROOT)
    except OSError:
        pass
    return None, None


def calib_copy(torch, dev):
    """A plain copy of known size (reads 1 GiB, writes 1 GiB): the reference
    dispatch for scaling FETCH_SIZE / WRITE_SIZE counters."""
    src = torch.ones(1 << 28, dtype=torch.float32, device=dev)
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    del src, dst


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="cluster", choices=["cluster", "plain", "sedov"])
    ap.add_argument("--n", type=int, default=0, help="override the particle count")
    ap.add_argument("--npix", type=int, default=0)
    ap.add_argument("--exact-norm", action="store_true")
    ap.add_argument("--cpu-fraction", type=float, default=1.0 / 16,
                    help="particle fraction of the timed CPU-baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the other workloads (3D, all-sky, TSC, plain, sedov)")
    ap.add_argument("--no-tune", action="store_true",
                    help="skip the autotuned timings of the plain and batched extras")
    ap.add_argument("--no-oracle", action="store_true",
                    help="skip the oracle comparisons inside the extras")
    ap.add_argument("--calib-copy", action="store_true",
                    help="also run one 1 GiB device copy (counter calibration)")
    ap.add_argument("--only", nargs="+", default=None, choices=ONLY,
                    help="run only these extras (no headline step, no CPU baseline) and print "
                         "their records as one JSON line; for profiling one workload")
    args = ap.parse_args()
    if args.gpus != 1:
        msg = ("bench.py runs on one GPU only: the deposit is measured on a single MI355X "
               f"(--gpus {args.gpus} is not supported)")
        log(msg)
        print(msg)
        sys.exit(2)

    import torch

    import sph2grid as s2g
    from sph2grid import _lib, workloads

    if args.only and not {"cluster3d", "streaming"} & set(args.only):
        # the selected extras generate their own particle sets
        dev = torch.device("cuda:0")
        if args.calib_copy:
            calib_copy(torch, dev)
        print(json.dumps({"extra": extras(args, None, None, torch, s2g, _lib)}), flush=True)
        return

    t0 = time.perf_counter()
    gen = {"cluster": workloads.cluster, "plain": workloads.plain, "sedov": workloads.sedov}
    kw = {}
    if args.n:
        kw["n"] = args.n
    if args.npix:
        kw["npix"] = args.npix
    d = gen[args.workload](**kw)
    p = prepare(d, s2g)
    n = p["x"].shape[0]
    log(f"workload {args.workload}: {n} particles, {p['npix']}^2, {p['kernel']}, "
        f"nq={p['nq']}, generated in {time.perf_counter() - t0:.1f} s")

    dev = torch.device("cuda:0")
    tdt = torch.float32 if p["x"].dtype == np.float32 else torch.float64
    T = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev, dtype=tdt)
    dx, dh, dm, drho, dw, dbq = (T(p[k]) for k in ("x", "hsml", "m", "rho", "w", "bq"))
    if args.only:
        if args.calib_copy:
            calib_copy(torch, dev)
        ex = extras(args, p, (dx, dh, dm, drho, dw, dbq), torch, s2g, _lib)
        if "streaming" in args.only:
            ex["subfiles_cluster_16x2e20"] = streaming_extra(args, d, torch, s2g, _lib)
        print(json.dumps({"extra": ex}), flush=True)
        return
    npix, nq = p["npix"], p["nq"]
    nd = npix * npix
    d_img = torch.empty(nd * (nq + 1), dtype=torch.float64, device=dev)
    d_map = torch.empty(nd * nq, dtype=torch.float64, device=dev)
    torch.cuda.synchronize()

    ctx = s2g.default_context()
    params = _lib.Params(npixels=npix, len2pix=p["len2pix"])
    flags = (_lib.S2G_FLAG_CALC_MEAN if p["calc_mean"] else 0) | \
        (_lib.S2G_FLAG_EXACT_NORM if args.exact_norm else 0)
    rflags = _lib.S2G_FLAG_REDUCE if p["reduce"] else 0
    dtype_id = _lib.S2G_F32 if tdt == torch.float32 else _lib.S2G_F64
    stream = torch.cuda.current_stream().cuda_stream
    vp = lambda t: C.c_void_p(t.data_ptr())
    st = _lib.Stats()

    def step():
        rc = ctx.lib.s2g_deposit_device(ctx.handle, 2, dtype_id, vp(dx), vp(dh), vp(dm),
                                        vp(drho), vp(dbq), nq, vp(dw), n, C.byref(params),
                                        KIDS[p["kernel"]], 2, flags, vp(d_img),
                                        C.c_void_p(stream), C.byref(st))
        ctx.check(rc, "s2g_deposit_device")
        rc = ctx.lib.s2g_reduce_device(ctx.handle, 2, npix, nq, rflags, vp(d_img), vp(d_map),
                                       C.c_void_p(stream))
        ctx.check(rc, "s2g_reduce_device")

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    stage = {k: 0.0 for k in ("ms_device", "ms_setup", "ms_norm", "ms_sort", "ms_gather",
                              "ms_finalize", "ms_small")}
    t1 = time.perf_counter()
    for i in range(args.steps):
        step()
        for k in stage:
            stage[k] += getattr(st, k)
        log(f"step {i}: device {st.ms_device:.1f} ms (gather {st.ms_gather:.1f} ms)")
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    ms_step = (t2 - t1) * 1e3 / args.steps
    if args.calib_copy:
        calib_copy(torch, dev)
    for k in stage:
        stage[k] /= args.steps
    value = n / (ms_step / 1e3)

    # full-size property checks: conservation of the weight plane
    img = d_img.view(nq + 1, nd)
    w_sum = float(img[nq].sum().item())
    # the same step with S2G_FLAG_FP32_ACCUM (opt-in FP32 per-item sums in the
    # gather), reported beside the FP64 default; not the headline value
    fp32 = None
    if not args.no_extras:
        map64 = d_map.clone()
        flags_keep = flags
        flags |= _lib.S2G_FLAG_FP32_ACCUM
        step()
        torch.cuda.synchronize()
        t3 = time.perf_counter()
        for _ in range(3):
            step()
        torch.cuda.synchronize()
        ms32 = (time.perf_counter() - t3) * 1e3 / 3
        fp32 = {"ms_per_step": ms32, "particles_per_s": n / (ms32 / 1e3),
                "gather_ms": st.ms_gather,
                "rel_l1_vs_fp64": float((d_map - map64).abs().sum().item() /
                                        map64.abs().sum().item())}
        flags = flags_keep
        step()
        torch.cuda.synchronize()
    x64 = {k: p[k].astype(np.float64) for k in ("hsml", "m", "rho", "w")}
    wi, wj = footprints(p["x"], p["hsml"], npix, p["len2pix"])
    dep = (wi > 0) & (wj > 0)
    if not p["calc_mean"]:
        bq = p["bq"] if p["bq"].ndim == 1 else p["bq"].any(axis=1)
        dep &= (bq != 0)
    w_expect = float((x64["w"] * x64["m"] / x64["rho"])[dep].sum() * p["len2pix"] ** 3)
    cons = abs(w_sum / w_expect - 1.0) if w_expect else 0.0

    # roofline of the dominant kernel (the tile gather): algorithmic flops of
    # the pixel-updates it performs / its measured duration
    small = (wi <= 34) & (wj <= 34)
    fp_large = float((wi * wj)[dep & ~small].sum())
    fp_all = float((wi * wj)[dep].sum())
    fk = 10 + 2 * nq + K_FLOPS[p["kernel"]]
    gather_s = stage["ms_gather"] / 1e3
    achieved = fp_large * fk / gather_s / 1e12 if gather_s > 0 else 0.0
    s_in = p["x"].dtype.itemsize
    b_alg = n * (6 + nq) * s_in + nd * nq * 8
    traffic, traffic_src = measured_traffic(args, n, npix)
    roofline = {"bound": "valu", "achieved": achieved, "peak": PEAK_FP32_TFLOPS,
                "unit": "TFLOP/s", "frac": achieved / PEAK_FP32_TFLOPS, "traffic": traffic,
                "traffic_source": traffic_src,
                "kernel": "k2d_gather", "compute_unit": "FP32 vector ALU (no MFMA: "
                "scatter/gather); peak = MI355X FP32 vector rate with packed FP32",
                "peak_nonpacked": PEAK_FP32_NONPACKED_TFLOPS,
                "frac_nonpacked": achieved / PEAK_FP32_NONPACKED_TFLOPS,
                "alg_flops_per_launch": fp_large * fk,
                "alg_pixel_updates_per_launch": fp_large,
                "ms_per_launch": stage["ms_gather"],
                # HBM bytes the gather moves per launch (rocprofv3 FETCH_SIZE +
                # WRITE_SIZE, calibrated) over its measured duration
                "hbm_measured_GBps": (traffic / (stage["ms_gather"] / 1e3) / 1e9
                                      if traffic and stage["ms_gather"] else None),
                "hbm_peak_GBps": 8000.0}

    cpu = None
    if not args.no_cpu:
        try:
            aff = len(os.sched_getaffinity(0))
        except AttributeError:
            aff = os.cpu_count() or 1
        # one GPU's share of the host is 16 cores, whatever the affinity mask
        # shows; fewer if fewer are visible
        threads = max(1, min(16, aff))
        cpu = cpu_baseline(p, threads, args.cpu_fraction)
        cpu["cores_visible"] = aff
        cpu["cores_note"] = ("threads = min(16, visible cores): 16 host cores are one GPU's "
                             "share of this box")

    out = {
        "metric": "particles/s (2D 4096^2 WendlandC6 map, cluster 2^24)"
        if args.workload == "cluster" else f"particles/s ({args.workload})",
        "value": value, "unit": "particles/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_step, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"{args.workload}: {n} particles, {npix}^2 2D map, "
                               f"{p['kernel']}, Nq={nq}, FP32 inputs in HBM, deposit + "
                               f"reduce_image per step",
                   "particles": n, "npix": npix, "kernel": p["kernel"], "nq": nq,
                   "exact_norm": bool(args.exact_norm)},
        "hbm_alg_GBps": b_alg / (ms_step / 1e3) / 1e9,
        "pixel_updates_per_s": fp_all / (ms_step / 1e3),
        "stage_ms": stage,
        "counts": {"small": st.n_small, "large": st.n_large, "large_exact": st.n_large_exact,
                   "pairs": st.n_pairs},
        "conservation_rel_err": cons,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "fp32_accum": fp32,
    }
    if cpu:
        out["speedup_vs_cpu"] = value / cpu["value"]
    if not args.no_extras and args.workload == "cluster":
        out["extra"] = extras(args, p, (dx, dh, dm, drho, dw, dbq), torch, s2g, _lib)
        if not args.n and not args.npix:
            out["extra"]["subfiles_cluster_16x2e20"] = streaming_extra(args, d, torch, s2g,
                                                                        _lib)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
