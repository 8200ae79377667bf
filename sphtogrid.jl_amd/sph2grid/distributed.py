"""``distributed_cic_map`` / ``distributed_allsky_map`` with the sum kept in
device memory across subfiles (distributed_mapping/cic.jl:24-110,
distributed_mapping/healpix.jl:16-87, distributed_mapping/main.jl:2-26).

The reference hands ``mapping_function(subfile)`` to Distributed workers
through job / result channels, adds every returned map into one host image
while skipping NaN and Inf values (cic.jl:63-73, healpix.jl:50-58), then
reduces and saves it.  Here:

* the sum lives in HBM for the whole run: one flat image of
  N_distr x (Nimages + 1) doubles (the maps, then the weights), or the two
  all-sky maps;
* up to ``prefetch`` subfiles are produced at once by a thread pool (mapping
  functions usually read files); each result is staged in pinned memory and
  uploaded on a copy stream while the GPU works on the previous subfile, and
  the compute stream waits for the upload's event;
* ``mapping_function`` may return what the reference's returns -- a tuple
  ``(maps, weight_map)``, uploaded and added with the same NaN / Inf rule
  (``s2g_accumulate_device``) -- or a particle set (``Subfile`` /
  ``AllskySubfile``), which is centred and filtered on the host and deposited
  straight into device memory, so no per-subfile map goes through the host.
"""
from __future__ import annotations

import ctypes as C
import os
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np

from . import _lib
from .healpix import _prepare_healpix
from .io import save_healpix_fits, write_fits_image, write_vtk_image
from .mapping import _prepare
from .parameters import ErrorException, MappingParameters


class Subfile(dict):
    """Particles of one subfile for ``distributed_cic_map``: the arguments
    of ``sphMapping`` (Pos, HSML, M, Rho, Bin_Quant, Weights=Rho, kernel,
    calc_mean=False, exact_norm=False).  Pos is centred in place."""

    def __init__(self, Pos, HSML, M, Rho, Bin_Quant, Weights=None, *, kernel,
                 calc_mean: bool = False, exact_norm: bool = False):
        super().__init__(Pos=Pos, HSML=HSML, M=M, Rho=Rho, Bin_Quant=Bin_Quant,
                         Weights=Weights, kernel=kernel, calc_mean=calc_mean,
                         exact_norm=exact_norm)


class AllskySubfile(dict):
    """Particles of one subfile for ``distributed_allsky_map``: the arguments
    of ``healpix_map`` (Float64 arrays; Pos is mutated as healpix_map does)."""

    def __init__(self, Pos, Hsml, M, Rho, Bin_q, Weights, *, kernel,
                 center=(0.0, 0.0, 0.0), radius_limits=(0.0, np.inf), calc_mean: bool = True):
        super().__init__(Pos=Pos, Hsml=Hsml, M=M, Rho=Rho, Bin_q=Bin_q, Weights=Weights,
                         kernel=kernel, center=center, radius_limits=radius_limits,
                         calc_mean=calc_mean)


class _Staged:
    """One subfile's data on the device (or nothing), with its upload event."""
    __slots__ = ("kind", "tensors", "event", "meta", "t_produce")


def _upload(arrays, dev, stream, torch):
    """Pinned staging + asynchronous copies on ``stream``; returns the device
    tensors and an event recorded after the last copy."""
    out = []
    with torch.cuda.stream(stream):
        for a in arrays:
            h = torch.from_numpy(np.ascontiguousarray(a)).pin_memory()
            out.append(h.to(dev, non_blocking=True))
        ev = torch.cuda.Event()
        ev.record(stream)
    return out, ev


def _summary(image):
    v = np.asarray(image)
    v = v[~np.isnan(v)]
    if v.size == 0:
        return
    print("Image values")
    print(f"    Min:    {v.min()}")
    print(f"    Max:    {v.max()}")
    print(f"    Mean:   {v.mean()}")
    print(f"    Median: {np.median(v)}")


def _run(Nsubfiles, produce, consume, prefetch):
    """Produce subfiles on a thread pool (at most ``prefetch`` ahead) and
    consume them in subfile order on the calling thread."""
    wait = 0.0
    with ThreadPoolExecutor(max_workers=max(1, prefetch)) as pool:
        futures = {k: pool.submit(produce, k) for k in range(min(prefetch, Nsubfiles))}
        for k in range(Nsubfiles):
            t0 = time.perf_counter()
            item = futures.pop(k).result()
            wait += time.perf_counter() - t0
            if k + prefetch < Nsubfiles:
                futures[k + prefetch] = pool.submit(produce, k + prefetch)
            consume(k, item)
    return wait


def distributed_cic_map(cic_filename: str, Nsubfiles: int, mapping_function,
                        param: MappingParameters, Nimages: int = 1, *,
                        reduce_image: bool = True, Ndim: int = 2, snap: int = 0,
                        units: str = "", vtk: bool = True, context=None, prefetch: int = 2,
                        stats: dict | None = None):
    """distributed_mapping/cic.jl:24-110.  Returns the reduced image (as
    ``reduce_image_2D`` / ``reduce_image_3D`` lay it out) after writing it to
    ``cic_filename`` (FITS; VTK for 3D with ``vtk``)."""
    import torch
    if Ndim not in (2, 3):
        raise ErrorException("Only 2D and 3D are possible")
    if Ndim == 3 and Nimages != 1:
        raise ValueError("the 3D deposit maps one quantity (Nimages == 1)")
    ctx = context or _lib.default_context()
    dev = torch.device("cuda", ctx.device)
    npix = int(param.Npixels[0])
    nd = npix ** Ndim
    ncol = Nimages + 1
    print("allocating image")
    total = torch.zeros(nd * ncol, dtype=torch.float64, device=dev)
    scratch = torch.empty(nd * ncol, dtype=torch.float64, device=dev)
    torch.cuda.synchronize(dev)
    # explicit streams (default priority): uploads on `copy`; deposits, sums
    # and the reduce on `compute`, which waits for each upload's event
    compute = torch.cuda.Stream(dev)
    copy = torch.cuda.Stream(dev)
    cstream = C.c_void_p(compute.cuda_stream)
    vp = lambda t: C.c_void_p(t.data_ptr())
    st = {"subfiles": 0, "particles": 0, "ms_device": 0.0}

    def produce(k):
        t0 = time.perf_counter()
        r = mapping_function(k)
        s = _Staged()
        s.kind, s.tensors, s.event, s.meta = None, None, None, None
        if isinstance(r, Subfile):
            kw = dict(r)
            args = [kw.pop(x) for x in ("Pos", "HSML", "M", "Rho", "Bin_Quant", "Weights")]
            pm = _prepare(*args, param=param, kernel=kw["kernel"], show_progress=False,
                          reduce_image=False, dimensions=Ndim, calc_mean=kw["calc_mean"],
                          exact_norm=kw["exact_norm"])
            if pm.nq != Nimages:
                raise ValueError(f"subfile {k}: {pm.nq} quantities, Nimages = {Nimages}")
            s.kind = "particles"
            s.meta = pm
            s.tensors, s.event = _upload(pm.arrays, dev, copy, torch)
        elif r is not None and r[0] is not None:
            maps, wmap = np.asarray(r[0], dtype=np.float64), np.asarray(r[1], dtype=np.float64)
            if maps.size != nd * Nimages or wmap.size != nd:
                raise ValueError(f"subfile {k}: maps of {maps.size} and {wmap.size} values, "
                                 f"expected {nd * Nimages} and {nd}")
            flat = np.concatenate([maps.ravel(order="F"), wmap.ravel(order="F")])
            s.kind = "maps"
            s.tensors, s.event = _upload([flat], dev, copy, torch)
        s.t_produce = time.perf_counter() - t0
        return s

    def consume(k, s):
        if s.kind is None:
            return
        compute.wait_event(s.event)
        for t in s.tensors:
            t.record_stream(compute)
        if s.kind == "maps":
            src = s.tensors[0]
        else:
            pm = s.meta
            x, h, m, rho, bq, w = s.tensors
            one = _lib.Stats()
            ctx.check(ctx.lib.s2g_deposit_device(ctx.handle, Ndim, pm.dtype_id, vp(x), vp(h),
                                                 vp(m), vp(rho), vp(bq), pm.nq, vp(w), pm.n,
                                                 C.byref(pm.params), pm.kid, pm.kdim,
                                                 pm.flags, vp(scratch), cstream, C.byref(one)),
                      "s2g_deposit_device")
            st["particles"] += pm.n
            st["ms_device"] += one.ms_device
            src = scratch
        ctx.check(ctx.lib.s2g_accumulate_device(ctx.handle, vp(src), vp(total), src.numel(),
                                                cstream), "s2g_accumulate_device")
        st["subfiles"] += 1

    print("running")
    t0 = time.perf_counter()
    wait = _run(Nsubfiles, produce, consume, prefetch)
    print("reducing image")
    flags = _lib.S2G_FLAG_REDUCE if (reduce_image or Ndim == 3) else 0
    nout = Nimages if Ndim == 2 else 1
    out = torch.empty(nd * nout, dtype=torch.float64, device=dev)
    ctx.check(ctx.lib.s2g_reduce_device(ctx.handle, Ndim, npix, Nimages, flags, vp(total),
                                        vp(out), cstream), "s2g_reduce_device")
    compute.synchronize()
    shape = (npix, npix, nout) if Ndim == 2 else (npix, npix, npix)
    image = out.cpu().numpy().reshape(shape, order="F")
    st["wall_s"] = time.perf_counter() - t0
    st["wait_s"] = wait
    print("saving")
    if Ndim == 3 and vtk:
        write_vtk_image(cic_filename, image, "map", param, snap=snap, units=units)
    else:
        write_fits_image(cic_filename, image, param, snap=snap, units=units)
    _summary(image)
    if stats is not None:
        stats.update(st)
    return image


def distributed_allsky_map(allsky_filename: str, Nside: int, Nsubfiles: int, mapping_function,
                           *, reduce_image: bool = True, context=None, prefetch: int = 2,
                           stats: dict | None = None):
    """distributed_mapping/healpix.jl:16-87.  Returns (allsky, weights) after
    writing the all-sky map to ``allsky_filename``."""
    import torch
    ctx = context or _lib.default_context()
    dev = torch.device("cuda", ctx.device)
    npx = 12 * int(Nside) ** 2
    print("allocating images")
    total = torch.zeros(2 * npx, dtype=torch.float64, device=dev)
    scratch = torch.empty(2 * npx, dtype=torch.float64, device=dev)
    torch.cuda.synchronize(dev)
    compute = torch.cuda.Stream(dev)
    copy = torch.cuda.Stream(dev)
    cstream = C.c_void_p(compute.cuda_stream)
    vp = lambda t: C.c_void_p(t.data_ptr())
    st = {"subfiles": 0, "particles": 0, "ms_device": 0.0}

    def produce(k):
        t0 = time.perf_counter()
        r = mapping_function(k)
        s = _Staged()
        s.kind, s.tensors, s.event, s.meta = None, None, None, None
        if isinstance(r, AllskySubfile):
            kw = dict(r)
            args = [kw.pop(x) for x in ("Pos", "Hsml", "M", "Rho", "Bin_q", "Weights")]
            prep = _prepare_healpix(*args, **kw)
            if prep is not None:
                arrs, n, kid, kdim, flags = prep
                s.kind = "particles"
                s.meta = (n, kid, kdim, flags)
                s.tensors, s.event = _upload(arrs, dev, copy, torch)
        elif r is not None:
            a, w = np.asarray(r[0], dtype=np.float64), np.asarray(r[1], dtype=np.float64)
            if a.size != npx or w.size != npx:
                raise ValueError(f"subfile {k}: maps of {a.size} / {w.size} pixels, "
                                 f"Nside {Nside} has {npx}")
            s.kind = "maps"
            s.tensors, s.event = _upload([np.concatenate([a.ravel(), w.ravel()])], dev, copy,
                                         torch)
        s.t_produce = time.perf_counter() - t0
        return s

    def consume(k, s):
        if s.kind is None:
            return
        compute.wait_event(s.event)
        for t in s.tensors:
            t.record_stream(compute)
        if s.kind == "maps":
            src = s.tensors[0]
        else:
            n, kid, kdim, flags = s.meta
            x, h, m, rho, bq, w = s.tensors
            one = _lib.Stats()
            ctx.check(ctx.lib.s2g_healpix_deposit_device(
                ctx.handle, _lib.S2G_F64, vp(x), vp(h), vp(m), vp(rho), vp(bq), vp(w), n,
                int(Nside), kid, kdim, flags, vp(scratch), C.c_void_p(scratch.data_ptr() + 8 * npx),
                cstream, C.byref(one)), "s2g_healpix_deposit_device")
            st["particles"] += n
            st["ms_device"] += one.ms_device
            src = scratch
        ctx.check(ctx.lib.s2g_accumulate_device(ctx.handle, vp(src), vp(total), 2 * npx,
                                                cstream), "s2g_accumulate_device")
        st["subfiles"] += 1

    print("running")
    t0 = time.perf_counter()
    wait = _run(Nsubfiles, produce, consume, prefetch)
    if reduce_image:
        print("reducing image")
        ctx.check(ctx.lib.s2g_healpix_reduce_device(ctx.handle, vp(total),
                                                    C.c_void_p(total.data_ptr() + 8 * npx),
                                                    npx, cstream), "s2g_healpix_reduce_device")
    compute.synchronize()
    both = total.cpu().numpy()
    st["wall_s"] = time.perf_counter() - t0
    st["wait_s"] = wait
    allsky, weights = both[:npx].copy(), both[npx:].copy()
    print("saving")
    if os.path.isfile(allsky_filename):
        os.remove(allsky_filename)
    save_healpix_fits(allsky_filename, allsky, int(Nside))
    _summary(allsky)
    v = allsky[~np.isnan(allsky)]
    print(f"    Sum:    {v.sum()}")
    if stats is not None:
        stats.update(st)
    return allsky, weights
