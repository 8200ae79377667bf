"""Many independent 2D maps in one device pass (``s2g_deposit_batch``).

The reference makes a set of maps by calling ``sphMapping`` once per map
(cic_interpolation.jl:35-273), or per subfile in ``distributed_cic_map``
(distributed_mapping/cic.jl:24-110).  Here the maps of a batch share one pass
of the deposit pipeline: their particles and tiles are concatenated, so small
maps fill the GPU together and the host synchronises a few times per batch
instead of per map.

``map_batch`` takes a list of ``sphMapping`` argument sets and returns what
``sphMapping`` would return for each.  ``deposit_batch_device`` is the thin
wrapper of the C ABI for inputs that are already in device memory (torch
tensors or raw device addresses).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from .kernels import kernel_args
from .mapping import _prepare


def _addr(a) -> int:
    if a is None:
        return 0
    if isinstance(a, int):
        return a
    if hasattr(a, "data_ptr"):
        return int(a.data_ptr())
    raise TypeError(f"expected a device tensor or address, got {type(a).__name__}")


def deposit_batch_device(maps, *, dtype: str = "f32", context=None, stream=None,
                         stats: dict | None = None):
    """``s2g_deposit_batch`` on device inputs.

    ``maps``: sequence of dicts with keys ``pos`` (3 n, interleaved xyz),
    ``hsml``, ``m``, ``rho``, ``bin_q`` (n nq, quantity fastest), ``weights``,
    ``out`` (flat image of (nq+1) npix^2 doubles, or with ``reduce_image`` the
    reduced map of nq npix^2 doubles), ``n``, ``nq``, ``npix``, ``len2pix``,
    ``kernel`` (kernel object or id) and optionally ``calc_mean``,
    ``reduce_image``, ``exact_norm``, ``fp32_accum``.  Arrays are torch CUDA tensors or device
    addresses.  Particles must already be centred and filtered."""
    ctx = context or _lib.default_context()
    descs = _descs(maps)
    st = _lib.Stats()
    rc = ctx.lib.s2g_deposit_batch(ctx.handle, _dtype_id(dtype), descs, len(maps),
                                   C.c_void_p(stream) if stream else None, C.byref(st))
    ctx.check(rc, "s2g_deposit_batch")
    if stats is not None:
        stats.update(st.as_dict())


def _dtype_id(dtype) -> int:
    return _lib.S2G_F32 if dtype in ("f32", "float32", np.float32) else _lib.S2G_F64


def _descs(maps):
    """s2g_map_desc array of deposit_batch_device's map dicts."""
    descs = (_lib.MapDesc * max(len(maps), 1))()
    for k, d in enumerate(maps):
        kid = d["kernel"] if isinstance(d["kernel"], int) else kernel_args(d["kernel"])[0]
        flags = ((_lib.S2G_FLAG_CALC_MEAN if d.get("calc_mean", False) else 0) |
                 (_lib.S2G_FLAG_REDUCE if d.get("reduce_image", False) else 0) |
                 (_lib.S2G_FLAG_EXACT_NORM if d.get("exact_norm", False) else 0) |
                 (_lib.S2G_FLAG_FP32_ACCUM if d.get("fp32_accum", False) else 0))
        descs[k] = _lib.MapDesc(
            dims=2, nq=int(d["nq"]), n=int(d["n"]),
            params=_lib.Params(npixels=int(d["npix"]), len2pix=float(d["len2pix"])),
            kernel=kid, flags=flags,
            d_pos=_addr(d["pos"]), d_hsml=_addr(d["hsml"]), d_m=_addr(d["m"]),
            d_rho=_addr(d["rho"]), d_bin_q=_addr(d["bin_q"]), d_weights=_addr(d["weights"]),
            d_out=_addr(d["out"]))
    return descs


class BatchPlan:
    """A batch captured once in a hipGraph and replayed (``s2g_plan_*``).

    ``maps`` as for ``deposit_batch_device``; the device arrays must stay
    allocated (and at the same addresses) for the plan's lifetime, their
    contents may change between runs.  Capacities are the creation run's
    counts times ``headroom``; ``run`` raises ``LibraryError`` with code
    ``S2G_ECAPACITY`` when the inputs outgrow them (the outputs are then
    incomplete; build a new plan)."""

    def __init__(self, maps, *, dtype: str = "f32", headroom: float = 1.25, context=None):
        self.ctx = context or _lib.default_context()
        self._maps = list(maps)  # keeps the tensors alive
        self._descs = _descs(self._maps)
        h = C.c_void_p()
        rc = self.ctx.lib.s2g_plan_create(self.ctx.handle, _dtype_id(dtype), self._descs,
                                          len(self._maps), float(headroom), C.byref(h))
        self.ctx.check(rc, "s2g_plan_create")
        self.handle = h

    def run(self, stream=None, stats: dict | None = None):
        st = _lib.Stats()
        rc = self.ctx.lib.s2g_plan_run(self.handle, C.c_void_p(stream) if stream else None,
                                       C.byref(st))
        self.ctx.check(rc, "s2g_plan_run")
        if stats is not None:
            stats.update(st.as_dict())

    def close(self):
        if getattr(self, "handle", None):
            self.ctx.lib.s2g_plan_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def map_batch(calls, *, context=None, stats: dict | None = None):
    """Run several 2D ``sphMapping`` calls as one batch.

    ``calls``: sequence of dicts of ``sphMapping`` arguments (``Pos``,
    ``HSML``, ``M``, ``Rho``, ``Bin_Quant``, optional ``Weights``, ``param``,
    ``kernel``, ``reduce_image``, ``return_both_maps``, ``calc_mean``,
    ``sort_z``, ``exact_norm``).  Each ``Pos`` is centred in place, as
    sphMapping does.  Returns the list of results in call order, each laid out
    as sphMapping returns it.  Calls with Float32 inputs and with Float64
    inputs go to separate device passes.  Device memory comes from torch."""
    import torch

    dev = torch.device("cuda", (context or _lib.default_context()).device)
    prepared = []
    for kw in calls:
        kw = dict(kw)
        if kw.pop("dimensions", 2) != 2:
            raise ValueError("map_batch maps 2D images only")
        for k in ("RM", "show_progress", "parallel", "stokes"):
            kw.pop(k, None)
        args = [kw.pop(k) for k in ("Pos", "HSML", "M", "Rho", "Bin_Quant")]
        w = kw.pop("Weights", None)
        prepared.append(_prepare(*args, w, show_progress=False, **kw))
    results = [None] * len(prepared)
    total = {}
    for dtype in (np.float32, np.float64):
        idx = [i for i, pm in enumerate(prepared) if pm.dtype == dtype]
        if not idx:
            continue
        tdt = torch.float32 if dtype == np.float32 else torch.float64
        keep, descs, outs = [], [], []
        for i in idx:
            pm = prepared[i]
            t = [torch.from_numpy(a).to(dev, dtype=tdt) for a in pm.arrays]
            keep.append(t)
            nd = pm.npix * pm.npix
            # reduce_image=true: the pass writes the reduced map itself;
            # reduce_image=false: flat image, transposed afterwards without
            # division (reduce_image.jl:8-31); return_both_maps: flat image
            reduced = bool(pm.flags & _lib.S2G_FLAG_REDUCE) and not pm.return_both_maps
            out = torch.empty(nd * (pm.nq if reduced else pm.nq + 1), dtype=torch.float64,
                              device=dev)
            outs.append((i, out, reduced))
            descs.append(dict(pos=t[0], hsml=t[1], m=t[2], rho=t[3], bin_q=t[4], weights=t[5],
                              out=out, n=pm.n, nq=pm.nq, npix=pm.npix,
                              len2pix=pm.params.len2pix, kernel=pm.kid,
                              calc_mean=bool(pm.flags & _lib.S2G_FLAG_CALC_MEAN),
                              exact_norm=bool(pm.flags & _lib.S2G_FLAG_EXACT_NORM),
                              reduce_image=reduced))
        st = {}
        stream = torch.cuda.current_stream(dev).cuda_stream
        deposit_batch_device(descs, dtype="f32" if dtype == np.float32 else "f64",
                             context=context, stream=stream, stats=st)
        torch.cuda.synchronize(dev)
        for k, v in st.items():
            total[k] = total.get(k, 0) + v
        ctx = context or _lib.default_context()
        for i, out, reduced in outs:
            pm = prepared[i]
            nd = pm.npix * pm.npix
            if not reduced and not pm.return_both_maps:
                tr = torch.empty(nd * pm.nq, dtype=torch.float64, device=dev)
                ctx.check(ctx.lib.s2g_reduce_device(ctx.handle, 2, pm.npix, pm.nq, 0,
                                                    C.c_void_p(out.data_ptr()),
                                                    C.c_void_p(tr.data_ptr()),
                                                    C.c_void_p(stream)), "s2g_reduce_device")
                torch.cuda.synchronize(dev)
                out, reduced = tr, True
            a = out.cpu().numpy()
            if reduced:
                results[i] = a.reshape((pm.npix, pm.npix, pm.nq), order="F")
            else:
                results[i] = a.reshape((nd, pm.nq + 1), order="F")
    if stats is not None:
        stats.update(total)
    return results
