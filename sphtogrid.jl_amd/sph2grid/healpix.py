"""``healpix_map`` -- host side of the all-sky deposit, mirroring
``healpix_interpolation/main.jl:92-227`` and ``filter_particles.jl:6-54``.

Returns two RING-ordered Float64 arrays of 12*Nside^2 pixels (0-based pixel
index): the quantity map and the weight map, unreduced, as the reference
returns its two ``HealpixMap``s.  The particle loop runs on the GPU through
``s2g_healpix_deposit``.
"""
from __future__ import annotations

import ctypes as C
import time

import numpy as np

from . import _lib
from .kernels import kernel_args


class BoundsError(IndexError):
    """Julia's BoundsError, raised where the reference raises it."""


def find_in_shell(dx, radius_limits):
    """filter_particles.jl:6-8."""
    return (radius_limits[0] <= dx) & (dx <= radius_limits[1])


def filter_sort_particles(Pos, Hsml, M, Rho, Bin_q, Weights, center, radius_limits,
                          calc_mean):
    """filter_particles.jl:17-54.  Subtracts ``center`` from ``Pos`` IN PLACE.

    The selection mirrors the reference exactly: ``sorted[sel]`` applies the
    shell mask, which is in the original particle order, to the radially
    sorted order; with ``calc_mean=False`` the reference only runs when no
    particle is removed and every ``Bin_q > 0`` (otherwise it raises a
    BoundsError, as here).  When the mask keeps every particle the selection
    is a permutation and the radial sort is skipped: the deposit does not
    depend on particle order beyond floating-point summation order.
    """
    for d in range(3):
        Pos[d] -= center[d]
    dx = np.sqrt(Pos[0] ** 2 + Pos[1] ** 2 + Pos[2] ** 2)
    sel = find_in_shell(dx, radius_limits)
    n = sel.shape[0]
    if not calc_mean:
        k = int(sel.sum())
        if k != n:
            raise BoundsError("BoundsError: sel[Bin_q[sel] .> 0.0] (filter_particles.jl:28-30)")
        if not np.all(np.asarray(Bin_q) > 0.0):
            raise BoundsError("BoundsError: sorted[sel] (filter_particles.jl:36)")
    if sel.all():
        idx = slice(None)
    else:
        order = np.argsort(dx, kind="stable")[::-1]
        idx = order[sel]
    return (Pos[:, idx], np.asarray(Hsml)[idx], np.asarray(M)[idx], np.asarray(Rho)[idx],
            np.asarray(Bin_q)[idx], np.asarray(Weights)[idx])


def _prepare_healpix(Pos, Hsml, M, Rho, Bin_q, Weights, *, center=(0.0, 0.0, 0.0),
                     radius_limits=(0.0, np.inf), kernel, calc_mean: bool = True):
    """The host steps of healpix_map before the particle loop (main.jl:100-143):
    type check, the early return of :114-121 (None), filter_sort_particles
    (mutates Pos) and the interleaved layout of the C ABI.  Returns
    (arrays, n, kernel id, kernel dim, flags) or None."""
    Pos = np.asarray(Pos)
    arrays = [Pos] + [np.asarray(a) for a in (Hsml, M, Rho, Bin_q, Weights)]
    if any(a.dtype != np.float64 for a in arrays):
        # calculate_weights is Float64-only (pixel_weights.jl:87-91)
        raise TypeError("healpix_map takes Float64 arrays (MethodError in the reference "
                        "for other element types)")
    if Pos.ndim != 2 or Pos.shape[0] != 3:
        raise ValueError("Pos must be a 3 x N array")
    kid, kdim = kernel_args(kernel)
    # main.jl:114-121
    if not calc_mean and np.sum(Bin_q) == 0:
        return None
    pos, hsml, m, rho, bq, w = filter_sort_particles(Pos, Hsml, M, Rho, Bin_q, Weights,
                                                     np.asarray(center, dtype=np.float64),
                                                     radius_limits, calc_mean)
    n = hsml.shape[0]
    cv = lambda a: np.ascontiguousarray(a, dtype=np.float64)
    arrs = (np.ascontiguousarray(pos.T),) + tuple(cv(a) for a in (hsml, m, rho, bq, w))
    flags = _lib.S2G_FLAG_CALC_MEAN if calc_mean else 0
    return arrs, n, kid, kdim, flags


def healpix_map(Pos, Hsml, M, Rho, Bin_q, Weights, *, center=(0.0, 0.0, 0.0),
                radius_limits=(0.0, np.inf), Nside: int = 1024, kernel,
                show_progress: bool = True, output_from_all_workers: bool = False,
                calc_mean: bool = True, context=None, stats: dict | None = None):
    """main.jl:92-227 on the GPU.  ``Pos`` (3 x N, Float64) is mutated."""
    npix = 12 * int(Nside) ** 2
    allsky = np.zeros(npix)
    weight = np.zeros(npix)
    t0 = time.perf_counter()
    prep = _prepare_healpix(Pos, Hsml, M, Rho, Bin_q, Weights, center=center,
                            radius_limits=radius_limits, kernel=kernel, calc_mean=calc_mean)
    if prep is None:
        return allsky, weight
    (x_il, hsml, m, rho, bq, w), n, kid, kdim, flags = prep
    ctx = context or _lib.default_context()
    st = _lib.Stats()
    ptr = lambda a: a.ctypes.data_as(C.c_void_p)
    rc = ctx.lib.s2g_healpix_deposit(ctx.handle, _lib.S2G_F64, ptr(x_il), ptr(hsml), ptr(m),
                                     ptr(rho), ptr(bq), ptr(w), n, int(Nside), kid, kdim,
                                     flags, ptr(allsky), ptr(weight), C.byref(st))
    ctx.check(rc, "s2g_healpix_deposit")
    if stats is not None:
        stats.update(st.as_dict())
        stats["wall_s"] = time.perf_counter() - t0
    return allsky, weight


def healpix_maps(Pos, Hsml, M, Rho, quantities, *, center=(0.0, 0.0, 0.0),
                 radius_limits=(0.0, np.inf), Nside: int = 1024, kernel,
                 calc_mean: bool = True, context=None, stats: dict | None = None):
    """Several ``healpix_map`` calls on one particle set in one deposit.

    ``quantities`` is a sequence of ``(Bin_q, Weights)`` pairs; returns the
    list of ``(map, weight_map)`` that ``healpix_map(Pos, Hsml, M, Rho, Bin_q,
    Weights, ...)`` returns for each (main.jl:92-227), with the same checks
    and early returns per map.  ``Pos`` is shifted by ``center`` once, as
    each reference call would shift its own copy.  The disc queries, the
    normalisation sums and the per-ring lists depend only on positions and
    hsml, so the GPU computes them once (``s2g_healpix_deposit_maps``)."""
    npix = 12 * int(Nside) ** 2
    Pos = np.asarray(Pos)
    arrays = [Pos] + [np.asarray(a) for a in (Hsml, M, Rho)]
    qs = [(np.asarray(q), np.asarray(w)) for q, w in quantities]
    if any(a.dtype != np.float64 for a in arrays + [x for p in qs for x in p]):
        raise TypeError("healpix_map takes Float64 arrays (MethodError in the reference "
                        "for other element types)")
    if Pos.ndim != 2 or Pos.shape[0] != 3:
        raise ValueError("Pos must be a 3 x N array")
    kid, kdim = kernel_args(kernel)
    out = [(np.zeros(npix), np.zeros(npix)) for _ in qs]
    # main.jl:114-121 per map; filter_sort_particles' checks per map (the
    # selection itself depends only on the positions)
    live = [k for k, (q, _) in enumerate(qs) if calc_mean or np.sum(q) != 0]
    if not live:
        return out
    for k in live:
        if not calc_mean:
            q = qs[k][0]
            dx = np.sqrt(((Pos - np.asarray(center, dtype=np.float64)[:, None]) ** 2).sum(0))
            if not find_in_shell(dx, radius_limits).all():
                raise BoundsError("BoundsError: sel[Bin_q[sel] .> 0.0] (filter_particles.jl:28-30)")
            if not np.all(q > 0.0):
                raise BoundsError("BoundsError: sorted[sel] (filter_particles.jl:36)")
    pos, hsml, m, rho, *rest = filter_sort_particles(
        Pos, Hsml, M, Rho, np.stack([qs[k][0] for k in live]).T,
        np.stack([qs[k][1] for k in live]).T, np.asarray(center, dtype=np.float64),
        radius_limits, True)
    bq, w = rest
    n = hsml.shape[0]
    cv = lambda a: np.ascontiguousarray(a, dtype=np.float64)
    x_il, hsml, m, rho = np.ascontiguousarray(pos.T), cv(hsml), cv(m), cv(rho)
    qk = [cv(bq[:, j]) for j in range(len(live))]
    wk = [cv(w[:, j]) for j in range(len(live))]
    ctx = context or _lib.default_context()
    st = _lib.Stats()
    vp = lambda a: a.ctypes.data_as(C.c_void_p)
    arr = lambda xs: (C.c_void_p * len(xs))(*[x.ctypes.data for x in xs])
    maps = [out[k][0] for k in live]
    wmaps = [out[k][1] for k in live]
    flags = _lib.S2G_FLAG_CALC_MEAN if calc_mean else 0
    rc = ctx.lib.s2g_healpix_deposit_maps(ctx.handle, _lib.S2G_F64, vp(x_il), vp(hsml), vp(m),
                                          vp(rho), len(live), arr(qk), arr(wk), n, int(Nside),
                                          kid, kdim, flags, arr(maps), arr(wmaps),
                                          C.byref(st))
    ctx.check(rc, "s2g_healpix_deposit_maps")
    if stats is not None:
        stats.update(st.as_dict())
    return out
