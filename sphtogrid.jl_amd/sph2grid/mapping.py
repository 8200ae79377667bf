"""``sphMapping`` -- host side of the deposition path, mirroring
``cic_interpolation/cic_interpolation.jl:35-273`` and
``cic_interpolation/filter_shift.jl:6-67``.

Array conventions follow the Julia surface: ``Pos`` is 3 x N (``Pos[:, i]`` is
particle i), ``Bin_Quant`` is N or Nq x N, and the result is indexed like the
Julia array it replaces (2D: ``[x, y, q]``, 3D: ``[z, y, x]``), returned in
Fortran order so that the memory layout matches Julia's as well.

The deposit itself (``cic_mapping_2D`` / ``cic_mapping_3D``) and the reduction
(``reduce_image_2D`` / ``reduce_image_3D``) run on the GPU through
libsph2grid's C ABI.  ``parallel=True`` routes to the same GPU call.
"""
from __future__ import annotations

import ctypes as C
import logging
import time

import numpy as np

from . import _lib
from .kernels import kernel_args
from .parameters import MappingParameters, mappingParameters

log = logging.getLogger("sph2grid")


def center_particles(x: np.ndarray, par: MappingParameters):
    """filter_shift.jl:6-32.  Shifts ``x`` (3 x N) IN PLACE, as the reference
    does: each coordinate is computed in Float64 and stored back in the
    array's own element type; the periodic shift is boxsize/2 (:18-22)."""
    cen = par.center.copy()
    for d in range(3):
        v = x[d].astype(np.float64) - cen[d]
        x[d] = v.astype(x.dtype)
        if par.periodic:
            cur = x[d].astype(np.float64)
            big = np.abs(cur) > par.boxsize / 2
            if np.any(big):
                shifted = np.where(cur > 0, cur - par.boxsize / 2, cur + par.boxsize / 2)
                x[d] = np.where(big, shifted, cur).astype(x.dtype)
    xlim = par.x_lim - cen[0]
    ylim = par.y_lim - cen[1]
    zlim = par.z_lim - cen[2]
    par2 = mappingParameters(center=[0.0, 0.0, 0.0], x_lim=xlim, y_lim=ylim, z_lim=zlim,
                             Npixels=int(np.max(par.Npixels)), boxsize=par.boxsize)
    return x, par2


def filter_particles_in_image(pos: np.ndarray, par: MappingParameters, sort_z: bool = False):
    """filter_shift.jl:40-67: keep particles whose centre lies in
    [center-halfsize, center+halfsize] (inclusive) in all three dimensions.
    With ``sort_z`` the reference returns ``reverse(sortperm(z))[mask]`` --
    the original-order mask applied to the sorted order -- which is mirrored."""
    lo = par.center - par.halfsize
    hi = par.center + par.halfsize
    keep = np.ones(pos.shape[1], dtype=bool)
    for d in range(3):
        v = pos[d].astype(np.float64)
        keep &= (lo[d] <= v) & (v <= hi[d])
    if sort_z:
        order = np.argsort(pos[2].astype(np.float64), kind="stable")[::-1]
        return order[keep]
    return keep


def _common_dtype(arrays):
    return np.float64 if any(a.dtype == np.float64 for a in arrays) else np.float32


def _check_float(name, a):
    if not np.issubdtype(a.dtype, np.floating) or a.dtype not in (np.float32, np.float64):
        raise TypeError(f"{name} must be a Float32 or Float64 array (got {a.dtype}); "
                        "strip units before mapping")


class _Prepared:
    """A sphMapping call after the host steps: centred, filtered, converted."""
    __slots__ = ("arrays", "nq", "npix", "n", "flags", "kid", "kdim", "params", "dtype",
                 "dtype_id", "dimensions", "return_both_maps")


def _prepare(Pos, HSML, M, Rho, Bin_Quant, Weights=None, *, param: MappingParameters, kernel,
             show_progress: bool = True, reduce_image: bool = True, dimensions: int = 2,
             calc_mean: bool = False, stokes: bool = False, sort_z: bool = False,
             exact_norm: bool = False, return_both_maps: bool = False) -> _Prepared:
    """The host steps of sphMapping before the deposit (cic_interpolation.jl:
    60-150): argument checks, center_particles (mutates Pos), the image
    filter, Float32/Float64 promotion and the interleaved layouts of the C ABI."""
    Pos = np.asarray(Pos)
    if Pos.ndim != 2 or Pos.shape[0] != 3:
        raise ValueError("Pos must be a 3 x N array")
    arrays = {"Pos": Pos, "HSML": np.asarray(HSML), "M": np.asarray(M),
              "Rho": np.asarray(Rho), "Bin_Quant": np.asarray(Bin_Quant)}
    if Weights is None:
        Weights = Rho
    arrays["Weights"] = np.asarray(Weights)
    for k, a in arrays.items():
        _check_float(k, a)
    if dimensions not in (2, 3):
        raise ValueError("dimensions must be 2 or 3")
    kid, kdim = kernel_args(kernel)
    N_in = arrays["M"].shape[0]

    Pos, par = center_particles(Pos, param)
    if stokes:
        sort_z = True
    p_in_image = filter_particles_in_image(Pos, par, sort_z)
    x = Pos[:, p_in_image]
    hsml = arrays["HSML"][p_in_image]
    m = arrays["M"][p_in_image]
    rho = arrays["Rho"][p_in_image]
    bq = arrays["Bin_Quant"]
    bin_q = bq[p_in_image] if bq.ndim == 1 else bq[:, p_in_image]
    weights = arrays["Weights"][p_in_image]
    N_map = m.shape[0]
    if show_progress:
        log.info("Particles in image: %d / %d", N_map, N_in)

    nq = 1 if bin_q.ndim == 1 else bin_q.shape[0]
    if dimensions == 3:
        if nq != 1:
            raise ValueError("the 3D deposit maps a scalar Bin_Quant (cic_3D.jl:140)")
        # calc_mean is not forwarded to cic_mapping_3D (cic_interpolation.jl:219-221)
        calc_mean = False
    dt = _common_dtype([x, hsml, m, rho, bin_q, weights])
    cvt = lambda a: np.ascontiguousarray(a, dtype=dt)
    pm = _Prepared()
    # interleaved xyz per particle (Julia's 3 x N column-major memory)
    pm.arrays = (np.ascontiguousarray(x.T, dtype=dt), cvt(hsml), cvt(m), cvt(rho),
                 cvt(bin_q.T if nq > 1 else bin_q), cvt(weights))
    pm.nq = nq
    pm.npix = int(param.Npixels[0])
    pm.n = N_map
    pm.flags = ((_lib.S2G_FLAG_CALC_MEAN if calc_mean else 0) |
                (_lib.S2G_FLAG_EXACT_NORM if exact_norm else 0) |
                (_lib.S2G_FLAG_REDUCE if reduce_image else 0))
    pm.kid, pm.kdim = kid, kdim
    pm.params = _lib.Params(npixels=int(par.Npixels[0]), len2pix=float(par.len2pix))
    pm.dtype = dt
    pm.dtype_id = _lib.S2G_F32 if dt == np.float32 else _lib.S2G_F64
    pm.dimensions = dimensions
    pm.return_both_maps = return_both_maps
    return pm


def sphMapping(Pos, HSML, M, Rho, Bin_Quant, Weights=None, RM=None, *,
               param: MappingParameters, kernel, show_progress: bool = True,
               parallel: bool = False, reduce_image: bool = True,
               return_both_maps: bool = False, dimensions: int = 2,
               calc_mean: bool = False, stokes: bool = False, sort_z: bool = False,
               exact_norm: bool = False, context=None, stats: dict | None = None):
    """cic_interpolation.jl:35-273 on the GPU.

    ``Pos`` is mutated (centred) exactly as the reference mutates it.  ``RM``
    is accepted and, as in the reference, has no effect on the maps (the
    Faraday rotation is never enabled on this path, cic_interpolation.jl:152-155,
    and is dropped on the parallel path, :191-196).  ``stokes`` implies
    ``sort_z`` (:74-81).  ``exact_norm`` disables the closed-form per-particle
    normalisation of large unclipped footprints (it changes results by less
    than 1e-9 relative).
    """
    t0 = time.perf_counter()
    pm = _prepare(Pos, HSML, M, Rho, Bin_Quant, Weights, param=param, kernel=kernel,
                  show_progress=show_progress, reduce_image=reduce_image,
                  dimensions=dimensions, calc_mean=calc_mean, stokes=stokes, sort_z=sort_z,
                  exact_norm=exact_norm)
    x_il, hsml_c, m_c, rho_c, bq_il, w_c = pm.arrays
    nq, npix, N_map, flags = pm.nq, pm.npix, pm.n, pm.flags
    kid, kdim = pm.kid, pm.kdim
    p = pm.params
    dtype_id = pm.dtype_id
    ctx = context or _lib.default_context()
    st = _lib.Stats()
    ptr = lambda a: a.ctypes.data_as(C.c_void_p)
    args = (ctx.handle, dimensions, dtype_id, ptr(x_il), ptr(hsml_c), ptr(m_c), ptr(rho_c),
            ptr(bq_il), nq, ptr(w_c), N_map, C.byref(p), kid, kdim, flags)
    nd = npix ** dimensions
    if dimensions == 2 and return_both_maps:
        out = np.empty(nd * (nq + 1), dtype=np.float64)
        ctx.check(ctx.lib.s2g_deposit(*args, ptr(out), C.byref(st)), "s2g_deposit")
        result = out.reshape((nd, nq + 1), order="F")
    elif dimensions == 2:
        out = np.empty(nd * nq, dtype=np.float64)
        ctx.check(ctx.lib.s2g_map(*args, ptr(out), C.byref(st)), "s2g_map")
        result = out.reshape((npix, npix, nq), order="F")
    else:
        out = np.empty(nd, dtype=np.float64)
        ctx.check(ctx.lib.s2g_map(*args, ptr(out), C.byref(st)), "s2g_map")
        result = out.reshape((npix, npix, npix), order="F")
    if stats is not None:
        stats.update(st.as_dict())
        stats["wall_s"] = time.perf_counter() - t0
    if show_progress:
        log.info("Mapping done: device %.3f ms, %d small / %d large particles",
                 st.ms_device, st.n_small, st.n_large)
    return result


def deposit(x, hsml, m, rho, bin_q, weights, *, npix: int, len2pix: float, kernel,
            dimensions: int = 2, calc_mean: bool = True, exact_norm: bool = False,
            fp32_accum: bool = False,
            context=None, stats: dict | None = None) -> np.ndarray:
    """The boundary itself: ``cic_mapping_2D`` / ``cic_mapping_3D`` on already
    centred and filtered particles.  ``x`` is N x 3 (interleaved), ``bin_q`` N
    or N x Nq.  Returns the flat unreduced image (N_distr x (Nq+1)), Fortran
    order, deposit order, last column = weights (cic_2D.jl:103-244).
    ``fp32_accum`` (2D): the tile gather sums each work item's pixels in FP32
    (S2G_FLAG_FP32_ACCUM), faster with three or four planes at rel-L1 about
    1e-6 instead of 1e-8."""
    x = np.asarray(x)
    dt = _common_dtype([x, np.asarray(hsml), np.asarray(m), np.asarray(rho),
                        np.asarray(bin_q), np.asarray(weights)])
    cvt = lambda a: np.ascontiguousarray(a, dtype=dt)
    x_il, hsml_c, m_c, rho_c, w_c, bq = (cvt(a) for a in (x, hsml, m, rho, weights, bin_q))
    n = x_il.shape[0]
    nq = 1 if bq.ndim == 1 else bq.shape[1]
    kid, kdim = kernel_args(kernel)
    ctx = context or _lib.default_context()
    p = _lib.Params(npixels=int(npix), len2pix=float(len2pix))
    flags = (_lib.S2G_FLAG_CALC_MEAN if calc_mean else 0) | \
        (_lib.S2G_FLAG_EXACT_NORM if exact_norm else 0) | \
        (_lib.S2G_FLAG_FP32_ACCUM if fp32_accum else 0)
    nd = int(npix) ** dimensions
    ncol = nq + 1 if dimensions == 2 else 2
    out = np.empty(nd * ncol, dtype=np.float64)
    st = _lib.Stats()
    ptr = lambda a: a.ctypes.data_as(C.c_void_p)
    rc = ctx.lib.s2g_deposit(ctx.handle, dimensions,
                             _lib.S2G_F32 if dt == np.float32 else _lib.S2G_F64,
                             ptr(x_il), ptr(hsml_c), ptr(m_c), ptr(rho_c), ptr(bq), nq,
                             ptr(w_c), n, C.byref(p), kid, kdim, flags, ptr(out),
                             C.byref(st))
    ctx.check(rc, "s2g_deposit")
    if stats is not None:
        stats.update(st.as_dict())
    return out.reshape((nd, ncol), order="F")
