"""TSC grid deposit -- the reference's `sphMapping(Pos, Bin_Quant; ...)` TSC
flavour (tsc_interpolation/tsc_interpolation.jl:40-178), which is entirely
commented out in SPHtoGrid.jl and called an external
`TSCInterpolation(bin_q, pos_tsc, Npixels, average=true)` (:131-133).

Definition used here (textbook TSC, Hockney & Eastwood):
* particles are kept if x_lim[1] <= x < x_lim[2] in every dimension
  (filter_particles_tsc, :1-6 of the commented file);
* grid units g = (x - center) * len2pix + Npix/2, as the 3D deposit;
* per axis the nearest cell and its two neighbours get the quadratic TSC
  weights, which sum to one; cells outside the grid are dropped, or wrapped
  when the parameters are periodic;
* ``average=True`` returns sum(q W) / sum(W) per cell (0 where no weight),
  ``average=False`` returns sum(q W);
* ``dimensions=3`` returns the grid indexed [z, y, x] (the 3D deposit's
  layout); ``dimensions=2`` sums it along z and returns [x, y].
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from .parameters import MappingParameters


def tsc_mapping(Pos, Bin_Quant, *, param: MappingParameters, dimensions: int = 3,
                average: bool = True, filter_particles: bool = True, periodic=None,
                context=None, stats: dict | None = None):
    Pos = np.asarray(Pos)
    q = np.asarray(Bin_Quant)
    if Pos.ndim != 2 or Pos.shape[0] != 3:
        raise ValueError("Pos must be a 3 x N array")
    if dimensions not in (2, 3):
        raise ValueError("dimensions must be 2 or 3")
    if filter_particles:
        keep = np.ones(Pos.shape[1], dtype=bool)
        for d, lim in enumerate((param.x_lim, param.y_lim, param.z_lim)):
            v = Pos[d].astype(np.float64)
            keep &= (lim[0] <= v) & (v < lim[1])
        Pos, q = Pos[:, keep], q[keep]
    dt = np.float64 if (Pos.dtype == np.float64 or q.dtype == np.float64) else np.float32
    x = (Pos.astype(np.float64) - param.center[:, None]).T
    x = np.ascontiguousarray(x, dtype=dt)
    qq = np.ascontiguousarray(q, dtype=dt)
    npix = int(param.Npixels[0])
    n = x.shape[0]
    per = param.periodic if periodic is None else bool(periodic)
    ctx = context or _lib.default_context()
    grid = np.empty(2 * npix ** 3)
    st = _lib.Stats()
    p = _lib.Params(npixels=npix, len2pix=float(param.len2pix))
    ptr = lambda a: a.ctypes.data_as(C.c_void_p)
    rc = ctx.lib.s2g_tsc3d(ctx.handle, _lib.S2G_F32 if dt == np.float32 else _lib.S2G_F64,
                           ptr(x), ptr(qq), n, C.byref(p),
                           _lib.S2G_FLAG_PERIODIC if per else 0, ptr(grid), C.byref(st))
    ctx.check(rc, "s2g_tsc3d")
    if stats is not None:
        stats.update(st.as_dict())
    nd = npix ** 3
    gq, gw = grid[:nd], grid[nd:]
    if average:
        out = np.zeros(nd)
        m = gw > 0
        out[m] = gq[m] / gw[m]
    else:
        out = gq.copy()
    out3 = out.reshape((npix, npix, npix), order="F")  # [z, y, x]
    if dimensions == 3:
        return out3
    return np.asfortranarray(out3.sum(axis=0).T)  # [x, y]
