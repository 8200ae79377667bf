"""``mass_density`` -- effects/mass_density.jl:30-99 on the GPU.

SPH density and smoothing length of an arbitrary particle distribution (the
step before mapping stars or dark matter, which carry no density of their
own): hsml_i is the distance to the farthest of the Nneighbors + 1 nearest
particles (the particle itself included; the reference asks its BallTree for
Nneighbors + 1 neighbours, :64), rho_i the kernel sum over them (:73-81) and
the Wendland kernels' bias correction of Dehnen & Aly (2012, eqs. 18-19;
SPHKernels.jl ``bias_correction``, :86).  The neighbour search runs on the
device (Morton-ordered particles, exact k-th distance); there is no tree on
the host.
"""
from __future__ import annotations

import ctypes as C
import logging

import numpy as np

from . import _lib
from .kernels import Cubic

log = logging.getLogger("sph2grid")


def mass_density(pos, mass, *, kernel=None, Nneighbors: int = 32, boxsize=None,
                 verbose: bool = False, context=None, stats: dict | None = None):
    """effects/mass_density.jl:30-99.  ``pos`` is 3 x N (Float32 or Float64),
    ``mass`` N.  ``boxsize`` all zero (default) selects Euclidean distances,
    positive values a periodic box.  Returns (rho, hsml), Float64."""
    if kernel is None:
        kernel = Cubic(3)
    pos = np.asarray(pos)
    mass = np.asarray(mass)
    if pos.ndim != 2 or pos.shape[0] != 3:
        raise ValueError("pos must be a 3 x N array")
    n = mass.shape[0]
    if pos.shape[1] != n:
        raise ValueError("pos and mass disagree on the particle count")
    box = np.zeros(3) if boxsize is None else np.asarray(boxsize, dtype=np.float64).ravel()
    if box.size != 3:
        raise ValueError("boxsize must have three entries")
    if verbose:
        log.info("Periodic box" if np.any(box != 0) else "Non-periodic box")
    dt = np.float64 if (pos.dtype == np.float64 or mass.dtype == np.float64) else np.float32
    x = np.ascontiguousarray(pos.T, dtype=dt)
    m = np.ascontiguousarray(mass, dtype=dt)
    rho, hsml = np.empty(n), np.empty(n)
    ctx = context or _lib.default_context()
    st = _lib.Stats()
    ptr = lambda a: a.ctypes.data_as(C.c_void_p)
    rc = ctx.lib.s2g_mass_density(ctx.handle, _lib.S2G_F32 if dt == np.float32 else _lib.S2G_F64,
                                  ptr(x), ptr(m), n, kernel.kernel_id,
                                  kernel.dim if kernel.dim in (2, 3) else 3, int(Nneighbors),
                                  ptr(np.ascontiguousarray(box)), ptr(rho), ptr(hsml),
                                  C.byref(st))
    ctx.check(rc, "s2g_mass_density")
    if stats is not None:
        stats.update(st.as_dict())
    return rho, hsml
