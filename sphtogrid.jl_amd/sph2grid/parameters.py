"""``mappingParameters`` -- mirror of SPHtoGrid.jl's parameter object.

Follows ``shared/parameters.jl:14-27`` (fields) and ``:44-125`` (constructor),
including its sentinels ([-1,-1] limits, [-1,-1,-1] centre, -1 sizes and
pixel length, 0 pixels) and its two error messages.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np


class ErrorException(Exception):
    """Julia's ``ErrorException``: raised with the reference's exact message."""


@dataclass
class MappingParameters:
    x_lim: np.ndarray
    y_lim: np.ndarray
    z_lim: np.ndarray
    center: np.ndarray
    halfsize: np.ndarray
    len2pix: float
    pixelSideLength: float
    Npixels: np.ndarray = field(default_factory=lambda: np.zeros(3, dtype=np.int64))
    boxsize: float = -1.0
    periodic: bool = False


def _vec(v, n):
    a = np.array(v, dtype=np.float64).reshape(-1)
    if a.shape[0] != n:
        raise ValueError(f"expected {n} values, got {a.shape[0]}")
    return a


def mappingParameters(*, x_lim=(-1.0, -1.0), y_lim=(-1.0, -1.0), z_lim=(-1.0, -1.0),
                      center=(-1.0, -1.0, -1.0), x_size=-1.0, y_size=-1.0, z_size=-1.0,
                      pixelSideLength=-1.0, Npixels=0, boxsize=-1.0) -> MappingParameters:
    """shared/parameters.jl:44-125."""
    x_lim, y_lim, z_lim = _vec(x_lim, 2), _vec(y_lim, 2), _vec(z_lim, 2)
    center = _vec(center, 3)
    unset2 = lambda v: v[0] == -1.0 and v[1] == -1.0
    c_unset = bool(np.all(center == -1.0))

    # parameters.jl:58-70
    if unset2(x_lim) and unset2(y_lim) and unset2(z_lim):
        if (not c_unset) and x_size != -1.0 and y_size != -1.0 and z_size != -1.0:
            x_lim = np.array([center[0] - 0.5 * x_size, center[0] + 0.5 * x_size])
            y_lim = np.array([center[1] - 0.5 * y_size, center[1] + 0.5 * y_size])
            z_lim = np.array([center[2] - 0.5 * z_size, center[2] + 0.5 * z_size])
        else:
            raise ErrorException(
                "Giving a center position requires extent in x, y and z direction.")
    # parameters.jl:73-81
    if x_size == -1.0:
        x_size = float(x_lim[1] - x_lim[0])
    if y_size == -1.0:
        y_size = float(y_lim[1] - y_lim[0])
    if z_size == -1.0:
        z_size = float(z_lim[1] - z_lim[0])
    # parameters.jl:83-87
    if c_unset:
        center = np.array([x_lim[0] + 0.5 * x_size, y_lim[0] + 0.5 * y_size,
                           z_lim[0] + 0.5 * z_size])
    # parameters.jl:90-100
    max_size = max(float(x_size), float(y_size))
    if pixelSideLength == -1.0 and Npixels != 0:
        pixelSideLength = max_size / Npixels
    elif pixelSideLength != -1.0 and Npixels == 0:
        Npixels = int(np.floor(max_size / pixelSideLength))
        pixelSideLength = max_size / Npixels
    else:
        raise ErrorException("Please specify pixelSideLength or number of pixels!")
    Npix = np.array([Npixels, Npixels, Npixels], dtype=np.int64)  # parameters.jl:105
    periodic = boxsize != -1.0  # parameters.jl:107-111
    halfsize = 0.5 * np.array([x_size, y_size, z_size], dtype=np.float64)
    len2pix = 1.0 / pixelSideLength  # parameters.jl:115
    return MappingParameters(x_lim=np.asarray(x_lim, dtype=np.float64),
                             y_lim=np.asarray(y_lim, dtype=np.float64),
                             z_lim=np.asarray(z_lim, dtype=np.float64),
                             center=np.asarray(center, dtype=np.float64),
                             halfsize=halfsize, len2pix=float(len2pix),
                             pixelSideLength=float(pixelSideLength), Npixels=Npix,
                             boxsize=float(boxsize), periodic=bool(periodic))
