"""``Weights`` presets (cic_interpolation/weight_functions.jl:10-63).

Host-side, elementwise input preparation: they produce the ``Weights`` array
passed to ``sphMapping``.
"""
from __future__ import annotations

import numpy as np

# weight_functions.jl:1-3
ev2K = 1.160451812e4
K2eV = 1.0 / 1.160451812e4
kB = 1.38066e-16
# effects/constants.jl:11,13 (conversion used by part_weight_XrayBand)
eV2cgs = 1.60218e-12
cgs2eV = 1.0 / eV2cgs


def part_weight_one(N: int) -> np.ndarray:
    """weight_functions.jl:10-12."""
    return np.ones(int(N))


def part_weight_physical(N: int, par=None, x_cgs: float = 3.085678e21) -> np.ndarray:
    """weight_functions.jl:20-22 (with ``par``) and :30-32 (without)."""
    if par is None:
        return np.ones(int(N)) * x_cgs
    if not hasattr(par, "pixelSideLength"):
        # part_weight_physical(N, x_cgs) positional form
        return np.ones(int(N)) * float(par)
    return np.ones(int(N)) * par.pixelSideLength * x_cgs


def part_weight_emission(rho, T_K) -> np.ndarray:
    """weight_functions.jl:39-41: rho^2 sqrt(T)."""
    rho = np.asarray(rho, dtype=np.float64)
    return rho ** 2 * np.sqrt(np.asarray(T_K, dtype=np.float64))


def part_weight_spectroscopic(rho, T_K) -> np.ndarray:
    """weight_functions.jl:49-51: rho^2 / T^(3/4) (Mazzotta+04)."""
    rho = np.asarray(rho, dtype=np.float64)
    return rho ** 2 / np.sqrt(np.sqrt(np.asarray(T_K, dtype=np.float64))) ** 3


def part_weight_XrayBand(T_K, Emin: float = 5.0e4, Emax: float = 1.0e10) -> np.ndarray:
    """weight_functions.jl:58-63."""
    T_eV = np.asarray(T_K, dtype=np.float64) * cgs2eV
    with np.errstate(over="ignore", under="ignore", divide="ignore"):
        return np.exp(-Emin / (kB * T_eV)) - np.exp(-Emax / (kB * T_eV))
