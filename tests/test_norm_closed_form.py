"""Bound of the closed-form footprint normalisation used for large
unclipped particles.

The reference normalises every particle by the discrete footprint sum
S = sum_pixels shape(u) * A (cic_2D.jl:49-69).  For a compactly supported
kernel of smoothness C^k the midpoint rule converges like h^-(k+2), so for
h >= h_min(kernel) S equals h^2 * I_k (I_k = integral of the shape over the
unit disc = 1/norm_2D) to better than 1e-9 relative -- far below the FP32
rounding of the deposit.  The thresholds and I_k are read from the library
source, so this test pins the values the GPU path uses.  CPU only.
"""
import os
import re

import numpy as np
import pytest
from scipy import integrate

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = open(os.path.join(ROOT, "sphtogrid.jl_amd", "csrc", "s2g_internal.h")).read()
NAMES = {"cubic": "CUBIC", "quintic": "QUINTIC", "wendlandc2": "WENDLAND_C2",
         "wendlandc4": "WENDLAND_C4", "wendlandc6": "WENDLAND_C6", "wendlandc8": "WENDLAND_C8"}


def _table(func):
    body = SRC[SRC.index(func):]
    body = body[:body.index("\n}\n")]
    out = {}
    for k, v in re.findall(r"case S2G_KERNEL_(\w+):\s*return ([^;]+);", body):
        out[k] = v
    out["WENDLAND_C8"] = re.search(r"default:\s*return ([^;]+);", body).group(1)
    return out


def _eval(expr):
    return float(eval(expr.replace("pi", "np.pi")))


H_MIN = {k: _eval(v) for k, v in _table("closed_form_min_h").items()}
I_K = {k: _eval(v) for k, v in _table("shape_disc_integral").items()}


@pytest.mark.parametrize("kernel", list(NAMES))
def test_disc_integral_constant(orc, kernel):
    ref = integrate.quad(lambda u: orc.kernel_shape(kernel, u) * 2 * np.pi * u, 0, 1,
                         points=[1 / 3, 0.5, 2 / 3], epsabs=1e-15, epsrel=1e-13)[0]
    assert abs(I_K[NAMES[kernel]] / ref - 1) < 1e-12
    # and it is 1 / norm_2D
    assert abs(I_K[NAMES[kernel]] * orc.kernel_norm(kernel, 2) - 1) < 1e-12


@pytest.mark.parametrize("kernel", list(NAMES))
def test_closed_form_matches_discrete_sum(orc, kernel):
    r = np.random.default_rng(42)
    hmin = H_MIN[NAMES[kernel]]
    shape = np.vectorize(lambda u: orc.kernel_shape(kernel, u))
    worst = 0.0
    for h in (hmin, hmin * 1.37, hmin * 2.9):
        for _ in range(6):
            x = 500.0 + r.random()
            y = 500.0 + r.random()
            i = np.arange(np.floor(x - h), np.floor(x + h) + 1)
            j = np.arange(np.floor(y - h), np.floor(y + h) + 1)
            ax = np.minimum(x + h, i + 1) - np.maximum(x - h, i)
            ay = np.minimum(y + h, j + 1) - np.maximum(y - h, j)
            u = np.sqrt((x - i - 0.5)[:, None] ** 2 + (y - j - 0.5)[None, :] ** 2) / h
            S = (shape(u) * ax[:, None] * ay[None, :]).sum()
            worst = max(worst, abs(S / (h * h * I_K[NAMES[kernel]]) - 1))
    assert worst < 2e-9, worst
