"""HEALPix RING geometry built independently of the oracle, for the tests'
closed forms (no pixel-index inversion; CPU only, test infrastructure).
"""
import numpy as np


# ---------------------------------------------------------------------------
# Per-pixel area weights (pixel_weights.jl:6-8, 34-76, 120-137), pinned by
# geometry built here ring by ring from Gorski et al. (2005, sec. 4): RING
# order runs over the 4*nside - 1 iso-latitude rings from north to south;
# polar ring i (1 <= i < nside) holds 4i pixels at z = 1 - i^2 / (3 nside^2),
# centres at phi = (j - 1/2) pi / (2i), j = 1..4i; equatorial ring i
# (nside <= i <= 3 nside) holds 4 nside pixels at z = 4/3 - 2i / (3 nside),
# centres at phi = (j - f) pi / (2 nside), j = 1..4 nside, with f = 1/2 when
# i + nside is even and f = 1 when it is odd (rings alternate between a pixel
# centred on phi = 0 and one starting there);
# the southern cap mirrors the northern one.  None of this goes through the
# oracle's closed-form index inversion.
# ---------------------------------------------------------------------------
def ring_centres(nside):
    """Unit vectors of every pixel centre in RING order and their ring numbers."""
    z, phi, ring = [], [], []
    for i in range(1, 4 * nside):
        if i < nside:
            zi, cnt, off = 1.0 - i * i / (3.0 * nside * nside), 4 * i, 0.5
            step = np.pi / (2 * i)
        elif i <= 3 * nside:
            zi, cnt = 4.0 / 3.0 - 2.0 * i / (3.0 * nside), 4 * nside
            off = 0.5 if (i + nside) % 2 == 0 else 1.0
            step = np.pi / (2 * nside)
        else:
            k = 4 * nside - i
            zi, cnt, off = -(1.0 - k * k / (3.0 * nside * nside)), 4 * k, 0.5
            step = np.pi / (2 * k)
        j = np.arange(1, cnt + 1)
        z.append(np.full(cnt, zi))
        phi.append((j - off) * step)
        ring.append(np.full(cnt, i))
    z, phi, ring = np.concatenate(z), np.concatenate(phi), np.concatenate(ring)
    st = np.sqrt(np.maximum(0.0, 1.0 - z * z))
    return np.stack([st * np.cos(phi), st * np.sin(phi), z], axis=1), ring


def area_closed_form(d, theta, ang):
    """contributing_area (pixel_weights.jl:6-8) as the piecewise function of
    the pixel-centre distance d it is: 1 for d <= theta - ang/2, falling
    linearly to 0 at d = theta + ang/2 (1/2 at d = theta), rising again to 1
    at d = theta + 3 ang/2 (the |.| in the reference), 1 beyond."""
    lo, mid, hi = theta - 0.5 * ang, theta + 0.5 * ang, theta + 1.5 * ang
    return np.where(d <= lo, 1.0,
                    np.where(d <= mid, (mid - d) / ang,
                             np.where(d <= hi, (d - mid) / ang, 1.0)))


