"""Synthetic particle sets of the five benchmark workloads (seeded, numpy PCG64).

Each generator returns a dict with the ``sphMapping`` / ``healpix_map``
arguments: ``Pos`` (3 x N), ``HSML``, ``M``, ``Rho``, ``Bin_Quant``, ``Weights``
plus the mapping parameters / options of that workload.  ``n`` overrides the
particle count (parity tests use scaled-down versions of the same shapes).

Choices the workload descriptions leave open are fixed here:

* plain (65,536 particles, 256^2, Cubic, Float64): as described.
* sedov: the 64-neighbour relation h = (3*64*m / (4 pi rho))^(1/3) with
  m ~ rho * V_cell gives a constant h = (48 V_cell / pi)^(1/3) (about 25 pixels
  at 1024^2, not "0.2-5 pixels"); the literal relation is implemented.  One
  map with Bin_Quant = [rho; T] (Nq = 2), Weights = Rho, calc_mean, reduce.
* cluster: Rho = 64 m / (4/3 pi h^3); Bin_Quant = T, Weights = Rho, reduce.
* batched: 64 sedov-like maps, N log-uniform [1e4, 4e6], side from
  {128, 256, 512, 1024, 2048}, Nq = 3 (rho, T, |B|), Weights = Rho.
* allsky: shell 0.05-1 around the origin, p(r) ~ r (density ~ 1/r),
  Rho = 64 m / (4/3 pi h^3), T log-normal; the TSC grid covers [-1, 1]^3.
"""
from __future__ import annotations

import numpy as np

SEED_BASE = 0x5348544F47520000


def _rng(config_id: int, map_id: int = 0):
    return np.random.default_rng(SEED_BASE + config_id + (map_id << 8))


def plain(n: int = 65536, npix: int = 256, seed_offset: int = 0):
    r = _rng(1 + seed_offset)
    pos = r.random((3, n))
    m = np.full(n, 1.0 / n)
    h = 2.0 * n ** (-1.0 / 3.0) * np.exp(0.3 * r.standard_normal(n))
    rho = 64.0 * m / (4.0 / 3.0 * np.pi * h ** 3)
    return dict(Pos=pos, HSML=h, M=m, Rho=rho, Bin_Quant=rho.copy(), Weights=rho.copy(),
                lims=([0.0, 1.0], [0.0, 1.0], [0.0, 1.0]), Npixels=npix, kernel="cubic",
                reduce_image=True, calc_mean=False)


def _lattice(n: int, r, jitter=0.01):
    side = int(np.ceil(n ** (1.0 / 3.0) - 1e-9))
    g = (np.arange(side) + 0.5) / side
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    pts = np.stack([X.ravel(), Y.ravel(), Z.ravel()])
    extra = pts.shape[1] - n
    if extra > 0:
        drop = r.choice(pts.shape[1], size=extra, replace=False)
        keep = np.ones(pts.shape[1], dtype=bool)
        keep[drop] = False
        pts = pts[:, keep]
    pts = pts + jitter / side * r.uniform(-1.0, 1.0, size=pts.shape)
    return pts, 1.0 / side ** 3


def sedov(n: int = 1 << 20, npix: int = 1024, seed_offset: int = 0, nq: int = 2,
          dtype=np.float32, map_id: int = 0, config_id: int = 2):
    r = _rng(config_id + seed_offset, map_id)
    pos, vcell = _lattice(n, r)
    rad = np.sqrt(((pos - 0.5) ** 2).sum(axis=0))
    rho = 1.0 + 4.0 * np.exp(-(rad - 0.3) ** 2 / 0.002)
    m = rho * vcell
    h = (3.0 * 64.0 * m / (4.0 * np.pi * rho)) ** (1.0 / 3.0)
    T = 1e4 * (1.0 + 100.0 * np.exp(-rad ** 2 / 0.01))
    q = [rho, T]
    if nq >= 3:
        q.append(np.exp(np.log(1e-6) + 0.8 * r.standard_normal(n)))
    bq = np.stack(q[:nq]) if nq > 1 else q[0]
    c = lambda a: np.asarray(a, dtype=dtype)
    return dict(Pos=c(pos), HSML=c(h), M=c(m), Rho=c(rho), Bin_Quant=c(bq), Weights=c(rho),
                lims=([0.0, 1.0], [0.0, 1.0], [0.0, 1.0]), Npixels=npix,
                kernel="wendlandc6", reduce_image=True, calc_mean=True)


def cluster(n: int = 1 << 24, npix: int = 4096, seed_offset: int = 0, dtype=np.float32):
    r = _rng(3 + seed_offset)
    n_halo = int(round(0.8 * n))
    n_bg = n - n_halo
    rad = r.uniform(1e-3, 0.5, n_halo)  # rho ~ r^-2  =>  dN/dr = const
    v = r.standard_normal((3, n_halo))
    v /= np.sqrt((v ** 2).sum(axis=0))
    halo = 0.5 + v * rad
    bg = r.random((3, n_bg))
    pos = np.concatenate([halo, bg], axis=1)
    h = np.exp(r.uniform(np.log(1e-4), np.log(1e-1), n))
    m = np.full(n, 1.0 / n)
    rho = 64.0 * m / (4.0 / 3.0 * np.pi * h ** 3)
    T = np.exp(np.log(1e7) + 1.0 * r.standard_normal(n))
    c = lambda a: np.asarray(a, dtype=dtype)
    return dict(Pos=c(pos), HSML=c(h), M=c(m), Rho=c(rho), Bin_Quant=c(T), Weights=c(rho),
                lims=([0.0, 1.0], [0.0, 1.0], [0.0, 1.0]), Npixels=npix,
                kernel="wendlandc6", reduce_image=True, calc_mean=False)


def batched(nmaps: int = 64, scale: float = 1.0, seed_offset: int = 0, dtype=np.float32):
    r = _rng(4 + seed_offset)
    sides = [128, 256, 512, 1024, 2048]
    out = []
    for k in range(nmaps):
        n = int(np.exp(r.uniform(np.log(1e4), np.log(4e6))) * scale)
        n = max(n, 64)
        npix = int(sides[r.integers(0, len(sides))])
        d = sedov(n=n, npix=npix, nq=3, dtype=dtype, map_id=k + 1, config_id=4 + seed_offset)
        out.append(d)
    return out


def allsky(n: int = 1 << 22, nside: int = 1024, seed_offset: int = 0):
    r = _rng(5 + seed_offset)
    rmin, rmax = 0.05, 1.0
    rad = np.sqrt(r.random(n) * (rmax ** 2 - rmin ** 2) + rmin ** 2)  # p(r) ~ r
    v = r.standard_normal((3, n))
    v /= np.sqrt((v ** 2).sum(axis=0))
    pos = v * rad
    h = np.exp(r.uniform(np.log(1e-3), np.log(5e-2), n))
    m = np.full(n, 1.0 / n)
    rho = 64.0 * m / (4.0 / 3.0 * np.pi * h ** 3)
    T = np.exp(np.log(1e7) + 1.0 * r.standard_normal(n))
    return dict(Pos=pos, HSML=h, M=m, Rho=rho, T=T, Nside=nside, kernel="wendlandc4",
                tsc_npix=512, tsc_lims=([-1.0, 1.0], [-1.0, 1.0], [-1.0, 1.0]))
