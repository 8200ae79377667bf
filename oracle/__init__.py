"""FP64 CPU oracle for SPHtoGrid.jl's deposition path (ctypes binding).

TEST INFRASTRUCTURE ONLY: the checker for the HIP product path and the timed
CPU baseline.  Only tests/, ``__graft_entry__.smoke()`` and bench.py's
``cpu_baseline`` leg import this module.  The product package (sph2grid)
never imports it.

The C library (``sph2grid_oracle.c``) restates, function by function:

* ``shared/parameters.jl:44-125``           -> :func:`mapping_parameters`
* ``cic_interpolation/filter_shift.jl``     -> :func:`center_particles`,
  :func:`filter_particles`
* ``cic_interpolation/cic_2D.jl:11-244``    -> :func:`cic_mapping_2D`
* ``cic_interpolation/cic_3D.jl:13-209``    -> :func:`cic_mapping_3D`
* ``cic_interpolation/reduce_image.jl``     -> :func:`reduce_image_2D`,
  :func:`reduce_image_3D`
* ``parallel/domain_decomp.jl:7-17`` + the ``parallel=true`` fan-out of
  ``cic_interpolation.jl:171-215`` -> ``*_parallel``
* ``healpix_interpolation/*.jl``            -> :func:`healpix_map`
* SPHKernels.jl shapes (external, not vendored) -> :func:`kernel_value`
* Healpix.jl RING geometry (external)       -> ``hp_*``
* TSC (reference fully commented out)       -> :func:`tsc_3d` (textbook)

Pinning (what checks each function against something other than itself) is
listed in tests/test_oracle_pins.py.  The pieces that only an import of the
Julia reference could pin -- the exact SPHKernels coefficient spelling
(WendlandC8 in particular), Healpix.jl's disc-edge membership, and the
reference's ``@fastmath`` ulp-level rounding -- are "parity unpinned" at the
ulp level; every other function is pinned by a known-answer test, a closed
form or brute force.
"""
from __future__ import annotations

import ctypes as C
import math
import os
import subprocess
import threading

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libsph2grid_oracle.so")
_lib = None
_lock = threading.Lock()

KERNELS = {"cubic": 0, "quintic": 1, "wendlandc2": 2, "wendlandc4": 3,
           "wendlandc6": 4, "wendlandc8": 5}

_d = C.POINTER(C.c_double)
_f = C.POINTER(C.c_float)
_i64 = C.c_int64


class Params(C.Structure):
    """Mirror of ``orc_params`` (shared/parameters.jl:14-27)."""
    _fields_ = [("x_lim", C.c_double * 2), ("y_lim", C.c_double * 2),
                ("z_lim", C.c_double * 2), ("center", C.c_double * 3),
                ("halfsize", C.c_double * 3), ("len2pix", C.c_double),
                ("pixelSideLength", C.c_double), ("Npixels", C.c_int64 * 3),
                ("boxsize", C.c_double), ("periodic", C.c_int32),
                ("_pad", C.c_int32)]


def build() -> str:
    """Compile the oracle with its Makefile (gcc, no fast-math)."""
    subprocess.run(["make", "-s", "-C", _HERE, "libsph2grid_oracle.so"],
                   check=True)
    return _LIB_PATH


def lib():
    global _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(_LIB_PATH):
                build()
            L = C.CDLL(_LIB_PATH)
            L.orc_kernel_shape.restype = C.c_double
            L.orc_kernel_shape.argtypes = [C.c_int, C.c_double]
            L.orc_kernel_norm.restype = C.c_double
            L.orc_kernel_norm.argtypes = [C.c_int, C.c_int]
            L.orc_kernel_value.restype = C.c_double
            L.orc_kernel_value.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double]
            L.orc_bias_correction.restype = C.c_double
            L.orc_bias_correction.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double,
                                              C.c_double, _i64]
            L.orc_mass_density.restype = C.c_int
            L.orc_mass_density.argtypes = [_d, _d, _i64, C.c_int, C.c_int, _i64, _d, _d, _d]
            L.orc_mass_density_subset.restype = C.c_int
            L.orc_mass_density_subset.argtypes = [_d, _d, _i64, C.c_void_p, _i64, C.c_int,
                                                  C.c_int, _i64, _d, _d, _d]
            L.orc_mapping_parameters.restype = C.c_int
            L.orc_mapping_parameters.argtypes = [
                C.POINTER(Params), _d, _d, _d, _d, C.c_double, C.c_double,
                C.c_double, C.c_double, _i64, C.c_double]
            L.orc_center_particles.restype = None
            L.orc_center_particles.argtypes = [C.c_void_p, C.c_int, _i64,
                                               C.POINTER(Params), C.POINTER(Params)]
            L.orc_filter_particles.restype = None
            L.orc_filter_particles.argtypes = [C.c_void_p, C.c_int, _i64,
                                               C.POINTER(Params), C.c_void_p]
            L.orc_cic_mapping_2D.restype = C.c_int
            L.orc_cic_mapping_2D.argtypes = [_d, _d, _d, _d, _d, C.c_int, _d, _i64,
                                             _i64, C.c_double, C.c_int, C.c_int,
                                             C.c_int, _d]
            L.orc_cic_mapping_3D.restype = C.c_int
            L.orc_cic_mapping_3D.argtypes = [_d, _d, _d, _d, _d, _d, _i64, _i64,
                                             C.c_double, C.c_int, C.c_int, C.c_int,
                                             _d, _d, _d]
            L.orc_cic_mapping_2D_parallel.restype = C.c_int
            L.orc_cic_mapping_2D_parallel.argtypes = [C.c_int] + \
                L.orc_cic_mapping_2D.argtypes
            L.orc_cic_mapping_3D_parallel.restype = C.c_int
            L.orc_cic_mapping_3D_parallel.argtypes = [
                C.c_int, _d, _d, _d, _d, _d, _d, _i64, _i64, C.c_double, C.c_int,
                C.c_int, C.c_int, _d]
            L.orc_domain_decomposition.restype = None
            L.orc_domain_decomposition.argtypes = [_i64, C.c_int, C.c_int,
                                                   C.POINTER(_i64), C.POINTER(_i64)]
            L.orc_reduce_image_2D.restype = None
            L.orc_reduce_image_2D.argtypes = [_d, _i64, C.c_int, C.c_int, _d]
            L.orc_reduce_image_3D.restype = None
            L.orc_reduce_image_3D.argtypes = [_d, _i64, _d]
            L.orc_hp_npix.restype = _i64
            L.orc_hp_npix.argtypes = [_i64]
            L.orc_hp_pix2ang_ring.restype = None
            L.orc_hp_pix2ang_ring.argtypes = [_i64, _i64, _d, _d]
            L.orc_hp_ang2pix_ring.restype = _i64
            L.orc_hp_ang2pix_ring.argtypes = [_i64, C.c_double, C.c_double]
            L.orc_hp_pix2vec_ring.restype = None
            L.orc_hp_pix2vec_ring.argtypes = [_i64, _i64, _d]
            L.orc_hp_vec2ang.restype = None
            L.orc_hp_vec2ang.argtypes = [_d, _d, _d]
            L.orc_hp_ring_above.restype = _i64
            L.orc_hp_ring_above.argtypes = [_i64, C.c_double]
            L.orc_hp_query_disc_ring.restype = _i64
            L.orc_hp_query_disc_ring.argtypes = [_i64, C.c_double, C.c_double,
                                                 C.c_double, C.POINTER(_i64), _i64]
            L.orc_footprints.restype = C.c_int
            L.orc_footprints.argtypes = [C.c_int, _d, _d, _d, C.c_int, _i64, _i64, C.c_double,
                                         C.c_int, C.c_uint32, C.c_void_p, C.c_void_p,
                                         C.c_void_p]
            L.orc_hp_particle_weights.restype = _i64
            L.orc_hp_particle_weights.argtypes = [_d, C.c_double, _i64, C.c_int, C.c_int, _i64,
                                                  C.POINTER(_i64), _d, _d, _d]
            L.orc_healpix_map.restype = C.c_int
            L.orc_healpix_map.argtypes = [_d, _d, _d, _d, _d, _d, _i64, _d, _d, _i64,
                                          C.c_int, C.c_int, C.c_int, _d, _d]
            L.orc_healpix_deposit_parallel.restype = C.c_int
            L.orc_healpix_deposit_parallel.argtypes = [
                C.c_int, _d, _d, _d, _d, _d, _d, _i64, _i64, C.c_int, C.c_int,
                C.c_int, _d, _d]
            L.orc_tsc_3d.restype = C.c_int
            L.orc_tsc_3d.argtypes = [_d, _d, _i64, _i64, C.c_double, C.c_int, _d]
            _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(_d)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _kid(kernel):
    return KERNELS[kernel] if isinstance(kernel, str) else int(kernel)


# ---------------------------------------------------------------- kernels
def kernel_shape(kernel, u: float) -> float:
    return lib().orc_kernel_shape(_kid(kernel), float(u))


def kernel_norm(kernel, dim: int) -> float:
    return lib().orc_kernel_norm(_kid(kernel), int(dim))


def kernel_value(kernel, dim: int, u: float, h_inv: float) -> float:
    return lib().orc_kernel_value(_kid(kernel), int(dim), float(u), float(h_inv))


# ---------------------------------------------------------------- parameters
class OracleParameterError(Exception):
    pass


_PARAM_ERRORS = {
    1: "Giving a center position requires extent in x, y and z direction.",
    2: "Please specify pixelSideLength or number of pixels!",
}


def mapping_parameters(x_lim=(-1.0, -1.0), y_lim=(-1.0, -1.0), z_lim=(-1.0, -1.0),
                       center=(-1.0, -1.0, -1.0), x_size=-1.0, y_size=-1.0,
                       z_size=-1.0, pixelSideLength=-1.0, Npixels=0,
                       boxsize=-1.0) -> Params:
    out = Params()
    arrs = [np.asarray(v, dtype=np.float64).copy() for v in (x_lim, y_lim, z_lim, center)]
    rc = lib().orc_mapping_parameters(C.byref(out), *[_p(a) for a in arrs],
                                      float(x_size), float(y_size), float(z_size),
                                      float(pixelSideLength), int(Npixels),
                                      float(boxsize))
    if rc:
        raise OracleParameterError(_PARAM_ERRORS[rc])
    return out


def center_particles(pos: np.ndarray, par: Params) -> Params:
    """In place on ``pos`` (shape (n,3) C-order == Julia 3 x n column-major)."""
    assert pos.flags.c_contiguous and pos.dtype in (np.float32, np.float64)
    out = Params()
    lib().orc_center_particles(pos.ctypes.data, int(pos.dtype == np.float32),
                               pos.shape[0], C.byref(par), C.byref(out))
    return out


def filter_particles(pos: np.ndarray, par: Params) -> np.ndarray:
    assert pos.flags.c_contiguous and pos.dtype in (np.float32, np.float64)
    keep = np.zeros(pos.shape[0], dtype=np.uint8)
    lib().orc_filter_particles(pos.ctypes.data, int(pos.dtype == np.float32),
                               pos.shape[0], C.byref(par), keep.ctypes.data)
    return keep.astype(bool)


# ---------------------------------------------------------------- deposits
def _bin_q_layout(bin_q, n):
    b = np.asarray(bin_q)
    if b.ndim == 1:
        return _f64(b), 1
    # Julia Nq x N column-major == numpy (N, Nq) C-order
    assert b.shape[0] == n
    return _f64(b), b.shape[1]


def cic_mapping_2D(pos, hsml, m, rho, bin_q, weights, npix, len2pix, kernel,
                   kernel_dim=2, calc_mean=True, nworkers=0) -> np.ndarray:
    """Returns the flat unreduced image, shape (npix^2, nq+1), Fortran order."""
    pos = _f64(pos)
    n = pos.shape[0]
    hsml, m, rho, weights = (_f64(a) for a in (hsml, m, rho, weights))
    bq, nq = _bin_q_layout(bin_q, n)
    img = np.zeros(npix * npix * (nq + 1), dtype=np.float64)
    args = (_p(pos), _p(hsml), _p(m), _p(rho), _p(bq), nq, _p(weights), n, npix,
            float(len2pix), _kid(kernel), int(kernel_dim), int(bool(calc_mean)),
            _p(img))
    rc = (lib().orc_cic_mapping_2D_parallel(int(nworkers), *args) if nworkers
          else lib().orc_cic_mapping_2D(*args))
    if rc:
        raise MemoryError("oracle 2D mapping failed")
    return img.reshape((npix * npix, nq + 1), order="F")


def cic_mapping_3D(pos, hsml, m, rho, bin_q, weights, npix, len2pix, kernel,
                   kernel_dim=3, calc_mean=False, nworkers=0, diagnostics=False):
    pos = _f64(pos)
    n = pos.shape[0]
    hsml, m, rho, weights, bq = (_f64(a) for a in (hsml, m, rho, weights, bin_q))
    img = np.zeros(npix ** 3 * 2, dtype=np.float64)
    gm = C.c_double(0.0)
    pm = C.c_double(0.0)
    if nworkers:
        rc = lib().orc_cic_mapping_3D_parallel(
            int(nworkers), _p(pos), _p(hsml), _p(m), _p(rho), _p(bq), _p(weights), n,
            npix, float(len2pix), _kid(kernel), int(kernel_dim), int(bool(calc_mean)),
            _p(img))
    else:
        rc = lib().orc_cic_mapping_3D(
            _p(pos), _p(hsml), _p(m), _p(rho), _p(bq), _p(weights), n, npix,
            float(len2pix), _kid(kernel), int(kernel_dim), int(bool(calc_mean)),
            _p(img), C.byref(gm), C.byref(pm))
    if rc:
        raise MemoryError("oracle 3D mapping failed")
    img = img.reshape((npix ** 3, 2), order="F")
    if diagnostics:
        return img, gm.value, pm.value
    return img


def reduce_image_2D(image, npix, reduce_image=True) -> np.ndarray:
    """Julia im_plot (npix, npix, nq) [x, y, q], returned Fortran-ordered."""
    img = np.asfortranarray(image, dtype=np.float64)
    nq = img.shape[1] - 1
    out = np.zeros(npix * npix * nq, dtype=np.float64)
    lib().orc_reduce_image_2D(img.ctypes.data_as(_d), npix, nq, int(bool(reduce_image)),
                              _p(out))
    return out.reshape((npix, npix, nq), order="F")


def reduce_image_3D(image, npix) -> np.ndarray:
    img = np.asfortranarray(image, dtype=np.float64)
    out = np.zeros(npix ** 3, dtype=np.float64)
    lib().orc_reduce_image_3D(img.ctypes.data_as(_d), npix, _p(out))
    return out.reshape((npix, npix, npix), order="F")


def domain_decomposition(n, nworkers):
    out = []
    for i in range(nworkers):
        s = _i64(0)
        ln = _i64(0)
        lib().orc_domain_decomposition(n, nworkers, i, C.byref(s), C.byref(ln))
        out.append((s.value, ln.value))
    return out


# ---------------------------------------------------------------- full sphMapping
def sph_mapping(Pos, HSML, M, Rho, Bin_Quant, Weights, par: Params, kernel,
                kernel_dim=2, reduce_image=True, return_both_maps=False,
                dimensions=2, calc_mean=False, nworkers=0):
    """cic_interpolation.jl:35-273 (no RM / stokes / Unitful).  Mutates Pos."""
    Pos_, par2 = Pos, center_particles(Pos, par)
    keep = filter_particles(Pos_, par2)
    x = Pos_[keep].astype(np.float64)
    f = lambda a: np.asarray(a)[keep].astype(np.float64)
    hsml, m, rho, w = f(HSML), f(M), f(Rho), f(Weights)
    bq = np.asarray(Bin_Quant)
    bq = bq[keep].astype(np.float64)
    npix = int(par.Npixels[0])
    if dimensions == 2:
        img = cic_mapping_2D(x, hsml, m, rho, bq, w, npix, par2.len2pix, kernel,
                             kernel_dim, calc_mean, nworkers)
        if return_both_maps:
            return img
        return reduce_image_2D(img, npix, reduce_image)
    img = cic_mapping_3D(x, hsml, m, rho, bq, w, npix, par2.len2pix, kernel,
                         kernel_dim, False, nworkers)
    if not reduce_image:
        img[:, 1] = 1.0
    return reduce_image_3D(img, npix)


# ---------------------------------------------------------------- integer work
CLS_SKIP, CLS_SMALL, CLS_LARGE_CLOSED, CLS_LARGE_EXPLICIT = 0, 1, 2, 3


def footprints(dims, pos, hsml, bin_q, npix, len2pix, kernel, calc_mean=True,
               exact_norm=False):
    """Per particle: footprint box (n x 2*dims int32), class and fallback flag
    (orc_footprints).  pos is (n, 3) centred, bin_q n or (n, nq)."""
    x = _f64(pos)
    n = x.shape[0]
    bq = _f64(bin_q)
    nq = 1 if bq.ndim == 1 else bq.shape[1]
    box = np.zeros((n, 2 * dims), dtype=np.int32)
    cls = np.zeros(n, dtype=np.int32)
    fb = np.zeros(n, dtype=np.int32)
    flags = (1 if calc_mean else 0) | (2 if exact_norm else 0)
    lib().orc_footprints(int(dims), _p(x), _p(_f64(hsml)), _p(bq), nq, n, int(npix),
                         float(len2pix), _kid(kernel), flags, box.ctypes.data_as(C.c_void_p),
                         cls.ctypes.data_as(C.c_void_p), fb.ctypes.data_as(C.c_void_p))
    return box, cls, fb


def large_stride(L):
    """Stride of the order in which the tile lists take the L large
    particles: the integer nearest to L / phi (golden ratio, L * 0.618... in
    double precision, rounded half up), raised to the first value coprime to
    L; 1 for L <= 2."""
    if L <= 2:
        return 1
    r = int(L * 0.6180339887498949 + 0.5)
    while math.gcd(r, L) != 1:
        r += 1
    return r


def tile_lists(dims, box, cls, npix):
    """The deposit's (tile, particle) lists, restated from their definition:
    tiles of 64^2 pixels (2D) or 16^3 voxels (3D) in row-major order; the
    large particles (class 2 or 3) e_0 .. e_{L-1} in particle order are taken
    as f_j = e_{(j r) mod L}, r = large_stride(L), and tile t lists, in that
    order, every one whose box meets it.  Returns (tile_off, pairs) as CSR
    arrays."""
    T = 64 if dims == 2 else 16
    nt = (npix + T - 1) // T
    ntiles = nt ** dims
    lists = [[] for _ in range(ntiles)]
    large = np.flatnonzero((cls == CLS_LARGE_CLOSED) | (cls == CLS_LARGE_EXPLICIT))
    L = large.size
    r = large_stride(L)
    for p in large[(np.arange(L, dtype=np.int64) * r) % max(L, 1)]:
        r = [range(box[p, 2 * d] // T, box[p, 2 * d + 1] // T + 1) for d in range(dims)]
        if dims == 2:
            for a in r[0]:
                for b in r[1]:
                    lists[a * nt + b].append(p)
        else:
            for a in r[0]:
                for b in r[1]:
                    for c in r[2]:
                        lists[(a * nt + b) * nt + c].append(p)
    off = np.zeros(ntiles + 1, dtype=np.int64)
    off[1:] = np.cumsum([len(x) for x in lists])
    pairs = np.array([p for x in lists for p in x], dtype=np.int32)
    return off, pairs


def small_buckets(dims, box, cls, npix):
    """The small particles (class 1): in 2D bucketed by the 32 x 32 tile
    holding their box's first pixel (row-major), particle order inside a
    bucket; in 3D one bucket in particle order.  Returns (off, order)."""
    small = np.flatnonzero(cls == CLS_SMALL).astype(np.int32)
    if dims == 3:
        return np.array([0, small.size], dtype=np.int64), small
    S = 32
    ns = (npix + S - 1) // S
    key = (box[small, 0] // S) * ns + box[small, 2] // S
    order = small[np.argsort(key, kind="stable")]
    off = np.searchsorted(np.sort(key), np.arange(ns * ns + 1), side="left").astype(np.int64)
    return off, order


# ---------------------------------------------------------------- HEALPix
def hp_npix(nside):
    return 12 * nside * nside


def hp_pix2ang_ring(nside, pix):
    t = C.c_double()
    p = C.c_double()
    lib().orc_hp_pix2ang_ring(nside, pix, C.byref(t), C.byref(p))
    return t.value, p.value


def hp_ang2pix_ring(nside, theta, phi):
    return lib().orc_hp_ang2pix_ring(nside, float(theta), float(phi))


def hp_pix2vec_ring(nside, pix):
    v = np.zeros(3)
    lib().orc_hp_pix2vec_ring(nside, pix, _p(v))
    return v


def hp_vec2ang(v):
    v = _f64(v)
    t = C.c_double()
    p = C.c_double()
    lib().orc_hp_vec2ang(_p(v), C.byref(t), C.byref(p))
    return t.value, p.value


def hp_ring_above(nside, z):
    return lib().orc_hp_ring_above(nside, float(z))


def hp_query_disc_ring(nside, theta, phi, radius):
    cap = 1024
    while True:
        out = np.zeros(cap, dtype=np.int64)
        cnt = lib().orc_hp_query_disc_ring(nside, float(theta), float(phi), float(radius),
                                           out.ctypes.data_as(C.POINTER(_i64)), cap)
        if cnt <= cap:
            return out[:cnt].copy()
        cap = int(cnt)


def hp_particle_weights(pos, hsml, nside, kernel="wendlandc4", kernel_dim=2):
    """One particle's pixel list, areas A and kernel weights wk
    (pixel_weights.jl:6-140) and its info: dict with Dx, proj_hsml, ang_pix,
    weight_per_pix, n_distr and fallback.  Empty arrays if skipped."""
    p = _f64(pos)
    info = np.zeros(6)
    cap = 4096
    while True:
        pix = np.zeros(cap, dtype=np.int64)
        A = np.zeros(cap)
        wk = np.zeros(cap)
        cnt = lib().orc_hp_particle_weights(_p(p), float(hsml), int(nside), _kid(kernel),
                                            int(kernel_dim), cap,
                                            pix.ctypes.data_as(C.POINTER(_i64)), _p(A), _p(wk),
                                            _p(info))
        if cnt < 0:
            raise MemoryError("oracle hp_particle_weights failed")
        if cnt <= cap:
            break
        cap = int(cnt)
    d = {"Dx": info[0], "proj_hsml": info[1], "ang_pix": info[2], "weight_per_pix": info[3],
         "n_distr": int(info[4]), "fallback": bool(info[5])}
    return pix[:cnt].copy(), A[:cnt].copy(), wk[:cnt].copy(), d


class OracleBoundsError(IndexError):
    pass


def healpix_map(Pos, Hsml, M, Rho, Bin_q, Weights, center=(0.0, 0.0, 0.0),
                radius_limits=(0.0, np.inf), Nside=1024, kernel="wendlandc6",
                kernel_dim=2, calc_mean=True):
    """healpix_interpolation/main.jl:92-227.  Mutates Pos (Float64, (n,3))."""
    assert Pos.dtype == np.float64 and Pos.flags.c_contiguous
    n = Pos.shape[0]
    arrs = [_f64(a) for a in (Hsml, M, Rho, Bin_q, Weights)]
    c = _f64(center)
    rl = _f64(radius_limits)
    npix = hp_npix(Nside)
    mp = np.zeros(npix)
    wm = np.zeros(npix)
    rc = lib().orc_healpix_map(_p(Pos), *[_p(a) for a in arrs], n, _p(c), _p(rl),
                               int(Nside), _kid(kernel), int(kernel_dim),
                               int(bool(calc_mean)), _p(mp), _p(wm))
    if rc == -1:
        raise OracleBoundsError("BoundsError in filter_sort_particles")
    if rc:
        raise MemoryError("oracle healpix failed")
    return mp, wm


def healpix_deposit_parallel(pos, hsml, m, rho, bin_q, weights, nside, kernel,
                             kernel_dim=2, calc_mean=True, nworkers=1):
    """Deposit of already centred particles, threaded (CPU baseline only)."""
    pos = _f64(pos)
    arrs = [_f64(a) for a in (hsml, m, rho, bin_q, weights)]
    npix = hp_npix(nside)
    mp = np.zeros(npix)
    wm = np.zeros(npix)
    rc = lib().orc_healpix_deposit_parallel(int(nworkers), _p(pos), *[_p(a) for a in arrs],
                                            pos.shape[0], int(nside), _kid(kernel),
                                            int(kernel_dim), int(bool(calc_mean)),
                                            _p(mp), _p(wm))
    if rc:
        raise MemoryError("oracle healpix deposit failed")
    return mp, wm


# ---------------------------------------------------------------- TSC
def tsc_3d(pos, bin_q, npix, len2pix, periodic=False):
    pos = _f64(pos)
    bq = _f64(bin_q)
    g = np.zeros(2 * npix ** 3)
    lib().orc_tsc_3d(_p(pos), _p(bq), pos.shape[0], npix, float(len2pix), int(periodic),
                     _p(g))
    return g.reshape((npix ** 3, 2), order="F")


# ---------------------------------------------------------------- mass_density
def bias_correction(kernel, dim, density, m, h_inv, nneighbors):
    """Dehnen & Aly (2012) eqs. 18-19 (SPHKernels.jl bias_correction)."""
    return lib().orc_bias_correction(_kid(kernel), int(dim), float(density), float(m),
                                     float(h_inv), int(nneighbors))


def mass_density(pos, mass, kernel="cubic", dim=3, nneighbors=32, boxsize=(0.0, 0.0, 0.0)):
    """effects/mass_density.jl:30-99 by brute force.  ``pos`` is 3 x N (the
    reference's layout).  Returns (rho, hsml)."""
    x = np.ascontiguousarray(np.asarray(pos, dtype=np.float64).T)
    m = _f64(mass)
    n = m.shape[0]
    box = np.ascontiguousarray(np.asarray(boxsize, dtype=np.float64))
    rho, hsml = np.zeros(n), np.zeros(n)
    rc = lib().orc_mass_density(_p(x), _p(m), n, _kid(kernel), int(dim), int(nneighbors),
                                _p(box), _p(rho), _p(hsml))
    if rc:
        raise ValueError("oracle mass_density: fewer particles than neighbours, or no memory")
    return rho, hsml


def mass_density_subset(pos, mass, idx, kernel="cubic", dim=3, nneighbors=32,
                        boxsize=(0.0, 0.0, 0.0)):
    """mass_density of the particles ``idx`` only, against the whole set."""
    x = np.ascontiguousarray(np.asarray(pos, dtype=np.float64).T)
    m = _f64(mass)
    ii = np.ascontiguousarray(np.asarray(idx, dtype=np.int64))
    box = np.ascontiguousarray(np.asarray(boxsize, dtype=np.float64))
    rho, hsml = np.zeros(ii.size), np.zeros(ii.size)
    rc = lib().orc_mass_density_subset(_p(x), _p(m), m.shape[0], ii.ctypes.data_as(C.c_void_p),
                                       ii.size, _kid(kernel), int(dim), int(nneighbors),
                                       _p(box), _p(rho), _p(hsml))
    if rc:
        raise ValueError("oracle mass_density: fewer particles than neighbours, or no memory")
    return rho, hsml
