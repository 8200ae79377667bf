"""ctypes binding of libsph2grid.so (the C ABI declared in include/sph2grid.h).

The library is the product: every deposit runs its HIP kernels on the GPU.
There is no CPU fallback; if the library is missing or no GPU is visible the
calls raise :class:`LibraryError`.
"""
from __future__ import annotations

import ctypes as C
import os
import threading

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsph2grid.so")
CSRC = os.path.join(os.path.dirname(_HERE), "csrc")

S2G_OK = 0
S2G_EINVAL = -1
S2G_EHIP = -2
S2G_ENOMEM = -3
S2G_EUNSUPPORTED = -4
S2G_ECAPACITY = -5

S2G_F32 = 0
S2G_F64 = 1

S2G_FLAG_CALC_MEAN = 0x1
S2G_FLAG_EXACT_NORM = 0x2
S2G_FLAG_PERIODIC = 0x4
S2G_FLAG_REDUCE = 0x8
S2G_FLAG_FP32_ACCUM = 0x10

# every symbol include/sph2grid.h declares
EXPORTS = (
    "s2g_create", "s2g_destroy", "s2g_last_error", "s2g_version",
    "s2g_deposit", "s2g_deposit_device", "s2g_reduce", "s2g_reduce_device",
    "s2g_map", "s2g_healpix_deposit", "s2g_healpix_deposit_device", "s2g_tsc3d",
    "s2g_tsc3d_device", "s2g_deposit_batch", "s2g_accumulate_device",
    "s2g_healpix_reduce_device", "s2g_mass_density", "s2g_mass_density_device",
    "s2g_footprints", "s2g_tsc3d_device_noalloc", "s2g_healpix_deposit_maps",
    "s2g_healpix_deposit_maps_device", "s2g_set_tuning", "s2g_get_tuning",
    "s2g_plan_create", "s2g_plan_run", "s2g_plan_destroy",
)


class LibraryError(RuntimeError):
    """Raised when libsph2grid is missing or a call returns an error code."""

    def __init__(self, msg, code=None):
        super().__init__(msg)
        self.code = code


class Params(C.Structure):
    _fields_ = [("npixels", C.c_int64), ("len2pix", C.c_double)]


class Stats(C.Structure):
    _fields_ = [("ms_device", C.c_double), ("ms_h2d", C.c_double), ("ms_d2h", C.c_double),
                ("n_particles", C.c_int64), ("n_skipped", C.c_int64),
                ("n_small", C.c_int64), ("n_large", C.c_int64),
                ("n_large_exact", C.c_int64), ("n_pairs", C.c_int64),
                ("sum_footprint", C.c_double), ("ms_setup", C.c_double),
                ("ms_norm", C.c_double), ("ms_sort", C.c_double), ("ms_gather", C.c_double),
                ("ms_finalize", C.c_double), ("ms_small", C.c_double),
                ("n_gather_launches", C.c_int64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class Tuning(C.Structure):
    """s2g_tuning: the 2D deposit's work split (no effect on the arithmetic)."""
    _fields_ = [("small_side", C.c_int32), ("small_item_min", C.c_int32),
                ("small_items_per_cu", C.c_int32), ("gather_item_min", C.c_int32),
                ("gather_items_per_cu", C.c_int32), ("gather_planes", C.c_int32),
                ("small_planes", C.c_int32), ("gather_item_max", C.c_int32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


# the library's defaults (kDefaultTuning in csrc/s2g_api.hip)
DEFAULT_TUNING = dict(small_side=34, small_item_min=256, small_items_per_cu=16,
                      gather_item_min=64, gather_items_per_cu=16, gather_planes=4,
                      small_planes=4, gather_item_max=4096)


class MapDesc(C.Structure):
    _fields_ = [("dims", C.c_int32), ("nq", C.c_int32), ("n", C.c_int64),
                ("params", Params), ("kernel", C.c_int32), ("flags", C.c_uint32),
                ("d_pos", C.c_void_p), ("d_hsml", C.c_void_p), ("d_m", C.c_void_p),
                ("d_rho", C.c_void_p), ("d_bin_q", C.c_void_p), ("d_weights", C.c_void_p),
                ("d_out", C.c_void_p)]


_lib = None
_lock = threading.RLock()  # default_context() -> Context() -> load() re-enters


def build_library(jobs: int = 8) -> str:
    """Compile the HIP sources for gfx950 (hipcc, via csrc/Makefile)."""
    import subprocess
    subprocess.run(["make", "-s", "-C", CSRC, f"-j{jobs}", "-B"], check=True)
    return LIB_PATH


def load(path: str | None = None):
    """Load libsph2grid.so and declare the prototypes (no GPU call is made)."""
    global _lib
    with _lock:
        if _lib is not None and path is None:
            return _lib
        p = path or LIB_PATH
        if not os.path.exists(p):
            raise LibraryError(f"libsph2grid is not built ({p} missing); "
                               "run __graft_entry__.build() or make -C csrc")
        # One HIP runtime per process: torch's wheel ships its own
        # libamdhip64 (soname libamdhip64.so.7, file name libamdhip64.so).
        # Loaded first, it also satisfies this library's dependency; loaded
        # after the system copy, it is a second runtime and finds no GPU.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = C.CDLL(p)
        vp = C.c_void_p
        i64 = C.c_int64
        L.s2g_create.restype = C.c_int
        L.s2g_create.argtypes = [C.POINTER(vp), C.c_int]
        L.s2g_destroy.restype = None
        L.s2g_destroy.argtypes = [vp]
        L.s2g_last_error.restype = C.c_char_p
        L.s2g_last_error.argtypes = [vp]
        L.s2g_version.restype = C.c_char_p
        L.s2g_version.argtypes = []
        dep = [vp, C.c_int, C.c_int, vp, vp, vp, vp, vp, C.c_int, vp, i64,
               C.POINTER(Params), C.c_int, C.c_int, C.c_uint32, vp]
        L.s2g_deposit.restype = C.c_int
        L.s2g_deposit.argtypes = dep + [C.POINTER(Stats)]
        L.s2g_deposit_device.restype = C.c_int
        L.s2g_deposit_device.argtypes = dep + [vp, C.POINTER(Stats)]
        L.s2g_map.restype = C.c_int
        L.s2g_map.argtypes = dep + [C.POINTER(Stats)]
        L.s2g_reduce.restype = C.c_int
        L.s2g_reduce.argtypes = [vp, C.c_int, i64, C.c_int, C.c_uint32, vp, vp]
        L.s2g_reduce_device.restype = C.c_int
        L.s2g_reduce_device.argtypes = [vp, C.c_int, i64, C.c_int, C.c_uint32, vp, vp, vp]
        L.s2g_healpix_deposit.restype = C.c_int
        L.s2g_healpix_deposit.argtypes = [vp, C.c_int, vp, vp, vp, vp, vp, vp, i64, i64,
                                          C.c_int, C.c_int, C.c_uint32, vp, vp,
                                          C.POINTER(Stats)]
        L.s2g_healpix_deposit_device.restype = C.c_int
        L.s2g_healpix_deposit_device.argtypes = [vp, C.c_int, vp, vp, vp, vp, vp, vp, i64, i64,
                                                 C.c_int, C.c_int, C.c_uint32, vp, vp, vp,
                                                 C.POINTER(Stats)]
        L.s2g_tsc3d.restype = C.c_int
        L.s2g_tsc3d.argtypes = [vp, C.c_int, vp, vp, i64, C.POINTER(Params), C.c_uint32, vp,
                                C.POINTER(Stats)]
        L.s2g_tsc3d_device.restype = C.c_int
        L.s2g_tsc3d_device.argtypes = [vp, C.c_int, vp, vp, i64, C.POINTER(Params), C.c_uint32,
                                       vp, vp]
        L.s2g_deposit_batch.restype = C.c_int
        L.s2g_deposit_batch.argtypes = [vp, C.c_int, C.POINTER(MapDesc), C.c_int, vp,
                                        C.POINTER(Stats)]
        L.s2g_accumulate_device.restype = C.c_int
        L.s2g_accumulate_device.argtypes = [vp, vp, vp, i64, vp]
        L.s2g_healpix_reduce_device.restype = C.c_int
        L.s2g_healpix_reduce_device.argtypes = [vp, vp, vp, i64, vp]
        L.s2g_mass_density.restype = C.c_int
        L.s2g_mass_density.argtypes = [vp, C.c_int, vp, vp, i64, C.c_int, C.c_int, i64, vp, vp,
                                       vp, C.POINTER(Stats)]
        L.s2g_mass_density_device.restype = C.c_int
        L.s2g_mass_density_device.argtypes = [vp, C.c_int, vp, vp, i64, C.c_int, C.c_int, i64,
                                              vp, vp, vp, vp, C.POINTER(Stats)]
        pp = C.POINTER(vp)
        L.s2g_healpix_deposit_maps.restype = C.c_int
        L.s2g_healpix_deposit_maps.argtypes = [vp, C.c_int, vp, vp, vp, vp, C.c_int, pp, pp, i64,
                                               i64, C.c_int, C.c_int, C.c_uint32, pp, pp,
                                               C.POINTER(Stats)]
        L.s2g_healpix_deposit_maps_device.restype = C.c_int
        L.s2g_healpix_deposit_maps_device.argtypes = [vp, C.c_int, vp, vp, vp, vp, C.c_int, pp,
                                                      pp, i64, i64, C.c_int, C.c_int, C.c_uint32,
                                                      pp, pp, vp, C.POINTER(Stats)]
        L.s2g_tsc3d_device_noalloc.restype = C.c_int
        L.s2g_tsc3d_device_noalloc.argtypes = [vp, C.c_int, vp, vp, i64, C.POINTER(Params),
                                               C.c_uint32, vp, vp]
        L.s2g_footprints.restype = C.c_int
        L.s2g_footprints.argtypes = [vp, C.c_int, C.c_int, vp, vp, vp, vp, vp, C.c_int, vp, i64,
                                     C.POINTER(Params), C.c_int, C.c_uint32, vp, vp, vp, vp, vp,
                                     i64, vp, vp, vp]
        L.s2g_plan_create.restype = C.c_int
        L.s2g_plan_create.argtypes = [vp, C.c_int, C.POINTER(MapDesc), C.c_int, C.c_double,
                                      C.POINTER(vp)]
        L.s2g_plan_run.restype = C.c_int
        L.s2g_plan_run.argtypes = [vp, vp, C.POINTER(Stats)]
        L.s2g_plan_destroy.restype = None
        L.s2g_plan_destroy.argtypes = [vp]
        L.s2g_set_tuning.restype = C.c_int
        L.s2g_set_tuning.argtypes = [vp, C.POINTER(Tuning)]
        L.s2g_get_tuning.restype = C.c_int
        L.s2g_get_tuning.argtypes = [vp, C.POINTER(Tuning)]
        if path is None:
            _lib = L
        return L


class Context:
    """Owns one ``s2g_ctx`` (device buffers + stream) on one GPU."""

    def __init__(self, device: int = 0):
        self.lib = load()
        h = C.c_void_p()
        rc = self.lib.s2g_create(C.byref(h), int(device))
        if rc != S2G_OK:
            raise LibraryError(f"s2g_create(device={device}) failed with code {rc} "
                               "(no usable GPU?)", rc)
        self.handle = h
        self.device = device

    def check(self, rc: int, what: str):
        if rc != S2G_OK:
            msg = self.lib.s2g_last_error(self.handle)
            msg = msg.decode() if msg else ""
            raise LibraryError(f"{what} failed ({rc}): {msg}", rc)

    def get_tuning(self) -> dict:
        t = Tuning()
        self.check(self.lib.s2g_get_tuning(self.handle, C.byref(t)), "s2g_get_tuning")
        return t.as_dict()

    def set_tuning(self, tuning: dict | None = None):
        """Set the 2D work split (a dict of s2g_tuning fields; missing fields
        keep the library defaults); None restores the defaults."""
        if tuning is None:
            self.check(self.lib.s2g_set_tuning(self.handle, None), "s2g_set_tuning")
            return
        unknown = set(tuning) - set(DEFAULT_TUNING)
        if unknown:
            raise ValueError(f"unknown tuning fields {sorted(unknown)}")
        t = Tuning(**{**DEFAULT_TUNING, **tuning})
        self.check(self.lib.s2g_set_tuning(self.handle, C.byref(t)), "s2g_set_tuning")

    def close(self):
        if getattr(self, "handle", None):
            self.lib.s2g_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_default_ctx = None


def default_context() -> Context:
    global _default_ctx
    with _lock:
        if _default_ctx is None:
            _default_ctx = Context(0)
        return _default_ctx


def version() -> str:
    return load().s2g_version().decode()
