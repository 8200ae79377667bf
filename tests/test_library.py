"""The C-ABI library loads and exports every symbol include/sph2grid.h
declares (no compute call is made, so this runs without a GPU)."""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "sph2grid.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(s2g_[a-z0-9_]+)\s*\(", src)))


def test_header_declares_the_boundary():
    names = _declared()
    for n in ("s2g_create", "s2g_destroy", "s2g_last_error", "s2g_deposit",
              "s2g_deposit_device", "s2g_reduce", "s2g_map", "s2g_healpix_deposit",
              "s2g_tsc3d", "s2g_deposit_batch"):
        assert n in names


def test_library_exports_every_declared_symbol(s2g):
    from sph2grid import _lib
    L = _lib.load()
    for name in _declared():
        assert hasattr(L, name), f"{name} declared in include/sph2grid.h but not exported"
    assert set(_declared()) == set(_lib.EXPORTS)


def test_version_string(s2g):
    v = s2g.version()
    assert v.startswith("sph2grid ") and "gfx950" in v


def test_null_context_is_rejected(s2g):
    from sph2grid import _lib
    L = _lib.load()
    assert L.s2g_deposit(None, 2, 0, None, None, None, None, None, 1, None, 0, None, 0, 2, 0,
                         None, None) == _lib.S2G_EINVAL
    assert L.s2g_last_error(None) == b"null context"


def test_no_gpu_create_fails_cleanly(s2g):
    """Without a visible GPU s2g_create returns S2G_EHIP instead of crashing."""
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    from sph2grid import _lib
    L = _lib.load()
    h = C.c_void_p()
    assert L.s2g_create(C.byref(h), 0) == _lib.S2G_EHIP
    with pytest.raises(_lib.LibraryError):
        _lib.Context(0)


def test_product_does_not_import_the_oracle():
    pkg = os.path.join(ROOT, "sphtogrid.jl_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".hip", ".h", ".cpp", ".jl")):
                txt = open(os.path.join(dirpath, f)).read()
                assert "import oracle" not in txt and "sph2grid_oracle" not in txt, f


def test_default_context_does_not_deadlock(s2g):
    """default_context() creates the context while holding the module lock;
    the nested load() must re-enter it (regression: a non-reentrant lock
    blocked forever here)."""
    import threading

    import torch
    from sph2grid import _lib
    box = {}

    def run():
        try:
            box["ctx"] = _lib.default_context()
        except _lib.LibraryError as e:
            box["err"] = e

    t = threading.Thread(target=run, daemon=True)
    t.start()
    t.join(timeout=60)
    assert not t.is_alive(), "default_context() deadlocked"
    if torch.cuda.is_available():
        assert "ctx" in box
    else:
        assert "err" in box
