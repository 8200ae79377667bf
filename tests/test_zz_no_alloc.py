"""s2g_tsc3d_device_noalloc (the allocation-free variant of s2g_tsc3d_device,
whose work space is the library's fixed per-GPU device array) against
s2g_tsc3d_device through the C ABI: two calls in a row, the second with
fewer particles and a smaller grid, each equal to the allocating entry
point's grid bit for bit (same kernels, deterministic gather); plus the
limits, rejected before any device work."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _grid(ctx, fn, dtype_id, dpos, dq, n, par, flags, grid, stream=None):
    vp = lambda t: C.c_void_p(t.data_ptr())
    rc = fn(ctx.handle, dtype_id, vp(dpos), vp(dq), n, C.byref(par), flags, vp(grid), stream)
    return rc


def test_noalloc_matches_allocating_entry_point(s2g):
    import torch
    from sph2grid import _lib
    ctx = s2g.default_context()
    L = ctx.lib
    dev = torch.device("cuda:0")
    r = np.random.default_rng(2024)
    cases = [(1 << 20, 256, np.float64, 0), (30000, 40, np.float32, _lib.S2G_FLAG_PERIODIC)]
    for n, npix, dt, flags in cases:
        pos = r.uniform(-0.55, 0.55, (n, 3)).astype(dt)
        q = r.uniform(-1.0, 3.0, n).astype(dt)
        dpos = torch.from_numpy(pos).to(dev)
        dq = torch.from_numpy(q).to(dev)
        par = _lib.Params(npixels=npix, len2pix=float(npix))
        did = _lib.S2G_F32 if dt == np.float32 else _lib.S2G_F64
        want = torch.empty(2 * npix ** 3, dtype=torch.float64, device=dev)
        got = torch.full((2 * npix ** 3,), np.nan, dtype=torch.float64, device=dev)
        ctx.check(_grid(ctx, L.s2g_tsc3d_device, did, dpos, dq, n, par, flags, want),
                  "s2g_tsc3d_device")
        ctx.check(_grid(ctx, L.s2g_tsc3d_device_noalloc, did, dpos, dq, n, par, flags, got),
                  "s2g_tsc3d_device_noalloc")
        torch.cuda.synchronize()
        print(f"n={n} npix={npix} {dt.__name__} flags={flags}: sum W "
              f"{float(got.view(2, -1)[1].sum().item()):.6f}", flush=True)
        assert torch.equal(got, want)
        # every particle deposits at most unit weight; periodic grids keep it all
        sw = float(got.view(2, -1)[1].sum().item())
        assert sw <= n * (1 + 1e-12) and (abs(sw / n - 1) < 1e-12 if flags else sw > 0.5 * n)


def test_noalloc_limits(s2g):
    import torch
    from sph2grid import _lib
    ctx = s2g.default_context()
    L = ctx.lib
    dev = torch.device("cuda:0")
    t = torch.zeros(8, dtype=torch.float64, device=dev)
    g = torch.zeros(2 * 8 ** 3, dtype=torch.float64, device=dev)
    vp = lambda x: C.c_void_p(x.data_ptr())
    big = _lib.Params(npixels=513, len2pix=513.0)
    rc = L.s2g_tsc3d_device_noalloc(ctx.handle, _lib.S2G_F64, vp(t), vp(t), 1, C.byref(big), 0,
                                    vp(g), None)
    assert rc == _lib.S2G_EUNSUPPORTED and b"512" in L.s2g_last_error(ctx.handle)
    ok = _lib.Params(npixels=8, len2pix=8.0)
    rc = L.s2g_tsc3d_device_noalloc(ctx.handle, _lib.S2G_F64, vp(t), vp(t), (1 << 20) + 1,
                                    C.byref(ok), 0, vp(g), None)
    assert rc == _lib.S2G_EUNSUPPORTED and b"2^20" in L.s2g_last_error(ctx.handle)
    # an empty call writes a zero grid
    g.fill_(7.0)
    rc = L.s2g_tsc3d_device_noalloc(ctx.handle, _lib.S2G_F64, vp(t), vp(t), 0, C.byref(ok), 0,
                                    vp(g), None)
    assert rc == _lib.S2G_OK
    torch.cuda.synchronize()
    assert float(g.abs().max().item()) == 0.0
