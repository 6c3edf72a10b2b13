"""sw_smooth_noalloc, the smoothing entry point that allocates no device
memory (its intermediates live in the library's fixed per-GPU array),
called through the C ABI and checked against sw_smooth bit for bit on two
calls in a row, the second on a smaller grid (the array keeps the first
call's contents), and its rejection of grids beyond the array."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _field_context(seiswg, nx, nz, seed):
    c = seiswg.Context(nx, nz, 4, 10.0, 10.0, 1e-3, sh=False, psv=True)
    v = np.random.default_rng(seed).standard_normal(nx * nz).astype(np.float32)
    return c, v


def test_noalloc_smoothing_equals_sw_smooth_on_two_calls(seiswg):
    L = seiswg.load()
    F = seiswg.abi.F_K_MU
    for nx, nz, sigma, seed in ((640, 480, 3.0, 1), (300, 200, 5.0, 2)):
        print(f"case {nx}x{nz} sigma={sigma}", flush=True)
        c, v = _field_context(seiswg, nx, nz, seed)
        c.write_field(F, v)
        c.smooth(F, sigma)
        ref = c.field(F)
        c.write_field(F, v)
        rc = L.sw_smooth_noalloc(c.h, F, C.c_float(sigma))
        assert rc == 0, L.sw_last_error().decode()
        got = c.field(F)
        assert np.array_equal(got.view(np.uint32), ref.view(np.uint32)), (nx, nz)
        assert np.any(got != v)
        c.close()


def test_noalloc_smoothing_rejects_grids_beyond_its_scratch(seiswg):
    L = seiswg.load()
    c = seiswg.Context(2400, 2400, 4, 10.0, 10.0, 1e-3, sh=False, psv=True)
    rc = L.sw_smooth_noalloc(c.h, seiswg.abi.F_K_MU, C.c_float(3.0))
    assert rc == -2
    assert "beyond its fixed scratch" in L.sw_last_error().decode()
    c.close()
