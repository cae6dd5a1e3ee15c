"""The allocation-free operator entry point (sfdx_apply_operator_noalloc)
against the allocating one (sfdx_apply_operator), through the C ABI, on two
calls in a row -- the second with a smaller grid, so the shared scratch holds
the first call's reciprocal widths when the second starts.  FP64 kernel: the
results must agree to 1e-14 relative (same stencil kernel, same reciprocals).
The test, as the caller, allocates the device buffers with the HIP runtime the
library itself loaded."""
import ctypes

import numpy as np
import pytest

from sfd_excmg_amd import sfdx
from test_oracle import tiny_model

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def _hip():
    """The HIP runtime that sees the GPU.  Another copy may be mapped too (torch
    ships its own, which does not see the device once the library's runtime
    owns it): take the one that reports a device."""
    sfdx.load()
    paths = []
    with open("/proc/self/maps") as f:
        for line in f:
            if "libamdhip64" in line and line.split()[-1] not in paths:
                paths.append(line.split()[-1])
    for path in sorted(paths, key=lambda p: "torch" in p):
        lib = ctypes.CDLL(path)
        n = ctypes.c_int(0)
        if lib.hipGetDeviceCount(ctypes.byref(n)) == 0 and n.value > 0:
            lib.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
            lib.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
            lib.hipFree.argtypes = [ctypes.c_void_p]
            return lib
    raise RuntimeError(f"no loaded HIP runtime sees a device: {paths}")


class _Dev:
    """Device copies of host arrays for one call (freed on close)."""

    def __init__(self, hip):
        self.hip, self.ptrs = hip, []

    def put(self, arr):
        arr = np.ascontiguousarray(arr)
        p = ctypes.c_void_p()
        assert self.hip.hipMalloc(ctypes.byref(p), arr.nbytes) == 0
        self.ptrs.append(p)
        assert self.hip.hipMemcpy(p, arr.ctypes.data_as(ctypes.c_void_p), arr.nbytes, H2D) == 0
        return p

    def get(self, p, like):
        out = np.empty_like(like)
        assert self.hip.hipDeviceSynchronize() == 0
        assert self.hip.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), p, out.nbytes, D2H) == 0
        return out

    def close(self):
        for p in self.ptrs:
            self.hip.hipFree(p)


def _noalloc_apply(ctx, hip, a, b, c, sig, f, x):
    dv = _Dev(hip)
    try:
        ad, bd, cd = dv.put(np.asarray(a, np.float64)), dv.put(np.asarray(b, np.float64)), dv.put(np.asarray(c, np.float64))
        sd = dv.put(sfdx.sigma_flat(sig))
        xd = dv.put(np.asarray(x, np.complex128))
        yd = dv.put(np.full(len(x), np.nan, np.complex128))
        g = sfdx.Grid(len(a), len(b), len(c), 0, ad, bd, cd)
        assert hip.hipDeviceSynchronize() == 0
        rc = ctx.lib.sfdx_apply_operator_noalloc(ctx.h, ctypes.byref(g), sd, float(f), xd, yd, None)
        assert rc == 0, ctx.lib.sfdx_last_error().decode()
        return dv.get(yd, np.asarray(x, np.complex128))
    finally:
        dv.close()


def test_noalloc_operator_matches_allocating_entry_point():
    ctx = sfdx.Context(0)
    hip = _hip()
    try:
        for shape, aniso, f, seed in (((40, 17, 26), True, 0.37, 1), ((9, 7, 8), False, 3.0, 2)):
            a, b, c, nair, sig = tiny_model(*shape, 2, seed=seed, aniso=aniso)
            rng = np.random.default_rng(seed)
            nl = sfdx.nedges(len(a), len(b), len(c))
            x = rng.standard_normal(nl) + 1j * rng.standard_normal(nl)
            print(f"noalloc operator: shape {shape} aniso {aniso} f {f}", flush=True)
            y_ref = ctx.apply_operator(a, b, c, sig, f, x)
            y = _noalloc_apply(ctx, hip, a, b, c, sig, f, x)
            assert np.all(np.isfinite(y))
            assert np.abs(y - y_ref).max() <= 1e-14 * np.abs(y_ref).max()
        # sizes beyond the scratch bound are refused before any launch
        big = sfdx.Grid(5000, 4, 4, 0, ctypes.c_void_p(8), ctypes.c_void_p(8), ctypes.c_void_p(8))
        rc = ctx.lib.sfdx_apply_operator_noalloc(ctx.h, ctypes.byref(big), ctypes.c_void_p(8), 1.0,
                                                 ctypes.c_void_p(8), ctypes.c_void_p(8), None)
        assert rc == 1 and b"4096" in ctx.lib.sfdx_last_error()
    finally:
        ctx.close()
