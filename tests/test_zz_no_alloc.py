"""sp_filter_noalloc against sp_filter (marked gpu).

The no-allocation variant filters a caller's device grid in place with the
library's static workspace; the context entry point filters the context's
grid with its own work grid.  Both run the same kernels, so the results must
be bit-identical -- on two calls in a row, the second on a smaller grid (the
workspace keeps the first call's contents), with one launch (<= 8 passes,
result copied back from the workspace) and two (> 8 passes, result back in the
caller's grid), in FP32 and FP64.  The caller's buffers come from hipMalloc of
the HIP runtime the library itself links."""
import ctypes as ct

import numpy as np
import pytest

from spectral_rs_amd import capi

pytestmark = pytest.mark.gpu


def hip():
    capi.load()  # the library's own HIP runtime is loaded first
    L = ct.CDLL("libamdhip64.so.7")
    L.hipMalloc.argtypes = [ct.POINTER(ct.c_void_p), ct.c_size_t]
    L.hipFree.argtypes = [ct.c_void_p]
    L.hipMemcpy.argtypes = [ct.c_void_p, ct.c_void_p, ct.c_size_t, ct.c_int]
    L.hipDeviceSynchronize.argtypes = []
    for f in (L.hipMalloc, L.hipFree, L.hipMemcpy, L.hipDeviceSynchronize):
        f.restype = ct.c_int
    return L


H2D, D2H = 1, 2


def noalloc_filter(L, a, passes, prec):
    ny, nx = a.shape[0] - 2, a.shape[1] - 2
    buf = ct.c_void_p()
    assert L.hipMalloc(ct.byref(buf), a.nbytes) == 0
    try:
        assert L.hipMemcpy(buf, a.ctypes.data, a.nbytes, H2D) == 0
        rc = capi.load().sp_filter_noalloc(buf, nx, ny, capi.SP_F64 if prec == 8 else capi.SP_F32, passes, None)
        assert rc == capi.SP_OK, capi.load().sp_last_error(None)
        assert L.hipDeviceSynchronize() == 0
        out = np.empty_like(a)
        assert L.hipMemcpy(out.ctypes.data, buf, a.nbytes, D2H) == 0
        return out
    finally:
        L.hipFree(buf)


def ctx_filter(a, passes, prec):
    ny, nx = a.shape[0] - 2, a.shape[1] - 2
    g = capi.GpuSim(size_x=nx, size_y=ny, delta=2, dens=1, prec=prec)
    g.set_grid("jx", a)
    g.filter("jx", passes)
    return g.grid("jx")


@pytest.mark.parametrize("prec", [4, 8])
@pytest.mark.parametrize("passes", [4, 11])
def test_noalloc_filter_equals_the_context_filter(prec, passes):
    L = hip()
    ft = np.float64 if prec == 8 else np.float32
    rng = np.random.default_rng(passes + prec)
    # a padded grid whose ghost ring is not periodic (the first pass reads it
    # as it stands, like the reference after the deposit fold)
    for ny, nx in [(96, 384), (24, 48)]:
        a = rng.standard_normal((ny + 2, nx + 2)).astype(ft)
        got = noalloc_filter(L, a, passes, prec)
        want = ctx_filter(a, passes, prec)
        assert np.array_equal(got, want), (ny, nx)
        assert not np.array_equal(got, a)


def test_noalloc_filter_rejects_what_does_not_fit():
    L = hip()
    lib = capi.load()
    buf = ct.c_void_p()
    assert L.hipMalloc(ct.byref(buf), 1024) == 0
    try:
        # 8192 x 1024 doubles need 67 MB, beyond the 33.7 MB workspace; the
        # call returns before touching the grid
        assert lib.sp_filter_noalloc(buf, 8192, 1024, capi.SP_F64, 4, None) == capi.SP_EINVAL
        assert b"exceeds" in lib.sp_last_error(None)
        assert lib.sp_filter_noalloc(buf, 7, 8, capi.SP_F32, 4, None) == capi.SP_EINVAL
        assert lib.sp_filter_noalloc(None, 8, 8, capi.SP_F32, 4, None) == capi.SP_EINVAL
        assert lib.sp_filter_noalloc(buf, 8, 8, 5, 4, None) == capi.SP_EINVAL
    finally:
        L.hipFree(buf)
    # the 4096 x 1024 grid fits in both precisions
    a = np.random.default_rng(3).standard_normal((1026, 4098)).astype(np.float64)
    got = noalloc_filter(L, a, 4, 8)
    assert np.array_equal(got, ctx_filter(a, 4, 8))
