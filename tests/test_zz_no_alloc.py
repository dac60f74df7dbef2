"""bsrg_mrk_cell_noalloc (the R^k band of one multipole into caller device
memory, work arrays in the library's static device array, no allocation)
against bsrg_mrk_cell on the same grid: two calls in a row, the second on a
smaller grid and on a caller stream, and the rejection of a grid whose work
arrays exceed the array.  Tolerance 1e-12 relative (the kernels are the
same, so the bits are expected to agree as well)."""
import ctypes as C

import numpy as np
import pytest

from tests.conftest import HAS_GPU, rel_err

pytestmark = pytest.mark.gpu


def _hip():
    hip = C.CDLL("libamdhip64.so")
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipFree.argtypes = [C.c_void_p]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
    hip.hipStreamSynchronize.argtypes = [C.c_void_p]
    hip.hipStreamDestroy.argtypes = [C.c_void_p]
    return hip


def _band_noalloc(hip, ctx, k, n, stream=None):
    from bsr_amd import lib as L
    d = C.c_void_p()
    assert hip.hipMalloc(C.byref(d), n * 8) == 0
    try:
        L.check(L.lib().bsrg_mrk_cell_noalloc(ctx._p, k, d, stream))
        if stream is not None:
            assert hip.hipStreamSynchronize(stream) == 0
        out = np.empty(n)
        assert hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), d, n * 8, 2) == 0   # device to host
        return out
    finally:
        hip.hipFree(d)


def test_noalloc_band_matches_mrk_cell():
    if not HAS_GPU:
        pytest.skip("no HIP device")
    from bsr_amd import Context, grid as G
    from bsr_amd.lib import BsrGpuError
    hip = _hip()
    with Context(0) as ctx:
        # first call: ks=9 grid with a geometric region (scaling law)
        t, ns, ml, me, h = G.mkgrid(1, 9, 3.0, 0.125, 0.4, 30.0)
        ctx.set_grid(9, ns, ml, me, h, t)
        n = ns * ns * 81
        print(f"[noalloc] call 1: ks=9 ns={ns} me={me} k=4", flush=True)
        got = _band_noalloc(hip, ctx, 4, n)
        ref = ctx.mrk_cell(4).ravel(order="F")
        assert rel_err(got, ref) < 1e-12
        assert np.array_equal(got, ref)
        # second call: a smaller grid, on a caller stream
        g = G.fixed_ns_grid(8, 40, 4.0, 12.0, 0.25)
        ctx.set_grid(8, g.ns, g.ml, g.me, g.h, g.t)
        s = C.c_void_p()
        assert hip.hipStreamCreate(C.byref(s)) == 0
        try:
            print(f"[noalloc] call 2: ks=8 ns={g.ns} k=2 on a caller stream", flush=True)
            got = _band_noalloc(hip, ctx, 2, g.ns * g.ns * 64, stream=s)
        finally:
            hip.hipStreamDestroy(s)
        ref = ctx.mrk_cell(2).ravel(order="F")
        assert rel_err(got, ref) < 1e-12
        assert np.array_equal(got, ref)
        # a grid whose work arrays exceed the 16 MB array is rejected
        ns = 900   # uniform knots: ns - 7 intervals of 0.05
        t = np.concatenate([np.zeros(8), 0.05 * np.arange(1, ns - 7), np.full(8, 0.05 * (ns - 7))])
        assert len(t) == ns + 8 and np.all(np.diff(t) >= 0)
        ctx.set_grid(8, ns, 0, 0, 0.05, t)
        print(f"[noalloc] rejection: ks=8 ns={ns}", flush=True)
        from bsr_amd import lib as L
        d = C.c_void_p()
        assert hip.hipMalloc(C.byref(d), ns * ns * 64 * 8) == 0   # room for the band: only the work arrays are too big
        try:
            with pytest.raises(BsrGpuError) as e:
                L.check(L.lib().bsrg_mrk_cell_noalloc(ctx._p, 0, d, None))
        finally:
            hip.hipFree(d)
        assert e.value.code == -6
