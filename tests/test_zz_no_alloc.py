"""wm_getrs_batched_noalloc: the allocation-free triangular solves (scratch in
a library-owned device array) give the same results as wm_getrs_batched, bit
for bit, on two calls in a row (the second smaller, after the first left its
data in the scratch array), and reject sizes beyond the array."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

vp, i32, i64 = C.c_void_p, C.c_int, C.c_int64


def _fn(wm):
    f = wm.lib().wm_getrs_batched_noalloc
    f.argtypes = [vp, i32, i32, i32, vp, i32, i64, vp, i64, vp, i32, i64, i32]
    f.restype = i32
    return f


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_noalloc_matches_getrs_twice(wm, ctx, prec):
    noalloc = _fn(wm)
    rng = np.random.RandomState(11)
    for n, b, nr in [(320, 3, 3), (192, 2, 1)]:
        print(f"noalloc: prec={prec} n={n} batch={b} nrhs={nr}", flush=True)
        A = rng.randn(b, n, n) + 1j * rng.randn(b, n, n)
        B = rng.randn(b, n, nr) + 1j * rng.randn(b, n, nr)
        dev, piv, info = wm.getrf_batched(ctx, A, prec)
        assert not info.any()
        ref = wm.getrs_batched(ctx, dev, piv, B, prec)
        Bd = wm.DeviceArray(ctx, np.ascontiguousarray(B.transpose(0, 2, 1)))
        rc = noalloc(ctx.h, wm.PREC_NAMES[prec], n, nr, dev.ptr, n, n * n, piv.ptr, n, Bd.ptr, n, n * nr, b)
        assert rc == 0, wm.lib().wm_last_error()
        X = Bd.download().transpose(0, 2, 1)
        assert np.array_equal(X, ref)
        for i in range(b):
            r = np.abs(A[i] @ X[i] - B[i]).max() / (np.abs(A[i]).max() * np.abs(X[i]).max() * n)
            assert r <= (1e-15 if prec == "f64" else 1e-6), r


def test_noalloc_rejects_oversize(wm, ctx):
    """A size whose buffers exceed the scratch array is refused before any launch."""
    noalloc = _fn(wm)
    L = wm.lib()
    small = [C.c_void_p() for _ in range(3)]
    for p in small:
        assert L.wm_dev_alloc(ctx.h, 4096, C.byref(p)) == 0
    n, nr, b = 8192, 8, 64
    rc = noalloc(ctx.h, wm.PREC_F64, n, nr, small[0], n, n * n, small[1], n, small[2], n, n * nr, b)
    assert rc == wm.ERR_NOMEM
    assert b"scratch" in L.wm_last_error()
    for p in small:
        L.wm_dev_free(ctx.h, p)
