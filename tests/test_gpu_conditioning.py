"""FP64 parity with the oracle across the lightning structure's band, with the
condition number in view, and the condition estimate itself.

Below the structure's resonances the EFIE matrix's condition number grows as
1/(k dl)^2: on the lightning structure at 10 kHz kappa_1(Z) exceeds 1e9 at
N = 1024 and 1e11 at N = 8192.  Two FP64 solutions of Z x = b whose normwise
backward errors are e1 and e2 (relative to ||Z|| ||x||) differ by up to
kappa * (e1 + e2) relative, and the GPU fill differs from the oracle's by an
ulp or so per entry (its sin/cos, square root and reciprocals are not the host
library's), so agreement to 1e-10 is a property of the matrix, not of the
solver, once kappa * u64 (u64 = 2^-53) approaches 1e-10.  The tests therefore
assert, per frequency:

  backward  the GPU solution solves the ORACLE's system to LAPACK's backward
            error criterion: ||Z_o (x_gpu - x_o)||_1 / (||Z_o||_1 ||x_o||_1 n eps)
            <= 30 (the residual difference taken in extended precision), which
            a wrong entry, pivot or update fails by orders of magnitude;
  forward   |Zin_gpu - Zin_o| / |Zin_o| <= 1e-10 wherever kappa_1 u64 <= 1e-11,
            and everywhere within the kappa-scaled bound
            ||Z_o^-1||_1 ||Z_o (x_gpu - x_o)||_1 / |x_o[gap]|;
  estimate  the GPU condition estimate (Hager-Higham with the GPU's factors)
            lies within [kappa / 3, kappa (1 + 1e-6)] of the exact kappa_1
            (numpy at N = 1024) or of LAPACK zgecon's estimate (N = 8192).
"""
import ctypes as C
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

U64 = 2.0 ** -53
LAPACK_RATIO = 30.0


def _light(G, geo_dir, n):
    p = os.path.join(geo_dir, f"light{n}.geo")
    if not os.path.exists(p):
        G.lightning(p, n)
    return p


def _check_freq(n, f, Zo, xo, xg, zin_o, zin_g, gap, kappa_gpu, kappa_ref, inv_norm):
    d = xg - xo
    dr = Zo.astype(np.clongdouble) @ d.astype(np.clongdouble)
    dr1 = float(np.abs(dr).sum())
    an = float(np.abs(Zo).sum(axis=0).max())
    ratio = dr1 / (an * np.abs(xo).sum() * n * U64)
    rel = abs(zin_g - zin_o) / abs(zin_o)
    bound = inv_norm * dr1 / abs(xo[gap - 1])
    ku = kappa_ref * U64
    print(f"N={n} f={f:.4g} MHz kappa1={kappa_ref:.3e} (gpu est {kappa_gpu:.3e}) kappa*u={ku:.2e} "
          f"rel(Zin)={rel:.3e} rel/(kappa u)={rel / ku:.3f} backward ratio={ratio:.3f}")
    assert ratio <= LAPACK_RATIO, (f, ratio)
    if ku <= 1e-11:
        assert rel <= 1e-10, (f, rel, ku)
    assert rel <= max(1e-10, bound * (1 + 1e-6)), (f, rel, bound)
    assert kappa_ref / 3 <= kappa_gpu <= kappa_ref * (1 + 1e-6), (f, kappa_gpu, kappa_ref)


def test_fp64_lightning_band_n1024_vs_oracle(wm, geometries, oracle, ctx, geo_dir):
    """All 12 log-spaced frequencies of 10 kHz - 10 MHz on the lightning structure
    at N = 1024: FP64 GPU path against the oracle (fill, LAPACK-semantics LU)."""
    p = _light(geometries, geo_dir, 1024)
    t = oracle.load(p)
    g = wm.Geometry.load(ctx, p)
    gap = geometries.nearest_basis(t, (0, 0, 19.9))
    f = np.logspace(-2, 1, 12).astype(np.float32)
    zin, I, st = wm.solve_sweep(g, f, [gap], "f64", cond_est=True)
    n = t.npls
    for i, fr in enumerate(f):
        zo, Xo, Zo = oracle.solve_freq(t, float(fr), gaps=(gap,))
        inv_norm = np.abs(np.linalg.inv(Zo)).sum(axis=0).max()
        kappa = np.abs(Zo).sum(axis=0).max() * inv_norm
        assert abs(st[i]["anorm1"] - np.abs(Zo).sum(axis=0).max()) <= 1e-12 * st[i]["anorm1"]
        _check_freq(n, float(fr), Zo, Xo[:, 0], I[i, 0], zo[0], zin[i, 0], gap, st[i]["cond1"], kappa, inv_norm)
    # the condition number grows toward low frequency (1/(k dl)^2 below resonance)
    k = [s["cond1"] for s in st]
    assert k[0] > 1e3 * k[-1]


def test_fp64_lightning_n8192_vs_oracle(wm, geometries, oracle, ctx, geo_dir):
    """The metric structure at full size, N = 8192, at 10 kHz, 1 MHz and 10 MHz:
    the oracle's fill and LAPACK's zgetrf/zgetrs/zgecon (scipy's LAPACK: the
    routines the reference calls through MKL, "wersja referencyjna/Wire_MoM.cu":
    372, 382) against the FP64 GPU sweep."""
    sl = pytest.importorskip("scipy.linalg")
    from scipy.linalg import lapack
    p = _light(geometries, geo_dir, 8192)
    t = oracle.load(p)
    g = wm.Geometry.load(ctx, p)
    gap = geometries.nearest_basis(t, (0, 0, 19.9))
    f = np.array([0.01, 1.0, 10.0], np.float32)
    zin, I, st = wm.solve_sweep(g, f, [gap], "f64", cond_est=True)
    n = t.npls
    b = oracle.excitation(n, [gap])[:, 0]
    for i, fr in enumerate(f):
        Zo = oracle.zmatrix(t, oracle.wavenumber(float(fr)))
        an = float(np.abs(Zo).sum(axis=0).max())
        lu = sl.lu_factor(Zo)
        xo = sl.lu_solve(lu, b)
        rcond, info = lapack.zgecon(lu[0], an, norm="1")
        assert info == 0
        kappa = 1.0 / rcond
        zo = b[gap - 1] / xo[gap - 1]
        _check_freq(n, float(fr), Zo, xo, I[i, 0], zo, zin[i, 0], gap, st[i]["cond1"], kappa, kappa / an * 3)
        del Zo, lu


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_conjugate_transpose_solve(wm, ctx, prec):
    """A^H X = B with the factors of A (the condition estimate's second solve),
    against numpy; n = 320 (five 64-row blocks), two matrices, three right-hand sides."""
    L = wm.lib()
    L.wmx_getrs_trans.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_int, C.c_int64,
                                  C.c_void_p, C.c_int64, C.c_void_p, C.c_int, C.c_int64, C.c_int]
    rng = np.random.RandomState(4)
    n, b, nr = 320, 2, 3
    A = rng.randn(b, n, n) + 1j * rng.randn(b, n, n)
    B = rng.randn(b, n, nr) + 1j * rng.randn(b, n, nr)
    dev, piv, info = wm.getrf_batched(ctx, A, prec)
    assert not info.any()
    Bd = wm.DeviceArray(ctx, np.ascontiguousarray(B.transpose(0, 2, 1)))
    assert L.wmx_getrs_trans(ctx.h, wm.PREC_NAMES[prec], n, nr, dev.ptr, n, n * n, piv.ptr, n, Bd.ptr, n, n * nr,
                             b) == 0
    X = Bd.download().transpose(0, 2, 1)
    for i in range(b):
        Ah = A[i].conj().T
        if prec == "f64":
            r = np.abs(Ah @ X[i] - B[i]).max() / (np.abs(Ah).max() * np.abs(X[i]).max() * n)
            assert r <= 1e-15, r
        else:  # complex64 factors of A: X solves a nearby system
            Xr = np.linalg.solve(Ah, B[i])
            assert np.abs(X[i] - Xr).max() <= 1e-3 * np.abs(Xr).max()


def test_mixed_sweep_without_fallback(wm, geometries, ctx, geo_dir):
    """fallback=False: frequencies the refinement rates route to FP64 factors are
    still solved (in FP64, identical to the FP64 sweep); only items that tried
    complex64 refinement and failed make the call report WM_ERR_NOCONV, after
    every other frequency has its result."""
    p = _light(geometries, geo_dir, 1024)
    g = wm.Geometry.load(ctx, p)
    t = wm.configwire(p, cwd=geo_dir)
    gap = geometries.nearest_basis(t, (0, 0, 19.9))
    f = np.logspace(-2, 1, 12).astype(np.float32)
    z64, _, s64 = wm.solve_sweep(g, f, [gap], "f64", cond_est=True)
    zin, _, st, rc = wm.solve_sweep(g, f, [gap], "f32", batch=4, size_policy=False, fallback=False,
                                    return_status=True)
    assert rc in (0, wm.ERR_NOCONV), rc
    routed = [i for i, s in enumerate(st) if s["fell_back"] == 2]
    assert routed, st
    for i in routed:
        assert zin[i, 0] == z64[i, 0], (f[i], zin[i, 0], z64[i, 0])
    conv = [i for i, s in enumerate(st) if s["fell_back"] == 0]
    assert conv and all(st[i]["prec"] != wm.PREC_F64 for i in conv)
    for i in conv:  # two FP64-accurate solutions: within 1e-10, or kappa_1 u64 where that is larger
        bound = max(1e-10, s64[i]["cond1"] * 2.0 ** -53)
        assert abs(zin[i, 0] - z64[i, 0]) <= bound * abs(z64[i, 0]), (f[i], s64[i]["cond1"])
    failed = [s for s in st if s["fell_back"] == -1]
    assert (rc == wm.ERR_NOCONV) == bool(failed)


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_grouped_solve_equals_per_item(wm, ctx, prec):
    """Solves that share factors run as the right-hand sides of one launch
    member (the recycled sweep's preconditioner): every column equals the same
    column solved on its own with its factors, bit for bit, whatever it is
    grouped with; n = 384, three factors, nine items of two columns each
    (groups of 10, 6 and 2 columns: members of 8 columns, the last padded)."""
    L = wm.lib()
    L.wmx_getrs_grouped.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_int, C.c_void_p,
                                    C.c_void_p, C.c_int64, C.c_int, C.c_void_p, C.c_void_p, C.c_int, C.c_int]
    rng = np.random.RandomState(7)
    n, nf, nr = 384, 3, 2
    A = rng.randn(nf, n, n) + 1j * rng.randn(nf, n, n) + 4 * n ** 0.5 * np.eye(n)
    fac = np.array([0, 0, 1, 2, 0, 1, 0, 1, 0], np.int32)
    items = np.array([8, 1, 2, 3, 4, 5, 6, 7, 0], np.int32)  # item order differs from storage order
    V = rng.randn(9, n, nr) + 1j * rng.randn(9, n, nr)
    dev, piv, info = wm.getrf_batched(ctx, A, prec)
    assert not info.any()
    Vd = wm.DeviceArray(ctx, np.ascontiguousarray(V.transpose(0, 2, 1)))
    assert L.wmx_getrs_grouped(ctx.h, wm.PREC_NAMES[prec], n, nr, dev.ptr, nf, piv.ptr, Vd.ptr, n * nr, len(items),
                               items.ctypes.data, fac.ctypes.data, 0, 8) == 0
    got = Vd.download().transpose(0, 2, 1)
    LU = dev.download()
    P = piv.download()
    for q, (it, f) in enumerate(zip(items, fac)):
        d1 = wm.DeviceArray(ctx, np.ascontiguousarray(LU[f:f + 1]))
        p1 = wm.DeviceArray(ctx, np.ascontiguousarray(P[f:f + 1]))
        for r in range(nr):  # one column at a time: launch members of one right-hand side
            one = wm.getrs_batched(ctx, d1, p1, V[it][:, r:r + 1][None], prec)[0][:, 0]
            g = np.ascontiguousarray(got[it][:, r])
            rel = np.abs(g - one).max() / np.abs(one).max()
            assert np.array_equal(g.view(np.uint64), np.ascontiguousarray(one).view(np.uint64)), \
                (prec, q, int(it), int(f), r, rel)
        d1.free()
        p1.free()


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_grouped_solve_with_inverted_diagonal_blocks(wm, ctx, prec):
    """The preconditioner's solves: the 64 x 64 diagonal blocks of L and U
    inverted once (FP64) and applied as products.  Against numpy's solve with
    the same factors' matrix (FP64 factors: 1e-12 on these well-conditioned
    matrices; complex64 factors: the solution of the nearby system, 1e-3), and
    a column's values are the same grouped or alone, bit for bit."""
    L = wm.lib()
    L.wmx_getrs_grouped.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_int, C.c_void_p,
                                    C.c_void_p, C.c_int64, C.c_int, C.c_void_p, C.c_void_p, C.c_int, C.c_int]
    rng = np.random.RandomState(8)
    n, nf, nr = 384, 2, 1
    A = rng.randn(nf, n, n) + 1j * rng.randn(nf, n, n) + 4 * n ** 0.5 * np.eye(n)
    fac = np.array([0, 1, 0, 0, 1], np.int32)
    items = np.arange(5, dtype=np.int32)
    V = rng.randn(5, n, nr) + 1j * rng.randn(5, n, nr)
    dev, piv, info = wm.getrf_batched(ctx, A, prec)
    assert not info.any()
    p = wm.PREC_NAMES[prec]
    Vd = wm.DeviceArray(ctx, np.ascontiguousarray(V.transpose(0, 2, 1)))
    assert L.wmx_getrs_grouped(ctx.h, p, n, nr, dev.ptr, nf, piv.ptr, Vd.ptr, n * nr, 5, items.ctypes.data,
                               fac.ctypes.data, 1, 0) == 0
    got = Vd.download().transpose(0, 2, 1)
    for q in range(5):
        x = np.linalg.solve(A[fac[q]], V[q])
        tol = 1e-12 if prec == "f64" else 1e-3
        assert np.abs(got[q] - x).max() <= tol * np.abs(x).max(), (prec, q)
        one = wm.DeviceArray(ctx, np.ascontiguousarray(V.transpose(0, 2, 1)))
        it1, f1 = np.array([q], np.int32), np.array([fac[q]], np.int32)
        assert L.wmx_getrs_grouped(ctx.h, p, n, nr, dev.ptr, nf, piv.ptr, one.ptr, n * nr, 1, it1.ctypes.data,
                                   f1.ctypes.data, 1, 0) == 0
        alone = one.download().transpose(0, 2, 1)[q]
        one.free()
        assert np.array_equal(np.ascontiguousarray(alone).view(np.uint64),
                              np.ascontiguousarray(got[q]).view(np.uint64)), (prec, q)


def test_c3_full_size_sweeps_match_fp64(wm, geometries, ctx, geo_dir):
    """The benchmark workload itself: the lightning structure at N = 8192, all 64
    frequencies 10 kHz - 10 MHz.  The FP32 and split-BF16x3 sweeps (with the
    ladder) and the recycled sweep as the benchmark runs it (r = 2.5) against
    the FP64 sweep, Zin element by element: within 1e-10 where kappa_1 u64 <=
    1e-11 (no frequency of this band), else within kappa_1 u64 (kappa_1 from the
    FP64 sweep's estimate); the recycled anchors equal the FP64 sweep bit for bit."""
    p = _light(geometries, geo_dir, 8192)
    g = wm.Geometry.load(ctx, p)
    t = wm.configwire(p, cwd=geo_dir)
    assert g.npls == 8192
    gap = geometries.nearest_basis(t, (0.0, 0.0, 19.9))
    f = geometries.freqs_c3()
    z64, _, s64 = wm.solve_sweep(g, f, [gap], "f64", cond_est=True)
    ku = np.array([s["cond1"] for s in s64]) * 2.0 ** -53
    bound = np.where(ku <= 1e-11, 1e-10, np.maximum(1e-10, ku))
    for prec, kw in (("f32", {}), ("bf16x3", {}), ("f64", {"recycle": 2.5})):
        z, _, st = wm.solve_sweep(g, f, [gap], prec, **kw)
        rel = np.abs(z[:, 0] - z64[:, 0]) / np.abs(z64[:, 0])
        print(prec, kw, "max rel", rel.max(), "max rel / bound", (rel / bound).max())
        assert (rel <= bound).all(), (prec, kw, f[rel > bound], rel[rel > bound], bound[rel > bound])
        if kw:
            anchors = [i for i, s in enumerate(st) if s["anchor_mhz"] > 0 and s["iters"] == 0]
            assert 3 <= len(anchors) < 64
            for i in anchors:
                assert z[i, 0] == z64[i, 0], (f[i], z[i, 0], z64[i, 0])
    g.close()
