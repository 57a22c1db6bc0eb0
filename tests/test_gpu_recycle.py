"""Recycled sweeps (wm_solve_opts.recycle): FP64 factors at anchor frequencies,
FP64 GMRES refinement with the nearest anchor's factors everywhere else.

A recycled result is an FP64 solve of its own frequency's matrix, refined until
the correction is below 1e-12 of the solution, so it must meet the same bars as
the FP64 path: LAPACK's backward-error criterion against the oracle's matrix
(||Z_o (x - x_o)||_1 / (||Z_o||_1 ||x_o||_1 n eps) <= 30), Zin within 1e-10 of
the FP64 sweep wherever kappa_1 u64 <= 1e-11, and elsewhere within the
kappa-limited distance that separates any two backward-stable FP64 solutions.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

U64 = 2.0 ** -53


def _light(G, d, n):
    p = os.path.join(d, f"light{n}.geo")
    if not os.path.exists(p):
        G.lightning(p, n)
    return p


def _covering(r, f_mhz, diam):
    """The anchor covering factor of a recycled sweep at f: r^(1 + log10(1 / (k D)))
    for k D <= 1, 1 + (r - 1) / (k D) above."""
    kd = 2.0 * np.pi * f_mhz * 1e6 / 2.99792448e8 * diam
    return 1.0 + (r - 1.0) / kd if kd > 1.0 else r ** (1.0 - np.log10(kd))


def test_recycled_lightning_band_vs_oracle(wm, geometries, oracle, ctx, geo_dir):
    """12 log-spaced frequencies over 10 kHz - 10 MHz, r = 2 (a factor 2 at k D = 1,
    4 a decade below): every frequency within its covering factor of its anchor and
    backward-stable against the oracle's own matrix."""
    p = _light(geometries, geo_dir, 1024)
    t = oracle.load(p)
    g = wm.Geometry.load(ctx, p)
    gap = geometries.nearest_basis(t, (0, 0, 19.9))
    f = np.logspace(-2, 1, 12).astype(np.float32)
    z64, I64, s64 = wm.solve_sweep(g, f, [gap], "f64", cond_est=True)
    zr, Ir, sr = wm.solve_sweep(g, f, [gap], "f64", recycle=2.0)
    for fr, s in zip(f, sr):
        print(f"f={fr:.4g} anchor={s['anchor_mhz']:.4g} iters={s['iters']} fell_back={s['fell_back']} "
              f"dx_rel={s['dx_rel']:.2e}")
    anchors = [i for i, s in enumerate(sr) if s["anchor_mhz"] == float(f[i])]
    deps = [i for i, s in enumerate(sr) if s["anchor_mhz"] != float(f[i])]
    assert 3 <= len(anchors) <= 8 and deps, [s["anchor_mhz"] for s in sr]
    n = t.npls
    pts = np.stack([t.x[:t.npnt], t.y[:t.npnt], t.z[:t.npnt]]).astype(np.float64)
    diam = float(np.sqrt((np.ptp(pts, axis=1) ** 2).sum()))
    for i, fr in enumerate(f):
        s = sr[i]
        assert s["fell_back"] == 0 and s["anchor_mhz"] > 0
        ratio_f = max(s["anchor_mhz"] / fr, fr / s["anchor_mhz"])
        # never beyond the covering factor at the higher of the two frequencies
        assert ratio_f <= _covering(2.0, max(float(fr), s["anchor_mhz"]), diam) * 1.01, (fr, s["anchor_mhz"])
        zo, Xo, Zo = oracle.solve_freq(t, float(fr), gaps=(gap,))
        d = Ir[i, 0] - Xo[:, 0]
        dr = np.abs(Zo.astype(np.clongdouble) @ d.astype(np.clongdouble)).sum()
        ratio = float(dr / (np.abs(Zo).sum(axis=0).max() * np.abs(Xo[:, 0]).sum() * n * U64))
        rel = abs(zr[i, 0] - z64[i, 0]) / abs(z64[i, 0])
        ku = s64[i]["cond1"] * U64
        print(f"f={fr:.4g} anchor={s['anchor_mhz']:.4g} iters={s['iters']} kappa*u={ku:.2e} "
              f"rel vs fp64={rel:.2e} backward ratio={ratio:.3f}")
        assert ratio <= 30.0, (fr, ratio)
        if ku <= 1e-11:
            assert rel <= 1e-10, (fr, rel)
        if i in deps:
            assert 1 <= s["iters"] <= 96, s


def test_recycled_grid_sweep_matches_fp64(wm, geometries, ctx, geo_dir):
    """A planar grid (N = 1024, two feeds, 10 x 10 m) over 12 frequencies linearly
    spaced in 1 - 5 MHz, where it is electrically small to moderate (k D <= 1.5):
    anchors within a factor 1.5, every Zin within 1e-10 of the FP64 sweep (the
    grid is well conditioned there)."""
    p = os.path.join(geo_dir, "rgrid1024.geo")
    geometries.planar_grid(p, 1024, seed=7)
    g = wm.Geometry.load(ctx, p)
    f = np.linspace(1.0, 5.0, 12).astype(np.float32)
    z64, I64, _ = wm.solve_sweep(g, f, [5, 700], "f64")
    zr, Ir, sr = wm.solve_sweep(g, f, [5, 700], "f64", recycle=1.5)
    rel = (np.abs(zr - z64) / np.abs(z64)).max(axis=1)
    print([(float(a), s["anchor_mhz"], s["iters"], float(e)) for a, s, e in zip(f, sr, rel)])
    assert (rel <= 1e-10).all(), rel
    relI = np.abs(Ir - I64).max(axis=2) / np.abs(I64).max(axis=2)
    assert (relI <= 1e-9).all(), relI
    assert any(s["anchor_mhz"] != float(x) for x, s in zip(f, sr))
    # an electrically large sweep (1 - 100 MHz on the 10 m grid, k D up to 29):
    # the covering factor shrinks to 1 + 0.5 / (k D), below the spacing, so every
    # frequency is factored itself and the result is the FP64 sweep's
    f2 = np.linspace(1.0, 100.0, 12).astype(np.float32)
    z2, _, s2 = wm.solve_sweep(g, f2, [5, 700], "f64", recycle=1.5)
    z2_64, _, _ = wm.solve_sweep(g, f2, [5, 700], "f64")
    assert all(s["anchor_mhz"] == float(x) for x, s in zip(f2, s2))
    assert np.array_equal(z2, z2_64)


def test_recycled_result_independent_of_batch(wm, geometries, ctx, geo_dir):
    """An item's GMRES refinement depends only on its own matrix and its anchor's
    factors: the same frequencies in batches of 4 (anchors chosen within each
    batch) agree with one batch wherever the anchor is the same, bit for bit."""
    p = _light(geometries, geo_dir, 1024)
    g = wm.Geometry.load(ctx, p)
    t = wm.configwire(p, cwd=geo_dir)
    gap = geometries.nearest_basis(t, (0, 0, 19.9))
    f = np.logspace(0, 1, 8).astype(np.float32)
    za, _, sa = wm.solve_sweep(g, f, [gap], "f64", recycle=1.4)
    zb, _, sb = wm.solve_sweep(g, f, [gap], "f64", recycle=1.4, batch=4)
    same = [i for i in range(len(f)) if sa[i]["anchor_mhz"] == sb[i]["anchor_mhz"]]
    assert same
    for i in same:
        assert za[i, 0] == zb[i, 0], (f[i], za[i, 0], zb[i, 0])


def test_recycled_fill_on_second_stream_bit_identical(wm, geometries, ctx, geo_dir):
    """The GMRES items' matrices are filled on a second stream while the anchors
    are factored; the result is the same bit for bit as with every fill on the
    launch stream, over three sweeps in a row (the streams' order holds across
    calls), in free space and over the PEC ground."""
    rods = os.path.join(geo_dir, "rec_rods.geo")
    geometries.lightning(rods, 1024, depth=0.0, ring=False)
    f = np.logspace(-2, 1, 16).astype(np.float32)
    try:
        for p, ienv in ((_light(geometries, geo_dir, 1024), 1), (rods, 2)):
            g = wm.Geometry.load(ctx, p, ienv=ienv)
            gap = [40]
            ctx.set_fill_overlap(False)
            z0, I0, _ = wm.solve_sweep(g, f, gap, "f64", recycle=2.0)
            ctx.set_fill_overlap(True)
            for _ in range(3):
                z1, I1, _ = wm.solve_sweep(g, f, gap, "f64", recycle=2.0)
                assert np.array_equal(z0, z1) and np.array_equal(I0, I1)
    finally:
        ctx.set_fill_overlap(True)


def test_recycled_pec_loads_generators_plane_wave(wm, geometries, ctx, geo_dir):
    """Recycling with everything else the sweep supports: the PEC ground (image
    terms; the anchor radius uses the diameter with the images), lumped loads,
    two generators and a plane wave over three right-hand sides, 10 frequencies
    0.2 - 20 MHz: every generator's V/I and every current within 1e-10 of the
    FP64 sweep wherever kappa_1 u64 <= 1e-11, else within kappa_1 u64."""
    p = os.path.join(geo_dir, "rec_rods.geo")
    geometries.lightning(p, 1024, depth=0.0, ring=False)
    g = wm.Geometry.load(ctx, p, ienv=2)
    f = np.logspace(np.log10(0.2), np.log10(20.0), 10).astype(np.float32)
    sources = [(40, 0, 1.0, 0.0), (700, 1, 0.5, 45.0)]
    loads = [(300, 25.0, 0.0, 0.0), (500, 0.0, 0.5, 0.0), (900, 5.0, 0.0, 1e-3)]
    waves = [(2, 60.0, 30.0, 1.0 + 0.5j, 0.0)]
    z64, I64, s64 = wm.solve_sweep_ex(g, f, 3, sources, loads, "f64", plane_waves=waves, cond_est=True)
    zr, Ir, sr = wm.solve_sweep_ex(g, f, 3, sources, loads, "f64", plane_waves=waves, recycle=2.0)
    ku = np.array([s["cond1"] for s in s64]) * U64
    bound = np.where(ku <= 1e-11, 1e-10, np.maximum(1e-10, ku))
    anchors = sum(1 for s in sr if s["anchor_mhz"] > 0 and s["iters"] == 0)
    assert 1 <= anchors < len(f)
    for i in range(len(f)):
        ez = (np.abs(zr[i] - z64[i]) / np.abs(z64[i])).max()
        eI = (np.abs(Ir[i] - I64[i]).max(axis=1) / np.abs(I64[i]).max(axis=1)).max()
        print(float(f[i]), sr[i]["anchor_mhz"], sr[i]["iters"], ez, eI, bound[i])
        assert ez <= bound[i] and eI <= 10 * bound[i], (float(f[i]), ez, eI, bound[i])


def test_recycled_reference_127_frequency_sweep(wm, geometries, ctx, geo_dir):
    """The reference's own lightning sweep grid ("wersja referencyjna/freq1.txt":
    127 frequencies, 1 kHz - 10 MHz in piecewise-uniform steps of 1 kHz, 10 kHz,
    50 kHz and 100 kHz) on the lightning structure at N = 1024, recycled with the
    benchmark's r = 2.5: far fewer anchors than frequencies (the quasi-static
    decades share anchors), every Zin within max(1e-10, kappa_1 u64) of the FP64
    sweep, and the anchors equal to it bit for bit."""
    f = np.concatenate([np.arange(1, 11) * 1e-3, np.arange(2, 11) * 1e-2, 0.1 + np.arange(1, 19) * 0.05,
                        1.0 + np.arange(1, 11) * 0.1, 2.0 + np.arange(1, 81) * 0.1]).astype(np.float32)
    assert len(f) == 127
    p = _light(geometries, geo_dir, 1024)
    g = wm.Geometry.load(ctx, p)
    t = wm.configwire(p, cwd=geo_dir)
    gap = geometries.nearest_basis(t, (0, 0, 19.9))
    z64, _, s64 = wm.solve_sweep(g, f, [gap], "f64", cond_est=True)
    zr, _, sr = wm.solve_sweep(g, f, [gap], "f64", recycle=2.5)
    ku = np.array([s["cond1"] for s in s64]) * U64
    bound = np.maximum(1e-10, ku)
    rel = np.abs(zr[:, 0] - z64[:, 0]) / np.abs(z64[:, 0])
    anchors = [i for i, s in enumerate(sr) if s["anchor_mhz"] == float(f[i])]
    print("anchors", [float(f[i]) for i in anchors], "max rel / bound", (rel / bound).max(),
          "iterations", sum(s["iters"] for s in sr))
    assert 2 <= len(anchors) <= 20
    assert all(s["fell_back"] == 0 for s in sr)
    assert (rel <= bound).all(), (f[rel > bound], rel[rel > bound], bound[rel > bound])
    for i in anchors:
        assert zr[i, 0] == z64[i, 0], (f[i], zr[i, 0], z64[i, 0])
