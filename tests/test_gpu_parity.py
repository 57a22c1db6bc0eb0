"""GPU parity: every kernel of the HIP path against the oracle on the same inputs.

Tolerances (north star: Zin within 1e-10 relative of the reference's FP64
path, which the oracle restates):
  fill          max|Z_gpu - Z_oracle| <= 1e-12 max|Z|     (FP64 rounding, FMA)
  FP64 LU       identical pivots; factors within 1e-11 relative
  Zin (FP64)    <= 1e-10 relative to the oracle
  Zin (F32+IR)  <= 1e-10 relative to the FP64 GPU path (also with split-BF16x3 updates)
  trailing update vs numpy, relative to max(|L| |U|): FP64 1e-14, FP32 2e-6,
                split-BF16x3 6e-5 (about 2^-16 per product, 64-term sums)
"""
import ctypes as C
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _dipole(G, d, nseg, length=25.0, radius=0.0005):
    p = os.path.join(d, f"gdip_{nseg}_{length}_{radius}.geo")
    if not os.path.exists(p):
        G.dipole(p, nseg, length, radius)
    return p


def test_device_is_gfx950(ctx):
    name, cu, hbm = ctx.info()
    print(name, cu, hbm)
    assert "gfx950" in name
    assert cu >= 200


@pytest.mark.parametrize("case", ["dipole_c1", "grid_1024"])
def test_fill_matches_oracle(wm, geometries, oracle, ctx, geo_dir, case):
    if case == "dipole_c1":
        p = _dipole(geometries, geo_dir, 512, 1.0, 0.001)
        freqs = [150.0]
    else:
        p = os.path.join(geo_dir, "g1024.geo")
        geometries.planar_grid(p, 1024, seed=7)
        freqs = [1.0, 37.5, 100.0]
    t = oracle.load(p)
    g = wm.Geometry.from_topology(ctx, wm.configwire(p, cwd=geo_dir))
    Zg = g.fill(freqs)
    for i, f in enumerate(freqs):
        Zo = oracle.zmatrix(t, oracle.wavenumber(f))
        err = np.abs(Zg[i] - Zo).max() / np.abs(Zo).max()
        print(case, f, err)
        assert err <= 1e-12, (case, f, err)


def _pec_case(G, d, case):
    """Structures on and above the PEC ground plane (ienv 2)."""
    p = os.path.join(d, f"pec_{case}.geo")
    if case == "monopole":
        G.monopole(p, 200, 1.0, 0.001)
    elif case == "lightning":  # ring lying in the ground plane: shorted, zero columns (fill only)
        G.lightning(p, 1024, depth=0.0)
    elif case == "rods":  # down-conductors ending on the ground plane
        G.lightning(p, 1024, depth=0.0, ring=False)
    else:  # bent wire 1-2.5 m above the ground, no ground contact
        G.write_geo(p, [(0.0, 0.0, 1.0), (1.0, 0.0, 1.0), (1.0, 0.5, 2.0), (0.3, 0.8, 2.5)],
                    [(1, 2, 0.002, 40), (2, 3, 0.002, 45), (3, 4, 0.0015, 35)])
    return p


@pytest.mark.parametrize("case", ["monopole", "lightning", "rods", "raised"])
def test_pec_fill_matches_oracle(wm, geometries, oracle, ctx, geo_dir, case):
    """PEC ground (ienv 2): fill with the image terms, the self rule on in-plane
    segments and the virtual ground wires, against the oracle.  Five
    frequencies in one request: every frequency slot of the image kernel's
    launches (four frequencies per launch, then one) is compared."""
    p = _pec_case(geometries, geo_dir, case)
    t = oracle.load(p, ienv=2)
    g = wm.Geometry.load(ctx, p, ienv=2)
    assert g.npls == t.npls
    freqs = [40.0, 25.0, 10.0, 2.0, 0.5]
    Zg = g.fill(freqs)
    for i, f in enumerate(freqs):
        Zo = oracle.zmatrix(t, oracle.wavenumber(f), ienv=2)
        err = np.abs(Zg[i] - Zo).max() / np.abs(Zo).max()
        print(case, f, err)
        assert err <= 1e-12, (case, f, err)


@pytest.mark.parametrize("ienv", [1, 2])
def test_fill_slots_bit_identical(wm, geometries, ctx, geo_dir, ienv):
    """An entry of Z does not depend on which slot of a launch its frequency
    takes or how many frequencies share the launch: seven frequencies filled
    together (launches of four and three) equal each frequency filled alone,
    bit for bit, with and without the PEC image terms."""
    p = _pec_case(geometries, geo_dir, "rods")
    g = wm.Geometry.load(ctx, p, ienv=ienv)
    freqs = [40.0, 25.0, 10.0, 2.0, 0.5, 7.5, 0.02]
    Zall = g.fill(freqs)
    for i, f in enumerate(freqs):
        Z1 = g.fill([f])[0]
        a, b = np.ascontiguousarray(Zall[i]), np.ascontiguousarray(Z1)
        assert np.array_equal(a.view(np.uint64), b.view(np.uint64)), (ienv, i, f)


def test_pec_zin_matches_oracle_f64(wm, geometries, oracle, ctx, geo_dir):
    """FP64 GPU path on the PEC ground vs the oracle: Zin within 1e-10; the
    monopole's Zin is half the free-space dipole's (image theory)."""
    for case, gap_at, freqs in [("monopole", (0, 0, 0), [50.0, 75.0, 120.0]),
                                ("rods", (0, 0, 20), [0.5, 3.0, 25.0])]:
        p = _pec_case(geometries, geo_dir, case)
        t = oracle.load(p, ienv=2)
        g = wm.Geometry.load(ctx, p, ienv=2)
        gap = geometries.nearest_basis(t, gap_at)
        zin, I, _ = wm.solve_sweep(g, freqs, [gap], "f64")
        for i, f in enumerate(freqs):
            zo, Io, _ = oracle.solve_freq(t, f, gaps=(gap,), ienv=2)
            rel = abs(zin[i, 0] - zo[0]) / abs(zo[0])
            print(case, f, zin[i, 0], rel)
            assert rel <= 1e-10, (case, f, rel)
            assert np.abs(I[i, 0] - Io[:, 0]).max() <= 1e-9 * np.abs(Io).max()
    pd = os.path.join(geo_dir, "pec_dipole.geo")
    geometries.write_geo(pd, [(0.0, 0.0, -1.0), (0.0, 0.0, 1.0)], [(1, 2, 0.001, 400)])
    gd = wm.Geometry.load(ctx, pd)
    zd, _, _ = wm.solve_sweep(gd, [50.0, 75.0, 120.0], [200], "f64")
    gm = wm.Geometry.load(ctx, _pec_case(geometries, geo_dir, "monopole"), ienv=2)
    zm, _, _ = wm.solve_sweep(gm, [50.0, 75.0, 120.0], [1], "f64")
    assert (np.abs(zm[:, 0] - zd[:, 0] / 2) <= 2e-7 * np.abs(zd[:, 0])).all(), (zm, zd)


@pytest.mark.parametrize("prec", ["f32", "bf16x3"])
def test_pec_mixed_precision_matches_f64(wm, geometries, ctx, geo_dir, prec):
    """Complex64 factors + FP64 refinement on a PEC-ground structure reach the
    FP64 path's Zin within 1e-10 (or route to FP64 when they cannot)."""
    p = _pec_case(geometries, geo_dir, "rods")
    g = wm.Geometry.load(ctx, p, ienv=2)
    f = [2.0, 10.0, 25.0, 40.0]
    z64, _, _ = wm.solve_sweep(g, f, [1, 500], "f64")
    zm, _, st = wm.solve_sweep(g, f, [1, 500], prec, size_policy=False)
    print(st)
    rel = np.abs(zm - z64).max(axis=1) / np.abs(z64).max(axis=1)
    print(rel)
    assert (rel <= 1e-10).all(), rel


def test_pec_dropin_zmatrix(wm, geometries, oracle, ctx, geo_dir):
    """The drop-in zmatrix with the reference's PEC arguments (ienv 2, signs
    1, 1, -1, -1) and CalculateByFreq on the PEC ground."""
    p = _pec_case(geometries, geo_dir, "monopole")
    t = wm.configwire(p, ienv=2, cwd=geo_dir)
    to = oracle.load(p, ienv=2)
    ka = wm.wavenumber(75.0)
    Z = wm.zmatrix(t, ka, ienv=2)
    Zo = oracle.zmatrix(to, oracle.wavenumber(75.0), ienv=2)
    assert np.abs(Z - Zo).max() <= 1e-12 * np.abs(Zo).max()
    zin, _, _ = wm.CalculateByFreq(t, 75.0, 1, ienv=2)
    zo, _, _ = oracle.solve_freq(to, 75.0, gaps=(1,), ienv=2)
    assert abs(zin - zo[0]) <= 1e-10 * abs(zo[0])


def test_fill_f32_copy_is_rounded_f64(wm, geometries, ctx, geo_dir):
    p = _dipole(geometries, geo_dir, 512, 1.0, 0.001)
    g = wm.Geometry.load(ctx, p)
    Z64, Z32 = g.fill([150.0, 300.0], f32_copy=True)
    assert np.array_equal(Z32, Z64.astype(np.complex64))


@pytest.mark.parametrize("n", [64, 128, 320, 576])
def test_getrf_f64_matches_lapack(wm, ctx, oracle, n):
    rng = np.random.RandomState(n)
    A = rng.randn(2, n, n) + 1j * rng.randn(2, n, n)
    dev, piv, info = wm.getrf_batched(ctx, A, "f64")
    LU = dev.download().transpose(0, 2, 1)
    P = piv.download()
    assert list(info) == [0, 0]
    for b in range(2):
        LUo, pivo, _ = oracle.zgetrf(A[b])
        assert np.array_equal(P[b], pivo), b
        assert np.abs(LU[b] - LUo).max() <= 1e-11 * np.abs(LUo).max()
    B = rng.randn(2, n, 3) + 1j * rng.randn(2, n, 3)
    X = wm.getrs_batched(ctx, dev, piv, B, "f64")
    for b in range(2):
        r = np.abs(A[b] @ X[b] - B[b]).max() / (np.abs(A[b]).max() * np.abs(X[b]).max() * n)
        assert r <= 1e-15, r


@pytest.mark.parametrize("prec,bound", [("f64", 1e-14), ("f32", 2e-6), ("bf16x3", 6e-5)])
def test_trailing_update_step(wm, ctx, prec, bound):
    """One trailing update A22 -= L21 U12 on the matrix cores against numpy, with
    edge tiles (n = 448, k = 64: a 320 x 320 update)."""
    L = wm.lib()
    L.wmx_gemm_step.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_int]
    n, k = 448, 64
    rng = np.random.RandomState(11)
    dt = np.complex128 if prec == "f64" else np.complex64
    A = (rng.randn(n, n) + 1j * rng.randn(n, n)).astype(dt)
    d = wm.DeviceArray(ctx, np.ascontiguousarray(A.T))  # column-major
    assert L.wmx_gemm_step(ctx.h, wm.PREC_NAMES[prec], n, d.ptr, k) == 0
    got = d.download().T.astype(np.complex128)
    A64 = A.astype(np.complex128)
    r = k + 64
    L21, U12 = A64[r:, k:k + 64], A64[k:k + 64, r:]
    want = A64.copy()
    want[r:, r:] -= L21 @ U12
    scale = (np.abs(L21) @ np.abs(U12)).max()
    err = np.abs(got - want).max() / scale
    print(prec, err)
    assert err <= bound, err
    assert np.array_equal(got[:r], A64[:r]) and np.array_equal(got[:, :r], A64[:, :r])


@pytest.mark.parametrize("panels", [1, 2, 8, 16])
def test_getrf_outer_block_depths(wm, ctx, oracle, panels):
    """Every outer-block depth (panels per trailing update) gives LAPACK's pivots."""
    n = 576
    rng = np.random.RandomState(panels)
    A = rng.randn(2, n, n) + 1j * rng.randn(2, n, n)
    ctx.set_getrf_block("f64", n, panels)
    try:
        assert ctx.getrf_block("f64", n) == panels
        dev, piv, info = wm.getrf_batched(ctx, A, "f64")
    finally:
        ctx.set_getrf_block("f64", n, 0)
    LU, P = dev.download().transpose(0, 2, 1), piv.download()
    for b in range(2):
        LUo, pivo, _ = oracle.zgetrf(A[b])
        assert np.array_equal(P[b], pivo), (panels, b)
        assert np.abs(LU[b] - LUo).max() <= 1e-11 * np.abs(LUo).max()


def test_tune_getrf(wm, ctx):
    """The tuner times every depth on synthetic matrices and keeps the fastest."""
    try:
        best, ms = ctx.tune_getrf("f32", 1024)
        print("tuned", best, ms)
        assert best in (1, 2, 4, 8, 16) and ms > 0
        assert ctx.getrf_block("f32", 1024) == best
    finally:
        ctx.set_getrf_block("f32", 1024, 0)


def test_tune_fill_and_cache(wm, geometries, ctx, geo_dir, tmp_path):
    """The fill tuner times 1-4 frequencies per launch and keeps the fastest
    (the matrices are the same bit for bit whatever it picks); the settings
    survive a JSON cache written by one context and loaded by another, entries
    of another architecture are skipped."""
    p = os.path.join(geo_dir, "tune_grid.geo")
    geometries.planar_grid(p, 2000, seed=5)
    g = wm.Geometry.load(ctx, p)
    freqs = [30.0, 12.0, 4.0, 1.0, 0.3]
    Z0 = g.fill(freqs)
    c2 = wm.Context(0)
    try:
        best, ms = ctx.tune_fill(g)
        print("fill tuned", best, ms)
        assert best in (1, 2, 3, 4) and ms > 0 and ctx.fill_freqs(1, g.ld) == best
        for f in (1, 3):
            ctx.set_fill_freqs(1, g.ld, f)
            Z1 = g.fill(freqs)
            assert np.array_equal(np.ascontiguousarray(Z1).view(np.uint64), np.ascontiguousarray(Z0).view(np.uint64))
        ctx.set_fill_freqs(1, g.ld, best)
        ctx.set_getrf_block("bf16x3", 8192, 16)
        path = str(tmp_path / "tune.json")
        assert ctx.tune_save(path) == 2
        with open(path, "a") as fh:  # another architecture's entry is not applied
            fh.write('{"arch": "gfx000", "knob": "getrf_block", "prec": 0, "size_class": 8, "value": 1}\n')
        assert c2.tune_load(path) == 2
        assert c2.fill_freqs(1, g.ld) == best and c2.getrf_block("bf16x3", 8192) == 16
        assert c2.getrf_block("f64", 8192) == 8  # default kept
    finally:
        ctx.set_fill_freqs(1, g.ld, 0)
        ctx.set_getrf_block("bf16x3", 8192, 0)
        c2.close()


@pytest.mark.parametrize("prec", ["f64", "bf16x3"])
def test_getrf_graph_replays_match_direct(wm, ctx, prec):
    """A captured factorization replayed on the same buffers (with eager copies
    and synchronizations between replays) gives the direct launches' result
    every time, bit for bit."""
    n, b = 448, 3
    rng = np.random.RandomState(17)
    dt = np.complex128 if prec == "f64" else np.complex64
    A = (rng.randn(b, n, n) + 1j * rng.randn(b, n, n)).astype(dt)
    colmajor = np.ascontiguousarray(A.transpose(0, 2, 1))
    dev = wm.DeviceArray(ctx, colmajor)
    piv = wm.DeviceArray(ctx, np.zeros((b, n), np.int32))
    info = (C.c_int * b)()
    p = wm.PREC_NAMES[prec]

    def factor():
        dev.upload(colmajor)
        assert wm.lib().wm_getrf_batched(ctx.h, p, n, dev.ptr, n, n * n, piv.ptr, n, b, info) == 0
        ctx.sync()
        return dev.download(), piv.download()

    ctx.set_graphs(False)
    ref_lu, ref_piv = factor()
    ctx.set_graphs(True)
    try:
        for _ in range(3):  # the first call captures, the others replay
            lu, pv = factor()
            assert np.array_equal(pv, ref_piv)
            assert np.array_equal(lu, ref_lu)
    finally:
        ctx.set_graphs(True)


@pytest.mark.parametrize("prec", ["f64", "f32", "bf16x3"])
@pytest.mark.parametrize("indexed", [False, True])
def test_getrf_split_panel_launches_match(wm, ctx, prec, indexed):
    """Members of a batch step through the panels together, each panel in as
    many launches as keep the clusters co-resident at its height.  With the
    co-resident count capped (tall panels in 3 launches, short ones in 1) the
    factors and pivots equal those of one launch per panel, bit for bit, also
    when members are picked through an index table."""
    n, b = 576, 5
    rng = np.random.RandomState(23)
    dt = np.complex128 if prec == "f64" else np.complex64
    A = (rng.randn(b, n, n) + 1j * rng.randn(b, n, n)).astype(dt)
    colmajor = np.ascontiguousarray(A.transpose(0, 2, 1))
    dev = wm.DeviceArray(ctx, colmajor)
    piv = wm.DeviceArray(ctx, np.zeros((b, n), np.int32))
    info = (C.c_int * b)()
    idx = (C.c_int * 3)(4, 1, 3)
    L = wm.lib()
    L.wmx_set_panel_resident.argtypes = [C.c_void_p, C.c_int]
    L.wmx_getrf_batched_idx.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_int, C.c_int64, C.c_void_p,
                                        C.c_int64, C.c_int, C.c_void_p, C.c_void_p]
    p = wm.PREC_NAMES[prec]

    def factor():
        dev.upload(colmajor)
        piv.upload(np.zeros((b, n), np.int32))
        if indexed:
            rc = L.wmx_getrf_batched_idx(ctx.h, p, n, dev.ptr, n, n * n, piv.ptr, n, len(idx), info, idx)
        else:
            rc = L.wm_getrf_batched(ctx.h, p, n, dev.ptr, n, n * n, piv.ptr, n, b, info)
        assert rc == 0
        ctx.sync()
        return dev.download(), piv.download()

    ref_lu, ref_piv = factor()
    # complex128 panels hold 128 rows per workgroup, complex64 256: the first
    # panels need 5 (3) workgroups per member, the last 1
    cap = 12 if prec == "f64" else 6
    assert L.wmx_set_panel_resident(ctx.h, cap) == 0
    try:
        lu, pv = factor()
    finally:
        assert L.wmx_set_panel_resident(ctx.h, 0) == 0
    assert np.array_equal(pv, ref_piv)
    assert np.array_equal(lu, ref_lu)
    if indexed:  # members outside the table are untouched
        for m in (0, 2):
            assert np.array_equal(lu[m], colmajor[m])


def test_getrf_f64_zero_pivot_info(wm, ctx):
    A = np.eye(128, dtype=complex)[None].repeat(2, axis=0)
    A[1, :, 70] = 0
    _, _, info = wm.getrf_batched(ctx, A, "f64")
    assert list(info) == [0, 71]


@pytest.mark.parametrize("prec,bound", [("f32", 4e-5), ("bf16x3", 1e-3)])
def test_getrf_f32_backward_error(wm, ctx, prec, bound):
    """Complex64 factors: backward error of P A = L U at the working precision
    (FP32 updates, or split-BF16x3 updates with about 16-bit products).  The FP32
    updates take three real products per complex product, whose imaginary part
    sums (Ar + Ai)(Br + Bi): in a numpy emulation of the same blocked LU this
    about doubles the backward error of four products (7e-6 -> 1.0-1.4e-5 here,
    n u = 2.3e-5)."""
    n = 384
    rng = np.random.RandomState(5)
    A = (rng.randn(3, n, n) + 1j * rng.randn(3, n, n)).astype(np.complex64)
    dev, piv, info = wm.getrf_batched(ctx, A, prec)
    LU = dev.download().transpose(0, 2, 1).astype(np.complex128)
    P = piv.download()
    for b in range(3):
        L = np.tril(LU[b], -1) + np.eye(n)
        U = np.triu(LU[b])
        PA = A[b].astype(np.complex128).copy()
        for i in range(n):
            p = P[b, i] - 1
            PA[[i, p]] = PA[[p, i]]
        err = np.abs(PA - L @ U).max() / np.abs(A[b]).max()
        print(prec, b, err)
        assert err <= bound, err
        # pivoting on |re|+|im| (izamax) bounds the modulus of l_ij by sqrt(2)
        assert np.abs(L).max() <= np.sqrt(2) + 1e-6


def test_zin_matches_oracle_f64(wm, geometries, oracle, ctx, geo_dir):
    """FP64 GPU path vs oracle: Zin within 1e-10 (dipoles and a junction grid)."""
    for nseg, f in [(512, [100.0, 600.0, 1200.0]), (768, [300.0])]:
        p = _dipole(geometries, geo_dir, nseg)
        t = oracle.load(p)
        g = wm.Geometry.load(ctx, p)
        zin, I, st = wm.solve_sweep(g, f, [100], "f64")
        for i, fr in enumerate(f):
            zo, Io, _ = oracle.solve_freq(t, fr, gaps=(100,))
            rel = abs(zin[i, 0] - zo[0]) / abs(zo[0])
            print(nseg, fr, zin[i, 0], rel)
            assert rel <= 1e-10, (nseg, fr, rel)
            assert np.abs(I[i, 0] - Io[:, 0]).max() <= 1e-9 * np.abs(Io).max()
    p = os.path.join(geo_dir, "g1024.geo")
    geometries.planar_grid(p, 1024, seed=7)
    t = oracle.load(p)
    g = wm.Geometry.load(ctx, p)
    gap = geometries.nearest_basis(t, (5, 5, 0))
    zin, _, _ = wm.solve_sweep(g, [10.0, 55.0], [gap, 17], "f64")
    for i, fr in enumerate([10.0, 55.0]):
        zo, _, _ = oracle.solve_freq(t, fr, gaps=(gap, 17))
        rel = np.abs(zin[i] - zo).max() / np.abs(zo).max()
        assert rel <= 1e-10, (fr, rel)


@pytest.mark.parametrize("band", [(0, 4096), (4097, 20480)])
def test_zin_known_answers_gpu(wm, geometries, ctx, geo_dir, band):
    """The FP64 GPU path reproduces every logged Zin (4 s.f.) of the reference
    for the 25 m dipole: all 308 cases, N = 512 ... 20480 segments (180 up to
    4096, 128 above), e.g. "wersja referencyjna/results_1_6_8192.txt":44
    (100 MHz: 96.37 -j462.6) and results_1_2_20480.txt:44 (98.67 -j1019)."""
    cases = json.load(open(os.path.join(GOLD, "zin_kat.json")))["cases"]
    by_n = {}
    for c in cases:
        if band[0] <= c["nseg"] <= band[1]:
            by_n.setdefault(c["nseg"], []).append(c)
    assert sum(len(v) for v in by_n.values()) == (180 if band[0] == 0 else 128)
    bad = []
    for nseg, cs in sorted(by_n.items()):
        g = wm.Geometry.load(ctx, _dipole(geometries, geo_dir, nseg))
        zin, _, _ = wm.solve_sweep(g, [float(c["freq_mhz"]) for c in cs], [100], "f64")
        for c, z in zip(cs, zin[:, 0]):
            if ("%.4g" % z.real, "%.4g" % z.imag) != (c["re"], c["im"]):
                bad.append((nseg, c["freq_mhz"], z, c["re"], c["im"]))
        print(nseg, len(cs), "ok" if not bad else bad[-1])
    assert not bad, bad


@pytest.mark.parametrize("prec", ["f32", "bf16x3"])
def test_mixed_precision_matches_f64(wm, geometries, ctx, geo_dir, prec):
    """Complex64 LU (FP32 or split-BF16x3 updates) + FP64 refinement reaches the
    FP64 path's Zin within 1e-10."""
    p = os.path.join(geo_dir, "g1024.geo")
    geometries.planar_grid(p, 1024, seed=7)
    g = wm.Geometry.load(ctx, p)
    f = [1.0, 20.0, 60.0, 100.0]
    z64, I64, _ = wm.solve_sweep(g, f, [5, 700], "f64")
    z32, I32, st = wm.solve_sweep(g, f, [5, 700], prec, size_policy=False)
    print(st)
    rel = np.abs(z32 - z64).max(axis=1) / np.abs(z64).max(axis=1)
    print(rel)
    assert (rel <= 1e-10).all(), rel
    assert all(s["iters"] >= 1 for s in st)
    relI = np.abs(I32 - I64).max(axis=2) / np.abs(I64).max(axis=2)
    assert (relI <= 1e-9).all(), relI


def test_dropin_zmatrix_and_lu(wm, geometries, oracle, ctx, geo_dir):
    p = _dipole(geometries, geo_dir, 512, 1.0, 0.001)
    t = wm.configwire(p, cwd=geo_dir)
    ka = wm.wavenumber(150.0)
    Z = wm.zmatrix(t, ka)
    Zo = oracle.zmatrix(oracle.load(p), oracle.wavenumber(150.0))
    assert np.abs(Z - Zo).max() <= 1e-12 * np.abs(Zo).max()
    b = np.zeros(t.npls, complex)
    b[255] = 1.0
    x = wm.LU_cuBLAS(Z, b)
    xo = np.linalg.solve(Zo, b)
    assert np.abs(x - xo).max() <= 1e-10 * np.abs(xo).max()
    zin, I, Zc = wm.CalculateByFreq(t, 150.0, 256)
    zo, _, _ = oracle.solve_freq(oracle.load(p), 150.0, gaps=(256,))
    assert abs(zin - zo[0]) <= 1e-10 * abs(zo[0])


def test_server_result_same_alone_and_in_mix(wm, geometries, ctx, geo_dir):
    """A request gives bit-identical results solved alone and batched with others,
    including another request on the same structure whose frequency joins its
    fill launch (two matrices per launch alone, three in the mix)."""
    p1 = os.path.join(geo_dir, "s1.geo")
    p2 = os.path.join(geo_dir, "s2.geo")
    geometries.planar_grid(p1, 512, seed=21)
    geometries.planar_grid(p2, 500, seed=22)  # same padded size, different structure
    g1, g2 = wm.Geometry.load(ctx, p1), wm.Geometry.load(ctx, p2)
    s = wm.Server(ctx, "f32", size_policy=False)  # keep the mixed path at this size
    rid = s.submit(g1, [3.0, 40.0], [7, 100])
    while s.step() == 0:
        pass
    alone = s.result(rid)
    ids = [s.submit(g2, [5.0, 9.0, 11.0], [3]), s.submit(g1, [3.0, 40.0], [7, 100]), s.submit(g2, [1.0], [44, 45, 46]),
           s.submit(g1, [5.0], [7])]
    got = {}
    while len(got) < len(ids):
        s.step()
        for i in ids:
            if i not in got:
                try:
                    got[i] = s.result(i)
                except wm.WireMoMError:
                    pass
    mix = got[ids[1]]
    assert np.array_equal(alone[0], mix[0])
    assert np.array_equal(alone[1], mix[1])
    s.close()


def test_getrs_alternating_layouts(wm, ctx):
    """Regression: solves of different (n, batch, nrhs) back to back must not see
    each other's completion flags."""
    rng = np.random.RandomState(9)
    for n, b, nr in [(320, 2, 3), (512, 3, 1), (128, 1, 8), (512, 3, 1), (320, 2, 3), (64, 4, 2)]:
        A = rng.randn(b, n, n) + 1j * rng.randn(b, n, n)
        dev, piv, info = wm.getrf_batched(ctx, A, "f64")
        B = rng.randn(b, n, nr) + 1j * rng.randn(b, n, nr)
        X = wm.getrs_batched(ctx, dev, piv, B, "f64")
        for i in range(b):
            r = np.abs(A[i] @ X[i] - B[i]).max() / (np.abs(A[i]).max() * np.abs(X[i]).max() * n)
            assert r <= 1e-15, (n, b, nr, r)


def test_mixed_precision_fallback_band(wm, geometries, ctx, geo_dir):
    """Across a 10 kHz - 10 MHz band the low frequencies are too ill-conditioned
    for an FP32 factorization; those items are refactored in FP64 inside the same
    sweep, and every Zin still matches the FP64 path within 1e-10, or within
    kappa_1 u64 where the matrix's condition makes two FP64-accurate solutions
    differ by more (kappa_1 from the FP64 sweep's estimate)."""
    p = os.path.join(geo_dir, "light1024.geo")
    geometries.lightning(p, 1024)
    g = wm.Geometry.load(ctx, p)
    t = wm.configwire(p, cwd=geo_dir)
    gap = geometries.nearest_basis(t, (0, 0, 19.9))
    f = np.logspace(-2, 1, 12).astype(np.float32)
    z64, _, s64 = wm.solve_sweep(g, f, [gap], "f64", cond_est=True)
    bound = np.array([max(1e-10, s["cond1"] * 2.0 ** -53) for s in s64])
    z32, _, st = wm.solve_sweep(g, f, [gap], "f32", size_policy=False)
    rel = np.abs(z32[:, 0] - z64[:, 0]) / np.abs(z64[:, 0])
    print([(float(fr), s["iters"], s["fell_back"], float(r), float(b)) for fr, s, r, b in zip(f, st, rel, bound)])
    assert (rel <= bound).all(), (rel, bound)
    assert any(s["fell_back"] for s in st)
    # batches of 4: the 4 highest frequencies are refined first and their
    # contraction rates route the lowest ones to FP64 factors directly
    z32b, _, stb = wm.solve_sweep(g, f, [gap], "f32", batch=4, size_policy=False)
    relb = np.abs(z32b[:, 0] - z64[:, 0]) / np.abs(z64[:, 0])
    print([(float(fr), s["iters"], s["fell_back"], float(r)) for fr, s, r in zip(f, stb, relb)])
    assert (relb <= bound).all(), (relb, bound)
    routed = [s["fell_back"] == 2 for s in stb]
    assert routed[0] and not any(routed[-4:])  # lowest routed, highest refined
    for a, b in zip(st, stb):  # routing only skips refinements that fail or take most of max_iter
        if b["fell_back"] == 2:
            assert a["fell_back"] == 1 or a["iters"] >= 12, (a, b)
    # precision ladder: split-BF16x3 factors where they refine, then FP32, then FP64
    zb, _, stl = wm.solve_sweep(g, f, [gap], "bf16x3", batch=4, size_policy=False)
    rell = np.abs(zb[:, 0] - z64[:, 0]) / np.abs(z64[:, 0])
    print([(float(fr), s["iters"], s["prec"], s["fell_back"], float(r)) for fr, s, r in zip(f, stl, rell)])
    assert (rell <= bound).all(), (rell, bound)
    assert stl[-1]["prec"] == wm.PREC_BF16X3 and stl[0]["prec"] == wm.PREC_F64


def test_serving_replay_matches_sweeps(wm, geometries, ctx, geo_dir):
    """A Poisson stream on the server: every request's Zin equals the same
    precision's sweep of its structure (to 1e-12: each item's arithmetic is its
    own), is within the refinement's own error estimate of the FP64 sweep, and
    each latency is positive."""
    import wiremom.serving as S
    reqs = S.poisson_requests(6, 40.0, seed=3, sizes=(512, 1024), fmin=1.0, fmax=50.0, max_freqs=4, max_rhs=2)
    geoms, feeds = [], []
    for i, r in enumerate(reqs):
        p = os.path.join(geo_dir, f"srv{i}.geo")
        geometries.planar_grid(p, r["n"], seed=r["seed"])
        geoms.append(wm.Geometry.load(ctx, p))
        feeds.append(np.random.RandomState(r["seed"]).choice(geoms[-1].npls, r["nrhs"], replace=False) + 1)
    srv = wm.Server(ctx, "f32", size_policy=False)  # keep the mixed path at these sizes
    out = S.replay(srv, reqs, geoms, feeds)
    srv.close()
    assert all(t > 0 for t in out["latency"])
    for r, g, fd, (zin, I, st) in zip(reqs, geoms, feeds, out["results"]):
        z32, _, _ = wm.solve_sweep(g, r["freqs"], fd, "f32", size_policy=False)
        rel32 = np.abs(zin - z32) / np.abs(z32)
        assert rel32.max() <= 1e-12, (r["n"], rel32.max())
        z64, _, _ = wm.solve_sweep(g, r["freqs"], fd, "f64")
        rel = (np.abs(zin - z64) / np.abs(z64)).max(axis=1)
        # two FP64-accurate solutions differ by about cond(Z) * 1e-16; the
        # refinement's last correction bounds that from the mixed side
        for f, e, s_ in zip(r["freqs"], rel, st):
            bound = max(1e-10, 1e3 * s_["dx_rel"])
            print(r["n"], float(f), s_["iters"], s_["fell_back"], s_["dx_rel"], float(e))
            assert e <= bound, (r["n"], float(f), float(e), s_)


def test_large_cage_multi_rhs(wm, geometries, ctx, geo_dir):
    """The large single-GPU case: a 3-D cage at N = 14336 (panel clusters of 112
    workgroups, 4 matrices per launch) with 8 feeds; mixed precision (FP32 and
    split-BF16x3) matches the FP64 path, Zin within 1e-10 and currents within 1e-9."""
    p = os.path.join(geo_dir, "cage.geo")
    geometries.cage(p, 14336, seed=4)
    g = wm.Geometry.load(ctx, p)
    t = wm.configwire(p, cwd=geo_dir)
    assert g.npls == 14336
    gaps = list(np.random.RandomState(4).choice(t.npls, 8, replace=False) + 1)
    f = [10.0, 25.0, 40.0, 50.0]
    z64, I64, _ = wm.solve_sweep(g, f, gaps, "f64")
    for prec in ("f32", "bf16x3"):
        z, I, st = wm.solve_sweep(g, f, gaps, prec)
        rel = np.abs(z - z64).max() / np.abs(z64).max()
        relI = (np.abs(I - I64).max(axis=2) / np.abs(I64).max(axis=2)).max()
        print(prec, [(s["prec"], s["iters"]) for s in st], rel, relI)
        assert rel <= 1e-10 and relI <= 1e-9, (prec, rel, relI)
        assert all(s["prec"] != wm.PREC_F64 for s in st)


def test_large_cage_full_band(wm, geometries, ctx, geo_dir):
    """The large case over its whole band: the N = 14336 cage at 16 frequencies,
    0.1 - 50 MHz linearly spaced, with 8 feeds.  The mixed sweeps route the
    ill-conditioned low end to FP64 factors and refine the rest; every Zin of
    every feed is within 1e-10 of the FP64 sweep's, element by element."""
    p = os.path.join(geo_dir, "cage.geo")
    geometries.cage(p, 14336, seed=4)
    g = wm.Geometry.load(ctx, p)
    t = wm.configwire(p, cwd=geo_dir)
    gaps = list(np.random.RandomState(4).choice(t.npls, 8, replace=False) + 1)
    f = np.linspace(0.1, 50.0, 16).astype(np.float32)
    z64, _, _ = wm.solve_sweep(g, f, gaps, "f64")
    for prec in ("f32", "bf16x3"):
        z, _, st = wm.solve_sweep(g, f, gaps, prec)
        rel = np.abs(z - z64) / np.abs(z64)
        print(prec, [(float(fr), s["prec"], s["iters"], s["fell_back"]) for fr, s in zip(f, st)], rel.max())
        assert z.shape == (16, 8) and (rel <= 1e-10).all(), (prec, rel.max(axis=1))
        assert st[-1]["prec"] != wm.PREC_F64  # the top of the band refines


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_sweep_with_generators_and_loads_matches_oracle(wm, geometries, oracle, ctx, geo_dir, prec):
    """wm_solve_sweep_ex: three generators over two right-hand sides (two in
    one), loads with and without a capacitor (two on one basis, summed).  FP64
    currents and V/I at each generator within 1e-10 of the oracle; FP32
    factors with FP64 refinement within 1e-10 of FP64."""
    p = os.path.join(geo_dir, "gl1024.geo")
    geometries.planar_grid(p, 1024, seed=7)
    t = oracle.load(p)
    g = wm.Geometry.from_topology(ctx, wm.configwire(p, cwd=geo_dir))
    freqs = [5.0, 40.0, 90.0]
    sources = [(100, 0, 1.0, 0.0), (600, 0, 0.5, 30.0), (901, 1, 2.0, -60.0)]
    loads = [(250, 50.0, 0.0, 0.0), (250, 0.0, 0.1, 0.0), (777, 10.0, 0.2, 1e-4), (901, 5.0, 0.0, 0.0)]
    z64, I64, _ = wm.solve_sweep_ex(g, freqs, 2, sources, loads, "f64")
    if prec == "f64":
        for i, f in enumerate(freqs):
            zo, Xo = oracle.solve_freq_ex(t, f, 2, sources, loads)
            e = np.abs(I64[i].T - Xo).max() / np.abs(Xo).max()
            ez = np.abs(z64[i] - zo) / np.abs(zo)
            print(f, e, ez.max())
            assert e <= 1e-10 and ez.max() <= 1e-10, (f, e, ez)
    else:
        zm, Im, st = wm.solve_sweep_ex(g, freqs, 2, sources, loads, prec, size_policy=False)
        ez = np.abs(zm - z64) / np.abs(z64)
        print([s["prec"] for s in st], ez.max())
        assert ez.max() <= 1e-10
    g.close()


def test_sweep_loads_reproduce_reference_load_block(wm, geometries, ctx, geo_dir):
    """The reference's hard-wired loads (iload == 1: 1 + j 2 pi f 4.5 on bases
    592..2581, 30 on 2613, 1e12 on 2582; "wersja referencyjna/Wire_MoM.cu":346-363)
    as a load list give the currents of the CalculateByFreq drop-in with
    iload = 1, which adds them to its host copy of Z."""
    p = _dipole(geometries, geo_dir, 2700, 25.0, 0.0005)
    t = wm.configwire(p, cwd=geo_dir)
    g = wm.Geometry.from_topology(ctx, t)
    npls, ngap = g.npls, 1350
    loads = [(b, 1.0, 4.5, 0.0) for b in range(592, 2582)] + [(2613, 30.0, 0.0, 0.0), (2582, 1e12, 0.0, 0.0)]
    f = 55.0
    zs, Is, _ = wm.solve_sweep_ex(g, [f], 1, [(ngap, 0, 1.0, 0.0)], loads, "f64")
    g.close()
    L = wm.lib()
    cz = np.zeros(npls * npls, np.complex128)
    cu = np.zeros(npls, np.complex128)
    ci = np.zeros(npls, np.complex128)
    fp = C.POINTER(C.c_float)
    ip = C.POINTER(C.c_int)
    L.CalculateByFreq.restype = None
    L.CalculateByFreq(0, npls, 1, C.c_float(f), t.x.ctypes.data_as(fp), t.y.ctypes.data_as(fp),
                      t.z.ctypes.data_as(fp), t.rad.ctypes.data_as(fp), t.ipc.ctypes.data_as(ip),
                      t.ips.ctypes.data_as(ip), t.ise.ctypes.data_as(ip), t.nsmax, 1, C.c_float(1.0), C.c_float(1.0),
                      C.c_float(1.0), C.c_float(1.0), 1, C.c_float(0.0), ngap, C.c_float(1.0),
                      cz.ctypes.data_as(C.c_void_p), cu.ctypes.data_as(C.c_void_p), ci.ctypes.data_as(C.c_void_p), 0)
    e = np.abs(Is[0, 0] - ci).max() / np.abs(ci).max()
    print(npls, e, zs[0, 0], cu[ngap - 1] / ci[ngap - 1])
    assert e <= 1e-12


def test_transient_matches_oracle(wm, ctx, oracle):
    """wm_transient against the oracle's transform on random transfers over a
    non-uniform grid (the piecewise-uniform steps of the reference's freq1.txt)."""
    f = np.concatenate([np.arange(1, 10) * 1e-3, np.arange(1, 10) * 1e-2, np.arange(1, 101) * 1e-1]).astype(np.float32)
    rng = np.random.RandomState(3)
    H = rng.randn(len(f), 5) + 1j * rng.randn(len(f), 5)
    S = wm.spectrum_dexp(f, 30e3, 1.4e4, 6e6)
    assert np.abs(S - oracle.spectrum_dexp(f, 30e3, 1.4e4, 6e6)).max() <= 1e-12 * np.abs(S).max()
    t0, dt, nt = 1e-7, 2e-8, 700
    r = wm.transient(ctx, f, H, S, t0, dt, nt)
    ro = oracle.transient(f, H, S, t0 + dt * np.arange(nt))
    err = np.abs(r - ro).max() / np.abs(ro).max()
    print(err)
    assert err <= 1e-11


def test_current_injection_transient_from_sweep(wm, geometries, oracle, ctx, geo_dir):
    """Currents of an FP64 sweep divided by the feed current (a current
    injection): at the feed the transfer is 1 and the transient is the source
    waveform's transform; at another basis it equals the oracle's transform of
    the same transfer."""
    p = _dipole(geometries, geo_dir, 256, 1.0, 0.001)
    g = wm.Geometry.load(ctx, p)
    f = (np.arange(1, 129) * 2.0).astype(np.float32)  # 2 - 256 MHz
    gap = 128
    _, I, _ = wm.solve_sweep(g, f, [gap], "f64")
    g.close()
    H = I[:, 0, :] / I[:, 0, gap - 1][:, None]
    S = wm.spectrum_dexp(f, 1.0, 2e7, 2e8)
    t0, dt, nt = 0.0, 1e-10, 256
    r = wm.transient(ctx, f, H[:, [gap - 1, 40]], S, t0, dt, nt)
    t = t0 + dt * np.arange(nt)
    r_src = oracle.transient(f, np.ones((len(f), 1)), S, t)[0]
    assert np.abs(r[0] - r_src).max() <= 1e-12 * np.abs(r_src).max()
    r40 = oracle.transient(f, H[:, 40], S, t)[0]
    assert np.abs(r[1] - r40).max() <= 1e-11 * np.abs(r40).max()


def test_planar_grid_sweep_c2(wm, geometries, oracle, ctx, geo_dir):
    """The 10 x 10 m planar grid at N = 4096 (seed 2) over 12 frequencies
    1 - 100 MHz, fed at the basis nearest the centre: FP64 Zin within 1e-10 of
    the oracle at one frequency, and FP32 / split-BF16x3 factors with FP64
    refinement within 1e-10 of the FP64 sweep at all twelve."""
    p = os.path.join(geo_dir, "c2.geo")
    geometries.planar_grid(p, 4096, seed=2)
    t = oracle.load(p)
    gap = geometries.nearest_basis(t, (5.0, 5.0, 0.0))
    g = wm.Geometry.load(ctx, p)
    assert g.npls == 4096
    f = geometries.freqs_c2()
    z64, _, _ = wm.solve_sweep(g, f, [gap], "f64")
    zo = oracle.solve_freq(t, float(f[4]), gaps=(gap,))[0][0]
    e = abs(z64[4, 0] - zo) / abs(zo)
    print("c2 oracle", f[4], z64[4, 0], zo, e)
    assert e <= 1e-10
    for prec in ("f32", "bf16x3"):
        z, _, st = wm.solve_sweep(g, f, [gap], prec)
        rel = (np.abs(z - z64) / np.abs(z64)).max()
        print(prec, [(s["prec"], s["iters"]) for s in st], rel)
        assert rel <= 1e-10
    g.close()


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_sweep_with_plane_waves_matches_oracle(wm, geometries, oracle, ctx, geo_dir, prec):
    """Plane-wave excitation (an oblique wave in one right-hand side; a generator
    plus a second wave in the other) on the planar grid: FP64 currents within
    1e-10 of the oracle; FP32 factors with FP64 refinement within 1e-10 of FP64."""
    p = os.path.join(geo_dir, "pw1024.geo")
    geometries.planar_grid(p, 1024, seed=7)
    t = oracle.load(p)
    g = wm.Geometry.from_topology(ctx, wm.configwire(p, cwd=geo_dir))
    freqs = [10.0, 55.0]
    waves = [(0, 35.0, 20.0, 1.0 + 0.2j, -0.4j), (1, 120.0, 250.0, 0.0, 2.0)]
    sources = [(300, 1, 1.0, 0.0)]
    loads = [(640, 25.0, 0.0, 0.0)]
    z64, I64, _ = wm.solve_sweep_ex(g, freqs, 2, sources, loads, "f64", plane_waves=waves)
    if prec == "f64":
        for i, f in enumerate(freqs):
            _, Xo = oracle.solve_freq_ex(t, f, 2, sources, loads, plane_waves=waves)
            e = np.abs(I64[i].T - Xo).max() / np.abs(Xo).max()
            print(f, e)
            assert e <= 1e-10, (f, e)
    else:
        _, Im, st = wm.solve_sweep_ex(g, freqs, 2, sources, loads, prec, size_policy=False, plane_waves=waves)
        e = (np.abs(Im - I64).max(axis=2) / np.abs(I64).max(axis=2)).max()
        print([s["prec"] for s in st], e)
        assert e <= 1e-9
    g.close()
