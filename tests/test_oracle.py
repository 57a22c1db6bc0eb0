"""Pins the CPU oracle (oracle/wm_oracle.c) against the reference's own data and
against what the mathematics fixes.  CPU only."""
import hashlib
import json
import os

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _dipole(G, d, nseg, length=25.0, radius=0.0005):
    p = os.path.join(d, f"dipole_{nseg}_{length}_{radius}.geo")
    G.dipole(p, nseg, length, radius)
    return p


def _fmt4(v):
    # the reference prints Zin with a sticky std::setprecision(4) (default float format)
    return "%.4g" % v


def test_configwire_matches_reference_files(oracle, geometries, geo_dir):
    """configwire.txt written by the oracle is byte-identical to the files the reference shipped."""
    kat = json.load(open(os.path.join(GOLD, "configwire_kat.json")))
    assert len(kat) == 4
    for k in kat:
        p = _dipole(geometries, geo_dir, k["nseg"], 25.0, 0.0005)
        out = os.path.join(geo_dir, f"cw_{k['nseg']}.txt")
        t = oracle.load(p, wirefile=out)
        data = open(out, "rb").read()
        assert t.npls == k["npls"]
        assert len(data) == k["bytes"]
        assert hashlib.sha256(data).hexdigest() == k["sha256"], k["variant"]


def test_zin_known_answers(oracle, geometries, geo_dir):
    """Zin of the 25 m dipole (feed basis 100) equals every logged 4-s.f. value for N <= 1024."""
    cases = json.load(open(os.path.join(GOLD, "zin_kat.json")))["cases"]
    assert len(cases) == 308
    small = [c for c in cases if c["nseg"] <= 1024]
    assert len(small) == 36
    topo = {}
    bad = []
    for c in small:
        if c["nseg"] not in topo:
            topo[c["nseg"]] = oracle.load(_dipole(geometries, geo_dir, c["nseg"]))
        zin, _, _ = oracle.solve_freq(topo[c["nseg"]], float(c["freq_mhz"]), gaps=(100,))
        z = zin[0]
        if (_fmt4(z.real), _fmt4(z.imag)) != (c["re"], c["im"]):
            bad.append((c, z))
    assert not bad, bad


def test_gauss_legendre_rules_are_textbook():
    """The reference's quadrature constants are the 2- and 4-point Gauss-Legendre rules on [0, b]."""
    x2, w2 = np.polynomial.legendre.leggauss(2)
    assert abs(x2[1] / 2 - 0.288675134594812882) < 1e-16
    assert abs(w2[0] / 2 - 0.5) < 1e-16
    x4, w4 = np.polynomial.legendre.leggauss(4)
    assert abs(x4[3] / 2 - 0.43056815579702629) < 1e-16
    assert abs(x4[2] / 2 - 0.16999052179242813) < 1e-16
    assert abs(w4[3] / 2 - 0.17392742256872693) < 1e-16
    assert abs(w4[2] / 2 - 0.32607257743127307) < 1e-16


def test_elliptic_k_polynomial_matches_ellipk(oracle):
    """delick is the A&S 17.3.34 polynomial: |K_poly - K| <= 2e-8 against scipy's exact K."""
    from scipy.special import ellipk
    for m in np.concatenate([np.linspace(0.0, 0.999, 200), 1 - np.logspace(-3, -8, 40)]):
        bet = np.sqrt(m)
        assert abs(oracle.delick(bet) - ellipk(m)) <= 2.2e-8, m


def test_self_integral_against_adaptive_quadrature(oracle):
    """The singularity-extracted self term equals the integral of
    w(s) [ (exp(-jk rho) - 1)/rho + beta K(beta)/(pi a) ] over the segment: the
    log term integrated by the 4-point rules and the analytic term cancel
    (Zmatrix.cpp:279-323, 695-728).  Brute force with scipy quad + exact K.

    What is left is the reference's own 4-point quadrature error on a smooth
    feature of width ~a, so the gap must shrink as dl/a falls: <= 1e-4 at
    dl/a <= 10 and monotone in dl/a.  A dropped analytic term or a sign error
    is off by O(1) or more at every aspect ratio."""
    from scipy.integrate import quad
    from scipy.special import ellipk

    def rel_err(ifun, ka, dl, a):
        def w(s):
            return s / dl if ifun == 1 else (dl - s) / dl if ifun == 2 else 1.0

        def f(s, part):
            r = 0.5 * dl - s
            rho = np.hypot(a, r)
            bet = 2 * a / np.sqrt(4 * a * a + r * r)
            v = (np.exp(-1j * ka * rho) - 1) / rho + bet * ellipk(bet * bet) / (np.pi * a)
            return w(s) * (v.real if part == 0 else v.imag)

        halves = [(0, dl / 2), (dl / 2, dl)]
        ref = complex(sum(quad(f, lo, hi, args=(0,), limit=400)[0] for lo, hi in halves),
                      sum(quad(f, lo, hi, args=(1,), limit=400)[0] for lo, hi in halves))
        got = oracle.self_integral(ifun, ka, dl, a)
        return abs(got - ref) / abs(ref)

    for ifun in (1, 2, 3):
        for ka, dl, a in [(0.3, 0.04, 0.004), (1.0, 0.02, 0.002), (1.0, 0.01, 0.002)]:
            assert rel_err(ifun, ka, dl, a) <= 1e-4, (ifun, ka, dl, a)
        errs = [rel_err(ifun, 2.0, 0.05, a) for a in (0.0005, 0.001, 0.0025, 0.005, 0.01)]
        assert all(e2 < e1 for e1, e2 in zip(errs, errs[1:])), errs
        assert errs[0] < 0.05, errs


def test_uniform_wire_is_toeplitz_and_symmetric(oracle, geometries, geo_dir):
    """Z of a straight uniform wire is symmetric Toeplitz up to float32 node rounding."""
    t = oracle.load(_dipole(geometries, geo_dir, 64, 1.0, 0.001))
    Z = oracle.zmatrix(t, oracle.wavenumber(150.0))
    scale = np.abs(Z).max()
    assert np.abs(Z - Z.T).max() <= 1e-5 * scale
    for d in range(0, 10):
        diag = np.diagonal(Z, d)
        assert np.abs(diag - diag[len(diag) // 2]).max() <= 1e-5 * scale


def test_half_wave_dipole_impedance(oracle, geometries, geo_dir):
    """A thin half-wave dipole has Zin near 73 + j42 ohm (thin-wire limit): physical sanity of the fill."""
    t = oracle.load(_dipole(geometries, geo_dir, 200, 1.0, 1e-4))
    f = 2.99792448e8 / 2.0 / 1e6  # length = lambda/2
    zin, _, _ = oracle.solve_freq(t, f, gaps=(100,))
    assert 60 < zin[0].real < 100 and 20 < zin[0].imag < 70, zin


def test_lu_matches_lapack(oracle):
    """zgetrf/zgetrs: same pivots as LAPACK (scipy), factors and solution to rounding."""
    from scipy.linalg import lapack
    rng = np.random.RandomState(0)
    for n in (1, 7, 64, 130):
        A = rng.randn(n, n) + 1j * rng.randn(n, n)
        LU, piv, info = oracle.zgetrf(A)
        lu2, piv2, info2 = lapack.zgetrf(A)
        assert info == info2 == 0
        assert np.array_equal(piv, piv2 + 1)
        assert np.abs(LU - lu2).max() <= 1e-10 * np.abs(lu2).max()
        b = rng.randn(n, 2) + 1j * rng.randn(n, 2)
        x = oracle.zgetrs(LU, piv, b)
        assert np.abs(A @ x - b).max() <= 1e-11 * np.abs(A).max() * np.abs(x).max() * n


def test_lu_zero_pivot_info(oracle):
    """An exactly zero column gives LAPACK-style info (1-based column)."""
    A = np.eye(5, dtype=complex)
    A[:, 2] = 0
    _, _, info = oracle.zgetrf(A)
    assert info == 3


def _two_wire_dipole(d, kind, nhalf=64, L=8.0, a=0.001):
    """Straight dipole along z built from two wires meeting at z=0, joined by
    connection `kind` (1: start-start, 2: start-end, 3: end-start, 4: end-end)."""
    import wiremom.geometries as G
    bot, mid, top = (0, 0, -L / 2), (0, 0, 0.0), (0, 0, L / 2)
    nodes = [bot, mid, top]
    w = {1: [(2, 3), (2, 1)], 2: [(1, 2), (2, 3)], 3: [(2, 3), (1, 2)], 4: [(1, 2), (3, 2)]}[kind]
    wires = [(a_, b_, a, nhalf) for a_, b_ in w]
    p = os.path.join(d, f"two_wire_{kind}.geo")
    G.write_geo(p, nodes, wires)
    return p


@pytest.mark.parametrize("kind", [1, 2, 3, 4])
def test_junction_connections_reduce_to_single_wire(oracle, geometries, geo_dir, kind, wm):
    """Each junction kind (Configuration.cpp connections 1-4) joining two collinear
    wires gives the same Zin and the same current magnitudes along the wire as
    one straight wire with the same segmentation: the junction basis must span
    the joint, with either orientation (Zin = V/I does not depend on it)."""
    L, nh = 8.0, 64  # node steps L/(2 nh) = 1/16 m are exact in float32
    single = oracle.load(_dipole(geometries, geo_dir, 2 * nh, L, 0.001))
    two = oracle.load(_two_wire_dipole(geo_dir, kind, nh, L, 0.001))
    assert two.npls == single.npls == 2 * nh - 1
    pk1 = single.z[np.abs(single.ipc[:single.npls]) - 1]
    pk2 = two.z[np.abs(two.ipc[:two.npls]) - 1]
    order1, order2 = np.argsort(pk1), np.argsort(pk2)
    assert np.array_equal(pk1[order1], pk2[order2])
    for zf in (-1.5, 0.0, 2.25):  # feed positions (peak-node z), 0.0 is the junction basis
        g1 = geometries.nearest_basis(single, (0, 0, zf))
        g2 = geometries.nearest_basis(two, (0, 0, zf))
        z1, I1, _ = oracle.solve_freq(single, 30.0, gaps=(g1,))
        z2, I2, _ = oracle.solve_freq(two, 30.0, gaps=(g2,))
        assert abs(z1[0] - z2[0]) <= 1e-9 * abs(z1[0]), (kind, zf, z1, z2)
        a1, a2 = np.abs(I1[order1, 0]), np.abs(I2[order2, 0])
        assert np.abs(a1 - a2).max() <= 1e-9 * a1.max(), (kind, zf)


def _raised_structure(d, G, mirrored=False):
    """Bent three-wire structure 1-2.5 m above z = 0 (every direction
    component non-zero somewhere); with mirrored=True the wires of its mirror
    image in z = 0 follow, with the same node order."""
    nodes = [(0.0, 0.0, 1.0), (1.0, 0.0, 1.0), (1.0, 0.5, 2.0), (0.3, 0.8, 2.5)]
    wires = [(1, 2, 0.002, 8), (2, 3, 0.002, 9), (3, 4, 0.0015, 7)]
    if mirrored:
        nodes = nodes + [(x, y, -z) for x, y, z in nodes]
        wires = wires + [(a + 4, b + 4, r, s) for a, b, r, s in wires]
    p = os.path.join(d, f"raised_{int(mirrored)}.geo")
    G.write_geo(p, nodes, wires)
    return p


def test_pec_ground_is_image_theory(oracle, geometries, geo_dir):
    """PEC ground (ienv = 2) by images: for a structure above the ground, the
    matrix equals the free-space matrix of the structure plus its mirror image
    with the image currents taking the mirrored basis functions with weight
    sgnenv = -1, i.e. Z_pec[m, n] = Zd[m, n] - Zd[m, n + N]."""
    t = oracle.load(_raised_structure(geo_dir, geometries), ienv=2)
    t2 = oracle.load(_raised_structure(geo_dir, geometries, True), ienv=1)
    N = t.npls
    assert t2.npls == 2 * N

    def peak(tt, n):
        i = abs(int(tt.ipc[n])) - 1
        return float(tt.x[i]), float(tt.y[i]), float(tt.z[i])

    for n in range(N):  # the doubled structure numbers the mirrored bases N..2N-1
        x, y, z = peak(t, n)
        assert peak(t2, n) == (x, y, z) and peak(t2, n + N) == (x, y, -z)
    for f in (30.0, 150.0):
        ka = oracle.wavenumber(f)
        Zp = oracle.zmatrix(t, ka, ienv=2)
        Zd = oracle.zmatrix(t2, ka, ienv=1)
        scale = np.abs(Zd).max()
        assert np.abs(Zp - (Zd[:N, :N] - Zd[:N, N:])).max() <= 1e-13 * scale
        assert np.abs(Zp - Zd[:N, :N]).max() >= 1e-5 * scale  # the images matter


def test_pec_wire_in_ground_plane_cancels(oracle, geometries, geo_dir):
    """A wire lying in the ground plane is cancelled by its image: every term,
    the self terms of segments in the plane included (Zmatrix.cpp:384), has an
    equal and opposite image, so Z vanishes to rounding."""
    p = os.path.join(geo_dir, "in_plane.geo")
    geometries.write_geo(p, [(0.0, 0.0, 0.0), (2.0, 0.5, 0.0)], [(1, 2, 0.001, 16)])
    t = oracle.load(p, ienv=2)
    assert t.npls == 17  # 15 wire bases + one per grounded end over the virtual ground wires
    ka = oracle.wavenumber(100.0)
    Zf = oracle.zmatrix(t, ka, ienv=1)
    Zp = oracle.zmatrix(t, ka, ienv=2)
    assert np.abs(Zp).max() <= 1e-14 * np.abs(Zf).max()


@pytest.mark.parametrize("nseg", [40, 80])
def test_pec_monopole_is_half_dipole(oracle, geometries, geo_dir, nseg):
    """Image theory for the whole path (front end with its virtual ground wire,
    fill, LU): a vertical monopole of height h on the PEC ground, fed at its
    base, has half the input impedance of the free-space dipole of length 2h
    fed at its centre."""
    h = 1.0
    pm, pd = os.path.join(geo_dir, f"mono_{nseg}.geo"), os.path.join(geo_dir, f"dip_{nseg}.geo")
    geometries.write_geo(pm, [(0.0, 0.0, 0.0), (0.0, 0.0, h)], [(1, 2, 0.001, nseg)])
    geometries.write_geo(pd, [(0.0, 0.0, -h), (0.0, 0.0, h)], [(1, 2, 0.001, 2 * nseg)])
    tm, td = oracle.load(pm, ienv=2), oracle.load(pd, ienv=1)
    gm, gd = geometries.nearest_basis(tm, (0, 0, 0)), geometries.nearest_basis(td, (0, 0, 0))
    for f in (50.0, 75.0, 120.0):
        zm, _, _ = oracle.solve_freq(tm, f, gaps=(gm,), ienv=2)
        zd, _, _ = oracle.solve_freq(td, f, gaps=(gd,))
        assert abs(zm[0] - zd[0] / 2) <= 2e-7 * abs(zd[0]), (nseg, f, zm, zd)


def test_load_impedance_units_match_reference_block(oracle):
    """The reference's channel load is Z_MAKE(1, 2*pi*freq*4.5) with freq the
    float32 MHz value ("wersja referencyjna/Wire_MoM.cu":349): 4.5 uH in the
    load's units, bit for bit; a capacitor adds -1/(2 pi f C)."""
    for f in (0.01, 1.0, 37.5, 150.0):
        ff = float(np.float32(f))
        assert oracle.load_impedance(f, 1.0, 4.5, 0.0) == complex(1.0, 2.0 * np.pi * ff * 4.5)
        z = oracle.load_impedance(f, 0.0, 0.0, 2e-6)
        assert abs(z.imag + 1.0 / (2 * np.pi * ff * 1e6 * 2e-12)) <= 1e-12 * abs(z.imag)


def test_series_load_at_feed_adds_to_zin(oracle, geometries, geo_dir):
    """A series load Z_L at the fed basis: (Z + Z_L e e^T) I = V e gives
    V/I_g = 1/(Z^-1)_gg + Z_L = Zin + Z_L exactly (circuit theory)."""
    t = oracle.load(_dipole(geometries, geo_dir, 121, 1.0, 0.001))
    g = 61
    for f, load in ((150.0, (g, 50.0, 0.0, 0.0)), (90.0, (g, 10.0, 0.05, 0.0)), (200.0, (g, 0.0, 0.0, 1e-5))):
        z0 = oracle.solve_freq_ex(t, f, 1, [(g, 0, 1.0, 0.0)])[0][0]
        z1 = oracle.solve_freq_ex(t, f, 1, [(g, 0, 1.0, 0.0)], [load])[0][0]
        zl = oracle.load_impedance(f, *load[1:])
        assert abs(z1 - (z0 + zl)) <= 1e-10 * abs(z1), (f, z1, z0, zl)


def test_load_elsewhere_is_sherman_morrison(oracle, geometries, geo_dir):
    """A load Z_L on basis l != feed g: with x_g = Z^-1 e_g and x_l = Z^-1 e_l
    (two generators, two right-hand sides), the loaded current is
    I' = x_g - x_l Z_L x_g[l] / (1 + Z_L x_l[l]) (Sherman-Morrison)."""
    p = os.path.join(geo_dir, "sm_grid.geo")
    geometries.planar_grid(p, 1024, seed=11)
    t = oracle.load(p)
    g, l, f = 200, 731, 60.0
    _, X = oracle.solve_freq_ex(t, f, 2, [(g, 0, 1.0, 0.0), (l, 1, 1.0, 0.0)])
    # two loads on basis l are summed: 75 ohm and 0.02 uH
    _, Xl = oracle.solve_freq_ex(t, f, 1, [(g, 0, 1.0, 0.0)], [(l, 75.0, 0.0, 0.0), (l, 0.0, 0.02, 0.0)])
    xg, xl = X[:, 0], X[:, 1]
    zl = oracle.load_impedance(f, 75.0, 0.0, 0.0) + oracle.load_impedance(f, 0.0, 0.02, 0.0)
    want = xg - xl * zl * xg[l - 1] / (1 + zl * xl[l - 1])
    assert np.abs(Xl[:, 0] - want).max() <= 1e-10 * np.abs(want).max()


def test_generators_superpose(oracle, geometries, geo_dir):
    """Two generators in one right-hand side give the sum of their separate
    solutions (linearity); each generator's V/I uses its own voltage."""
    t = oracle.load(_dipole(geometries, geo_dir, 121, 1.0, 0.001))
    f = 120.0
    z, X = oracle.solve_freq_ex(t, f, 1, [(30, 0, 1.0, 0.0), (90, 0, 2.0, 45.0)])
    _, Y = oracle.solve_freq_ex(t, f, 2, [(30, 0, 1.0, 0.0), (90, 1, 2.0, 45.0)])
    s = Y[:, 0] + Y[:, 1]
    assert np.abs(X[:, 0] - s).max() <= 1e-12 * np.abs(s).max()
    v2 = 2.0 * np.exp(1j * np.pi * float(np.float32(45.0)) / 180.0)
    assert abs(z[1] - v2 / X[89, 0]) <= 1e-12 * abs(z[1])


def _uniform_grid(df_mhz, fmax_mhz):
    return (np.arange(1, int(round(fmax_mhz / df_mhz)) + 1) * df_mhz).astype(np.float32)


def test_transient_reproduces_double_exponential(oracle):
    """The transform of the double-exponential spectrum (H = 1) returns
    i0 (e^{-a t} - e^{-b t}) to the truncation error of the grid (closed form)."""
    a, b, i0 = 1e5, 1e6, 10.0
    f = _uniform_grid(0.002, 50.0)
    t = np.linspace(0.2e-6, 60e-6, 120)
    r = oracle.transient(f, np.ones((len(f), 1)), oracle.spectrum_dexp(f, i0, a, b), t)[0]
    want = i0 * (np.exp(-a * t) - np.exp(-b * t))
    assert np.abs(r - want).max() <= 2e-3 * np.abs(want).max(), np.abs(r - want).max()


def test_transient_of_low_pass_is_convolution(oracle):
    """A first-order low-pass transfer 1/(1 + j 2 pi f tau) applied to the
    double exponential gives the convolution with e^{-t/tau}/tau in closed form."""
    a, b, tau = 1e5, 1e6, 2e-6
    f = _uniform_grid(0.002, 50.0)
    t = np.linspace(0.5e-6, 60e-6, 100)
    H = 1.0 / (1.0 + 2j * np.pi * f.astype(np.float64) * 1e6 * tau)
    r = oracle.transient(f, H[:, None], oracle.spectrum_dexp(f, 1.0, a, b), t)[0]
    c = 1.0 / tau
    want = (np.exp(-a * t) - np.exp(-c * t)) * c / (c - a) - (np.exp(-b * t) - np.exp(-c * t)) * c / (c - b)
    assert np.abs(r - want).max() <= 2e-3 * np.abs(want).max(), np.abs(r - want).max()


def test_trapezoid_weights_integrate_linear_functions_exactly(oracle):
    """Over [0, f_K-1] the weights plus f_0/2 at the DC node integrate linear functions exactly."""
    f = np.array([0.001, 0.002, 0.005, 0.01, 0.05, 0.1, 1.0, 2.5], np.float32)
    w = oracle.trapezoid_weights(f)
    fh = f.astype(np.float64) * 1e6
    g = 3.0 + 2e-6 * fh
    exact = 3.0 * fh[-1] + 1e-6 * fh[-1] ** 2
    assert abs((w * g).sum() + 0.5 * fh[0] * 3.0 - exact) <= 1e-12 * exact


def _segment_steps(t):
    """Per basis, the distance between its two far nodes (its support) projected on each axis."""
    n = t.npls
    f1 = t.ise[t.ips[:n] - 1] - 1
    f2 = t.ise[t.ips[t.nsmax:t.nsmax + n] - 1 + t.nsmax] - 1
    return (t.x[f2].astype(np.float64) - t.x[f1], t.y[f2].astype(np.float64) - t.y[f1],
            t.z[f2].astype(np.float64) - t.z[f1])


def test_plane_wave_broadside_is_field_times_support(oracle, geometries, geo_dir):
    """A wave from theta = 90 deg, phi = 0 with e_theta = -E0 has E = E0 z on the
    z axis and constant phase along a z-directed wire: each basis receives E0
    times half its support; the crossed polarization (e_phi) gives zero."""
    t = oracle.load(_dipole(geometries, geo_dir, 121, 1.0, 0.001))
    E0 = 2.5 - 1.0j
    b = oracle.plane_wave_rhs(t, 150.0, 90.0, 0.0, -E0, 0.0)
    _, _, dz = _segment_steps(t)
    want = E0 * dz[: t.npls] / 2
    assert np.abs(b - want).max() <= 1e-12 * abs(E0)
    assert np.abs(oracle.plane_wave_rhs(t, 150.0, 90.0, 0.0, 0.0, E0)).max() <= 1e-15


def test_plane_wave_phase_and_orientation(oracle, geometries, geo_dir):
    """An x-directed wire at y = 0.3 under a wave from theta = 90, phi = 90 (from +y)
    with e_phi = -E0 (E = E0 x): each basis receives E0 times half its support
    times exp(j k 0.3), the phase of the wave at the wire's line."""
    p = os.path.join(geo_dir, "xwire.geo")
    geometries.write_geo(p, [(-0.5, 0.3, 0.0), (0.5, 0.3, 0.0)], [(1, 2, 0.001, 61)], "x wire")
    t = oracle.load(p)
    f, E0 = 80.0, 1.0 + 0.5j
    b = oracle.plane_wave_rhs(t, f, 90.0, 90.0, 0.0, -E0)
    dx, _, _ = _segment_steps(t)
    y0 = float(np.float32(0.3))
    want = E0 * dx[: t.npls] / 2 * np.exp(1j * oracle.wavenumber(f) * y0)
    assert np.abs(b - want).max() <= 1e-12 * abs(E0)
