"""Host-side decision logic of libwiremom, called through internal test hooks
that touch no device (CPU only).

Refinement convergence (wm::refine_decide, via wmx_refine_decide): after each
iteration the host reads max|dx| and max|x| per right-hand side (the device
reduction reports non-finite values as +inf) and decides to continue, stop
converged or stop failed; a failed item is refactored in FP64.

Anchors of recycled sweeps (choose_anchors, via wmx_choose_anchors): closed
forms of the covering factor and a brute-force check that the greedy cover is
valid and uses the fewest anchors.
"""
import ctypes as C
import math

import pytest


@pytest.fixture(scope="module")
def decide(wm):
    L = wm.lib()
    f = L.wmx_refine_decide
    dp = C.POINTER(C.c_double)
    f.argtypes = [C.c_int, dp, dp, dp, C.c_int, C.c_double, C.c_int, dp, dp, dp]
    f.restype = C.c_int

    class Item:
        def __init__(self, nrhs=1):
            self.nrhs = nrhs
            self.prev = (C.c_double * nrhs)(*([1e300] * nrhs))
            self.first, self.rate, self.worst = C.c_double(0), C.c_double(-1), C.c_double(0)

        def step(self, it, dm, xm, tol=1e-12, max_iter=20):
            a = (C.c_double * self.nrhs)(*dm)
            b = (C.c_double * self.nrhs)(*xm)
            return f(self.nrhs, a, b, self.prev, it, tol, max_iter, C.byref(self.first), C.byref(self.rate),
                     C.byref(self.worst))
    return Item


def test_converges_on_geometric_contraction(decide):
    it = decide()
    dx = 1.0
    verdicts = []
    for k in range(20):
        verdicts.append(it.step(k, [dx], [1.0]))
        if verdicts[-1] != 0:
            break
        dx *= 0.05
    assert verdicts[-1] == 1
    assert abs(it.rate.value - 0.05) < 1e-9
    assert it.worst.value <= 1e-12


@pytest.mark.parametrize("bad", [float("nan"), float("inf")])
def test_non_finite_correction_fails(decide, bad):
    """A NaN or infinite correction must never count as converged (NaN would
    pass a `rel <= tol` test written the other way round)."""
    it = decide()
    assert it.step(0, [1e-3], [1.0]) == 0
    assert it.step(1, [bad], [1.0]) == -1
    assert it.rate.value == math.inf


def test_non_finite_iterate_fails(decide):
    it = decide(nrhs=2)
    assert it.step(0, [1e-20, 1e-3], [1.0, float("inf")]) == -1


def test_slow_contraction_stops_early(decide):
    """Contraction 0.9 from 1e-2 cannot reach 1e-12 within 20 iterations: the
    item stops after its second correction instead of using all 20."""
    it = decide()
    assert it.step(0, [1e-2], [1.0]) == 0
    assert it.step(1, [0.9e-2], [1.0]) == -1


def test_stall_at_floor_counts_as_converged(decide):
    it = decide()
    assert it.step(0, [1e-10], [1.0]) == 0
    assert it.step(1, [5e-11], [1.0]) == 0
    # the correction no longer halves near the FP64 floor
    assert it.step(2, [3e-11], [1.0]) == 1


def test_out_of_iterations_fails(decide):
    """Contraction predicted fast enough at first, then slowing: the item runs
    its max_iter iterations and fails."""
    it = decide()
    v = [it.step(k, [dx], [1.0], max_iter=4) for k, dx in enumerate([1e-3, 1e-7, 1e-9, 1e-11])]
    assert v == [0, 0, 0, -1]


def test_zero_solution_converges(decide):
    it = decide()
    assert it.step(0, [0.0], [0.0]) == 1


def test_build_rejects_spilling_fill_kernels():
    """build.py's guard on the fill kernels: resource-usage remarks with AGPRs,
    scratch or VGPR spills in a k_fill / k_self_terms instantiation fail the
    build; SGPR spills (to VGPR lanes) and other kernels do not."""
    import importlib.util
    import os
    here = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "praca-magisterska_amd")
    spec = importlib.util.spec_from_file_location("wm_build", os.path.join(here, "build.py"))
    b = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(b)

    def remarks(name, **f):
        r = [f"x.hip:1:1: remark: Function Name: {name} [-Rpass-analysis=kernel-resource-usage]"]
        r += [f"x.hip:1:1: remark:     {k}: {v} [-Rpass-analysis=kernel-resource-usage]" for k, v in f.items()]
        return "\n".join(r)

    ok = remarks("_Z6k_fillILi4E", VGPRs=189, AGPRs=0, **{"ScratchSize [bytes/lane]": 0, "VGPRs Spill": 0,
                                                        "SGPRs Spill": 52})
    other = remarks("_Z9k_gemm_b3", VGPRs=128, AGPRs=64)
    b.check_spill_free("wm_fill.hip", ok + "\n" + other)
    assert b.resource_usage(ok)["_Z6k_fillILi4E"]["VGPRs"] == 189
    for bad in ({"AGPRs": 106}, {"ScratchSize [bytes/lane]": 16}, {"VGPRs Spill": 3}):
        f = {"VGPRs": 256, "AGPRs": 0, "ScratchSize [bytes/lane]": 0, "VGPRs Spill": 0}
        f.update(bad)
        with pytest.raises(RuntimeError):
            b.check_spill_free("wm_fill.hip", ok + "\n" + remarks("_Z12k_self_terms", **f))
    with pytest.raises(RuntimeError):
        b.check_spill_free("wm_fill.hip", other)   # no fill remarks at all
    b.check_spill_free("wm_lu.hip", other)          # not a checked source


def test_phase_sincos_accuracy(tmp_path):
    """The fill's sin/cos of k R (Cody-Waite reduction, Taylor polynomials):
    against long-double sinl/cosl at 300k random arguments in [0, 1], [0, 100],
    [0, 1e4] and [0, 1.5e6], error within 2 ulp below 100 and 3 ulp up to the
    limit (ulp of the result); exact at 0; n = rint(x 2/pi) quadrants checked
    at multiples of pi/2, where one of the two is 0 to within the reduction's
    absolute error."""
    import os
    import subprocess
    hdr = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "praca-magisterska_amd", "csrc")
    src = tmp_path / "sc.cpp"
    src.write_text(r'''
#include <cstdio>
#include <cmath>
#include <random>
#include "wm_sincos.h"
static double ulps(double v, long double ref) {
    const double m = std::fmax(std::fabs((double)ref), 1e-300);
    return std::fabs((double)((v - ref) / (long double)std::ldexp(1.0, std::ilogb(m) - 52)));
}
int main() {
    std::mt19937_64 g(7);
    const double hi[4] = {1.0, 100.0, 1e4, 1.5e6};
    for (int k = 0; k < 4; ++k) {
        std::uniform_real_distribution<double> U(0.0, hi[k]);
        double w = 0;
        for (int i = 0; i < 300000; ++i) {
            const double x = U(g);
            double s, c;
            wm::sincos_phase(x, s, c);
            w = std::fmax(w, std::fmax(ulps(s, sinl((long double)x)), ulps(c, cosl((long double)x))));
        }
        printf("%g %.4f\n", hi[k], w);
    }
    double s, c;
    wm::sincos_phase(0.0, s, c);
    printf("zero %.17g %.17g\n", s, c);
    for (int q = 1; q <= 8; ++q) {
        const long double xq = q * 1.5707963267948966192313216916397514L;
        wm::sincos_phase((double)xq, s, c);
        printf("q %d %.17g %.17g %.17Lg %.17Lg\n", q, s, c, sinl((long double)(double)xq), cosl((long double)(double)xq));
    }
}
''')
    exe = tmp_path / "sc"
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", hdr, str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.splitlines()
    worst = {float(l.split()[0]): float(l.split()[1]) for l in out[:4]}
    assert worst[1.0] <= 2.0 and worst[100.0] <= 2.0, worst
    assert worst[1e4] <= 3.0 and worst[1.5e6] <= 3.0, worst
    assert out[4] == "zero 0 1", out[4]
    for l in out[5:]:
        _, q, s, c, sr, cr = l.split()
        assert abs(float(s) - float(sr)) <= 1e-15 and abs(float(c) - float(cr)) <= 1e-15, l


# Anchors of recycled sweeps (choose_anchors, via wmx_choose_anchors): the
# covering factor at f is r^(1 + log10(1/(kD))) for kD <= 1 and 1 + (r-1)/(kD)
# above (k = 2 pi f / c, f in MHz as float32, c = 2.99792448e8).

def _anchors(wm, f, r, diam):
    L = wm.lib()
    h = L.wmx_choose_anchors
    h.argtypes = [C.POINTER(C.c_float), C.c_int, C.c_double, C.c_double, C.POINTER(C.c_int)]
    h.restype = C.c_int
    n = len(f)
    fa = (C.c_float * n)(*f)
    out = (C.c_int * n)()
    assert h(fa, n, r, diam, out) == 0
    return [out[i] for i in range(n)], [fa[i] for i in range(n)]


def _kd(f_mhz, diam):
    return 2.0 * math.pi * f_mhz * 1e6 / 2.99792448e8 * diam


def _covers(fa, fk, r, diam):
    """Frequency fk may take the anchor fa: the ratio is within the covering
    factor at the higher of the two (the factor does not grow with f)."""
    hi = max(fa, fk)
    kd = _kd(hi, diam)
    fac = 1.0 + (r - 1.0) / kd if kd > 1.0 else r ** (1.0 - math.log10(kd))
    return max(fa / fk, fk / fa) <= fac * (1 + 1e-12)


def test_anchor_cover_special_cases(wm):
    """Closed forms of the covering factor: at kD = 1 exactly r, a decade below
    r^2, two decades below r^3, at kD = 2 the factor 1 + (r-1)/2; probed with two
    frequencies a hair inside and outside each factor."""
    r, diam = 2.5, 10.0
    f_kd1 = 1.0 / _kd(1.0, diam)  # MHz at kD = 1
    for kd_hi, fac in ((1.0, r), (0.1, r * r), (0.01, r ** 3), (2.0, 1.0 + (r - 1.0) / 2.0)):
        hi = kd_hi * f_kd1
        for scale, same in ((0.999, True), (1.001, False)):
            lo = hi / (fac * scale)
            anc, ff = _anchors(wm, [hi, lo], r, diam)
            # float32 frequencies: compare on the values the library saw
            expect = max(ff) / min(ff) <= fac * (1 + 1e-12)
            assert expect == same, (kd_hi, scale)
            assert (anc[1] == anc[0]) == same, (kd_hi, scale, anc)


def test_anchor_cover_is_valid_and_minimal(wm):
    """Every frequency lies within the covering factor of its anchor, anchors
    anchor themselves and their groups are contiguous; on small random sets the
    number of anchors equals the brute-force minimum of any cover."""
    import itertools
    import random
    rng = random.Random(3)
    for trial in range(60):
        n = rng.randint(1, 10)
        f = sorted((10 ** rng.uniform(-2, 1.5) for _ in range(n)), reverse=True)
        r = rng.choice([1.5, 2.0, 2.5, 3.0])
        diam = rng.choice([5.0, 25.0, 80.0])
        anc, ff = _anchors(wm, f, r, diam)
        for i, a in enumerate(anc):
            assert anc[a] == a, (trial, anc)
            assert _covers(ff[a], ff[i], r, diam), (trial, i, a, ff)
        assert all(anc[i] <= anc[i + 1] for i in range(n - 1))
        best = None
        for k in range(1, n + 1):
            for sub in itertools.combinations(range(n), k):
                if all(any(_covers(ff[a], ff[i], r, diam) for a in sub) for i in range(n)):
                    best = k
                    break
            if best:
                break
        assert len(set(anc)) == best, (trial, anc, best, ff, r, diam)
