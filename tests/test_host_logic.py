"""Host-side decision logic of libwiremom, called through internal test hooks
that touch no device (CPU only).

Refinement convergence (wm::refine_decide, via wmx_refine_decide): after each
iteration the host reads max|dx| and max|x| per right-hand side (the device
reduction reports non-finite values as +inf) and decides to continue, stop
converged or stop failed; a failed item is refactored in FP64.
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
