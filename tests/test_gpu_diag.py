"""GPU diagnostics of the reduction: MFMA operand layout and the spectrum of the tridiagonal."""
import numpy as np
import pytest
from scipy.linalg import eigvalsh_tridiagonal

import qudos_amd as q
from qudos_amd._lib import selftest_mfma, selftest_quad_sum

pytestmark = pytest.mark.gpu


def test_mfma_layout():
    err = selftest_mfma()
    assert err[0] == 0.0 and err[1] == 0.0, err


@pytest.mark.parametrize("n", [20, 40, 70, 130, 300])
@pytest.mark.parametrize("prec", [q.QD_FP64, q.QD_FP32_REFINE])
def test_tridiagonal_spectrum(n, prec):
    rng = np.random.default_rng(n)
    A = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    A = 0.5 * (A + A.conj().T)
    s = q.Solver(0)
    try:
        # a window the solver supports for both precisions (FP32 + refinement needs n >= 2 k), so
        # any error is a failure of this test
        s.solve_matrix(A, 1, min(n // 2, 12), prec)
        d, e = s.tridiagonal(n)
    finally:
        s.close()
    ref = np.linalg.eigvalsh(A)
    got = eigvalsh_tridiagonal(d, e)
    tol = 1e-10 if prec == q.QD_FP64 else 1e-4
    print(n, prec, np.abs(got - ref).max(), d[:4], e[:4])
    assert np.abs(got - ref).max() < tol * np.abs(ref).max()


def test_dpp_lane_reductions():
    """The DPP lane reductions of the band kernels: every lane of an aligned group of 4, 8 or 16
    lanes receives the group sum; lane 0's value is broadcast; every lane receives the 64-lane sum."""
    rng = np.random.default_rng(5)
    x = rng.normal(size=64)
    got = selftest_quad_sum(x)
    for row, w in enumerate([4, 8, 16]):
        ref = x.reshape(64 // w, w).sum(axis=1).repeat(w)
        assert np.abs(got[row] - ref).max() < 1e-13, w
    assert np.all(got[3] == x[0])
    assert np.abs(got[4] - x.sum()).max() < 1e-13
