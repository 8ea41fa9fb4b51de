"""Status reporting of the solve and the adjoint: a limit that stops bisection, inverse iteration,
the FP64 refinement or the Sternheimer solve short returns QD_E_NOCONV (results still fetchable),
and a bounded wait recorded as failed makes every later use of that factorisation return
QD_E_DEVICE (the kernels themselves run normally: only the flag the host reads is set)."""
import os

import numpy as np
import pytest

import qudos_amd as q
from helpers import prism_paths
from qudos_amd.autograd import DotProblem

pytestmark = pytest.mark.gpu

SMALL = ("-M_Gx=3", "-M_Gy=3", "-M_Gz=2", "-M_S_Gx=3", "-M_S_Gy=3", "-M_S_Gz=2")
CFG = os.path.join(os.path.dirname(q.lib_path()), "..", "configs", "lens_11x11x9")
ADJ = ("-M_Gx=2", "-M_Gy=2", "-M_Gz=2", "-M_S_Gx=2", "-M_S_Gy=2", "-M_S_Gz=2", "-M_C_Gx=2", "-M_C_Gy=2",
       "-M_C_Gz=2", "-N_VB=4", "-N_CB=4")


@pytest.fixture(scope="module")
def prism():
    pp, sp = prism_paths()
    inp = q.read_inputs(pp, sp, SMALL)
    s = q.Solver(0)
    s.upload(q.build_tables(inp))
    il, iu = q.window(inp)
    w_ok, _, _ = s.solve(il, iu, q.QD_FP64)
    yield s, il, iu, w_ok
    s.close()


@pytest.mark.parametrize("limits,prec", [
    (dict(refine_sweeps=0), q.QD_FP32_REFINE),
    (dict(invit_sweeps=0), q.QD_FP64),
    (dict(bisect_rounds=1), q.QD_FP64),
])
def test_limit_reports_noconv(prism, limits, prec):
    s, il, iu, w_ok = prism
    s.set_limits(**limits)
    try:
        with pytest.raises(q.QdError) as e:
            s.solve_device(il, iu, prec)
        assert e.value.code == q.QD_E_NOCONV, str(e.value)
        w, _ = s.fetch(iu - il + 1, vectors=False)      # results stay fetchable
        assert np.all(np.isfinite(w))
    finally:
        s.set_limits()
    w, _, st = s.solve(il, iu, prec)
    assert np.abs(w - w_ok).max() < 1e-9
    assert st["tri_residual"] < 1e-11


def test_marked_wait_failure_fails_later_uses(prism):
    s, il, iu, w_ok = prism
    s.solve(il, iu, q.QD_FP64)
    s.mark_wait_failure()
    X = np.eye(s.n, 2, dtype=np.complex128)
    with pytest.raises(q.QdError) as e:
        s.apply_q(X)
    assert e.value.code == q.QD_E_DEVICE
    w, _, _ = s.solve(il, iu, q.QD_FP64)                # a new solve clears the record
    assert np.abs(w - w_ok).max() < 1e-9
    s.apply_q(X)


def test_adjoint_noconv_and_wait_failure():
    pp, sp = os.path.join(CFG, "parameters.in"), os.path.join(CFG, "supercell.in")
    s = q.Solver(0)
    prob = DotProblem(pp, sp, overrides=ADJ, precision=q.QD_FP32_REFINE, solver=s)
    try:
        theta = np.array([5.0, 10.0, 0.9, 3.0, -2.0, 8.0])
        prob.forward(theta)
        s.set_limits(sternheimer_iters=0)
        with pytest.raises(q.QdError) as e:
            prob.gradient()
        assert e.value.code == q.QD_E_NOCONV, str(e.value)
        s.set_limits()
        g, info = prob.gradient()
        assert info["sternheimer_residual"] < 1e-12
        s.mark_wait_failure()
        with pytest.raises(q.QdError) as e:
            prob.gradient()
        assert e.value.code == q.QD_E_DEVICE, str(e.value)
        prob.forward(theta)
        g2, _ = prob.gradient()
        assert np.abs(g2 - g).max() <= 1e-9 * np.abs(g).max()
    finally:
        prob.close()
        s.close()
