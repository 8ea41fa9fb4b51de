"""Momentum matrix elements (qd_momentum): against the oracle's dH/dK on the same eigenvectors,
and against the reference's momentum_matrix_elements.dat of the Prism_111 and Truncated_pyramid
examples through Kramers-pair block sums of |p|^2 (invariant under the phase and the mixing
within each degenerate pair, which the files fix arbitrarily)."""
import os

import numpy as np
import pytest

import qudos_amd as q
from helpers import GOLDEN, momentum_columns, momentum_pair_sums, oracle_problem, prism_paths, read_momentum_file
from oracle import qd_oracle as qo

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def solver():
    s = q.Solver(0)
    yield s
    s.close()


@pytest.mark.parametrize("overrides", [
    ("-M_Gx=2", "-M_Gy=2", "-M_Gz=2", "-M_S_Gx=2", "-M_S_Gy=2", "-M_S_Gz=2"),
    ("-M_Gx=2", "-M_Gy=2", "-M_Gz=1", "-M_S_Gx=2", "-M_S_Gy=2", "-M_S_Gz=1", "-Kx=0.13", "-Kz=-0.2", "-Fz=12"),
])
def test_momentum_matches_oracle(solver, overrides):
    pp, sp = prism_paths()
    inp = q.read_inputs(pp, sp, overrides)
    solver.upload(q.build_tables(inp))
    il, iu = q.window(inp)
    w, Z, _ = solver.solve(il, iu, q.QD_FP64)
    nvb, ncb = 2 * inp.N_vb, 2 * inp.N_cb
    p, st = solver.momentum(inp, nvb, ncb)
    po, sh = oracle_problem(pp, sp, overrides)
    t = qo.coefficient_tables(po, sh)
    ref = qo.momentum_elements(t, Z[:, nvb - 1::-1], Z[:, nvb:nvb + ncb])
    assert p.shape == ref.shape == (3, ncb, nvb)
    assert np.abs(p - ref).max() <= 1e-9 * np.abs(ref).max(), np.abs(p - ref).max()
    # fewer states: the leading block of the same elements
    p2, _ = solver.momentum(inp, 2, 2)
    assert np.abs(p2 - p[:, :2, :2]).max() <= 1e-12 * np.abs(p).max()


@pytest.mark.parametrize("name", ["prism_111", "truncated_pyramid"])
def test_momentum_fixture(solver, name):
    """code/example/{Prism_111,Truncated_pyramid}/Electronic_structure/PWEM_test_ref/momentum_matrix_elements.dat"""
    d = os.path.join(GOLDEN, name)
    inp = q.read_inputs(os.path.join(d, "parameters.in"), os.path.join(d, "supercell.in"))
    solver.upload(q.build_tables(inp))
    il, iu = q.window(inp)
    solver.solve(il, iu, q.QD_FP64)
    nvb, ncb = 2 * inp.N_vb, 2 * inp.N_cb
    p, st = solver.momentum(inp, nvb, ncb)
    got = momentum_pair_sums(np.abs(momentum_columns(p)))
    _, mag = read_momentum_file(os.path.join(d, "momentum_matrix_elements.dat"))
    ref = momentum_pair_sums(mag)
    assert got.shape == ref.shape
    err = np.abs(got - ref).max()
    print(f"{name}: momentum block sums max |d| {err:.3e} of {ref.max():.3e}; {st}")
    assert err <= 1e-6 * ref.max()
