"""CPU check of the momentum matrix elements: the momentum kernel's element code (dH/dK from the
element formula at K +- s), run on the host, against the oracle's dH/dK (central difference of
its own Hamiltonian, pinned by the K-quadratic and Hellmann-Feynman tests) on random states."""
import numpy as np
import pytest

import qudos_amd as q
from helpers import oracle_problem, prism_paths
from oracle import qd_oracle as qo


@pytest.mark.parametrize("overrides", [
    ("-M_Gx=1", "-M_Gy=1", "-M_Gz=1", "-M_S_Gx=1", "-M_S_Gy=1", "-M_S_Gz=1"),
    ("-M_Gx=1", "-M_Gy=2", "-M_Gz=1", "-M_S_Gx=1", "-M_S_Gy=2", "-M_S_Gz=1", "-Kx=0.21", "-Ky=-0.13", "-Kz=0.08",
     "-Fz=15"),
])
def test_momentum_element_code_matches_oracle(overrides):
    pp, sp = prism_paths()
    inp = q.read_inputs(pp, sp, overrides)
    t = q.build_tables(inp)
    nh = 8 * inp.N[0] * inp.N[1] * inp.N[2]
    rng = np.random.default_rng(9)
    V = rng.standard_normal((nh, 3)) + 1j * rng.standard_normal((nh, 3))
    Cc = rng.standard_normal((nh, 2)) + 1j * rng.standard_normal((nh, 2))
    p = q.selftest_momentum_host(t, V, Cc)
    po, sh = oracle_problem(pp, sp, overrides)
    ref = qo.momentum_elements(qo.coefficient_tables(po, sh), V[:, ::-1], Cc)
    assert p.shape == ref.shape == (3, 2, 3)
    assert np.abs(p - ref).max() <= 1e-11 * np.abs(ref).max(), np.abs(p - ref).max() / np.abs(ref).max()
