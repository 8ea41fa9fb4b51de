"""CPU checks of the library's FP64 tables against the reference's fixtures: the full H built on the
host with the assembly kernels' element code from the library's tables (the same tables the GPU
path uploads), solved by LAPACK (zheevr through scipy) in the reference's eigen window, against
PWEM_test_ref/energies_*.dat.  The Truncated_pyramid example exercises the commensurate strain cell
(3 x 3 x 4, FFT downsampling of strain and piezoelectric potential), 100 graded slices and
second-order piezoelectricity.  The momentum file is checked the same way, with the momentum
kernel's element code run on the host on those eigenvectors, and the carrier localisation files
through the oracle's restatement on them."""
import os

import numpy as np
import pytest
import scipy.linalg as sla

import qudos_amd as q
from helpers import (GOLDEN, momentum_columns, momentum_pair_sums, oracle_problem, prism_paths,
                     read_momentum_file)
from oracle import qd_oracle as qo


def test_host_hamiltonian_matches_oracle():
    pp, sp = prism_paths()
    over = ("-M_Gx=2", "-M_Gy=1", "-M_Gz=2", "-M_S_Gx=2", "-M_S_Gy=1", "-M_S_Gz=2", "-Kx=0.1", "-Fz=9")
    H = q.selftest_hamiltonian_host(q.build_tables(q.read_inputs(pp, sp, over)))
    p, sh = oracle_problem(pp, sp, over)
    ref = qo.build_hamiltonian(qo.coefficient_tables(p, sh))
    assert np.abs(H - ref).max() <= 1e-13 * np.abs(ref).max()


@pytest.mark.slow
def test_truncated_pyramid_fixture_energies_on_the_host():
    d = os.path.join(GOLDEN, "truncated_pyramid")
    pp, sp = os.path.join(d, "parameters.in"), os.path.join(d, "supercell.in")
    inp = q.read_inputs(pp, sp)
    t = q.build_tables(inp)
    H = q.selftest_hamiltonian_host(t)
    il, iu = q.window(inp)
    w, Z = sla.eigh(H, lower=True, subset_by_index=[il - 1, iu - 1], driver="evr", overwrite_a=True,
                    check_finite=False)
    p, _ = oracle_problem(pp, sp)
    rows_vb, rows_cb = qo.energies_rows(w, Z, p)
    for got, name in ((rows_vb, "energies_vb.dat"), (rows_cb, "energies_cb.dat")):
        ref = np.loadtxt(os.path.join(d, name), ndmin=2)
        assert got.shape == ref.shape
        assert np.max(np.abs(got[:, 0] - ref[:, 0])) < 1e-9, (name, got[:, 0] - ref[:, 0])
        g2 = got[:, 1:].reshape(-1, 2, 4).sum(axis=1)
        r2 = ref[:, 1:].reshape(-1, 2, 4).sum(axis=1)
        assert np.max(np.abs(g2 - r2)) < 1e-6, (name, g2 - r2)
    nvb = 2 * inp.N_vb
    # carrier localisation per shape through the oracle's restatement on these eigenvectors
    shapes = qo.read_supercell(open(sp).read(), p)
    G = [qo.basis_1d(inp.N[a], [shapes[0].L1, shapes[0].L2, shapes[0].L3][a]) for a in range(3)]
    chis = [None] + [qo.chi(shapes, s, *qo.grid3(G), 0) for s in range(1, len(shapes))]
    loc = 100.0 * qo.carrier_localisation(chis, shapes, tuple(inp.N), Z)
    for got, name in ((loc[nvb - 1::-1], "carrier_localisation_vb.dat"), (loc[nvb:], "carrier_localisation_cb.dat")):
        ref = np.loadtxt(os.path.join(d, name), ndmin=2)
        assert got.shape == ref.shape
        assert np.max(np.abs(got - ref) / np.maximum(1.0, np.abs(ref))) < 1e-5, (name, np.abs(got - ref).max())
    pm = q.selftest_momentum_host(t, Z[:, :nvb], Z[:, nvb:])
    got = momentum_pair_sums(np.abs(momentum_columns(pm)))
    _, mag = read_momentum_file(os.path.join(d, "momentum_matrix_elements.dat"))
    ref = momentum_pair_sums(mag)
    assert got.shape == ref.shape
    assert np.abs(got - ref).max() <= 1e-6 * ref.max(), np.abs(got - ref).max() / ref.max()
