"""GPU parity: Hamiltonian kernel and windowed eigensolver (FP64 and FP32 + FP64 refinement)
against the FP64 oracle, synthetic spectra and the reference's Prism_111 fixture."""
import os

import numpy as np
import pytest

import qudos_amd as q
from helpers import GOLDEN, oracle_problem, prism_paths
from oracle import qd_oracle as qo

pytestmark = pytest.mark.gpu

SMALL = ("-M_Gx=3", "-M_Gy=3", "-M_Gz=2", "-M_S_Gx=3", "-M_S_Gy=3", "-M_S_Gz=2")
FIELD = SMALL + ("-Fx=12", "-Fy=-5", "-Fz=20", "-Kz=0.02")


@pytest.fixture(scope="module")
def solver():
    s = q.Solver(0)
    yield s
    s.close()


def _oracle_H(overrides):
    pp, sp = prism_paths()
    p, sh = oracle_problem(pp, sp, overrides)
    return p, qo.build_hamiltonian(qo.coefficient_tables(p, sh))


@pytest.mark.parametrize("overrides", [SMALL, FIELD])
def test_hamiltonian_matches_oracle(solver, overrides):
    pp, sp = prism_paths()
    t = q.build_tables(q.read_inputs(pp, sp, overrides))
    solver.upload(t)
    p, Href = _oracle_H(overrides)
    L = np.tril(Href)
    scale = np.abs(Href).max()
    H64 = np.tril(solver.hamiltonian(q.QD_FP64))
    assert np.abs(H64 - L).max() <= 1e-13 * scale
    H32 = np.tril(solver.hamiltonian(q.QD_FP32_REFINE))
    assert np.abs(H32 - L).max() <= 2e-7 * scale


def _random_hermitian(n, seed, pairs=True):
    rng = np.random.default_rng(seed)
    lam = np.sort(rng.uniform(-10, 10, size=n))
    if pairs:   # Kramers-like exact pairs
        lam[1::2] = lam[0::2][: n // 2]
        lam = np.sort(lam)
    X = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    Q, _ = np.linalg.qr(X)
    A = (Q * lam) @ Q.conj().T
    A = 0.5 * (A + A.conj().T)
    return A, np.linalg.eigvalsh(A)


@pytest.mark.parametrize("n,il,k", [(300, 101, 20), (777, 300, 40), (1100, 501, 12)])
@pytest.mark.parametrize("prec", [q.QD_FP64, q.QD_FP32_REFINE])
def test_eigensolver_random_matrix(solver, n, il, k, prec):
    A, ref = _random_hermitian(n, seed=n)
    iu = il + k - 1
    w, Z, st = solver.solve_matrix(A, il, iu, prec)
    assert np.max(np.abs(w - ref[il - 1:iu])) < 1e-10
    R = A @ Z - Z * w
    assert np.linalg.norm(R, axis=0).max() < 1e-8
    G = Z.conj().T @ Z
    assert np.abs(G - np.eye(k)).max() < 1e-8


@pytest.mark.parametrize("prec", [q.QD_FP64, q.QD_FP32_REFINE])
@pytest.mark.parametrize("overrides", [SMALL, FIELD])
def test_small_prism_end_to_end(solver, prec, overrides):
    pp, sp = prism_paths()
    inp = q.read_inputs(pp, sp, overrides)
    t = q.build_tables(inp)
    solver.upload(t)
    il, iu = q.window(inp)
    w, Z, st = solver.solve(il, iu, prec)
    p, H = _oracle_H(overrides)
    wref, Zref = qo.solve(H, p)
    tol = 1e-9 if prec == q.QD_FP64 else 1e-6
    assert np.max(np.abs(w - wref)) < tol, np.abs(w - wref)
    R = H @ Z - Z * w
    assert np.linalg.norm(R, axis=0).max() < (1e-8 if prec == q.QD_FP64 else 1e-7)
    # gauge-invariant: projector onto each Kramers pair
    for a in range(0, len(w), 2):
        P1 = Z[:, a:a + 2] @ Z[:, a:a + 2].conj().T
        P2 = Zref[:, a:a + 2] @ Zref[:, a:a + 2].conj().T
        assert np.abs(P1 - P2).max() < 1e-6


@pytest.mark.parametrize("prec", [q.QD_FP64, q.QD_FP32_REFINE])
def test_prism_fixture_energies(solver, prec):
    """reference output code/example/Prism_111/Electronic_structure/PWEM_test_ref/energies_*.dat"""
    pp, sp = prism_paths()
    inp = q.read_inputs(pp, sp)
    solver.upload(q.build_tables(inp))
    il, iu = q.window(inp)
    w, Z, st = solver.solve(il, iu, prec)
    nv = 2 * inp.N_vb
    ref_vb = np.loadtxt(os.path.join(GOLDEN, "prism_111", "energies_vb.dat"))
    ref_cb = np.loadtxt(os.path.join(GOLDEN, "prism_111", "energies_cb.dat"))
    assert np.max(np.abs(w[:nv][::-1] - ref_vb[:, 0])) < 1e-9
    assert np.max(np.abs(w[nv:] - ref_cb[:, 0])) < 1e-9
    print("stats", st)


@pytest.mark.parametrize("prec", [q.QD_FP64, q.QD_FP32_REFINE])
def test_truncated_pyramid_fixture_energies(solver, prec):
    """reference output code/example/Truncated_pyramid/Electronic_structure/PWEM_test_ref/
    energies_*.dat: 100 stacked pyramid slices, a commensurate strain cell 3 x 3 x 4 the electronic
    cell (FFT downsampling of strain and piezoelectric potential), second-order piezoelectricity."""
    d = os.path.join(GOLDEN, "truncated_pyramid")
    pp, sp = os.path.join(d, "parameters.in"), os.path.join(d, "supercell.in")
    inp = q.read_inputs(pp, sp)
    assert list(inp.RL) == [3, 3, 4] and inp.strain_type == 1
    solver.upload(q.build_tables(inp))
    il, iu = q.window(inp)
    w, Z, st = solver.solve(il, iu, prec)
    p, _ = oracle_problem(pp, sp)
    rows_vb, rows_cb = qo.energies_rows(w, Z, p)
    for got, name in ((np.array(rows_vb), "energies_vb.dat"), (np.array(rows_cb), "energies_cb.dat")):
        ref = np.loadtxt(os.path.join(d, name), ndmin=2)
        assert got.shape == ref.shape
        assert np.max(np.abs(got[:, 0] - ref[:, 0])) < 1e-9, (name, got[:, 0] - ref[:, 0])
        # band characters per Kramers pair (gauge invariant); the file holds 11 significant digits
        g2 = got[:, 1:].reshape(-1, 2, 4).sum(axis=1)
        r2 = ref[:, 1:].reshape(-1, 2, 4).sum(axis=1)
        assert np.max(np.abs(g2 - r2)) < 1e-6, (name, g2 - r2)
    print("truncated pyramid", prec, "max |dE| vs fixture", st)
