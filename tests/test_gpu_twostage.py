"""GPU parity of the two-stage tridiagonal reduction (dense -> band of width 32 -> tridiagonal)
through the C ABI: spectra of the tridiagonal, windowed eigenpairs, both precisions, sizes around
the band and panel widths, matrices with zero panels, and agreement with the one-stage path."""
import numpy as np
import pytest
from scipy.linalg import eigvalsh_tridiagonal

import qudos_amd as q

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def solver():
    s = q.Solver(0)
    yield s
    s.close()


def _hermitian(n, seed, kind="random"):
    rng = np.random.default_rng(seed)
    if kind == "random":
        A = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    elif kind == "block":     # two decoupled blocks: panels below the first block are zero
        A = np.zeros((n, n), complex)
        h = n // 3
        for a, b in [(0, h), (h, n)]:
            A[a:b, a:b] = rng.normal(size=(b - a, b - a)) + 1j * rng.normal(size=(b - a, b - a))
    elif kind == "diag":      # already tridiagonal (every reflector is trivial)
        A = np.diag(rng.normal(size=n)).astype(complex)
    elif kind == "banded":    # complex band of width 20 (< 32)
        A = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
        A = np.triu(np.tril(A, 20), -20)
    A = 0.5 * (A + A.conj().T)
    return A


def _window(n):
    k = min(n, 12)
    il = max(1, (n - k) // 2 + 1)
    return il, il + k - 1


@pytest.mark.parametrize("n", [2, 3, 4, 31, 32, 33, 34, 35, 63, 64, 65, 66, 67, 96, 97, 130, 257, 300])
@pytest.mark.parametrize("prec", [q.QD_FP64, q.QD_FP32_REFINE])
def test_two_stage_sizes(solver, n, prec):
    if prec == q.QD_FP32_REFINE and n < 31:
        pytest.skip("the refinement's Rayleigh-Ritz space [X E] needs n >= 2k")
    print(f"[case] two-stage n={n} prec={prec}", flush=True)
    A = _hermitian(n, seed=1000 + n)
    il, iu = _window(n)
    ref = np.linalg.eigvalsh(A)
    w, Z, st = solver.solve_matrix(A, il, iu, prec)
    d, e = solver.tridiagonal(n)
    tri = eigvalsh_tridiagonal(d, e)
    scale = max(1.0, np.abs(ref).max())
    assert np.abs(tri - ref).max() < (1e-12 if prec == q.QD_FP64 else 2e-5) * scale
    assert np.abs(w - ref[il - 1:iu]).max() < 1e-10 * scale
    R = A @ Z - Z * w
    # the refinement stops on |delta lambda| < 1e-10: residuals ~ sqrt(delta lambda * gap)
    assert np.linalg.norm(R, axis=0).max() < (1e-9 if prec == q.QD_FP64 else 1e-8) * scale
    assert np.abs(Z.conj().T @ Z - np.eye(iu - il + 1)).max() < 1e-9


@pytest.mark.parametrize("kind", ["block", "diag", "banded"])
@pytest.mark.parametrize("prec", [q.QD_FP64, q.QD_FP32_REFINE])
def test_two_stage_structured(solver, kind, prec):
    n = 200
    print(f"[case] two-stage structured {kind} prec={prec}", flush=True)
    A = _hermitian(n, seed=7, kind=kind)
    il, iu = 90, 109
    ref = np.linalg.eigvalsh(A)
    w, Z, st = solver.solve_matrix(A, il, iu, prec)
    scale = np.abs(ref).max()
    assert np.abs(w - ref[il - 1:iu]).max() < 1e-10 * scale
    R = A @ Z - Z * w
    assert np.linalg.norm(R, axis=0).max() < (1e-9 if prec == q.QD_FP64 else 1e-8) * scale


@pytest.mark.parametrize("prec", [q.QD_FP64, q.QD_FP32_REFINE])
def test_two_stage_matches_one_stage(solver, prec):
    n = 777
    print(f"[case] two-stage vs one-stage n={n} prec={prec}", flush=True)
    A = _hermitian(n, seed=3)
    il, iu = 350, 389
    try:
        solver.set_reduction(q.QD_REDUCE_ONE_STAGE)
        w1, Z1, _ = solver.solve_matrix(A, il, iu, prec)
    finally:
        solver.set_reduction(q.QD_REDUCE_TWO_STAGE)
    w2, Z2, st = solver.solve_matrix(A, il, iu, prec)
    assert np.abs(w1 - w2).max() < 1e-10 * np.abs(w1).max()
    # eigenvectors up to phase (the spectrum of a random matrix is simple)
    ov = np.abs(np.sum(Z1.conj() * Z2, axis=0))
    assert np.abs(ov - 1).max() < 1e-8
    assert st["t_band"] > 0


@pytest.mark.parametrize("reduction", [q.QD_REDUCE_TWO_STAGE, q.QD_REDUCE_ONE_STAGE])
@pytest.mark.parametrize("prec", [q.QD_FP64, q.QD_FP32_REFINE])
def test_factorization_q(solver, reduction, prec):
    """The unitary factor Q of the reduction: Q^H Q = I and Q^H A Q = T (tridiagonal (d, e))."""
    n = 230
    print(f"[case] Q of the reduction n={n} reduction={reduction} prec={prec}", flush=True)
    A = _hermitian(n, seed=11)
    try:
        solver.set_reduction(reduction)
        solver.solve_matrix(A, 100, 111, prec)
        d, e = solver.tridiagonal(n)
        Q = np.concatenate([solver.apply_q(np.eye(n, dtype=complex)[:, c:c + 50]) for c in range(0, n, 50)], axis=1)
        QH = np.concatenate([solver.apply_q(np.eye(n, dtype=complex)[:, c:c + 50], adjoint=True)
                             for c in range(0, n, 50)], axis=1)
    finally:
        solver.set_reduction(q.QD_REDUCE_TWO_STAGE)
    T = np.diag(d) + np.diag(e, 1) + np.diag(e, -1)
    tol = 1e-13 if prec == q.QD_FP64 else 2e-6
    scale = np.abs(A).max()
    print("Q^H Q - I", np.abs(Q.conj().T @ Q - np.eye(n)).max(), "QH - Q^H", np.abs(QH - Q.conj().T).max(),
          "Q^H A Q - T", np.abs(Q.conj().T @ A @ Q - T).max() / scale)
    assert np.abs(QH - Q.conj().T).max() < 1e-13
    assert np.abs(Q.conj().T @ Q - np.eye(n)).max() < tol
    assert np.abs(Q.conj().T @ A @ Q - T).max() < tol * 10 * scale


@pytest.mark.parametrize("prec", [q.QD_FP64, q.QD_FP32_REFINE])
def test_apply_q_widths(solver, prec):
    """Q (and Q^H) applied to blocks of 1, 16, 31, 32, 33 and 64 vectors (the most the back-transformation
    takes; every vector-count variant of the kernels) equals the same columns of Q applied 50 at a time:
    every column's arithmetic is independent of the block it travels in."""
    n = 230
    print(f"[case] apply_q widths n={n} prec={prec}", flush=True)
    A = _hermitian(n, seed=13)
    solver.solve_matrix(A, 100, 111, prec)
    E = np.eye(n, dtype=complex)
    for adjoint in (False, True):
        Q = np.concatenate([solver.apply_q(E[:, c:c + 50], adjoint=adjoint) for c in range(0, n, 50)], axis=1)
        for w in (1, 16, 31, 32, 33, 64):
            c0 = 7
            Qw = solver.apply_q(E[:, c0:c0 + w], adjoint=adjoint)
            assert Qw.shape == (n, w)
            assert np.abs(Qw - Q[:, c0:c0 + w]).max() < 1e-14, (adjoint, w)
