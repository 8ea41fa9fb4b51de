"""Algebra of the FP32 panel QR's paired step (k_pqr_panel in qudos_amd/csrc/hip/band.hip), on the CPU.

The kernel factorises columns i and i+1 after one grid-wide reduction: besides column i's norm and
inner products it reduces the Gram row of column i+1 over rows >= i *before* H_i is applied, and
obtains column i+1's norm and inner products after H_i by subtracting the contributions of rows i
and i+1 (H_i is unitary on rows >= i).  When the downdated norm falls below 1e-4 of the Gram
diagonal the kernel factorises column i alone.  This file restates that step in numpy and checks it
against the sequential one-column Householder QR and against numpy's QR (|R| is unique up to the
phases of R's rows), on random, scaled, rank-deficient and structured panels.  The GPU kernel itself
is checked end to end by the two-stage tests (test_gpu_twostage.py).
"""
import numpy as np
import pytest


def _refl(alpha, xn2):
    # zlarfg: H^H (alpha; x) = (beta; 0), H = I - tau v v^H, v = (1; scl x)
    if xn2 == 0.0 and alpha.imag == 0.0:
        return 0j, alpha.real, 0j
    nrm = np.sqrt(abs(alpha) ** 2 + xn2)
    beta = -nrm if alpha.real >= 0 else nrm
    return complex((beta - alpha.real) / beta, -alpha.imag / beta), beta, 1.0 / (alpha - beta)


def _apply(a, i, tau, beta, scl, svha):
    v = np.concatenate([[1.0], scl * a[i + 1:, i]])
    for k in range(i + 1, a.shape[1]):
        a[i:, k] -= np.conj(tau) * v * svha[k]
    a[i, i] = beta
    a[i + 1:, i] = v[1:]


def _single(a, i):
    x = a[i + 1:, i]
    tau, beta, scl = _refl(a[i, i], float(np.vdot(x, x).real))
    ssum = np.array([np.vdot(x, a[i + 1:, k]) for k in range(a.shape[1])])
    _apply(a, i, tau, beta, scl, a[i, :] + np.conj(scl) * ssum)
    return tau


def _pair(a, i, guard=1e-4):
    nb = a.shape[1]
    x = a[i + 1:, i]
    ssum1 = np.array([np.vdot(x, a[i + 1:, k]) for k in range(nb)])
    g2 = np.array([np.vdot(a[i:, i + 1], a[i:, k]) for k in range(nb)])   # before H_i
    piv_i, piv_1 = a[i, :].copy(), a[i + 1, :].copy()
    tau, beta, scl = _refl(piv_i[i], float(ssum1[i].real))
    svha = piv_i + np.conj(scl) * ssum1
    svha[: i + 1] = 0
    ri = piv_i - np.conj(tau) * svha                       # row i after H_i
    r1 = piv_1 - np.conj(tau) * (scl * piv_1[i]) * svha    # row i+1 after H_i
    G = g2[i + 1].real
    xn2 = G - abs(ri[i + 1]) ** 2 - abs(r1[i + 1]) ** 2
    if not xn2 >= guard * G:
        return None
    tau2, beta2, scl2 = _refl(r1[i + 1], xn2)
    svha2 = r1 + np.conj(scl2) * (g2 - np.conj(ri[i + 1]) * ri - np.conj(r1[i + 1]) * r1)
    _apply(a, i, tau, beta, scl, svha)
    _apply(a, i + 1, tau2, beta2, scl2, svha2)
    return tau, tau2


def _panel(a, pairs):
    a = a.astype(complex).copy()
    nb = a.shape[1]
    taus, i, npair = np.zeros(nb, complex), 0, 0
    while i < nb:
        r = _pair(a, i) if pairs and i + 1 < nb else None
        if r is not None:
            taus[i], taus[i + 1] = r
            i, npair = i + 2, npair + 1
        else:
            taus[i] = _single(a, i)
            i += 1
    return a, taus, npair


def _q_r(a, taus):
    m, nb = a.shape
    Q = np.eye(m, dtype=complex)
    for i in range(nb):
        v = np.zeros(m, complex)
        v[i] = 1
        v[i + 1:] = a[i + 1:, i]
        Q = Q @ (np.eye(m) - taus[i] * np.outer(v, v.conj()))
    return Q, np.triu(a[:nb, :])


def _panels():
    rng = np.random.default_rng(7)
    m, nb = 90, 32
    z = rng.standard_normal((m, nb)) + 1j * rng.standard_normal((m, nb))
    top = z.copy()
    top[nb:, :] *= 1e-6                          # little weight below the pivots
    zd = z.copy()
    zd[:, 5] = 0                                 # zero column
    zd[:, 9] = zd[:, 8]                          # duplicate column (rank deficient)
    band = np.zeros((m, nb), complex)            # upper bidiagonal: column i+1 zero below row i+1,
    for j in range(nb):                          # so every pair falls back (guard)
        band[max(j - 1, 0): j + 1, j] = rng.standard_normal(min(j, 1) + 1) + 1j * rng.standard_normal(min(j, 1) + 1)
    return {"random": z, "scaled": 1e5 * z, "top_heavy": top, "rank_deficient": zd, "banded": band,
            "real": z.real.astype(complex)}


@pytest.mark.parametrize("name", list(_panels()))
def test_pair_step_is_a_householder_qr(name):
    A = _panels()[name]
    a, taus, npair = _panel(A, pairs=True)
    Q, R = _q_r(a, taus)
    scale = np.abs(A).max()
    assert np.abs(Q.conj().T @ Q - np.eye(A.shape[0])).max() < 1e-13
    assert np.abs(Q[:, : A.shape[1]] @ R - A).max() < 1e-13 * scale
    # |R| is unique (up to row phases) where the panel has full column rank
    if name != "rank_deficient":
        Rn = np.linalg.qr(A, mode="r")
        assert np.abs(np.abs(R) - np.abs(Rn)).max() < 1e-12 * scale
    if name == "banded":
        assert npair == 0          # every pair falls back: nothing below row i+1 in column i+1
    else:
        assert npair >= 1


@pytest.mark.parametrize("name", ["random", "scaled", "top_heavy", "real"])
def test_pair_step_matches_one_column_steps(name):
    A = _panels()[name]
    a2, t2, npair = _panel(A, pairs=True)
    a1, t1, _ = _panel(A, pairs=False)
    # the tiny tail of the top-heavy panel may send a late pair to the guard
    assert npair >= A.shape[1] // 2 - (2 if name == "top_heavy" else 0)
    assert np.abs(t2 - t1).max() < 1e-12
    assert np.abs(a2 - a1).max() < 1e-12 * np.abs(A).max()
