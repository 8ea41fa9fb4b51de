"""Real-space envelope functions (envelope_functions.c) and their grids on the CPU.

  * The oracle's linescan, cross-section and 3-D grids reproduce the positions the reference
    wrote into the Prism_111 example's probability-density files (tests/golden/prism_111/
    envelopes.npz, made by tests/golden/make_envelope_fixture.py): columns x, y, z [nm].
  * Parseval: on a periodic grid of n >= N points per axis over one cell, the summed |F_b|^2
    times the cell volume over the point count equals the band's weight sum_G |c_b(G)|^2 (pins
    the 1/sqrt(volume) normalisation and that every coefficient meets a distinct plane wave).
  * A single plane wave in the coefficient vector gives exp(i (G + K) . r) / sqrt(volume) at
    every point (pins the sign of the exponent, the K shift and the row order of the bands and
    wave vectors).
  * Slow: the Prism_111 example solved by the oracle (its own tables, LAPACK zheevr) gives the
    reference's probability densities on every stored grid, summed over each Kramers pair
    (the only combination the reference's arbitrary eigenvector gauge leaves fixed).
"""
import math
import os

import numpy as np
import pytest

from helpers import GOLDEN, prism_oracle_solution
from oracle import qd_oracle as qo

FIX = os.path.join(GOLDEN, "prism_111", "envelopes.npz")
L_PRISM = (200.0, 200.0, 200.0)
LINES = {"001": 2, "110": 3, "111": 6}
CROSS = {"-110": 3, "111": 7}


def _grids(fix):
    """(name, x, y, z, rows of the fixture) for every stored grid, points in the fixture's order."""
    out = []
    for name, d in LINES.items():
        x, y, z = qo.linescan_grid(d, 300, L_PRISM)
        out.append((f"line_{name}", x, y, z, fix[f"line_{name}"]))
    for name, d in CROSS.items():
        x, y, z = qo.cross_section_grid(d, 40, L_PRISM)
        idx = fix[f"cross_{name}_index"]
        sel = idx[:, 0] * 40 + idx[:, 1]
        out.append((f"cross_{name}", x[sel], y[sel], z[sel], fix[f"cross_{name}"]))
    x, y, z = qo.grid_3d(20, L_PRISM)
    idx = fix["grid3d_index"]
    sel = (idx[:, 0] * 20 + idx[:, 1]) * 20 + idx[:, 2]
    out.append(("grid3d", x[sel], y[sel], z[sel], fix["grid3d"]))
    return out


def test_grids_match_reference_positions():
    fix = np.load(FIX)
    for name, x, y, z, rows in _grids(fix):
        got = 0.1 * np.stack([x, y, z], axis=1)
        ref = rows[0][:, :3]
        # the files hold 7 significant digits
        assert np.abs(got - ref).max() <= 6e-7 * max(1.0, np.abs(ref).max()), name


def test_grids_other_directions():
    L = (120.0, 90.0, 60.0)
    for d in range(7):
        x, y, z = qo.linescan_grid(d, 11, L)
        used = [np.ptp(v) > 0 for v in (x, y, z)]
        assert sum(used) == (1 if d < 3 else (2 if d < 6 else 3))
        if d >= 3:   # diagonals take the shortest axis' extent on every axis they use
            spans = [np.ptp(v) for v, u in zip((x, y, z), used) if u]
            assert np.allclose(spans, min(L[a] for a in range(3) if used[a]))
    for d in range(7):
        x, y, z = qo.cross_section_grid(d, 7, L)
        P = np.stack([x, y, z], axis=1) - np.array([x, y, z]).mean(axis=1)
        assert np.linalg.matrix_rank(P, tol=1e-9) == 2       # a plane through the centre
    x, y, z = qo.grid_3d(5, L)
    assert np.allclose([x.min(), y.min(), z.min()], [-60, -45, -30]) and np.allclose([x.max(), y.max(), z.max()], [60, 45, 30])


def test_parseval_on_a_periodic_grid():
    rng = np.random.default_rng(3)
    N, L, K = (3, 5, 4), (70.0, 90.0, 50.0), (0.01, -0.02, 0.03)
    NG = N[0] * N[1] * N[2]
    Z = rng.standard_normal((8 * NG, 2)) + 1j * rng.standard_normal((8 * NG, 2))
    n = (5, 7, 6)
    g = np.meshgrid(*[np.arange(n[a]) * L[a] / n[a] for a in range(3)], indexing="ij")
    F = qo.envelopes(Z, N, L, K, g[0].ravel(), g[1].ravel(), g[2].ravel())
    vol = L[0] * L[1] * L[2]
    got = (np.abs(F) ** 2).sum(axis=2) * vol / F.shape[2]
    want = np.stack([(np.abs(Z[b * NG:(b + 1) * NG]) ** 2).sum(axis=0) for b in range(8)], axis=1)
    assert np.abs(got - want).max() <= 1e-12 * want.max()


def test_single_plane_wave():
    N, L, K = (3, 3, 5), (60.0, 80.0, 40.0), (0.02, 0.0, -0.01)
    NG = N[0] * N[1] * N[2]
    band, (mx, my, mz) = 5, (2, 0, 3)
    Z = np.zeros((8 * NG, 1), complex)
    Z[band * NG + (mx * N[1] + my) * N[2] + mz, 0] = 1.0
    rng = np.random.default_rng(1)
    x, y, z = (rng.uniform(-30, 30, 9) for _ in range(3))
    F = qo.envelopes(Z, N, L, K, x, y, z)
    G = [2 * math.pi / L[a] * (m - (N[a] - 1) / 2) + K[a] for a, m in enumerate((mx, my, mz))]
    want = np.exp(1j * (G[0] * x + G[1] * y + G[2] * z)) / math.sqrt(L[0] * L[1] * L[2])
    assert np.abs(F[0, band] - want).max() < 1e-15
    assert np.abs(np.delete(F[0], band, axis=0)).max() == 0.0


def pair_sums(dens):
    """(states, points, columns) -> Kramers-pair sums (pairs, points, columns)."""
    return dens.reshape(dens.shape[0] // 2, 2, *dens.shape[1:]).sum(axis=1)


def fixture_order(F, nvb, grid3d=False):
    """Probability-density columns of every window state in the files' order (valence states
    highest first, then conduction states lowest first); the 3-D files of conduction states carry
    the total density times 10 instead of 1000, as the reference writes them."""
    P = qo.probability_columns(F)
    Pc = qo.probability_columns(F[nvb:], total_scale=10.0) if grid3d else P[nvb:]
    return np.concatenate([P[nvb - 1::-1], Pc])


@pytest.mark.slow
def test_prism_fixture_probability_densities():
    p, sh, t, w, Z = prism_oracle_solution()
    fix = np.load(FIX)
    nvb = 2 * p.N_vb
    for name, x, y, z, rows in _grids(fix):
        F = qo.envelopes(Z, tuple(p.N), L_PRISM, tuple(p.K), x, y, z)
        got = pair_sums(fixture_order(F, nvb, grid3d=name == "grid3d"))
        ref = pair_sums(rows[:, :, 4:])
        scale = ref.max(axis=(1, 2))[:, None, None]   # relative to the pair's largest density
        err = (np.abs(got - ref) / scale).max()
        assert err < 2e-6, (name, err)   # the files hold 7 significant digits
        energies = np.concatenate([w[nvb - 1::-1], w[nvb:]])
        assert np.abs(rows[:, 0, 3] - energies).max() < 1e-5   # column 4 is the state's energy


def driver_file_errors(path, nvb=8, ncb=4):
    """Largest relative error of the Kramers-pair sums of the probability densities that the
    qd_pwem driver wrote into `path` for the Prism_111 example, on every grid stored in the
    fixture, with the positions and energies of the files checked as well."""
    fix = np.load(FIX)
    names = [f"vb_{i}" for i in range(1, nvb + 1)] + [f"cb_{i}" for i in range(1, ncb + 1)]
    files = {"line_001": ("", "_001.dat", None), "line_110": ("", "_110.dat", None), "line_111": ("", "_111.dat", None),
             "cross_-110": ("_cross_section", "_-110.dat", (fix["cross_-110_index"], 40)),
             "cross_111": ("_cross_section", "_111.dat", (fix["cross_111_index"], 40)),
             "grid3d": ("_3D", ".dat", (fix["grid3d_index"], 20))}
    errs = {}
    for key, (kind, suffix, sub) in files.items():
        got = np.stack([np.loadtxt(os.path.join(path, f"probability_density{kind}_{n}{suffix}"), ndmin=2) for n in names])
        if sub is not None:
            idx, n = sub
            flat = idx[:, 0] * n + idx[:, 1] if idx.shape[1] == 2 else (idx[:, 0] * n + idx[:, 1]) * n + idx[:, 2]
            got = got[:, flat]
        ref = fix[key]
        assert got.shape == ref.shape, key
        assert np.abs(got[:, :, :4] - ref[:, :, :4]).max() <= 1e-5 * max(1.0, np.abs(ref[:, :, :4]).max()), key
        g, r = pair_sums(got[:, :, 4:]), pair_sums(ref[:, :, 4:])
        errs[key] = float((np.abs(g - r) / r.max(axis=(1, 2))[:, None, None]).max())
    return errs

