"""Carrier localisation and overlaps on the CPU: the oracle's restatement of carrier_localisation.c
(padded FFT convolution per band and shape) pinned by a real-space brute-force integral of
|psi(r)|^2 over the shape and by shapes that tile the cell summing to the norm; the library's
formula (the state density rho(q) contracted with chi_s(q), the GPU path's algebra) against it,
with chi_s from the library's own shape code."""
import numpy as np
import pytest

import qudos_amd as q
from helpers import prism_paths
from oracle import qd_oracle as qo


def _cell_with(shapes_text, N=(5, 5, 5)):
    text = """QD shape: 0
Material: 0
Compos x: 1.0
Compos y: 0.0
Length 1: 10.0
Length 2: 10.0
Length 3: 8.0
Strain length 1: 1
Strain length 2: 1
Strain length 3: 1
Coulomb length 1: 10.0
Coulomb length 2: 10.0
Coulomb length 3: 8.0
""" + shapes_text
    pp, _ = prism_paths()
    p = qo.read_parameters(open(pp).read() + f"\nN_shape: {1 + shapes_text.count('QD shape')}\n"
                           f"M_Gx: {N[0] // 2}\nM_Gy: {N[1] // 2}\nM_Gz: {N[2] // 2}\n")
    return p, qo.read_supercell(text, p), text


def _shape(kind, L, centre=(0.0, 0.0, 0.0), emb=0):
    return f"""Embedded: {emb}
QD shape: {kind}
Material: 0
Compos x: 0.0
Compos y: 0.0
Centre x: {centre[0]}
Centre y: {centre[1]}
Centre z: {centre[2]}
Length 1: {L[0]}
Length 2: {L[1]}
Length 3: {L[2]}
Length 4: 0.0
Length 5: 0.0
Euler alpha angle: 0.0
Euler beta angle: 0.0
Euler gamma angle: 0.0
"""


def _chis(shapes, N):
    G = [qo.basis_1d(N[a], [shapes[0].L1, shapes[0].L2, shapes[0].L3][a]) for a in range(3)]
    return [None] + [qo.chi(shapes, s, *qo.grid3(G), 0) for s in range(1, len(shapes))]


def _states(N, k, seed):
    rng = np.random.default_rng(seed)
    Z = rng.standard_normal((8 * np.prod(N), k)) + 1j * rng.standard_normal((8 * np.prod(N), k))
    return Z / np.linalg.norm(Z, axis=0)


def test_localisation_of_a_sphere_vs_real_space():
    N = (5, 5, 5)
    p, shapes, _ = _cell_with(_shape(1, (3.1, 0, 0), centre=(0.7, -0.4, 0.2)), N)
    Z = _states(N, 3, 1)
    L = qo.carrier_localisation(_chis(shapes, N), shapes, N, Z)
    # |psi(r)|^2 on a fine grid, psi_b(r) = sum_G c_b(G) e^{iG.r}; fraction = (1/V) int_s |psi|^2
    Lc = (100.0, 100.0, 80.0)
    n = 96
    xs = [(np.arange(n) + 0.5) / n * Lc[a] - Lc[a] / 2 for a in range(3)]
    G = [qo.basis_1d(N[a], Lc[a])[N[a] // 2: N[a] // 2 + N[a]] for a in range(3)]
    E = [np.exp(1j * np.outer(xs[a], G[a])) for a in range(3)]
    inside = ((xs[0][:, None, None] - 7.0) ** 2 + (xs[1][None, :, None] + 4.0) ** 2
              + (xs[2][None, None, :] - 2.0) ** 2) <= 31.0 ** 2
    NG = np.prod(N)
    for col in range(Z.shape[1]):
        dens = np.zeros((n, n, n))
        for b in range(8):
            c = Z[b * NG:(b + 1) * NG, col].reshape(N)
            psi = np.einsum("ijk,xi,yj,zk->xyz", c, E[0], E[1], E[2], optimize=True)
            dens += np.abs(psi) ** 2
        brute = dens[inside].sum() / dens.size
        assert abs(L[col, 1] - brute) < 1e-2 * brute, (L[col, 1], brute)
        assert abs(L[col, 0] - (1.0 - L[col, 1])) < 1e-15


def test_localisation_of_shapes_that_tile_the_cell():
    N = (3, 5, 3)
    # two cuboids splitting the cell at z = 0.5 nm: their localisations add to 1 for any state
    text = _shape(0, (10.0, 10.0, 4.5), centre=(0, 0, -1.75)) + _shape(0, (10.0, 10.0, 3.5), centre=(0, 0, 2.25))
    p, shapes, _ = _cell_with(text, N)
    L = qo.carrier_localisation(_chis(shapes, N), shapes, N, _states(N, 4, 2))
    assert np.abs(L[:, 1] + L[:, 2] - 1.0).max() < 1e-12
    assert np.abs(L[:, 0]).max() < 1e-12
    assert L[:, 1].min() > 0.05 and L[:, 2].min() > 0.05


def test_library_formula_matches_oracle(tmp_path):
    """sum_q chi_s(q) conj(rho(q)), the GPU path's contraction, with chi_s from the library."""
    N = (5, 5, 3)
    text = _shape(1, (2.5, 0, 0), centre=(1.0, 0.0, 0.0)) + _shape(9, (2.0, 3.0, 0.0), centre=(-2.0, 1.0, -0.5), emb=1)
    p, shapes, full = _cell_with(text, N)
    sp = tmp_path / "supercell.in"
    sp.write_text(full)
    pp, _ = prism_paths()
    inp = q.read_inputs(pp, str(sp), ("-N_shape=3", f"-M_Gx={N[0] // 2}", f"-M_Gy={N[1] // 2}", f"-M_Gz={N[2] // 2}"))
    Z = _states(N, 3, 4)
    ref = qo.carrier_localisation(_chis(shapes, N), shapes, N, Z)
    NT = tuple(2 * n - 1 for n in N)
    G = [qo.basis_1d(N[a], [100.0, 100.0, 80.0][a]) for a in range(3)]
    gx, gy, gz = qo.grid3(G)
    Gv = np.stack([gx.ravel(), gy.ravel(), gz.ravel()], axis=1)
    NG = np.prod(N)
    for col in range(Z.shape[1]):
        rho = np.zeros(NT, complex)        # rho(q) = sum_b sum_G conj(psi(G)) psi(G + q)
        for b in range(8):
            c = Z[b * NG:(b + 1) * NG, col].reshape(N)
            cp = np.zeros(tuple(2 * NT[a] for a in range(3)), complex)
            cp[:N[0], :N[1], :N[2]] = c
            f = np.fft.fftn(cp)
            corr = np.fft.ifftn(np.conj(f) * f)           # sum_g conj(c(g)) c(g + q), q mod 2 NT
            for ix in range(NT[0]):
                for iy in range(NT[1]):
                    for iz in range(NT[2]):
                        rho[ix, iy, iz] += corr[(ix - (N[0] - 1)) % (2 * NT[0]), (iy - (N[1] - 1)) % (2 * NT[1]),
                                                (iz - (N[2] - 1)) % (2 * NT[2])]
        for s in (1, 2):
            chi = q.selftest_shape_host(inp, s, Gv).reshape(NT)
            got = np.sum(chi * np.conj(rho)).real
            assert abs(got - ref[col, s]) < 1e-13, (col, s, got, ref[col, s])
