"""Independent pins of the oracle parts that the reference's fixtures leave open (the Prism_111
fixture has no field, no K offset and no InGaAs):

  * field term  (field_fourier_coefficient.c:23-55): the oracle's H(F) - H(0) equals the plane-wave
    matrix elements (1/L) int_{-L/2}^{L/2} F x exp(-i (G1 - G2) x) dx of the sawtooth potential,
    evaluated by adaptive quadrature, on every band's diagonal block and nowhere else; and the
    Stark shift dE/dF_z of a dot level (central differences of the oracle's eigenvalues) equals
    the level's dipole <z> from its real-space envelope (Hellmann-Feynman).
  * kinetic and K terms (hamiltonian_8_band_matrix_elements.c:27-969): with no shapes H is block
    diagonal in G, and each 8 x 8 block at k = G + K must have the eigenvalues of the bulk 8-band
    Kane model at k, written here independently in the Cartesian basis |S>, |X>, |Y>, |Z> x spin
    with the spin-orbit coupling (Delta/3) L.sigma built from the angular-momentum matrices.
  * InGaAs (material 3 with y = 0): renormalised s_c and Luttinger parameters, Delta_so, average
    valence energy, band edges and gap at x = 0.2 and 0.6 by hand from the reference's data table
    (code/source/parameters/material_parameters.h: GaAs and InAs columns, GaInAs bowings) and its
    formulas (material_parameters.c:356-400, 1063-1075, 1526-1540).
"""
import math
import os

import numpy as np
import pytest
from scipy.integrate import quad

from helpers import ROOT, oracle_problem, prism_paths
from oracle import qd_oracle as qo

HB2 = 0.5 * qo.HBSQM     # hbar^2 / 2 m0, eV Angstrom^2
LENS = (os.path.join(ROOT, "qudos_amd", "configs", "lens_9x9x9", "parameters.in"),
        os.path.join(ROOT, "qudos_amd", "configs", "lens_9x9x9", "supercell.in"))


# ---------------------------------------------------------------------------------------------- field
def _sawtooth_element(F, L, dm):
    """(1/L) int_{-L/2}^{L/2} F x exp(-i 2 pi dm x / L) dx by adaptive quadrature."""
    k = 2.0 * math.pi * dm / L
    re = quad(lambda x: F * x * math.cos(k * x), -L / 2, L / 2, limit=400, epsabs=1e-14, epsrel=1e-13)[0]
    im = quad(lambda x: -F * x * math.sin(k * x), -L / 2, L / 2, limit=400, epsabs=1e-14, epsrel=1e-13)[0]
    return (re + 1j * im) / L


@pytest.mark.filterwarnings("ignore::scipy.integrate.IntegrationWarning")
@pytest.mark.parametrize("F", [(0.002, 0.0, 0.0), (0.0, -0.0013, 0.0), (0.0, 0.0, 0.0007), (0.001, -0.002, 0.0015)])
def test_field_term_is_the_sawtooth_transform(F):
    base = ("-M_Gx=2", "-M_Gy=3", "-M_Gz=2", "-M_S_Gx=2", "-M_S_Gy=3", "-M_S_Gz=2")
    pp, sp = prism_paths()
    p0, sh0 = oracle_problem(pp, sp, base)
    H0 = qo.build_hamiltonian(qo.coefficient_tables(p0, sh0))
    # the override takes kV/cm: 1 V/Angstrom = 1e5 kV/cm
    pF, shF = oracle_problem(pp, sp, base + tuple(f"-F{a}={1e5 * v}" for a, v in zip("xyz", F)))
    dH = qo.build_hamiltonian(qo.coefficient_tables(pF, shF)) - H0
    N = p0.N
    NG = N[0] * N[1] * N[2]
    L = (sh0[0].L1, sh0[0].L2, sh0[0].L3)
    m = np.array(np.unravel_index(np.arange(NG), N)).T      # plane-wave index triples, row order
    want = np.zeros((NG, NG), dtype=complex)
    cache = {}
    for i in range(NG):
        for j in range(NG):
            d = m[i] - m[j]
            for a in range(3):
                o = [c for c in range(3) if c != a]
                if d[a] != 0 and d[o[0]] == 0 and d[o[1]] == 0 and F[a] != 0.0:
                    key = (a, int(d[a]))
                    if key not in cache:
                        cache[key] = _sawtooth_element(F[a], L[a], d[a])
                    want[i, j] += cache[key]
    scale = max(abs(v) for v in cache.values())
    for br in range(8):
        for bc in range(8):
            blk = dH[br * NG:(br + 1) * NG, bc * NG:(bc + 1) * NG]
            ref = want if br == bc else 0.0
            assert np.abs(blk - ref).max() <= 1e-10 * scale, (F, br, bc)


def _envelope_dipole_z(Z, p, sh, nr=(24, 24, 32)):
    """<z> of unit-norm plane-wave states from |psi(r)|^2 on a real-space grid of the cell,
    z measured from the cell centre in [-L/2, L/2)."""
    N = p.N
    NG = N[0] * N[1] * N[2]
    L = (sh[0].L1, sh[0].L2, sh[0].L3)
    G = [qo.basis_1d(N[a], L[a])[(2 * N[a] - 1) // 2 - N[a] // 2:(2 * N[a] - 1) // 2 - N[a] // 2 + N[a]]
         for a in range(3)]
    r = [-L[a] / 2 + (np.arange(nr[a]) + 0.5) * L[a] / nr[a] for a in range(3)]
    E = [np.exp(1j * np.outer(r[a], G[a])) for a in range(3)]        # (points, plane waves)
    out = []
    for c in range(Z.shape[1]):
        dens = np.zeros(nr)
        for b in range(8):
            cb = Z[b * NG:(b + 1) * NG, c].reshape(N)
            psi = np.einsum("xi,yj,zk,ijk->xyz", E[0], E[1], E[2], cb, optimize=True)
            dens += np.abs(psi) ** 2
        zc = r[2][None, None, :]
        out.append(float(np.sum(dens * zc) / np.sum(dens)))
    return np.array(out)


def test_stark_shift_equals_envelope_dipole():
    base = ("-M_Gx=3", "-M_Gy=3", "-M_Gz=4", "-M_S_Gx=3", "-M_S_Gy=3", "-M_S_Gz=4", "-N_VB=2", "-N_CB=2")
    F = 0.5     # kV/cm
    Es = {}
    for sign in (-1, 0, 1):
        p, sh = oracle_problem(*LENS, base + (f"-Fz={sign * F}",))
        w, Z = qo.solve(qo.build_hamiltonian(qo.coefficient_tables(p, sh)), p)
        Es[sign] = (w, Z, p, sh)
    w0, Z0, p, sh = Es[0]
    dEdF = (Es[1][0] - Es[-1][0]) / (2 * F * 1e-5)                 # eV per (V / Angstrom)
    dip = _envelope_dipole_z(Z0, p, sh)
    nvb = 2 * p.N_vb
    for lv in (nvb - 2, nvb):                                        # h1 and e1 Kramers pairs
        g = 0.5 * (dEdF[lv] + dEdF[lv + 1])
        d = 0.5 * (dip[lv] + dip[lv + 1])
        assert abs(d) > 1.0                                          # the lens sits off the cell centre
        assert abs(g - d) <= 2e-3 * abs(d), (lv, g, d)


# ---------------------------------------------------------------------------------------------- bulk
def _angular_momentum_p():
    """L_a on the real p orbitals (X, Y, Z): (L_a)_{bc} = -i eps_{abc} (units of hbar)."""
    eps = np.zeros((3, 3, 3))
    for a, b, c in [(0, 1, 2), (1, 2, 0), (2, 0, 1)]:
        eps[a, b, c], eps[a, c, b] = 1.0, -1.0
    return [-1j * eps[a] for a in range(3)]


def kane_bulk(k, Ec, Ev, Dso, sc, g1, g2, g3, P):
    """8-band Kane Hamiltonian in the basis (S, X, Y, Z) x (up, down)."""
    kx, ky, kz = k
    A = HB2 * sc
    Lp = -HB2 * (g1 + 4 * g2)
    M = -HB2 * (g1 - 2 * g2)
    Np = -HB2 * 6 * g3
    Evp = Ev - Dso / 3.0
    kk = np.array(k)
    h = np.zeros((4, 4), dtype=complex)
    h[0, 0] = Ec + A * kk @ kk
    for a in range(3):
        h[0, 1 + a] = 1j * P * k[a]
        h[1 + a, 0] = -1j * P * k[a]
        h[1 + a, 1 + a] = Evp + Lp * k[a] ** 2 + M * (kk @ kk - k[a] ** 2)
        for b in range(3):
            if b != a:
                h[1 + a, 1 + b] = Np * k[a] * k[b]
    H = np.kron(np.eye(2), h)                                        # spin (outer) x orbital (inner)
    sig = [np.array([[0, 1], [1, 0]], complex), np.array([[0, -1j], [1j, 0]]), np.array([[1, 0], [0, -1]], complex)]
    Lm = _angular_momentum_p()
    so = np.zeros((8, 8), dtype=complex)
    for a in range(3):
        Lfull = np.zeros((4, 4), dtype=complex)
        Lfull[1:, 1:] = Lm[a]
        so += np.kron(sig[a], Lfull)
    return H + (Dso / 3.0) * so


def test_kane_spin_orbit_splitting():
    """(Delta/3) L.sigma on the p states: j = 3/2 at +Delta/3 (4 states), j = 1/2 at -2 Delta/3 (2)."""
    w = np.linalg.eigvalsh(kane_bulk((0, 0, 0), 1.0, 0.0, 0.3, 0, 0, 0, 0, 0))
    assert np.allclose(np.sort(w), [-0.3, -0.3, 0, 0, 0, 0, 1.0, 1.0], atol=1e-14)


@pytest.mark.parametrize("K", [(0.0, 0.0, 0.0), (0.013, -0.007, 0.021), (-0.03, 0.02, 0.0)])
@pytest.mark.parametrize("material", ["1.0", "0.2"])
def test_bulk_blocks_follow_the_kane_dispersion(K, material):
    over = ("-N_shape=1", "-Strain=0", "-Piezo=0", "-M_Gx=2", "-M_Gy=2", "-M_Gz=3", "-M_S_Gx=2", "-M_S_Gy=2",
            "-M_S_Gz=3") + tuple(f"-K{a}={10 * v}" for a, v in zip("xyz", K))   # override in 1/nm
    p, sh = oracle_problem(*LENS, over)
    sh[0].x = float(material)                                        # GaAs, or In0.8Ga0.2As bulk
    t = qo.coefficient_tables(p, sh)
    H = qo.build_hamiltonian(t)
    m0 = t.params[0]
    assert abs(m0["E_hh"] - m0["E_lh"]) < 1e-15
    N = p.N
    NG = N[0] * N[1] * N[2]
    G = [t.G[a][(2 * N[a] - 1) // 2 - N[a] // 2:(2 * N[a] - 1) // 2 - N[a] // 2 + N[a]] for a in range(3)]
    idx = np.array(np.unravel_index(np.arange(NG), N)).T
    off = H.copy()
    for g in range(NG):
        rows = [b * NG + g for b in range(8)]
        blk = H[np.ix_(rows, rows)]
        off[np.ix_(rows, rows)] = 0
        k = [G[a][idx[g, a]] + p.K[a] for a in range(3)]
        ref = np.linalg.eigvalsh(kane_bulk(k, m0["E_cb"], m0["E_hh"], m0["E_hh"] - m0["E_so"], m0["sc"], m0["g1"],
                                           m0["g2"], m0["g3"], m0["P"]))
        got = np.linalg.eigvalsh(blk)
        assert np.abs(got - ref).max() <= 1e-11 * max(1.0, np.abs(ref).max()), (g, k)
    assert np.abs(off).max() == 0.0                                  # no coupling between plane waves


# ------------------------------------------------------------------------------------------- InGaAs
# GaAs / InAs columns of the reference's table: Eg(0 K), Dso, mc, Ep, gL1..3, Ev_avg, av, a, Varshni
GAAS = dict(Eg=1.518, Dso=0.340, mc=0.0667, Ep=25.7, gL1=7.1, gL2=2.02, gL3=2.91, Ev=0.000, av=0.220, a=5.6503,
            alpha=0.5405e-3, beta=204.0)
INAS = dict(Eg=0.405, Dso=0.380, mc=0.02226, Ep=21.846, gL1=19.67, gL2=8.40, gL3=9.30, Ev=0.173, av=1.000, a=6.0553,
            alpha=0.276e-3, beta=93.0)
BOW_EG, BOW_DSO = 0.477, 0.150      # GaInAs bowings (bow_Eg_tern[4], bow_Dso_tern[4])

# Hand evaluation (x = Ga fraction, T = 300 K, GaAs matrix):
#   s_c(bin)  = 1/mc - Ep/3 (2/Eg0 + 1/(Eg0 + Dso));   g1 = gL1 - Ep/(3 Eg0 + Dso);
#   g2 = gL2 - Ep/(6 Eg0 + 2 Dso);   g3 = gL3 - Ep/(6 Eg0 + 2 Dso)       (binaries, then linear in x)
#   GaAs: s_c = 14.9925037 - 8.5666667 (1.3175231 + 0.5382131) = -0.9049696
#         g1 = 7.1 - 25.7/4.894 = 1.8486718;  g2 = 2.02 - 25.7/9.788 = -0.6056641;  g3 = 0.2843359
#   InAs: s_c = 44.9236298 - 7.282 (4.9382716 + 1.2738854) = -0.3132971
#         g1 = 19.67 - 21.846/1.595 = 5.9734483;  g2 = 8.40 - 21.846/3.19 = 1.5517241;  g3 = 2.4517241
#   Delta_so(x) = 0.34 x + 0.38 (1-x) - 0.15 x (1-x)
#   Ev_avg(x)   = 0.173 (1-x) - x (1-x) b_v,  b_v = (3/5.6503)(0.22 - 1.00)(5.6503 - 6.0553) = 0.1677256
#   E_hh = Ev_avg + Dso/3 - (0 + 0.34/3);  E_so = E_hh - Dso;  E_cb = E_hh + Eg
#   Eg(x, 300 K) = x (1.518 - 0.5405e-3 9e4/504) + (1-x)(0.405 - 0.276e-3 9e4/393) - 0.477 x (1-x)
HAND = {
    0.2: dict(sc=-0.43163162, g1=5.14849299, g2=1.12024649, g3=2.01824649, Dso=0.348, Ev_avg=0.11156390,
              E_hh=0.11423057, E_so=-0.23376943, Eg=0.48141154),
    0.6: dict(sc=-0.66830061, g1=3.49858242, g2=0.25729121, g3=1.15129121, Dso=0.320, Ev_avg=0.02894585,
              E_hh=0.02227919, E_so=-0.29772081, Eg=0.87512684),
}


@pytest.mark.parametrize("x", [0.2, 0.6])
def test_ingaas_interpolation_hand_values(x):
    sh = [qo.Shape(material=3, x=1.0), qo.Shape(material=3, x=x)]
    d = qo.material_parameters(sh, 300.0, 2)[1]
    want = HAND[x]
    for key, v in want.items():
        assert abs(d[key] - v) < 5e-8, (x, key, d[key], v)
    assert abs(d["E_cb"] - (want["E_hh"] + want["Eg"])) < 1e-7
    assert abs(d["E_lh"] - d["E_hh"]) < 1e-15
    # the tabulated binaries themselves
    for frac, col in ((1.0, GAAS), (0.0, INAS)):
        b = qo.material_parameters([qo.Shape(material=3, x=1.0), qo.Shape(material=3, x=frac)], 300.0, 2)[1]
        sc = 1 / col["mc"] - col["Ep"] / 3 * (2 / col["Eg"] + 1 / (col["Eg"] + col["Dso"]))
        assert abs(b["sc"] - sc) < 1e-12 and abs(b["Dso"] - col["Dso"]) < 1e-15
        assert abs(b["P"] - math.sqrt(0.5 * qo.HBSQM * col["Ep"])) < 1e-13
