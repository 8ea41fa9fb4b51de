"""FP64 CPU oracle for the QUDOS single-particle solve (TEST INFRASTRUCTURE ONLY).

This module is a numpy restatement of the reference's coefficient pipeline, Hamiltonian
assembly and windowed eigensolve.  It exists so that tests, ``__graft_entry__.smoke()`` and
``bench.py``'s CPU-baseline leg can check the MI355X library against an independent
implementation.  Nothing in the shipped library imports it.

Every function names the reference file:line it restates (paths relative to the QUDOS
repository root ``code/source/8_band``).  Parity pins (what fixes each function independently
of this file) are listed in the module-level table ``PINS`` and exercised in ``tests/``.

Units follow the reference: lengths in Angstrom, energies in eV, fields in V/Angstrom.
"""
from __future__ import annotations

import math
import re
from dataclasses import dataclass, field

import numpy as np

HBSQM = 7.61996      # 8_band.h:51
ESQEPS = 180.9513    # 8_band.h:52
EPS = 1.0e-08        # 8_band.h:59

# Which independent check pins each part (see tests/test_oracle_*.py).
PINS = {
    "materials": "hand values from the published tables + binary identities; InGaAs s_c, gamma_1..3, "
                 "Delta_so, Ev_avg, band edges and gap at x = 0.2, 0.6 by hand (test_oracle_physics_pins)",
    "field": "H(F) - H(0) == quadrature of the sawtooth transform; Stark shift == envelope dipole",
    "kinetic_K_terms": "no-shape blocks at k = G + K == independent Cartesian 8-band Kane model",
    "chi_closed_forms": "G=0 volumes, limits between branches, brute-force real-space quadrature",
    "chi_quadrature": "scipy.integrate.quad on the same integrands; the reference's QAG (oracle/quadpack.py, "
                      "pinned by the Kronrod rule's exactness and closed-form integrals) agrees to rounding",
    "strain": "isotropic special case, small-G continuity, trace identity",
    "convolution": "FFT convolution == direct convolution on small grids",
    "hamiltonian+solve": "Prism_111 PWEM_test_ref energies (1e-9 eV)",
    "commensurate_downsampling": "FFT path == free-length box projection on a commensurate geometry "
                                 "(strain) / box transform (Coulomb); Truncated_pyramid energies on the GPU",
    "envelopes": "Prism_111 probability-density files (linescans, cross-sections, 3-D grid: positions and Kramers-"
                 "pair sums) through eigenvectors solved on the host; Parseval sum over a full periodic grid",
    "localisation": "column sums (shapes that tile the cell add to 1), brute-force |psi(r)|^2 "
                    "integrals over the shape; Truncated_pyramid files through the library's eigenvectors",
    "momentum": "H exactly quadratic in K (third differences vanish); <n|dH/dK|n> == dE_n/dK "
                "(Hellmann-Feynman); Prism_111 / Truncated_pyramid momentum files on the GPU",
}

# ---------------------------------------------------------------------------------------------
# Material data: published values (Vurgaftman et al., JAP 89, 5815 (2001); Caro et al.,
# PRB 91, 075203 (2015)) as tabulated in code/source/parameters/material_parameters.h:137-355.
# Binary order: AlN AlP AlAs AlSb AlBi GaN GaP GaAs GaSb GaBi InN InP InAs InSb InBi
# ---------------------------------------------------------------------------------------------
BIN = {
    "a":   [4.3800, 5.4672, 5.6611, 6.1355, 6.3417, 4.5000, 5.4505, 5.6503, 6.0959, 6.3240,
            4.9800, 5.8697, 6.0553, 6.4794, 6.6860],
    "C11": [304.00, 133.00, 125.00, 87.69, 72.20, 293.00, 140.50, 118.8, 88.42, 73.00,
            187.00, 101.10, 83.3, 68.47, 58.50],
    "C12": [160.00, 63.00, 53.40, 43.41, 36.20, 159.00, 62.03, 53.8, 40.26, 32.70,
            125.00, 56.10, 45.3, 37.35, 30.40],
    "C44": [193.00, 61.50, 54.20, 40.76, 33.40, 155.00, 70.33, 59.4, 43.22, 36.30,
            86.00, 45.60, 39.6, 31.11, 27.50],
    "e14": [0.548, 0.014, -0.055, -0.094, -0.094, 0.366, -0.121, -0.160, -0.216, -0.216,
            0.593, 0.016, -0.045, -0.161, -0.161],
    "e14_calc": [0.548, 0.014, -0.055, -0.094, -0.094, 0.366, -0.121, -0.230, -0.216, -0.216,
                 0.593, 0.016, -0.115, -0.161, -0.161],
    "B114": [-6.810, -2.020, -1.610, -0.760, -0.760, -5.380, -1.230, -0.439, -0.310, -0.310,
             -5.960, -1.540, -0.531, -0.620, -0.620],
    "B124": [-5.040, -2.760, -2.590, -1.990, -1.990, -6.730, -3.270, -3.765, -2.770, -2.770,
             -6.320, -3.620, -4.076, -4.040, -4.040],
    "B156": [-4.150, -1.430, -1.320, -0.820, -0.820, -3.180, -1.380, -0.492, -0.700, -0.700,
             -2.000, -1.020, -0.120, -0.160, -0.160],
    "ac": [-6.000, -5.700, -5.640, -4.500, -4.500, -2.200, -8.200, -8.013, -7.500, -7.500,
           -1.850, -6.000, -5.080, -6.940, -6.940],
    "av": [3.400, 3.000, 2.470, 1.400, 1.400, 5.200, 1.700, 0.220, 0.800, 0.800,
           1.500, 0.600, 1.000, 0.360, 0.360],
    "b": [-1.900, -1.500, -2.300, -1.350, -1.350, -2.200, -1.600, -1.824, -2.000, -2.000,
          -1.200, -2.000, -1.800, -2.000, -2.000],
    "d": [-10.00, -4.600, -3.400, -4.300, -4.300, -3.400, -4.600, -5.062, -4.700, -4.700,
          -9.300, -5.000, -3.600, -4.700, -4.700],
    "Ev_avg": [-2.640, -0.940, -0.530, 0.3900, 1.1000, -1.840, -0.470, 0.000, 0.7700, 1.1000,
               -1.580, -0.140, 0.173, 0.8000, 1.1000],
    "Eg": [4.9000, 3.6300, 3.0990, 2.3860, 0.042, 3.2990, 2.8860, 1.5180, 0.8120, -1.450,
           1.9400, 1.4240, 0.4050, 0.2350, -1.630],
    "Dso": [0.0190, 0.0700, 0.2800, 0.6760, 2.1240, 0.0170, 0.0800, 0.3400, 0.7600, 2.1500,
            0.0060, 0.1080, 0.3800, 0.8100, 2.1500],
    "mc": [0.2500, 0.2200, 0.1500, 0.1400, 99.999, 0.1500, 0.1300, 0.0667, 0.0390, 99.999,
           0.1200, 0.0795, 0.02226, 0.0135, 99.999],
    "Ep": [27.100, 17.700, 21.100, 18.700, 99.999, 25.000, 31.400, 25.7, 27.000, 99.999,
           25.000, 20.700, 21.846, 23.300, 99.999],
    "gL1": [1.9200, 3.3500, 3.7600, 5.1800, 99.999, 2.6700, 4.0500, 7.1, 13.400, 99.999,
            3.7200, 5.0800, 19.67, 34.800, 99.999],
    "gL2": [0.4700, 0.7100, 0.8200, 1.1900, 99.999, 0.7500, 0.4900, 2.0200, 4.7000, 99.999,
            1.2600, 1.6000, 8.4000, 15.000, 99.999],
    "gL3": [0.8500, 1.2300, 1.4200, 1.9700, 99.999, 1.1000, 2.9300, 2.9100, 6.0000, 99.999,
            1.6300, 2.1000, 9.3000, 16.500, 99.999],
    "nr": [2.1600, 2.7500, 3.0000, 3.1900, 3.1900, 2.4000, 2.9000, 3.6304, 3.7900, 3.7900,
           2.5900, 3.1000, 3.8210, 3.9500, 3.9500],
    "alpha_T": [0.5930, 0.5771, 0.8850, 0.4200, 0.4200, 0.5930, 0.0000, 0.5405, 0.4170, 0.4170,
                0.2450, 0.3630, 0.2760, 0.4200, 0.4200],
    "beta_T": [600.00, 372.00, 530.00, 140.00, 140.00, 600.00, 0.0000, 204.00, 140.00, 140.00,
               624.00, 162.00, 93.00, 170.00, 170.00],
}
BOW_EG_TERN = [0.000, 0.528, 0.566, 0.650, 0.477, 0.415, -0.48, 0.700, 0.430,
               0.220, 2.700, 0.800, 0.190, 2.700, 1.430, 0.100, 1.900, 0.670]
BOW_DSO_TERN = [0.000, 0.000, 0.300, 0.000, 0.150, 0.100, -0.19, 0.150, 0.250,
                0.000, 0.000, 0.150, 0.000, 0.000, 0.600, 0.160, 0.750, 1.200]
# ternary pairs (binary indices) used by the Ev_avg bowing (material_parameters.c:1058-1075)
TERN_EV_PAIRS = [(1, 6), (2, 7), (3, 8), (6, 11), (7, 12), (8, 13), (1, 11), (2, 12), (3, 13),
                 (1, 2), (1, 3), (2, 3), (6, 7), (6, 8), (7, 8), (11, 12), (11, 13), (12, 13)]
# material system -> (type, binaries, ternaries)   (material_parameters.c:36-331)
SYSTEMS = {
    0: (0, (1, 2, 6, 7), (0, 1, 9, 12)),
    1: (0, (1, 3, 6, 8), (0, 2, 10, 13)),
    2: (0, (2, 3, 7, 8), (1, 2, 11, 14)),
    3: (0, (6, 7, 11, 12), (3, 4, 12, 15)),
    4: (0, (6, 8, 11, 13), (3, 5, 13, 16)),
    5: (0, (7, 8, 12, 13), (4, 5, 14, 17)),
    6: (0, (1, 2, 11, 12), (6, 7, 9, 15)),
    7: (0, (1, 3, 11, 13), (6, 8, 10, 16)),
    8: (0, (2, 3, 12, 13), (7, 8, 11, 17)),
    9: (1, (1, 6, 11), (0, 6, 3)),
    10: (1, (2, 7, 12), (1, 2, 4)),
    11: (1, (3, 8, 13), (2, 8, 5)),
    12: (2, (1, 2, 3), (9, 10, 11)),
    13: (2, (6, 7, 8), (12, 13, 14)),
    14: (2, (11, 12, 13), (15, 16, 17)),
}


@dataclass
class Shape:
    """One entry of supercell.in (read_supercell.c:159-303); lengths already in Angstrom."""
    shape: int = 0
    embedded: int = 0
    material: int = 0
    x: float = 0.0
    y: float = 0.0
    Ox: float = 0.0
    Oy: float = 0.0
    Oz: float = 0.0
    L1: float = 0.0
    L2: float = 0.0
    L3: float = 0.0
    L4: float = 0.0
    L5: float = 0.0
    SL1: float = 0.0
    SL2: float = 0.0
    SL3: float = 0.0
    CL1: float = 0.0
    CL2: float = 0.0
    CL3: float = 0.0
    RL1: int = 1
    RL2: int = 1
    RL3: int = 1
    RCL1: int = 1
    RCL2: int = 1
    RCL3: int = 1
    alpha: float = 0.0
    beta: float = 0.0
    gamma: float = 0.0


@dataclass
class Params:
    """Subset of parameters.in used on the solve path (read_parameters.c:40-502)."""
    N_shape: int = 1
    N: tuple = (1, 1, 1)
    strain_type: int = 0
    N_S: tuple = (1, 1, 1)
    N_vb: int = 1
    N_cb: int = 1
    T: float = 300.0
    F: tuple = (0.0, 0.0, 0.0)   # V/Angstrom
    K: tuple = (0.0, 0.0, 0.0)   # 1/Angstrom
    strain: int = 1
    isotropic: int = 0
    piezo: int = 0
    cme: int = 0
    coulomb_type: int = 0
    N_C: tuple = (1, 1, 1)
    softening: int = 0
    sigma_softening: float = 1.0
    extra: dict = field(default_factory=dict)


def _yaml_pairs(text):
    for raw in text.splitlines():
        line = raw.split("#", 1)[0].strip()
        if not line or ":" not in line:
            continue
        k, v = line.split(":", 1)
        yield k.strip(), v.strip()


def read_parameters(text: str) -> Params:
    """parameters.in reader (read_parameters.c:123-345: N=2M+1, F kV/cm->V/A, K 1/nm->1/A)."""
    p = Params()
    kv = dict(_yaml_pairs(text))
    g = lambda k, d: kv.get(k, d)
    p.N_shape = int(g("N_shape", 1))
    p.N = tuple(2 * int(g(f"M_G{a}", 0)) + 1 for a in "xyz")
    p.strain_type = int(g("Strain_supercell_type", 0))
    p.N_S = tuple(2 * int(g(f"M_S_G{a}", 0)) + 1 for a in "xyz")
    p.N_vb = int(g("N_VB", 1))
    p.N_cb = int(g("N_CB", 1))
    p.T = float(g("T", 300.0))
    p.F = tuple(1.0e-5 * float(g(f"F{a}", 0.0)) for a in "xyz")
    p.K = tuple(0.1 * float(g(f"K{a}", 0.0)) for a in "xyz")
    p.strain = int(g("Strain", 0))
    p.isotropic = int(g("Isotropic", 0))
    p.piezo = int(g("Piezo", 0))
    p.cme = int(g("CME", 0))
    p.coulomb_type = int(g("Coulomb_supercell_type", 0))
    p.N_C = tuple(2 * int(g(f"M_C_G{a}", 0)) + 1 for a in "xyz")
    p.softening = int(g("Softening", 0))
    p.sigma_softening = 10.0 * float(g("Sigma_softening", 0.1))
    p.extra = kv
    return p


def read_supercell(text: str, p: Params) -> list:
    """supercell.in reader (read_supercell.c:90-303), nm -> Angstrom, relative strain cells."""
    shapes = [Shape()]
    cur = shapes[0]
    for k, v in _yaml_pairs(text):
        num = float(v)
        if k == "Embedded":
            cur.embedded = int(num)
        elif k == "QD shape":
            cur.shape = int(num)
        elif k == "Material":
            cur.material = int(num)
        elif k == "Compos x":
            cur.x = num
        elif k == "Compos y":
            cur.y = num
        elif k.startswith("Centre "):
            setattr(cur, "O" + k[-1], 10.0 * num)
        elif k.startswith("Length "):
            setattr(cur, "L" + k[-1], 10.0 * num)
        elif k.startswith("Strain length "):
            a = int(k[-1])
            if p.strain_type == 0:
                setattr(cur, f"SL{a}", 10.0 * num)
            else:
                setattr(cur, f"RL{a}", int(num))
                setattr(cur, f"SL{a}", getattr(cur, f"L{a}") * int(num))
        elif k.startswith("Coulomb length "):
            a = int(k[-1])
            if p.coulomb_type == 0:
                setattr(cur, f"CL{a}", 10.0 * num)
            else:
                setattr(cur, f"RCL{a}", int(num))
                setattr(cur, f"CL{a}", getattr(cur, f"L{a}") * int(num))
            if a == 3 and len(shapes) == 1 and cur is shapes[0]:
                cur = Shape()
                shapes.append(cur)
        elif k == "Euler alpha angle":
            cur.alpha = num
        elif k == "Euler beta angle":
            cur.beta = num
        elif k == "Euler gamma angle":
            cur.gamma = num
            cur = Shape()
            shapes.append(cur)
    shapes = shapes[:p.N_shape]
    if len(shapes) != p.N_shape:
        raise ValueError("N_shape does not match supercell.in")
    return shapes


# ---------------------------------------------------------------------------------------------
# Materials (material_parameters.c:23-1542, interpolate_quaternary.c:23-160)
# ---------------------------------------------------------------------------------------------
def _interp(kind, x, y, P, B=None):
    """interpolate_quaternary.c:23-160; P = binary values in system order, B = bowings."""
    if B is None:
        if kind == 0:
            return x * y * P[0] + x * (1 - y) * P[1] + (1 - x) * y * P[2] + (1 - x) * (1 - y) * P[3]
        if kind == 1:
            return x * P[0] + y * P[1] + (1 - x - y) * P[2]
        return x * P[0] + y * P[1] + (1 - x - y) * P[2]
    if kind == 0:
        if abs(x - 1) < EPS and abs(y - 1) < EPS:
            return P[0]
        if abs(x - 1) < EPS and abs(y) < EPS:
            return P[1]
        if abs(x) < EPS and abs(y - 1) < EPS:
            return P[2]
        if abs(x) < EPS and abs(y) < EPS:
            return P[3]
        bABC, bABD, bACD, bBCD, bQ = B
        pABC = x * P[0] + (1 - x) * P[2] - x * (1 - x) * bABC
        pABD = x * P[1] + (1 - x) * P[3] - x * (1 - x) * bABD
        pACD = y * P[0] + (1 - y) * P[1] - y * (1 - y) * bACD
        pBCD = y * P[2] + (1 - y) * P[3] - y * (1 - y) * bBCD
        r = x * (1 - x) * (y * pABC + (1 - y) * pABD) + y * (1 - y) * (x * pACD + (1 - x) * pBCD)
        r /= x * (1 - x) + y * (1 - y)
        return r - x * (1 - x) * y * (1 - y) * bQ
    b1, b2, b3, bQ = B
    z = 1 - x - y
    return x * P[0] + y * P[1] + z * P[2] - x * y * b1 - x * z * b2 - y * z * b3 - x * y * z * bQ


def material_parameters(shapes, T, piezo):
    """Per-shape parameter dicts (material_parameters.c:23-1542)."""
    bin_ = {k: np.array(v, dtype=np.float64) for k, v in BIN.items()}
    mc, Ep, Eg0, Dso = bin_["mc"], bin_["Ep"], bin_["Eg"], bin_["Dso"]
    # modified parameters from T=0 gaps (material_parameters.c:356-368)
    sc = 1.0 / mc - (Ep / 3.0) * (2.0 / Eg0 + 1.0 / (Eg0 + Dso))
    g1 = bin_["gL1"] - Ep / (3.0 * Eg0 + Dso)
    g2 = bin_["gL2"] - Ep / (6.0 * Eg0 + 2.0 * Dso)
    g3 = bin_["gL3"] - Ep / (6.0 * Eg0 + 2.0 * Dso)
    # Varshni (material_parameters.c:377-400)
    Eg = Eg0.copy()
    aT = bin_["alpha_T"] * 0.001
    if T > 0.001:
        for i in range(15):
            if i != 6:
                Eg[i] -= aT[i] * T * T / (T + bin_["beta_T"][i])
            else:
                Eg[i] += 0.1081 * (1.0 - math.cosh(164.0 / T) / math.sinh(164.0 / T))
    lin = {"a": bin_["a"], "C11": bin_["C11"], "C12": bin_["C12"], "C44": bin_["C44"],
           "e14": bin_["e14"] if piezo == 1 else bin_["e14_calc"],
           "B114": bin_["B114"], "B124": bin_["B124"], "B156": bin_["B156"],
           "ac": bin_["ac"], "av": bin_["av"], "b": bin_["b"], "d": bin_["d"], "nr": bin_["nr"],
           "sc": sc, "Ep": Ep, "g1": g1, "g2": g2, "g3": g3}
    out = []
    for s in shapes:
        if s.material not in SYSTEMS:
            raise ValueError("Requested material system not recognised")
        kind, bins, _ = SYSTEMS[s.material]
        d = {}
        for k, arr in lin.items():
            d[k] = _interp(kind, s.x, s.y, [arr[i] for i in bins])
        out.append(d)
    a0 = out[0]["a"]
    bow_ev = [(3.0 / a0) * (bin_["av"][i] - bin_["av"][j]) * (bin_["a"][i] - bin_["a"][j])
              for (i, j) in TERN_EV_PAIRS]
    for s, d in zip(shapes, out):
        kind, bins, terns = SYSTEMS[s.material]
        d["Ev_avg"] = _interp(kind, s.x, s.y, [bin_["Ev_avg"][i] for i in bins],
                              [bow_ev[t] for t in terns] + [0.0])
        d["Eg"] = _interp(kind, s.x, s.y, [Eg[i] for i in bins],
                          [BOW_EG_TERN[t] for t in terns] + [0.0])
        d["Dso"] = _interp(kind, s.x, s.y, [Dso[i] for i in bins],
                           [BOW_DSO_TERN[t] for t in terns] + [0.0])
        d["P"] = math.sqrt(0.5 * HBSQM * d["Ep"])
    for d in out:   # band edges (material_parameters.c:1526-1540)
        d["E_hh"] = d["Ev_avg"] + d["Dso"] / 3.0 - (out[0]["Ev_avg"] + out[0]["Dso"] / 3.0)
        d["E_so"] = d["E_hh"] - d["Dso"]
        d["E_lh"] = d["E_hh"]
        d["E_cb"] = d["E_hh"] + d["Eg"]
    return out


# ---------------------------------------------------------------------------------------------
# Basis (basis_set.c:23-90)
# ---------------------------------------------------------------------------------------------
def basis_1d(N, L):
    NT = 2 * N - 1
    return (2.0 * np.pi / L) * (0.5 * (1.0 - NT) + np.arange(NT, dtype=np.float64))


# ---------------------------------------------------------------------------------------------
# Quadrature: fixed high-order Gauss-Legendre (independent of the library's rule)
# ---------------------------------------------------------------------------------------------
_GL_X, _GL_W = np.polynomial.legendre.leggauss(200)


def _gl(f, a, b):
    """Vectorised integral of f(x) over [a, b] (a, b broadcastable arrays)."""
    a = np.asarray(a, dtype=np.float64)[..., None]
    b = np.asarray(b, dtype=np.float64)[..., None]
    x = 0.5 * (b - a) * _GL_X + 0.5 * (b + a)
    return 0.5 * (b[..., 0] - a[..., 0]) * np.sum(_GL_W * f(x), axis=-1)


def _j0(x):
    from scipy.special import j0
    return j0(x)


def _j1(x):
    from scipy.special import j1
    return j1(x)


# The 1-D integrals below default to a 200-point Gauss-Legendre rule (converged: the integrands are
# analytic on their intervals).  QUADRATURE = "qag" evaluates them instead as the reference does,
# with QUADPACK's QAG and the 61-point Gauss-Kronrod rule at its tolerances 1e-4 (oracle/quadpack.py),
# so that the tests can measure how far the reference's integrals sit from the converged ones.
QUADRATURE = "converged"


def _qag_map(integrand, lo, hi, is_complex, *params):
    """Elementwise QAG over broadcast parameters; integrand(x, *p) for scalar p and array x."""
    from . import quadpack
    arrs = [np.asarray(v, dtype=np.float64) for v in (lo, hi) + params]
    b = np.broadcast(*arrs)
    out = np.empty(b.shape, dtype=np.complex128 if is_complex else np.float64)
    for idx, vals in zip(np.ndindex(b.shape), b):
        l, h, *p = (float(v) for v in vals)
        f = lambda x, p=p: integrand(x, *p)
        out[idx] = quadpack.qag_complex(f, l, h) if is_complex else quadpack.qag(f, l, h)[0]
    return out


def I_J0(alpha, beta):
    """characteristic_function_fourier_integrals.c:107-152: int_0^1 x cos(a sqrt(1-x^2)) J0(b x)."""
    if QUADRATURE == "qag":
        return _qag_map(lambda x, a, b: x * np.cos(a * np.sqrt(np.clip(1.0 - x * x, 0, None))) * _j0(b * x),
                        0.0, 1.0, False, alpha, beta)
    alpha = np.asarray(alpha)[..., None]
    beta = np.asarray(beta)[..., None]
    return _gl(lambda x: x * np.cos(alpha * np.sqrt(np.clip(1.0 - x * x, 0, None))) * _j0(beta * x),
               np.zeros(np.broadcast(alpha[..., 0], beta[..., 0]).shape), 1.0)


def I_2_TC(alpha, R1, R2, beta):
    """characteristic_function_fourier_integrals.c:451-520 (complex: cos + i sin)."""
    if QUADRATURE == "qag":
        return _qag_map(lambda x, a, b: x * _j0((R2 - R1) * a * x) * np.exp(1j * b * x),
                        R1 / (R2 - R1), R2 / (R2 - R1), True, alpha, beta)
    alpha = np.asarray(alpha)[..., None]
    beta = np.asarray(beta)[..., None]
    lo = R1 / (R2 - R1)
    hi = R2 / (R2 - R1)
    shp = np.broadcast(alpha[..., 0], beta[..., 0]).shape
    f = lambda x: x * _j0((R2 - R1) * alpha * x) * np.exp(1j * beta * x)
    return _gl(f, np.full(shp, lo), np.full(shp, hi))


def I_2_TC_Gz(alpha, R1, R2):
    """characteristic_function_fourier_integrals.c:577-625."""
    if QUADRATURE == "qag":
        return _qag_map(lambda x, a: x * x * _j0((R2 - R1) * a * x), R1 / (R2 - R1), R2 / (R2 - R1), False, alpha)
    alpha = np.asarray(alpha)[..., None]
    lo = R1 / (R2 - R1)
    hi = R2 / (R2 - R1)
    shp = alpha[..., 0].shape
    return _gl(lambda x: x * x * _j0((R2 - R1) * alpha * x), np.full(shp, lo), np.full(shp, hi))


def I_2(alpha, beta):
    """characteristic_function_fourier_integrals.c:217-283: int_0^1 x J0(a x) e^{i b x} dx."""
    if QUADRATURE == "qag":
        return _qag_map(lambda x, a, b: x * _j0(a * x) * np.exp(1j * b * x), 0.0, 1.0, True, alpha, beta)
    alpha = np.asarray(alpha)[..., None]
    beta = np.asarray(beta)[..., None]
    shp = np.broadcast(alpha[..., 0], beta[..., 0]).shape
    return _gl(lambda x: x * _j0(alpha * x) * np.exp(1j * beta * x), np.zeros(shp), np.ones(shp))


def I_2_Gz(alpha):
    """characteristic_function_fourier_integrals.c:342-388: int_0^1 x^2 J0(a x) dx."""
    if QUADRATURE == "qag":
        return _qag_map(lambda x, a: x * x * _j0(a * x), 0.0, 1.0, False, alpha)
    alpha = np.asarray(alpha)[..., None]
    shp = alpha[..., 0].shape
    return _gl(lambda x: x * x * _j0(alpha * x), np.zeros(shp), np.ones(shp))


def I_2_ETC(alpha, beta, f):
    """characteristic_function_fourier_integrals.c:688-752: int_f^1 x J0(a x) e^{i b x / (1 - f)} dx."""
    if QUADRATURE == "qag":
        return _qag_map(lambda x, a, b: x * _j0(a * x) * np.exp(1j * b * x / (1.0 - f)), f, 1.0, True, alpha, beta)
    alpha = np.asarray(alpha)[..., None]
    beta = np.asarray(beta)[..., None]
    shp = np.broadcast(alpha[..., 0], beta[..., 0]).shape
    return _gl(lambda x: x * _j0(alpha * x) * np.exp(1j * beta * x / (1.0 - f)), np.full(shp, f), np.ones(shp))


def I_2_ETC_Gz(alpha, f):
    """characteristic_function_fourier_integrals.c:813-857: int_f^1 x^2 J0(a x) dx."""
    if QUADRATURE == "qag":
        return _qag_map(lambda x, a: x * x * _j0(a * x), f, 1.0, False, alpha)
    alpha = np.asarray(alpha)[..., None]
    shp = alpha[..., 0].shape
    return _gl(lambda x: x * x * _j0(alpha * x), np.full(shp, f), np.ones(shp))


# ---------------------------------------------------------------------------------------------
# Shape characteristic-function Fourier transforms (characteristic_function_fourier.c)
# ---------------------------------------------------------------------------------------------
def _sel(conds, vals):
    """First-true selection like an if/elif chain; vals are arrays evaluated everywhere."""
    out = np.zeros(np.broadcast(*conds).shape, dtype=np.complex128)
    done = np.zeros(out.shape, dtype=bool)
    for c, v in zip(conds, vals):
        m = c & ~done
        out[m] = np.broadcast_to(v, out.shape)[m]
        done |= m
    return out


def _E(a):
    return np.exp(-1j * a)


def _pyramid_segment(Gx0, Gy0, Gz0, idx_side, L1, L2, L3, n_side=0):
    """characteristic_function_fourier.c:1112-1271: segment idx_side of a rectangular-base pyramid
    (n_side == 0: half base lengths L1, L2, height L3) or of an n_side-sided one (apothem L1,
    height L2).  The reference turns segment k by 2 atan(L2/L1) k (k even) or 2 atan(L1/L2) k (k
    odd), which tiles the base only when L1 == L2; the restatement keeps that."""
    if n_side == 0 and idx_side in (0, 2):
        alpha = math.atan(L2 / L1)
        x0, y0 = L1, L2
    elif n_side == 0:
        alpha = math.atan(L1 / L2)
        x0, y0 = L2, L1
    else:
        alpha = math.pi / n_side
        x0, y0 = L1, math.tan(alpha) * L1
    h = L3 if n_side == 0 else L2
    ang = 2.0 * alpha * idx_side
    Gx = math.cos(ang) * Gx0 + math.sin(ang) * Gy0
    Gy = -math.sin(ang) * Gx0 + math.cos(ang) * Gy0
    Gz = Gz0 * np.ones_like(Gx)
    Gm = np.sqrt(Gx ** 2 + Gy ** 2 + Gz ** 2)
    ax, ay, az = np.abs(Gx) < EPS, np.abs(Gy) < EPS, np.abs(Gz) < EPS
    with np.errstate(all="ignore"):
        def Ie0(a, G):
            return np.where(np.abs(G) < EPS, a + 0j, (np.exp(1j * G * a) - 1.0) / (1j * G))
        v0 = x0 * y0 * h / 3.0
        v1 = (1j * 2 * y0 * x0 / Gz) * ((1.0 - 1j * Gz * h - _E(Gz * h)) / (Gz * Gz * h * h) - 0.5)
        v2 = (2 * h * x0 / (Gy * Gy * y0)) * (1.0 - np.sin(Gy * y0) / (Gy * y0))
        v3 = (2 * h * y0 / (Gx * Gx * x0)) * (-1.0 - _E(Gx * x0) - 2j / (Gx * x0)
                                              + 2j * _E(Gx * x0) / (Gx * x0))
        pre = (-1j * x0) / (Gz * Gy)
        tail = -2.0 * (np.cos(Gy * y0) - 1.0) / (Gy * y0)
        sp, sm = Gy * y0 + Gz * h, Gy * y0 - Gz * h
        v4a = pre * (_E(Gz * h) * (-1j + (np.exp(1j * sp) - 1.0) / sp) + tail)
        v4b = pre * (_E(Gz * h) * ((np.exp(-1j * sm) - 1.0) / sm + 1j) + tail)
        v4c = pre * (_E(Gz * h) * ((np.exp(-1j * sm) - 1.0) / sm + (np.exp(1j * sp) - 1.0) / sp) + tail)
        c4a = np.abs(sm) < EPS
        c4b = np.abs(sp) < EPS
        v4 = np.where(c4a, v4a, np.where(c4b, v4b, v4c))
        q = Gx - Gz * h / x0
        base = (_E(Gx * x0) * (1.0 + 1j * Gx * x0) - 1.0) / (Gx * Gx)
        v5a = (2j * y0 / (Gz * x0)) * (0.5 * _E(Gz * h) * x0 * x0 - base)
        v5b = (2j * y0 / (Gz * x0)) * (_E(Gz * h) * (_E(Gx * x0 - Gz * h) * (1.0 + 1j * (Gx * x0 - Gz * h)) - 1.0)
                                       / q ** 2 - base)
        v5 = np.where(np.abs(q) < EPS, v5a, v5b)
        pp, pm = Gx + Gy * y0 / x0, Gx - Gy * y0 / x0
        v6a = (h / Gy) * (1.0 / pp - (1j / (x0 * pp ** 2)) * (_E(x0 * pp) - 1.0) - 0.5j * x0)
        v6b = (h / Gy) * (-1.0 / pm + (1j / (x0 * pm ** 2)) * (_E(x0 * pm) - 1.0) + 0.5j * x0)
        v6c = (h / Gy) * (1.0 / pp - 1.0 / pm - (1j / (x0 * pp ** 2)) * (_E(x0 * pp) - 1.0)
                          + (1j / (x0 * pm ** 2)) * (_E(x0 * pm) - 1.0))
        v6 = np.where(np.abs(pm) < EPS, v6a, np.where(np.abs(pp) < EPS, v6b, v6c))
        v7 = (-1.0 / (Gy * Gz)) * (Ie0(x0, -Gx - Gy * y0 / x0 + Gz * h / x0) * _E(Gz * h)
                                   - Ie0(x0, -Gx + Gy * y0 / x0 + Gz * h / x0) * _E(Gz * h)
                                   - Ie0(x0, -Gx - Gy * y0 / x0)
                                   + Ie0(x0, -Gx + Gy * y0 / x0))
        return _sel([Gm < EPS, ax & ay, ax & az, ay & az, ax, ay, az, np.ones_like(ax)],
                    [v0, v1, v2, v3, v4, v5, v6, v7])


def _prism_segment(Gx0, Gy0, Gz0, idx_side, N_side, x0, h):
    """characteristic_function_fourier.c:1276-1396."""
    alpha = math.pi / N_side
    ang = 2.0 * math.pi * idx_side / N_side
    Gx = math.cos(ang) * Gx0 + math.sin(ang) * Gy0
    Gy = -math.sin(ang) * Gx0 + math.cos(ang) * Gy0
    Gz = Gz0 * np.ones_like(Gx)
    Gm = np.sqrt(Gx ** 2 + Gy ** 2 + Gz ** 2)
    ta = math.tan(alpha)
    ax, ay, az = np.abs(Gx) < EPS, np.abs(Gy) < EPS, np.abs(Gz) < EPS
    with np.errstate(all="ignore"):
        v0 = x0 * x0 * h * ta + 0j
        v1 = 2.0 * x0 * x0 * ta * np.sin(0.5 * Gz * h) / Gz + 0j
        v2 = (2.0 * h / (Gy * Gy * ta)) * (1.0 - np.cos(Gy * x0 * ta)) + 0j
        xterm = 1j * x0 * _E(Gx * x0) + (1.0 / Gx) * (_E(Gx * x0) - 1.0)
        v3 = (2.0 * h * ta / Gx) * xterm
        v4 = (4.0 * np.sin(0.5 * Gz * h) / (Gy * Gy * Gz * ta)) * (1.0 - np.cos(Gy * x0 * ta)) + 0j
        v5 = (4.0 * np.sin(0.5 * Gz * h) * ta / (Gz * Gx)) * xterm
        up, um = Gy * ta + Gx, -Gy * ta + Gx
        fa = 1j * x0 - (1.0 / um) * (1.0 - np.exp(1j * Gy * x0 * ta - 1j * Gx * x0))
        fb = (1.0 / up) * (1.0 - np.exp(-1j * Gy * x0 * ta - 1j * Gx * x0)) - 1j * x0
        fc = (1.0 / up) * (1.0 - np.exp(-1j * Gy * x0 * ta - 1j * Gx * x0)) \
            - (1.0 / um) * (1.0 - np.exp(1j * Gy * x0 * ta - 1j * Gx * x0))
        f = np.where(np.abs(up) < EPS, fa, np.where(np.abs(um) < EPS, fb, fc))
        v6 = (h / Gy) * f
        v7 = (2.0 / (Gy * Gz)) * np.sin(0.5 * Gz * h) * f
        return _sel([Gm < EPS, ax & ay, ax & az, ay & az, ax, ay, az, np.ones_like(ax)],
                    [v0, v1, v2, v3, v4, v5, v6, v7])


def _cylinder(Gm, Gp, Gz, R, h):
    """characteristic_function_fourier.c:556-588 (shape 9): radius R, height h centred on O."""
    with np.errstate(all="ignore"):
        return _sel([np.abs(Gm) < EPS, (np.abs(Gp) < EPS) & (np.abs(Gz) > EPS), (np.abs(Gz) < EPS) & (np.abs(Gp) > EPS),
                     np.ones(Gm.shape, bool)],
                    [np.pi * R * R * h + 0j, 2 * np.pi * R * R * np.sin(0.5 * Gz * h) / Gz + 0j,
                     2 * np.pi * R * h * _j1(Gp * R) / Gp + 0j,
                     4 * np.pi * R * np.sin(0.5 * Gz * h) * _j1(Gp * R) / (Gp * Gz) + 0j])


def _cone(Gm, gq, Gz, R, A, H, B):
    """characteristic_function_fourier.c:622-667 (cone, shape 11: R = radius, A = R^2, B = R^2) and
    :769-823 (elliptical cone, shape 14: gq = |(Gx L1, Gy L2)|, R = 1, A = B = L1 L2), height H;
    gq R is the Bessel argument."""
    with np.errstate(all="ignore"):
        a = gq * R
        gen = ~((np.abs(Gm) < EPS) | (np.abs(gq) < EPS) | (np.abs(Gz) < EPS))
        cz = ~((np.abs(Gm) < EPS) | (np.abs(gq) < EPS)) & (np.abs(Gz) < EPS)
        i2z = np.zeros(Gm.shape, complex)
        if cz.any():
            i2z[cz] = I_2_Gz(a[cz])
        i2 = np.zeros(Gm.shape, complex)
        if gen.any():
            i2[gen] = I_2(a[gen], Gz[gen] * H)
        v1 = 2j * np.pi * A * ((1.0 - 1j * Gz * H - _E(Gz * H)) / (Gz * H) ** 2 - 0.5) / Gz
        v2 = 2 * np.pi * A * H * (_j1(a) / a - i2z)
        v3 = 2 * np.pi * B * 1j * _E(Gz * H) * i2 / Gz - 2j * np.pi * B * _j1(a) / (a * Gz)
        return _sel([np.abs(Gm) < EPS, np.abs(gq) < EPS, np.abs(Gz) < EPS, np.ones(Gm.shape, bool)],
                    [np.pi * A * H / 3.0 + 0j, v1, v2, v3])


def _trunc_cone(Gm, Gp, Gz, R1, R2, H):
    """characteristic_function_fourier.c:671-718 (shape 12): base radius R1, top radius R2, height H."""
    with np.errstate(all="ignore"):
        gen = ~((np.abs(Gm) < EPS) | (np.abs(Gp) < EPS))
        cz = np.abs(Gz) < EPS
        v1 = (2j * np.pi / Gz) * (0.5 * R2 ** 2 * (_E(Gz * H) - 1.0)
                                 + (R1 - R2) ** 2 / (Gz * H) ** 2
                                 * (_E(Gz * H) * (1j * Gz * R2 * H / (R1 - R2) - 1.0)
                                    + (1.0 - 1j * R1 * Gz * H / (R1 - R2)))
                                 - 0.5 * (R1 ** 2 - R2 ** 2))
        i2z = np.zeros(Gm.shape, complex)
        m = gen & cz
        if m.any():
            i2z[m] = I_2_TC_Gz(Gp[m], R2, R1)
        i2 = np.zeros(Gm.shape, complex)
        m = gen & ~cz
        if m.any():
            i2[m] = I_2_TC(Gp[m], R2, R1, Gz[m] * H)
        jj = (R1 * _j1(Gp * R1) - R2 * _j1(Gp * R2)) / Gp
        v2 = (2 * np.pi * H) * (R2 * _j1(Gp * R2) / Gp - (R1 - R2) ** 2 * i2z + (R1 / (R1 - R2)) * jj)
        v3 = (2j * np.pi / Gz) * (R2 * (_E(Gz * H) - 1.0) * _j1(Gp * R2) / Gp
                                 + (R1 - R2) ** 2 * _E(Gz * H * R1 / (R1 - R2)) * i2 - jj)
        return _sel([np.abs(Gm) < EPS, np.abs(Gp) < EPS, cz, np.ones_like(cz)],
                    [np.pi * H * (R1 * R1 + R1 * R2 + R2 * R2) / 3.0 + 0j, v1, v2, v3])


def _trunc_econe(Gm, gq, Gz, L1, L2, f, H):
    """characteristic_function_fourier.c:825-884 (shape 15): base semi-axes L1, L2, top ratio f =
    L3 / L1, height H; gq = |(Gx L1, Gy L2)|."""
    with np.errstate(all="ignore"):
        gen = ~((np.abs(Gm) < EPS) | (np.abs(gq) < EPS) | (np.abs(Gz) < EPS))
        cz = ~((np.abs(Gm) < EPS) | (np.abs(gq) < EPS)) & (np.abs(Gz) < EPS)
        iz = np.zeros(Gm.shape, complex)
        if cz.any():
            iz[cz] = I_2_ETC_Gz(gq[cz], f)
        ig = np.zeros(Gm.shape, complex)
        if gen.any():
            ig[gen] = I_2_ETC(gq[gen], Gz[gen] * H, f)
        A = L1 * L2
        e = _E(Gz * H)
        v1 = 2 * np.pi * 1j * A * ((0.5 * f * f * (e - 1.0))
                                   + ((f - 1.0) / (Gz * Gz * H * H)) * (e * (-1j * Gz * H * f - f + 1.0) + (1j * Gz * H + f - 1.0))
                                   - 0.5 * (1.0 - f * f)) / Gz
        jb = (_j1(gq) - f * _j1(f * gq)) / gq
        v2 = 2 * np.pi * A * H * (f * _j1(f * gq) / gq - iz / (1.0 - f) + jb / (1.0 - f))
        v3 = (2j * np.pi * A / Gz) * (f * (e - 1.0) * _j1(f * gq) / gq + _E(Gz * H / (1.0 - f)) * ig - jb)
        return _sel([np.abs(Gm) < EPS, np.abs(gq) < EPS, np.abs(Gz) < EPS, np.ones(Gm.shape, bool)],
                    [np.pi * A * H * (1.0 + f + f * f) / 3.0 + 0j, v1, v2, v3])


def chi(shapes, idx, Gx0, Gy0, Gz0, cell_type, softening=0, sigma=0.0):
    """characteristic_function_fourier.c:58-1105 on broadcastable G arrays.

    cell_type: 0 electronic-structure cell, 1/2 strain cell, 3 Coulomb cell.
    """
    s = shapes[idx]
    c0 = shapes[0]
    Gx0, Gy0, Gz0 = np.broadcast_arrays(np.atleast_1d(np.asarray(Gx0, float)), np.atleast_1d(np.asarray(Gy0, float)),
                                        np.atleast_1d(np.asarray(Gz0, float)))
    if cell_type == 3:
        vol = c0.CL1 * c0.CL2 * c0.CL3
    elif cell_type == 0 or idx == 0:
        vol = c0.L1 * c0.L2 * c0.L3
    else:
        vol = c0.SL1 * c0.SL2 * c0.SL3
    res = np.exp(-1j * (Gx0 * s.Ox + Gy0 * s.Oy + Gz0 * s.Oz)) / vol
    sa, ca = math.sin(math.pi * s.alpha / 180), math.cos(math.pi * s.alpha / 180)
    sb, cb = math.sin(math.pi * s.beta / 180), math.cos(math.pi * s.beta / 180)
    sg, cg = math.sin(math.pi * s.gamma / 180), math.cos(math.pi * s.gamma / 180)
    Gx = (cb * ca * cg - sa * sg) * Gx0 + (cg * cb * sa + sg * ca) * Gy0 + (-sb * cg) * Gz0
    Gy = (-cb * ca * sg - sa * cg) * Gx0 + (-sg * cb * sa + cg * ca) * Gy0 + (sb * sg) * Gz0
    Gz = (sb * ca) * Gx0 + (sa * sb) * Gy0 + cb * Gz0
    Gm = np.sqrt(Gx * Gx + Gy * Gy + Gz * Gz)
    Gp = np.sqrt(Gx * Gx + Gy * Gy)
    if softening:
        res = res * np.exp(-0.5 * sigma ** 2 * (Gx ** 2 + Gy ** 2 + Gz ** 2))
    L1, L2, L3, L4 = s.L1, s.L2, s.L3, s.L4
    sinc = lambda G, L: L * np.sinc(0.5 * G * L / np.pi)
    with np.errstate(all="ignore"):
        if s.shape == 0:
            f = sinc(Gx, L1) * sinc(Gy, L2) * sinc(Gz, L3) + 0j
        elif s.shape == 4:
            gm = np.sqrt((Gx * L1) ** 2 + (Gy * L2) ** 2 + (Gz * L3) ** 2)
            gp = np.sqrt((Gx * L1) ** 2 + (Gy * L2) ** 2)
            sph = 2 * np.pi * L1 * L2 * L3 * (np.sin(gm) - gm * np.cos(gm)) / gm ** 3
            a = Gz * L3
            v1 = sph - (2j * np.pi * L1 * L2 / Gz) * (0.5 - (np.cos(a) + a * np.sin(a) - 1.0) / a ** 2)
            gen = ~((np.abs(gm) < EPS) | (np.abs(gp) < EPS) | (np.abs(Gz) < EPS))
            ij0 = np.zeros(gm.shape)
            if gen.any():
                ij0[gen] = I_J0(a[gen], gp[gen])
            v3 = sph - (2j * np.pi * L1 * L2 / Gz) * (_j1(gp) / gp - ij0)
            f = _sel([np.abs(gm) < EPS, np.abs(gp) < EPS, np.abs(Gz) < EPS, gen],
                     [2 * np.pi * L1 * L2 * L3 / 3.0 + 0j, v1, sph + 0j, v3])
        elif s.shape == 6:
            h_full = L4 * L1 / (L1 - L3)
            seg = sum(_pyramid_segment(Gx, Gy, Gz, k, L1, L2, h_full) for k in range(4))
            L1b, L2b, L3b = L3, L2 * L3 / L1, L4 * L3 / (L1 - L3)
            seg = seg - _E(Gz * L4) * sum(_pyramid_segment(Gx, Gy, Gz, k, L1b, L2b, L3b) for k in range(4))
            f = seg
        elif s.shape == 12:
            f = _trunc_cone(Gm, Gp, Gz, L1, L2, L3)
        elif s.shape == 5:
            f = sum(_pyramid_segment(Gx, Gy, Gz, k, L1, L2, L3) for k in range(4))
        elif s.shape == 7:
            n_side = int(round(0.1 * L3))
            f = sum(_pyramid_segment(Gx, Gy, Gz, k, L1, L2, 0.0, n_side) for k in range(n_side))
        elif s.shape == 8:
            n_side = int(round(0.1 * L4))
            f = sum(_pyramid_segment(Gx, Gy, Gz, k, L1, L1 * L3 / (L1 - L2), 0.0, n_side) for k in range(n_side))
            f = f - _E(Gz * L3) * sum(_pyramid_segment(Gx, Gy, Gz, k, L2, L3 * L2 / (L1 - L2), 0.0, n_side)
                                      for k in range(n_side))
        elif s.shape == 9:
            f = _cylinder(Gm, Gp, Gz, L1, L2)
        elif s.shape == 10:
            f = _cylinder(Gm, Gp, Gz, L1, L3) - _cylinder(Gm, Gp, Gz, L2, L3)
        elif s.shape == 11:
            f = _cone(Gm, Gp, Gz, L1, L1 * L1, L2, L1 * L1)
        elif s.shape == 13:
            f = _trunc_cone(Gm, Gp, Gz, L1, L2, L4) - _E(0.5 * Gz * L4) * _cylinder(Gm, Gp, Gz, L3, L4)
        elif s.shape == 14:
            gq = np.sqrt((Gx * L1) ** 2 + (Gy * L2) ** 2)
            f = _cone(Gm, gq, Gz, 1.0, L1 * L2, L3, L1 * L2)
        elif s.shape == 15:
            f = _trunc_econe(Gm, np.sqrt((Gx * L1) ** 2 + (Gy * L2) ** 2), Gz, L1, L2, L3 / L1, L4)
        elif s.shape == 16:
            n_side = int(round(0.1 * L3))
            f = sum(_prism_segment(Gx, Gy, Gz, k, n_side, L1, L2) for k in range(n_side))
        elif s.shape in (17, 18):
            d = int(round(0.1 * L1))
            if cell_type == 0:
                C = (c0.L1, c0.L2, c0.L3)
            else:
                C = (c0.SL1, c0.SL2, c0.SL3)
            Ls = list(C)
            if s.shape == 17:       # well of thickness L2 along d, spanning the cell otherwise
                Ls[d] = L2
            else:                   # wire along d, cross-section L2 x L3 over the other two axes
                o = [a for a in range(3) if a != d]
                Ls[o[0]], Ls[o[1]] = L2, L3
            f = sinc(Gx, Ls[0]) * sinc(Gy, Ls[1]) * sinc(Gz, Ls[2]) + 0j
        elif s.shape == 1:
            v = 4 * np.pi * (np.sin(Gm * L1) - Gm * L1 * np.cos(Gm * L1)) / Gm ** 3
            f = np.where(np.abs(Gm) < EPS, 4 * np.pi * L1 ** 3 / 3.0, v) + 0j
        elif s.shape == 2:      # ellipsoid, semi-axes L1 L2 L3 (characteristic_function_fourier.c:207-237)
            gg = np.sqrt((Gx * L1) ** 2 + (Gy * L2) ** 2 + (Gz * L3) ** 2)
            v = 4 * np.pi * L1 * L2 * L3 * (np.sin(gg) - gg * np.cos(gg)) / gg ** 3
            f = np.where(np.abs(gg) < EPS, 4 * np.pi * L1 * L2 * L3 / 3.0, v) + 0j
        elif s.shape == 3:      # hemisphere of radius L1 on its base (characteristic_function_fourier.c:243-283)
            R = L1
            sph = 2 * np.pi * (np.sin(Gm * R) - Gm * R * np.cos(Gm * R)) / Gm ** 3
            a = R * Gz
            v1 = sph - (2j * np.pi * R * R / Gz) * (0.5 - (a * np.sin(a) + np.cos(a) - 1.0) / a ** 2)
            c0 = np.abs(Gm) < EPS
            c1 = (np.abs(Gp) < EPS) & (np.abs(Gz) > EPS)
            c2 = (np.abs(Gz) < EPS) & (np.abs(Gp) > EPS)
            gen = ~(c0 | c1 | c2)
            ij0 = np.zeros(Gm.shape)
            if gen.any():
                ij0[gen] = I_J0(Gz[gen] * R, Gp[gen] * R)
            v3 = sph - 2j * np.pi * R * _j1(Gp * R) / (Gp * Gz) + 2j * np.pi * R * R * ij0 / Gz
            f = _sel([c0, c1, c2, np.ones(Gm.shape, dtype=bool)], [2 * np.pi * R ** 3 / 3.0 + 0j, v1, sph + 0j, v3])
        else:
            raise NotImplementedError(f"shape {s.shape} not restated in the oracle")
    return res * f


# ---------------------------------------------------------------------------------------------
# Strain (strain_fourier_coefficient.c:23-390; strain_region_reciprocal_space.c:24-418)
# ---------------------------------------------------------------------------------------------
COMP = [(0, 0), (1, 1), (2, 2), (1, 2), (0, 2), (0, 1)]   # xx yy zz yz xz xy


def strain_factor(i, j, G, m, isotropic):
    """Bracket of strain_fourier_coefficient.c:123-172 for component (i, j); G = (Gx,Gy,Gz)."""
    C11, C12, C44 = m["C11"], m["C12"], m["C44"]
    Gx, Gy, Gz = G
    G2 = Gx * Gx + Gy * Gy + Gz * Gz
    Gm = np.sqrt(G2)
    dij = 1.0 if i == j else 0.0
    with np.errstate(all="ignore"):
        if isotropic:
            f0 = dij * (1.0 - (3 * C12 + 2 * C44) / (3 * C12 + 6 * C44))
            f = dij - (3 * C12 + 2 * C44) * G[i] * G[j] / ((C12 + 2 * C44) * G2)
        else:
            an = C11 - C12 - 2 * C44
            sum0 = 1.0 / (C44 + an / 3.0)
            f0 = dij * (1.0 - (0.5 * (C11 + 2 * C12) / 3.0) * (2.0 / (C44 + an / 3.0)) / (1.0 + (C12 + C44) * sum0))
            D = [C44 * G2 + an * Ga * Ga for Ga in (Gx, Gy, Gz)]
            s = Gx * Gx / D[0] + Gy * Gy / D[1] + Gz * Gz / D[2]
            f = dij - (0.5 * (C11 + 2 * C12) * (2 * C44 * G2 + an * (G[i] ** 2 + G[j] ** 2)) * G[i] * G[j]) \
                / ((1.0 + (C12 + C44) * s) * D[i] * D[j])
    return np.where(Gm < EPS, f0, f)


def grid3(b):
    return np.meshgrid(b[0], b[1], b[2], indexing="ij")


def conv_trunc(a, b):
    """Truncated linear convolution on an odd grid, as the zero-padded (2N_T-1)^3 FFT products
    with central extraction of convolutions.c:78-324 / piezo_region_reciprocal_space.c:155-352."""
    n = a.shape
    sz = tuple(2 * k - 1 for k in n)
    A = np.fft.fftn(a, sz, axes=(0, 1, 2))
    B = np.fft.fftn(b, sz, axes=(0, 1, 2))
    c = np.fft.ifftn(A * B, axes=(0, 1, 2))
    sl = tuple(slice(k // 2, k // 2 + k) for k in n)
    return c[sl]


@dataclass
class Tables:
    """Coefficient arrays on the electronic-structure Delta-G grid (the reference's globals)."""
    N: tuple
    NT: tuple
    G: tuple                      # 1-D basis arrays (N_T each)
    chi: list                     # per shape (index 0 unused -> None)
    strain: dict = None           # 'xx'.. on es grid
    conv: list = None             # per shape dict of char_xx.., char_hydrostatic, char_biaxial
    phi: np.ndarray = None
    params: list = None
    shapes: list = None
    p: Params = None


def coefficient_tables(p: Params, shapes: list) -> Tables:
    """The FP64 coefficient pipeline of 8_band_main.c:336-720."""
    mp = material_parameters(shapes, p.T, p.piezo)
    if p.strain and p.isotropic:
        for d in mp:
            d["C44"] = 0.5 * (d["C11"] - d["C12"])
    c0 = shapes[0]
    N = p.N
    NT = tuple(2 * k - 1 for k in N)
    Lc = (c0.L1, c0.L2, c0.L3)
    G = tuple(basis_1d(N[a], Lc[a]) for a in range(3))
    Ges = grid3(G)
    same = False
    strain_type = p.strain_type
    if p.strain:
        if strain_type == 1 and (c0.RL1, c0.RL2, c0.RL3) == (1, 1, 1) and tuple(p.N_S) == tuple(N):
            strain_type = 0
        Ls = (c0.SL1, c0.SL2, c0.SL3)
        NS = p.N_S
        GS = tuple(basis_1d(NS[a], Ls[a]) for a in range(3))
        same = all(N[a] == NS[a] and Lc[a] == Ls[a] for a in range(3))
        commensurate = strain_type == 1
        same = same and not commensurate
        RL = (c0.RL1, c0.RL2, c0.RL3) if commensurate else (1, 1, 1)
        if commensurate:
            # 8_band_main.c:158-188: N_S_G/2 >= (N_G/2) RL/2 + ((N_G/2) % 2)(RL % 2)
            for a in range(3):
                if NS[a] // 2 < (N[a] // 2) * RL[a] // 2 + ((N[a] // 2) % 2) * (RL[a] % 2):
                    raise ValueError(f"M_S_G{'xyz'[a]} is too small for the commensurate strain supercell")
        Gs = grid3(GS)
        chiS = [None] + [chi(shapes, s, *Gs, 1 + strain_type, p.softening, p.sigma_softening)
                         for s in range(1, len(shapes))]
        chi0S = chi(shapes, 0, *Gs, 2, p.softening, p.sigma_softening) if commensurate else None
    chiE = [None]
    for s in range(1, len(shapes)):
        chiE.append(chiS[s].copy() if (p.strain and same)
                    else chi(shapes, s, *Ges, 0, p.softening, p.sigma_softening))
    t = Tables(N=N, NT=NT, G=G, chi=chiE, params=mp, shapes=shapes, p=p)
    if not p.strain:
        return t
    # strain on the strain cell
    eS = {}
    for (i, j), name in zip(COMP, ["xx", "yy", "zz", "yz", "xz", "xy"]):
        acc = np.zeros(Gs[0].shape, complex)
        for s in range(1, len(shapes)):
            eps0 = (mp[shapes[s].embedded]["a"] - mp[s]["a"]) / mp[0]["a"]
            acc += eps0 * chiS[s] * strain_factor(i, j, Gs, mp[0], p.isotropic)
        eS[name] = acc
    phiS = None
    if p.piezo:
        phiS = piezo_strain_cell(p, shapes, mp, Gs, eS, chiS)

    def down(arr):
        if commensurate:
            return commensurate_down(arr, chi0S, NT, RL)
        if same:
            return arr * (1.0 / (c0.L1 * c0.L2 * c0.L3)) * (c0.L1 * c0.L2 * c0.L3)
        return free_down(arr, GS, G, Lc)

    t.strain = {k: down(v) for k, v in eS.items()}
    if phiS is not None:
        t.phi = down(phiS)
    conv = [None]
    for s in range(1, len(shapes)):
        d = {f"char_{k}": conv_trunc(chiE[s], t.strain[k]) for k in ["xx", "yy", "zz", "yz", "xz", "xy"]}
        d["char_hydrostatic"] = d["char_xx"] + d["char_yy"] + d["char_zz"]
        d["char_biaxial"] = d["char_xx"] + d["char_yy"] - 2.0 * d["char_zz"]
        conv.append(d)
    t.conv = conv
    return t


def free_down(arr, GS, G, Lc):
    """strain_reciprocal_space.c:23-230 (free lengths): the strain-cell field projected on the
    electronic cell's wave vectors with the box transform I0(dG) = 2 sin(dG L/2)/dG, I0(0) = L, per
    axis, divided by the electronic cell's volume."""
    mats = []
    for a in range(3):
        dG = GS[a][None, :] - G[a][:, None]
        with np.errstate(all="ignore"):
            I0 = np.where(np.abs(dG) < EPS, Lc[a], 2 * np.sin(dG * Lc[a] / 2) / dG)
        mats.append(I0)
    return np.einsum("abc,ia,jb,kc->ijk", arr, *mats, optimize=True) / (Lc[0] * Lc[1] * Lc[2])


def commensurate_down(arr, chi0, NT, RL):
    """strain_reciprocal_space_fft.c:23-362 / piezo_reciprocal_space_fft.c: the field and the
    electronic cell's characteristic function on the strain grid, zero-padded to (2 N_S_T - 1)^3,
    transformed, multiplied and transformed back (scale 1/size), then sampled at
    RL idx + N_S_T - 1 - RL (N_T / 2) along every axis."""
    n = arr.shape
    sz = tuple(2 * k - 1 for k in n)
    c = np.fft.ifftn(np.fft.fftn(chi0, sz, axes=(0, 1, 2)) * np.fft.fftn(arr, sz, axes=(0, 1, 2)), axes=(0, 1, 2))
    idx = []
    for a in range(3):
        ia = RL[a] * np.arange(NT[a]) + n[a] - 1 - RL[a] * (NT[a] // 2)
        assert ia[0] >= 0 and ia[-1] < sz[a]
        idx.append(ia)
    return c[np.ix_(*idx)]


def piezo_strain_cell(p, shapes, mp, Gs, e, chiS):
    """piezo_region_reciprocal_space.c:23-787 + piezo_fourier_coefficient.c:24-132."""
    Gx, Gy, Gz = Gs
    G2 = Gx * Gx + Gy * Gy + Gz * Gz
    Gm = np.sqrt(G2)
    nr0 = mp[0]["nr"]

    def phi_of(P):
        with np.errstate(all="ignore"):
            v = (-100.0j * (Gx * P[0] + Gy * P[1] + Gz * P[2])) / (8.854188 * nr0 ** 2 * G2)
        return np.where(Gm > EPS, v, 0.0)

    e14 = mp[0]["e14"]
    phi = phi_of([2 * e14 * e["yz"], 2 * e14 * e["xz"], 2 * e14 * e["xy"]])
    second = p.piezo in (2, 3)
    if second:
        pr = {}
        for a, b in [("xx", "yz"), ("yy", "yz"), ("zz", "yz"), ("xy", "xz"), ("yy", "xz"), ("xx", "xz"),
                     ("zz", "xz"), ("xy", "yz"), ("zz", "xy"), ("xx", "xy"), ("yy", "xy"), ("xz", "yz")]:
            pr[a + "_" + b] = conv_trunc(e[a], e[b])

        def P2(B114, B124, B156, f):
            return [2 * B114 * f("xx_yz") + 2 * B124 * f("yy_yz") + 2 * B124 * f("zz_yz") + 4 * B156 * f("xy_xz"),
                    2 * B114 * f("yy_xz") + 2 * B124 * f("xx_xz") + 2 * B124 * f("zz_xz") + 4 * B156 * f("xy_yz"),
                    2 * B114 * f("zz_xy") + 2 * B124 * f("xx_xy") + 2 * B124 * f("yy_xy") + 4 * B156 * f("xz_yz")]
        m0 = mp[0]
        phi = phi + phi_of(P2(m0["B114"], m0["B124"], m0["B156"], lambda k: pr[k]))
    for s in range(1, len(shapes)):
        emb = mp[shapes[s].embedded]
        de14 = mp[s]["e14"] - emb["e14"]
        cc = {k: conv_trunc(chiS[s], e[k]) for k in ["yz", "xz", "xy"]}
        phi = phi + phi_of([2 * de14 * cc["yz"], 2 * de14 * cc["xz"], 2 * de14 * cc["xy"]])
        if second:
            cp = {k: conv_trunc(chiS[s], v) for k, v in pr.items()}
            phi = phi + phi_of(P2(mp[s]["B114"] - emb["B114"], mp[s]["B124"] - emb["B124"],
                                  mp[s]["B156"] - emb["B156"], lambda k: cp[k]))
    return phi


# ---------------------------------------------------------------------------------------------
# Hamiltonian (hamiltonian_8_band.c:23-500, hamiltonian_8_band_matrix_elements.c:27-969,
# parameter_matrix_element.c:23-48, field_fourier_coefficient.c:23-55)
# ---------------------------------------------------------------------------------------------
def build_hamiltonian(t: Tables) -> np.ndarray:
    """Full Hermitian H (lower triangle from the reference block table, upper by conjugation)."""
    p, shapes, mp = t.p, t.shapes, t.params
    Nx, Ny, Nz = t.N
    NTx, NTy, NTz = t.NT
    NG = Nx * Ny * Nz
    g = np.arange(NG)
    ix, iy, iz = g // (Ny * Nz), (g % (Ny * Nz)) // Nz, g % Nz
    # Delta-G index for every (row, col) pair
    dx = ix[:, None] - ix[None, :]
    dy = iy[:, None] - iy[None, :]
    dz = iz[:, None] - iz[None, :]
    didx = (NTz // 2 + dz) + NTz * (NTy // 2 + dy) + NTy * NTz * (NTx // 2 + dx)
    diag = (dx == 0) & (dy == 0) & (dz == 0)
    Kx, Ky, Kz = p.K
    Gx = t.G[0][NTx // 2 - Nx // 2 + ix]
    Gy = t.G[1][NTy // 2 - Ny // 2 + iy]
    Gz = t.G[2][NTz // 2 - Nz // 2 + iz]
    k1x, k2x = (Gx + Kx)[:, None], (Gx + Kx)[None, :]
    k1y, k2y = (Gy + Ky)[:, None], (Gy + Ky)[None, :]
    k1z, k2z = (Gz + Kz)[:, None], (Gz + Kz)[None, :]
    chif = [None] + [c.reshape(-1) for c in t.chi[1:]]

    def pme(key, fn=None):
        val = (lambda d: d[key]) if fn is None else fn
        out = val(mp[0]) * diag.astype(complex)
        for s in range(1, len(shapes)):
            out = out + (val(mp[s]) - val(mp[shapes[s].embedded])) * chif[s][didx]
        return out

    # field (only where the other two components coincide)
    Fx, Fy, Fz = p.F
    fld = np.zeros(didx.shape, complex)
    with np.errstate(all="ignore"):
        for F, d, o1, o2, Gv, NTa, Na in [(Fx, dx, dy, dz, t.G[0], NTx, Nx), (Fy, dy, dx, dz, t.G[1], NTy, Ny),
                                           (Fz, dz, dx, dy, t.G[2], NTz, Nz)]:
            if F == 0.0:
                continue
            dG = -(d) * (Gv[1] - Gv[0])
            term = 1j * F * np.where(d % 2 == 0, -1.0, 1.0) / dG
            fld = fld + np.where((d != 0) & (o1 == 0) & (o2 == 0), term, 0.0)

    S = {}
    if p.strain:
        ef = {k: v.reshape(-1)[didx] for k, v in t.strain.items()}
        cv = [None] + [{k: v.reshape(-1)[didx] for k, v in c.items()} for c in t.conv[1:]]
        ph = t.phi.reshape(-1)[didx] if (p.piezo and t.phi is not None) else 0.0
        hyd = ef["xx"] + ef["yy"] + ef["zz"]
        bia = ef["xx"] + ef["yy"] - 2.0 * ef["zz"]

        def shp(fn):
            out = 0.0
            for s in range(1, len(shapes)):
                out = out + fn(s, mp[s], mp[shapes[s].embedded], cv[s])
            return out
        m0 = mp[0]
        S["cb"] = m0["ac"] * hyd + shp(lambda s, a, e, c: (a["ac"] - e["ac"]) * c["char_hydrostatic"]) + ph
        S["hh"] = m0["av"] * hyd + 0.5 * m0["b"] * bia + shp(
            lambda s, a, e, c: (a["av"] - e["av"]) * c["char_hydrostatic"] + 0.5 * (a["b"] - e["b"]) * c["char_biaxial"]) + ph
        S["lh"] = m0["av"] * hyd - 0.5 * m0["b"] * bia + shp(
            lambda s, a, e, c: (a["av"] - e["av"]) * c["char_hydrostatic"] - 0.5 * (a["b"] - e["b"]) * c["char_biaxial"]) + ph
        S["so"] = m0["av"] * hyd + shp(lambda s, a, e, c: (a["av"] - e["av"]) * c["char_hydrostatic"]) + ph
        S["Q"] = math.sqrt(0.5) * m0["b"] * bia + shp(lambda s, a, e, c: math.sqrt(0.5) * (a["b"] - e["b"]) * c["char_biaxial"])
        Rb = -0.5 * math.sqrt(3) * m0["b"] * (ef["xx"] - ef["yy"]) + shp(
            lambda s, a, e, c: -0.5 * math.sqrt(3) * (a["b"] - e["b"]) * (c["char_xx"] - c["char_yy"]))
        Rd = m0["d"] * ef["xy"] + shp(lambda s, a, e, c: (a["d"] - e["d"]) * c["char_xy"])
        S["R"], S["Rc"] = Rb + 1j * Rd, Rb - 1j * Rd
        Sxz = m0["d"] * ef["xz"] + shp(lambda s, a, e, c: (a["d"] - e["d"]) * c["char_xz"])
        Syz = m0["d"] * ef["yz"] + shp(lambda s, a, e, c: (a["d"] - e["d"]) * c["char_yz"])
        S["S"] = -math.sqrt(0.5) * (Sxz - 1j * Syz)
        S["Sc"] = -math.sqrt(0.5) * (Sxz + 1j * Syz)
        Pij = {k: m0["P"] * ef[k] + shp(lambda s, a, e, c, k=k: (a["P"] - e["P"]) * c["char_" + k])
               for k in ["xx", "yy", "zz", "yz", "xz", "xy"]}
        sx, sy, sz = k1x + k2x, k1y + k2y, k1z + k2z
        S["Tp"] = (-0.5 / math.sqrt(6)) * ((Pij["xx"] + 1j * Pij["xy"]) * sx + (Pij["xy"] + 1j * Pij["yy"]) * sy
                                          + (Pij["xz"] + 1j * Pij["yz"]) * sz)
        S["Tm"] = (-0.5 / math.sqrt(6)) * ((Pij["xx"] - 1j * Pij["xy"]) * sx + (Pij["xy"] - 1j * Pij["yy"]) * sy
                                          + (Pij["xz"] - 1j * Pij["yz"]) * sz)
        S["U"] = (-0.5 / math.sqrt(3)) * (Pij["xz"] * sx + Pij["yz"] * sy + Pij["zz"] * sz)
    z = lambda k: S.get(k, 0.0)
    kk_xy = k1x * k2x + k1y * k2y
    kk_z = k1z * k2z
    ECB = pme("E_cb") + 0.5 * HBSQM * pme("sc") * (kk_xy + kk_z) + z("cb") + fld
    EHH = pme("E_hh") - 0.5 * HBSQM * pme(None, lambda d: d["g1"] + d["g2"]) * kk_xy \
        - 0.5 * HBSQM * pme(None, lambda d: d["g1"] - 2 * d["g2"]) * kk_z + z("hh") + fld
    ELH = pme("E_lh") - 0.5 * HBSQM * pme(None, lambda d: d["g1"] - d["g2"]) * kk_xy \
        - 0.5 * HBSQM * pme(None, lambda d: d["g1"] + 2 * d["g2"]) * kk_z + z("lh") + fld
    ESO = pme("E_so") - 0.5 * HBSQM * pme("g1") * (kk_xy + kk_z) + z("so") + fld
    g2, g3, P = pme("g2"), pme("g3"), pme("P")
    Q = -math.sqrt(0.5) * HBSQM * g2 * kk_xy + math.sqrt(2.0) * HBSQM * g2 * kk_z + z("Q")
    rxy = k1x * k2y + k1y * k2x
    R = 0.5 * math.sqrt(3) * HBSQM * g2 * (k1x * k2x - k1y * k2y) - 0.5j * math.sqrt(3) * HBSQM * g3 * rxy + z("R")
    Rc = 0.5 * math.sqrt(3) * HBSQM * g2 * (k1x * k2x - k1y * k2y) + 0.5j * math.sqrt(3) * HBSQM * g3 * rxy + z("Rc")
    sxz = k1x * k2z + k1z * k2x
    syz = k1y * k2z + k1z * k2y
    Sm = 0.5 * math.sqrt(1.5) * HBSQM * g3 * sxz - 0.5j * math.sqrt(1.5) * HBSQM * g3 * syz + z("S")
    Sc = 0.5 * math.sqrt(1.5) * HBSQM * g3 * sxz + 0.5j * math.sqrt(1.5) * HBSQM * g3 * syz + z("Sc")
    Tp = P * (k1x + k2x) / (2 * math.sqrt(6)) + 1j * P * (k1y + k2y) / (2 * math.sqrt(6)) + z("Tp")
    Tm = P * (k1x + k2x) / (2 * math.sqrt(6)) - 1j * P * (k1y + k2y) / (2 * math.sqrt(6)) + z("Tm")
    U = P * (k1z + k2z) / (2 * math.sqrt(3)) + z("U")
    r2, r3 = math.sqrt(2.0), math.sqrt(3.0)
    # lower block table (row band, col band) -> element, hamiltonian_8_band.c:78-492 (0-based)
    blocks = {
        (0, 0): ECB, (1, 0): -r3 * Tm, (2, 0): r2 * U, (3, 0): -U, (6, 0): -Tp, (7, 0): -r2 * Tp,
        (1, 1): EHH, (2, 1): r2 * Sc, (3, 1): -Sc, (6, 1): -Rc, (7, 1): -r2 * Rc,
        (2, 2): ELH, (3, 2): Q, (4, 2): Tp, (5, 2): Rc, (7, 2): r3 * Sc,
        (3, 3): ESO, (4, 3): r2 * Tp, (5, 3): r2 * Rc, (6, 3): -r3 * Sc,
        (4, 4): ECB, (5, 4): -r3 * Tp, (6, 4): r2 * U, (7, 4): -U,
        (5, 5): EHH, (6, 5): r2 * Sm, (7, 5): -Sm,
        (6, 6): ELH, (7, 6): Q,
        (7, 7): ESO,
    }
    NH = 8 * NG
    H = np.zeros((NH, NH), dtype=np.complex128)
    for (br, bc), blk in blocks.items():
        blk = np.broadcast_to(blk, (NG, NG))
        H[br * NG:(br + 1) * NG, bc * NG:(bc + 1) * NG] = blk
    # keep the reference's lower triangle (row >= col), Hermitian-complete the rest
    Lw = np.tril(H)
    H = Lw + np.conj(np.tril(Lw, -1)).T
    return H


def element_block_table():
    """The (row band, col band) entries that the reference writes (hamiltonian_8_band.c:78-492)."""
    return [(0, 0), (1, 0), (2, 0), (3, 0), (6, 0), (7, 0), (1, 1), (2, 1), (3, 1), (6, 1), (7, 1),
            (2, 2), (3, 2), (4, 2), (5, 2), (7, 2), (3, 3), (4, 3), (5, 3), (6, 3), (4, 4), (5, 4),
            (6, 4), (7, 4), (5, 5), (6, 5), (7, 5), (6, 6), (7, 6), (7, 7)]


def window(p: Params):
    """eigenstates.c:29-36: 1-based il..iu."""
    NG = p.N[0] * p.N[1] * p.N[2]
    cb_min = 6 * NG
    il = cb_min - 2 * p.N_vb + 1
    iu = cb_min + 2 * p.N_cb
    return il, iu


def solve(H, p: Params):
    """diagonalise_hamiltonian.c:55-77 (zheevr, RANGE='I', UPLO='L')."""
    import scipy.linalg as sla
    il, iu = window(p)
    w, Z = sla.eigh(H, lower=True, subset_by_index=[il - 1, iu - 1], driver="evr",
                    overwrite_a=False, check_finite=False)
    return w, Z


def bloch(Z, NG):
    """bloch_character.c:23-105: per-state weight of each of the 8 bands."""
    a = np.abs(Z) ** 2
    return np.stack([a[b * NG:(b + 1) * NG].sum(axis=0) for b in range(8)], axis=0)


def energies_rows(w, Z, p: Params):
    """write_energies.c:40-84 rows: (E, %SO, %LH, %HH, %CB); VB highest first, CB lowest first."""
    NG = p.N[0] * p.N[1] * p.N[2]
    bl = bloch(Z, NG)
    nv = 2 * p.N_vb
    rows_vb, rows_cb = [], []
    for i in range(nv - 1, -1, -1):
        b = bl[:, i]
        rows_vb.append((w[i], 100 * (b[3] + b[7]), 100 * (b[2] + b[6]), 100 * (b[1] + b[5]), 100 * (b[0] + b[4])))
    for i in range(nv, nv + 2 * p.N_cb):
        b = bl[:, i]
        rows_cb.append((w[i], 100 * (b[3] + b[7]), 100 * (b[2] + b[6]), 100 * (b[1] + b[5]), 100 * (b[0] + b[4])))
    return np.array(rows_vb), np.array(rows_cb)


def dH_dK(t: Tables, axis: int, step: float = 0.1) -> np.ndarray:
    """dH/dK_axis (hamiltonian_derivative_8_band.c:23-500 with the element derivatives of
    hamiltonian_derivative_8_band_matrix_elements.c): every element of H is bilinear in the row and
    column wave vectors k = G + K, linear in them or independent of K, so the central difference
    [H(K + s e) - H(K - s e)] / 2s is the derivative for any step s."""
    import copy
    out = []
    for sgn in (1.0, -1.0):
        tt = copy.copy(t)
        tt.p = copy.copy(t.p)
        K = list(t.p.K)
        K[axis] += sgn * step
        tt.p.K = tuple(K)
        out.append(build_hamiltonian(tt))
    return (out[0] - out[1]) / (2.0 * step)


def momentum_elements(t: Tables, Zv, Zc, step: float = 0.1):
    """momentum_matrix_element.c:31-103: p[a, c, v] = <Zv[:, v]| dH/dK_a |Zc[:, c]> (eV Angstrom)."""
    return np.stack([(Zv.conj().T @ dH_dK(t, a, step) @ Zc).T for a in range(3)])


def carrier_localisation(chi_list, shapes, N, Z):
    """carrier_localisation.c:23-128 with 8_band_main.c:1427-1545: for each column n of Z and shape
    s >= 1, sum_b sum_G conj(psi_b(G)) (chi_s (*) psi_b)(G): the state's plane-wave grid (N per axis)
    and chi_s on the Delta-G grid (N_T), zero-padded to N + N_T - 1, transformed, multiplied,
    transformed back (scale 1/size), the block of the N plane waves taken; column 0 = 1 - the sum
    over the shapes embedded in the matrix (index 0).  chi_list[s]: chi_s on the electronic-structure
    Delta-G grid (index 0 unused).  Returns (ncols, N_shape) real."""
    NT = tuple(2 * n - 1 for n in N)
    NG = N[0] * N[1] * N[2]
    ns = len(shapes)
    sz = tuple(N[a] + NT[a] - 1 for a in range(3))
    off = tuple((N[a] + NT[a] - 1) // 2 - N[a] // 2 for a in range(3))
    chif = [None] + [np.fft.fftn(chi_list[s], sz, axes=(0, 1, 2)) for s in range(1, ns)]
    out = np.zeros((Z.shape[1], ns))
    for n in range(Z.shape[1]):
        for b in range(8):
            psi = Z[b * NG:(b + 1) * NG, n].reshape(N)
            fp = np.fft.fftn(psi, sz, axes=(0, 1, 2))
            for s in range(1, ns):
                conv = np.fft.ifftn(fp * chif[s], axes=(0, 1, 2))
                blk = conv[off[0]:off[0] + N[0], off[1]:off[1] + N[1], off[2]:off[2] + N[2]]
                out[n, s] += np.sum(blk * np.conj(psi)).real
    for n in range(Z.shape[1]):
        out[n, 0] = 1.0 - sum(out[n, s] for s in range(1, ns) if shapes[s].embedded == 0)
    return out


# ---- real-space envelope functions (secondary outputs) --------------------------------------------
LINESCAN_NAMES = ("100", "010", "001", "110", "101", "011", "111")
CROSS_NAMES = ("100", "010", "001", "-110", "110", "-101", "0-11", "111")


def _ends(n, L):
    """Endpoints and steps shared by the grid routines: -L/2 .. L/2 with n points per axis."""
    lo = [-0.5 * L[a] for a in range(3)]
    hi = [0.5 * L[a] for a in range(3)]
    d = [(hi[a] - lo[a]) / (n - 1.0) for a in range(3)]
    return lo, hi, d


def linescan_grid(direction, n, L):
    """real_space_grid_linescan.c:23-284: n points through the centre of the cell along [100],
    [010], [001], [110], [101], [011] or [111] (direction 0..6); a diagonal uses the shortest of its
    axes' steps and end points.  Points accumulate step by step from the first, the last is the end
    point itself.  Returns x, y, z (Angstrom)."""
    lo, hi, d = _ends(n, L)
    axes = {0: (0,), 1: (1,), 2: (2,), 3: (0, 1), 4: (0, 2), 5: (1, 2), 6: (0, 1, 2)}[direction]
    if len(axes) > 1:
        # the limiting axis (strictly smallest step, else the later one) sets every diagonal axis
        if len(axes) == 2:
            m = axes[0] if d[axes[0]] < d[axes[1]] else axes[1]
        else:
            m = 0 if (d[0] < d[1] and d[0] < d[2]) else (1 if d[1] < d[2] else 2)
        for a in axes:
            d[a], lo[a], hi[a] = d[m], lo[m], hi[m]
    xyz = np.zeros((3, n))
    for a in axes:
        v = lo[a]
        xyz[a, 0] = lo[a]
        for i in range(1, n - 1):
            v = v + d[a]
            xyz[a, i] = v
        xyz[a, n - 1] = hi[a]
    return xyz[0], xyz[1], xyz[2]


def cross_section_grid(direction, n, L):
    """real_space_grid_cross_section.c:23-306: n x n points (row index i1, column i2) of the plane
    through the centre with normal [100], [010], [001], [-110], [110], [-101], [0-11] or [111]
    (direction 0..7).  The (111) section is the expanded one of the reference: z over -L3 .. L3,
    y = -2 |(L1, L2)/2| / sqrt(2) - z/2 + 2 i1 dy, x = -y - z."""
    lo, hi, d = _ends(n, L)
    i1, i2 = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    i1, i2 = i1.ravel().astype(float), i2.ravel().astype(float)
    zero = np.zeros(n * n)

    def limit(a, b):
        if d[a] < d[b]:
            d[b], lo[b], hi[b] = d[a], lo[a], hi[a]
        else:
            d[a], lo[a], hi[a] = d[b], lo[b], hi[b]

    if direction == 0:
        return zero, lo[1] + i1 * d[1], lo[2] + i2 * d[2]
    if direction == 1:
        return lo[0] + i1 * d[0], zero, lo[2] + i2 * d[2]
    if direction == 2:
        return lo[0] + i1 * d[0], lo[1] + i2 * d[1], zero
    if direction == 3:
        limit(0, 1)
        return lo[0] + i1 * d[0], lo[1] + i1 * d[1], lo[2] + i2 * d[2]
    if direction == 4:
        limit(0, 1)
        return hi[0] - i1 * d[0], lo[1] + i1 * d[1], lo[2] + i2 * d[2]
    if direction == 5:
        limit(0, 2)
        return lo[0] + i1 * d[0], lo[1] + i2 * d[1], lo[2] + i1 * d[2]
    if direction == 6:
        limit(1, 2)
        return lo[0] + i2 * d[0], lo[1] + i1 * d[1], lo[2] + i1 * d[2]
    limit(0, 1)
    z = 2.0 * lo[2] + 2.0 * i2 * d[2]
    y = -2.0 * math.sqrt(0.25 * L[0] * L[0] + 0.25 * L[1] * L[1]) / math.sqrt(2.0) - z / 2.0 + 2.0 * i1 * d[1]
    return -y - z, y, z


def grid_3d(n, L):
    """real_space_grid_3D.c:23-66: n^3 points, x slowest, z fastest, -L/2 .. L/2 per axis."""
    lo, _, d = _ends(n, L)
    i1, i2, i3 = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    return lo[0] + i1.ravel() * d[0], lo[1] + i2.ravel() * d[1], lo[2] + i3.ravel() * d[2]


def envelopes(Z, N, L, K, x, y, z):
    """envelope_functions.c:23-220: F[state, band, point] = sum_G c_b(G) exp(i (G + K) . r) /
    sqrt(L1 L2 L3) for the columns of Z (a11 row order), G of the plane waves m (basis index
    N_T/2 - N/2 + m, basis_set.c)."""
    NG = N[0] * N[1] * N[2]
    g = [basis_1d(N[a], L[a])[(2 * N[a] - 1) // 2 - N[a] // 2:][:N[a]] + K[a] for a in range(3)]
    ex = np.exp(1j * np.outer(np.asarray(x), g[0]))
    ey = np.exp(1j * np.outer(np.asarray(y), g[1]))
    ez = np.exp(1j * np.outer(np.asarray(z), g[2]))
    F = np.zeros((Z.shape[1], 8, len(x)), complex)
    for n in range(Z.shape[1]):
        for b in range(8):
            c = Z[b * NG:(b + 1) * NG, n].reshape(N)
            F[n, b] = np.einsum("px,py,pz,xyz->p", ex, ey, ez, c, optimize=True)
    return F / math.sqrt(L[0] * L[1] * L[2])


def probability_columns(F, total_scale=1000.0):
    """write_probability_density.c:58-78 columns 5-9 (nm^-3): total, SO (bands 3, 7), LH (2, 6),
    HH (1, 5), CB (0, 4), from envelopes in Angstrom^-3/2.  total_scale: the factor of the total
    column, 1000 (nm^-3) except in the 3-D files of conduction states, where the reference writes
    10 (write_probability_density_3D.c:51)."""
    P = 1000.0 * np.abs(F) ** 2
    tot = (total_scale / 1000.0) * P.sum(axis=1)
    return np.stack([tot, P[:, 3] + P[:, 7], P[:, 2] + P[:, 6], P[:, 1] + P[:, 5], P[:, 0] + P[:, 4]], axis=2)


def read_energies_file(path):
    return np.loadtxt(path, ndmin=2)
