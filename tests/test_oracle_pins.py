"""Pins of the oracle that do not depend on the library: closed forms, limits, quadrature against
scipy, brute-force real-space transforms, convolution identities, and the reference's fixture."""
import os

import numpy as np
import pytest
from scipy import integrate, special

from helpers import GOLDEN, prism_oracle_solution, prism_paths
from oracle import qd_oracle as qo


def test_quadrature_against_scipy():
    for a, b in [(0.3, 0.7), (5.0, 12.0), (17.6, 34.0), (1e-3, 25.0)]:
        ref = integrate.quad(lambda x: x * np.cos(a * np.sqrt(1 - x * x)) * special.j0(b * x), 0, 1,
                             epsabs=1e-14, epsrel=1e-13, limit=200)[0]
        assert abs(qo.I_J0(np.array([a]), np.array([b]))[0] - ref) < 1e-12
    R1, R2 = 60.0, 120.0
    for al, be in [(0.05, 3.0), (0.2, 10.0)]:
        lo, hi = R1 / (R2 - R1), R2 / (R2 - R1)
        re = integrate.quad(lambda x: x * special.j0((R2 - R1) * al * x) * np.cos(be * x), lo, hi, epsabs=1e-14)[0]
        im = integrate.quad(lambda x: x * special.j0((R2 - R1) * al * x) * np.sin(be * x), lo, hi, epsabs=1e-14)[0]
        got = qo.I_2_TC(np.array([al]), R1, R2, np.array([be]))[0]
        assert abs(got - (re + 1j * im)) < 1e-11


def _cell(L=(200.0, 200.0, 160.0)):
    c = qo.Shape(L1=L[0], L2=L[1], L3=L[2], SL1=L[0], SL2=L[1], SL3=L[2])
    return c


def _brute(shape_fn, G, L, n=64):
    """(1/V) sum over a fine real-space grid of exp(-iG.r) inside the shape."""
    xs = [(np.arange(n) + 0.5) / n * L[a] - L[a] / 2 for a in range(3)]
    X, Y, Z = np.meshgrid(*xs, indexing="ij")
    inside = shape_fn(X, Y, Z)
    dV = np.prod([L[a] / n for a in range(3)])
    V = np.prod(L)
    return np.array([np.sum(inside * np.exp(-1j * (g[0] * X + g[1] * Y + g[2] * Z))) * dV / V for g in G])


@pytest.mark.parametrize("kind", ["lens", "cone", "prism", "qw", "pyramid"])
def test_shape_transforms_vs_brute_force(kind):
    L = (200.0, 200.0, 160.0)
    c = _cell(L)
    if kind == "lens":
        s = qo.Shape(shape=4, L1=60.0, L2=45.0, L3=30.0, Oz=-20.0)
        fn = lambda X, Y, Z: ((X / 60) ** 2 + (Y / 45) ** 2 + ((Z + 20) / 30) ** 2 <= 1) & (Z >= -20)
    elif kind == "cone":
        s = qo.Shape(shape=12, L1=60.0, L2=30.0, L3=40.0, Oz=-20.0)
        fn = lambda X, Y, Z: (np.hypot(X, Y) <= 60 - 30 * (Z + 20) / 40) & (Z >= -20) & (Z <= 20)
    elif kind == "prism":
        s = qo.Shape(shape=16, L1=40.0, L2=50.0, L3=60.0)   # hexagon, inradius 40, height 50
        ang = np.arange(6) * np.pi / 3
        fn = lambda X, Y, Z: np.all([X * np.cos(a) + Y * np.sin(a) <= 40 for a in ang], axis=0) & (np.abs(Z) <= 25)
    elif kind == "qw":
        s = qo.Shape(shape=17, L1=20.0, L2=30.0, Oz=10.0)
        fn = lambda X, Y, Z: np.abs(Z - 10) <= 15
    else:
        s = qo.Shape(shape=6, L1=60.0, L2=60.0, L3=30.0, L4=40.0, Oz=-20.0)
        fn = lambda X, Y, Z: (np.abs(X) <= 60 - 30 * (Z + 20) / 40) & (np.abs(Y) <= 60 - 30 * (Z + 20) / 40) \
            & (Z >= -20) & (Z <= 20)
    shapes = [c, s]
    rng = np.random.default_rng(3)
    k = 2 * np.pi / np.array(L)
    G = [np.zeros(3)] + [k * rng.integers(-3, 4, size=3) for _ in range(6)] + [k * np.array([2, 0, 0]),
                                                                           k * np.array([0, 0, 3])]
    got = np.array([qo.chi(shapes, 1, g[0], g[1], g[2], 0) for g in G]).ravel()
    ref = _brute(fn, G, L, n=128)
    # the voxelised boundary limits the brute force to ~1% of chi(0); sign, phase or factor
    # mistakes in a closed form show up as O(1) differences
    assert np.max(np.abs(got - ref)) < 1.2e-2 * abs(ref[0]), np.abs(got - ref)


def test_lens_branch_continuity():
    """the G_par = 0 and G_z = 0 closed forms are limits of the general quadrature branch"""
    c = _cell()
    s = qo.Shape(shape=4, L1=60.0, L2=60.0, L3=30.0)
    sh = [c, s]
    for g in [(0.0, 0.0, 0.07), (0.05, 0.02, 0.0)]:
        exact = qo.chi(sh, 1, *g, 0)
        near = qo.chi(sh, 1, g[0] + 3e-8, g[1] + 2e-8, g[2] + 4e-8, 0)
        assert abs(exact - near) < 1e-5 * abs(exact)


def test_isotropic_strain_trace():
    """isotropic continuum strain of an inclusion: trace bracket = 1 - (3C12+2C44)/(C12+2C44)"""
    m = {"C11": 120.0, "C12": 50.0, "C44": 35.0}
    G = (np.array([0.1]), np.array([0.05]), np.array([-0.02]))
    tr = sum(qo.strain_factor(i, i, G, m, True) for i in range(3))
    assert np.allclose(tr, 3 - (3 * 50 + 70) / (50 + 70))
    # anisotropic reduces to isotropic when C11 - C12 - 2 C44 = 0
    m2 = {"C11": 120.0, "C12": 50.0, "C44": 35.0}
    for (i, j) in qo.COMP:
        assert np.allclose(qo.strain_factor(i, j, G, m2, False), qo.strain_factor(i, j, G, m2, True))


def test_truncated_convolution_matches_direct():
    rng = np.random.default_rng(0)
    a = rng.normal(size=(5, 3, 7)) + 1j * rng.normal(size=(5, 3, 7))
    b = rng.normal(size=(5, 3, 7)) + 1j * rng.normal(size=(5, 3, 7))
    got = qo.conv_trunc(a, b)
    ref = np.zeros_like(a)
    c = [n // 2 for n in a.shape]
    for k in np.ndindex(*a.shape):
        s = 0
        for i in np.ndindex(*a.shape):
            j = tuple(k[d] + c[d] - i[d] for d in range(3))
            if all(0 <= j[d] < a.shape[d] for d in range(3)):
                s += a[i] * b[j]
        ref[k] = s
    assert np.allclose(got, ref, atol=1e-12)


def test_materials_hand_values():
    sh = [qo.Shape(material=3, x=1.0), qo.Shape(material=3, x=0.0), qo.Shape(material=3, x=0.2)]
    mp = qo.material_parameters(sh, 300.0, 2)
    gaas, inas = mp[0], mp[1]
    assert abs(gaas["a"] - 5.6503) < 1e-15 and abs(inas["a"] - 6.0553) < 1e-15
    # Varshni GaAs: 1.518 - 0.5405e-3 * 300^2 / (300 + 204)
    assert abs(gaas["Eg"] - (1.518 - 0.5405e-3 * 9e4 / 504)) < 1e-14
    assert abs(gaas["E_hh"]) < 1e-15
    # In0.8Ga0.2As gap with GaInAs bowing 0.477
    eg_gaas, eg_inas = 1.518 - 0.5405e-3 * 9e4 / 504, 0.405 - 0.276e-3 * 9e4 / 393
    assert abs(mp[2]["Eg"] - (0.2 * eg_gaas + 0.8 * eg_inas - 0.2 * 0.8 * 0.477)) < 1e-13
    assert abs(gaas["P"] - np.sqrt(0.5 * qo.HBSQM * 25.7)) < 1e-14


@pytest.mark.slow
def test_oracle_reproduces_prism_fixture():
    """reference fixture code/example/Prism_111/Electronic_structure/PWEM_test_ref/energies_*.dat"""
    p, sh, t, w, Z = prism_oracle_solution()
    vb, cb = qo.energies_rows(w, Z, p)
    ref_vb = np.loadtxt(os.path.join(GOLDEN, "prism_111", "energies_vb.dat"))
    ref_cb = np.loadtxt(os.path.join(GOLDEN, "prism_111", "energies_cb.dat"))
    assert np.max(np.abs(vb[:, 0] - ref_vb[:, 0])) < 1e-9
    assert np.max(np.abs(cb[:, 0] - ref_cb[:, 0])) < 1e-9
    # Bloch characters summed over each Kramers pair (gauge invariant)
    for got, ref in [(vb, ref_vb), (cb, ref_cb)]:
        g2 = got[:, 1:].reshape(-1, 2, 4).sum(axis=1)
        r2 = ref[:, 1:].reshape(-1, 2, 4).sum(axis=1)
        assert np.max(np.abs(g2 - r2)) < 1e-6
