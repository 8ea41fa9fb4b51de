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


def _more_shapes(kind):
    """(shape, membership) for the shapes 2, 3, 5, 7-11, 13-15, 18 (base planes at z = -20)."""
    r = lambda X, Y: np.hypot(X, Y)
    zp = lambda Z: Z + 20.0                       # height above the base plane
    hexa = np.arange(6) * np.pi / 3

    def polygon(X, Y, apothem):
        return np.all([X * np.cos(a) + Y * np.sin(a) <= apothem for a in hexa], axis=0)
    if kind == "rect_pyramid":                   # half base 55 x 55, height 40 (the reference's segment
        # rotation, 2 atan(L2/L1) k, tiles the base for square bases only)
        return (qo.Shape(shape=5, L1=55.0, L2=55.0, L3=40.0, Oz=-20.0),
                lambda X, Y, Z: (zp(Z) >= 0) & (np.abs(X) <= 55 * (1 - zp(Z) / 40)) & (np.abs(Y) <= 55 * (1 - zp(Z) / 40)))
    if kind == "hex_pyramid":                    # apothem 40, height 50
        return (qo.Shape(shape=7, L1=40.0, L2=50.0, L3=60.0, Oz=-20.0),
                lambda X, Y, Z: (zp(Z) >= 0) & (zp(Z) <= 50) & polygon(X, Y, 40 * (1 - zp(Z) / 50)))
    if kind == "hex_trunc_pyramid":              # apothems 40 -> 20, height 30
        return (qo.Shape(shape=8, L1=40.0, L2=20.0, L3=30.0, L4=60.0, Oz=-20.0),
                lambda X, Y, Z: (zp(Z) >= 0) & (zp(Z) <= 30) & polygon(X, Y, 40 - 20 * zp(Z) / 30))
    if kind == "cylinder":                       # radius 50, height 40 centred on O
        return (qo.Shape(shape=9, L1=50.0, L2=40.0, Oz=10.0), lambda X, Y, Z: (r(X, Y) <= 50) & (np.abs(Z - 10) <= 20))
    if kind == "ring":                           # radii 60 / 30, height 40
        return (qo.Shape(shape=10, L1=60.0, L2=30.0, L3=40.0),
                lambda X, Y, Z: (r(X, Y) <= 60) & (r(X, Y) >= 30) & (np.abs(Z) <= 20))
    if kind == "cone":                           # radius 50, height 40
        return (qo.Shape(shape=11, L1=50.0, L2=40.0, Oz=-20.0),
                lambda X, Y, Z: (zp(Z) >= 0) & (r(X, Y) <= 50 * (1 - zp(Z) / 40)))
    if kind == "conic_ring":                     # radii 60 -> 30, hole 15, height 40
        return (qo.Shape(shape=13, L1=60.0, L2=30.0, L3=15.0, L4=40.0, Oz=-20.0),
                lambda X, Y, Z: (zp(Z) >= 0) & (zp(Z) <= 40) & (r(X, Y) <= 60 - 30 * zp(Z) / 40) & (r(X, Y) >= 15))
    if kind == "elliptic_cone":                  # semi-axes 60 x 40, height 45
        return (qo.Shape(shape=14, L1=60.0, L2=40.0, L3=45.0, Oz=-20.0),
                lambda X, Y, Z: (zp(Z) >= 0) & (zp(Z) <= 45) & ((X / 60) ** 2 + (Y / 40) ** 2 <= (1 - zp(Z) / 45) ** 2))
    if kind == "elliptic_trunc_cone":            # semi-axes 60 x 40 -> 30 x 20, height 40
        return (qo.Shape(shape=15, L1=60.0, L2=40.0, L3=30.0, L4=40.0, Oz=-20.0),
                lambda X, Y, Z: (zp(Z) >= 0) & (zp(Z) <= 40)
                & ((X / 60) ** 2 + (Y / 40) ** 2 <= (1 - 0.5 * zp(Z) / 40) ** 2))
    if kind == "ellipsoid":                      # semi-axes 60 x 45 x 30 centred 10 above O
        return (qo.Shape(shape=2, L1=60.0, L2=45.0, L3=30.0, Oz=10.0),
                lambda X, Y, Z: (X / 60) ** 2 + (Y / 45) ** 2 + ((Z - 10) / 30) ** 2 <= 1)
    if kind == "hemisphere":                     # radius 50 on the base plane
        return (qo.Shape(shape=3, L1=50.0, Oz=-20.0),
                lambda X, Y, Z: (zp(Z) >= 0) & (X ** 2 + Y ** 2 + zp(Z) ** 2 <= 50 ** 2))
    if kind == "wire":                           # along z, cross-section 40.625 x 31.25 (voxel edges)
        return (qo.Shape(shape=18, L1=20.0, L2=40.625, L3=31.25, Oy=12.5),
                lambda X, Y, Z: (np.abs(X) <= 20.3125) & (np.abs(Y - 12.5) <= 15.625))
    raise KeyError(kind)


MORE_SHAPES = ["rect_pyramid", "hex_pyramid", "hex_trunc_pyramid", "cylinder", "ring", "cone", "conic_ring",
               "elliptic_cone", "elliptic_trunc_cone", "wire", "ellipsoid", "hemisphere"]


@pytest.mark.parametrize("kind", ["lens", "cone", "prism", "qw", "pyramid"] + ["more:" + k for k in MORE_SHAPES])
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
    elif kind.startswith("more:"):
        s, fn = _more_shapes(kind[5:])
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


@pytest.mark.parametrize("RL,NS", [((2, 3, 1), (6, 8, 4)), ((3, 3, 4), (8, 8, 10))])
def test_commensurate_strain_downsampling_equals_the_box_projection(RL, NS):
    """The FFT path (padded product, strided extraction) against the free-length box projection,
    an independent algorithm pinned by the Prism fixture: with the strain cell RL x the
    electronic cell they are the same sum, truncated only where |RL e - q| leaves the strain grid,
    so a field supported near the centre gives equal results."""
    N = (3, 3, 3)
    L = (60.0, 50.0, 35.0)
    SL = tuple(RL[a] * L[a] for a in range(3))
    NT = tuple(2 * n - 1 for n in N)
    NST = tuple(2 * n - 1 for n in NS)
    cell = qo.Shape(shape=0, L1=L[0], L2=L[1], L3=L[2], SL1=SL[0], SL2=SL[1], SL3=SL[2])
    GS = tuple(qo.basis_1d(NS[a], SL[a]) for a in range(3))
    G = tuple(qo.basis_1d(N[a], L[a]) for a in range(3))
    chi0 = qo.chi([cell], 0, *qo.grid3(GS), 2)
    rng = np.random.default_rng(7)
    eps = np.zeros(NST, complex)
    # support |q_a| <= N_S - 1 - RL (N - 1): no truncation for any electronic offset
    r = tuple(NS[a] - 1 - RL[a] * (N[a] - 1) for a in range(3))
    assert min(r) >= 0
    c = tuple(n // 2 for n in NST)
    sl = tuple(slice(c[a] - r[a], c[a] + r[a] + 1) for a in range(3))
    eps[sl] = rng.normal(size=eps[sl].shape) + 1j * rng.normal(size=eps[sl].shape)
    got = qo.commensurate_down(eps, chi0, NT, RL)
    want = qo.free_down(eps, GS, G, L)
    assert np.abs(got - want).max() <= 1e-13 * np.abs(want).max()
    # an off-by-one in the extraction start shifts every point: the same check must fail
    shifted = np.roll(want, 1, axis=2)
    assert np.abs(got - shifted).max() > 1e-3 * np.abs(want).max()


@pytest.mark.parametrize("RCL,NC", [((1, 1, 2), (3, 3, 5)), ((2, 1, 3), (5, 3, 8))])
def test_commensurate_coulomb_upsampling_equals_the_box_transform(RCL, NC):
    """B_coulomb_supercell_fft's FFT convolution against the free-length box transform
    (B_coulomb_supercell.c, pinned by the Prism fixture's Coulomb files) on a commensurate
    geometry: the box transform's cut-off |g| <= max G_C is the FFT path's kernel support."""
    from oracle import qd_coulomb as qc
    N = (3, 3, 3)
    L = (60.0, 60.0, 35.0)
    CL = tuple(RCL[a] * L[a] for a in range(3))
    rng = np.random.default_rng(3)
    B = rng.normal(size=(3,) + tuple(2 * n - 1 for n in N)) + 1j * rng.normal(size=(3,) + tuple(2 * n - 1 for n in N))
    got = qc.upsample_commensurate(B, N, NC, L, CL, RCL)
    want = qc.upsample_free(B, N, NC, L, CL)
    assert np.abs(got - want).max() <= 1e-13 * np.abs(want).max()


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


def _small_prism(K=(0.011, -0.017, 0.013)):
    pp, sp = prism_paths()
    over = ("-M_Gx=1", "-M_Gy=1", "-M_Gz=1", "-M_S_Gx=1", "-M_S_Gy=1", "-M_S_Gz=1",
            f"-Kx={10 * K[0]}", f"-Ky={10 * K[1]}", f"-Kz={10 * K[2]}")
    text = open(pp).read() + "".join(f"\n{o[1:].split('=')[0]}: {o.split('=')[1]}\n" for o in over)
    p = qo.read_parameters(text)
    sh = qo.read_supercell(open(sp).read(), p)
    return qo.coefficient_tables(p, sh)


def test_hamiltonian_is_quadratic_in_K():
    """The momentum operator of the oracle is a central difference in K; it is the derivative for
    every step only if H is at most quadratic in K: third differences vanish and two steps agree."""
    t = _small_prism()
    scale = np.abs(qo.build_hamiltonian(t)).max()
    for axis in range(3):
        Hs = []
        for j in (-1, 0, 1, 2):
            import copy
            tt = copy.copy(t)
            tt.p = copy.copy(t.p)
            K = list(t.p.K)
            K[axis] += 0.05 * j
            tt.p.K = tuple(K)
            Hs.append(qo.build_hamiltonian(tt))
        third = Hs[3] - 3 * Hs[2] + 3 * Hs[1] - Hs[0]
        assert np.abs(third).max() <= 1e-12 * scale
        d1, d2 = qo.dH_dK(t, axis, 0.1), qo.dH_dK(t, axis, 0.003)
        assert np.abs(d1 - d2).max() <= 1e-10 * np.abs(d1).max()
        # the derivative is not trivially zero: the Kane and kinetic terms depend on K
        assert np.abs(d1).max() > 1.0


def test_momentum_diagonal_is_the_band_slope():
    """Hellmann-Feynman: <n| dH/dK_a |n> = dE_n/dK_a for a non-degenerate level (K off the zone
    centre lifts the spin degeneracy)."""
    import copy
    t = _small_prism()
    w, Z = np.linalg.eigh(qo.build_hamiltonian(t))
    gaps = np.minimum(np.diff(w, prepend=-np.inf)[:], np.diff(w, append=np.inf)[:])
    picks = [n for n in np.argsort(-gaps)[:40] if 0 < n < len(w) - 1][:4]
    assert min(gaps[picks]) > 1e-3
    delta = 1e-6
    for axis in range(3):
        D = qo.dH_dK(t, axis)
        ws = []
        for sgn in (1.0, -1.0):
            tt = copy.copy(t)
            tt.p = copy.copy(t.p)
            K = list(t.p.K)
            K[axis] += sgn * delta
            tt.p.K = tuple(K)
            ws.append(np.linalg.eigvalsh(qo.build_hamiltonian(tt)))
        slope = (ws[0] - ws[1]) / (2 * delta)
        for n in picks:
            hf = (Z[:, n].conj() @ D @ Z[:, n]).real
            assert abs(hf - slope[n]) <= 1e-6 * max(1.0, abs(slope[n])), (axis, n, hf, slope[n])
