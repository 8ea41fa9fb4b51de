"""The reference's quadrature of the shape transforms' 1-D integrals against the library's.

The reference integrates with gsl_integration_qag (61-point Gauss-Kronrod, absolute and relative
tolerance 1e-4, characteristic_function_fourier_integrals.c with 8_band.h:61-63); the library
integrates to convergence (a 40/64-point Gauss-Legendre pair with interval halving to 1e-14).
oracle/quadpack.py restates QUADPACK's QAG; this file pins that restatement and then measures how
far the reference's integrals, and the shape transforms built on them, sit from the library's at the
wave vectors of the configurations.

  * The 61-point rule: positive weights, Kronrod nodes interlacing the 30 Gauss nodes, every Legendre
    polynomial up to degree 3N+1 = 91 integrated exactly (only degrees <= 60 fix the weights) and
    degree 92 not; the 5-point rule of the same construction equals the textbook G2K5 rule.
  * QAG: closed-form integrals, one that needs no subdivision and oscillatory ones that force it, all
    within the requested tolerance and within the returned error estimate.
  * Parity: every shape with quadratures (hemisphere, lens, cone, truncated cone, conic ring,
    elliptical and truncated elliptical cone) on a grid of wave vectors, and the lens of the bench's
    lens configuration and the truncated-cone slices of the cone and QD-molecule configurations on
    their electronic grids: the transforms with the reference's QAG integrals equal the library's to
    1e-12 of |chi(0)| (measured: at most 6e-16; at these wave vectors one 61-point pass meets the
    reference's tolerance and is already exact to rounding).
"""
import math
import os

import numpy as np
import pytest
from scipy.special import j0, j1

import qudos_amd as q
from helpers import prism_paths
from oracle import qd_oracle as qo
from oracle import quadpack as qp

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_kronrod_rule():
    leg = np.polynomial.legendre
    x, wk, wg = qp.gauss_kronrod(30)
    assert len(x) == 61 and np.all(wk > 0) and np.all(np.diff(x) > 0)
    xg, wgg = leg.leggauss(30)
    gauss = wg != 0
    assert gauss.sum() == 30 and np.allclose(x[gauss], xg, atol=1e-15) and np.allclose(wg[gauss], wgg, rtol=1e-14)
    # the Gauss nodes are every second node: Kronrod and Gauss nodes interlace
    assert np.array_equal(np.flatnonzero(gauss), np.arange(1, 61, 2))
    moments = [abs(np.dot(wk, leg.legval(x, np.eye(93)[k]))) for k in range(93)]
    assert abs(moments[0] - 2.0) < 1e-14
    assert max(moments[1:92]) < 1e-14
    assert moments[92] > 1e-8          # exactness stops at degree 3N + 1
    # G2K5: nodes 0, +-1/sqrt(3), +-sqrt(6/7); weights 308/495 (centre), 243/495, 98/495
    x5, w5, g5 = qp.gauss_kronrod(2)
    assert np.allclose(x5, [-math.sqrt(6 / 7), -1 / math.sqrt(3), 0.0, 1 / math.sqrt(3), math.sqrt(6 / 7)], atol=1e-15)
    assert np.allclose(w5, np.array([98, 243, 308, 243, 98]) / 495, atol=1e-15)
    assert np.allclose(g5, [0, 1, 0, 1, 0], atol=1e-15)


@pytest.mark.parametrize("a", [1.0, 30.0, 300.0])
def test_qag_closed_forms(a):
    # int_0^1 x J0(a x) dx = J1(a) / a
    exact = j1(a) / a
    got, err, ier, nint = qp.qag(lambda x: x * j0(a * x), 0.0, 1.0)
    assert ier == 0
    assert abs(got - exact) <= max(qp.EPS_ABS, qp.EPS_REL * abs(exact))
    assert abs(got - exact) <= err
    if a == 1.0:
        assert nint == 1 and abs(got - exact) < 1e-15     # one 61-point pass is exact to rounding
    if a == 300.0:
        assert nint > 1                                  # the oscillation forces subdivision


def test_qag_subdivides_oscillatory():
    got, err, ier, nint = qp.qag(np.cos, 0.0, 400.0)
    assert ier == 0 and nint > 1
    assert abs(got - math.sin(400.0)) <= max(1e-4, 1e-4 * abs(math.sin(400.0)))
    c = qp.qag_complex(lambda x: np.exp(1j * 250.0 * x), 0.0, 1.0)
    assert abs(c - (np.exp(250j) - 1) / 250j) <= 2e-4


SHAPES = {   # type: lengths L1..L4 (nm)
    3: (5.0, 0.0, 0.0, 0.0),
    4: (6.0, 5.0, 3.0, 0.0),
    11: (5.0, 4.0, 0.0, 0.0),
    12: (6.0, 3.0, 4.0, 0.0),
    13: (6.0, 3.0, 1.5, 4.0),
    14: (6.0, 4.0, 4.5, 0.0),
    15: (6.0, 4.0, 3.0, 4.0),
}


def _qag_chi(shapes, idx, G, cell):
    qo.QUADRATURE = "qag"
    try:
        return qo.chi(shapes, idx, G[:, 0], G[:, 1], G[:, 2], cell)
    finally:
        qo.QUADRATURE = "converged"


@pytest.mark.parametrize("shape", list(SHAPES))
def test_reference_quadrature_matches_library_shapes(tmp_path, shape):
    import test_shapes_cpu as ts
    sp, text = ts._files(tmp_path, shape, SHAPES[shape])
    pp, _ = prism_paths()
    inp = q.read_inputs(pp, sp, ("-N_shape=2",))
    p = qo.read_parameters(open(pp).read() + "\nN_shape: 2\n")
    shapes = qo.read_supercell(text, p)
    # wave vectors of a 9 x 9 x 9 grid of the 200 x 200 x 160 Angstrom cell (G up to 8 x 2 pi / L),
    # with the branch-covering vectors of the shape tests
    k = 2 * np.pi / np.array([200.0, 200.0, 160.0])
    m = np.arange(-8, 9, 2)
    grid = np.stack(np.meshgrid(m, m, m, indexing="ij"), axis=-1).reshape(-1, 3) * k
    G = np.concatenate([ts._gvecs(), grid])
    got = q.selftest_shape_host(inp, 1, G)
    ref = _qag_chi(shapes, 1, G, 0)
    scale = abs(got[0])
    assert np.abs(got - ref).max() <= 1e-12 * scale, (shape, np.abs(got - ref).max() / scale)


CONFIGS = [("lens_15x15x11", 4), ("cone_19x19x15", 12), ("molecule_17x17x21", 12)]


@pytest.mark.parametrize("name,shape", CONFIGS)
def test_reference_quadrature_matches_library_configs(name, shape):
    """Up to three shapes of the given type in the configuration, on every seventh wave vector of
    the electronic grid among those whose transform needs a quadrature."""
    d = os.path.join(ROOT, "qudos_amd", "configs", name)
    pp, sp = os.path.join(d, "parameters.in"), os.path.join(d, "supercell.in")
    inp = q.read_inputs(pp, sp)
    p = qo.read_parameters(open(pp).read())
    shapes = qo.read_supercell(open(sp).read(), p)
    g = [2 * np.pi / L * (np.arange(2 * n - 1) - (n - 1)) for L, n in
         zip((inp.shapes[0].L1, inp.shapes[0].L2, inp.shapes[0].L3), inp.N)]
    G = np.stack(np.meshgrid(*g, indexing="ij"), axis=-1).reshape(-1, 3)
    idxs = [i for i in range(1, inp.N_shape) if inp.shapes[i].shape == shape][:3]
    assert idxs
    for idx in idxs:
        # the quadratures enter where G has both an in-plane and a z component (the general branch)
        sel = (np.abs(G[:, 2]) > 1e-8) & (np.hypot(G[:, 0], G[:, 1]) > 1e-8)
        Gs = G[sel][::7]
        got = q.selftest_shape_host(inp, idx, Gs)
        ref = _qag_chi(shapes, idx, Gs, 0)
        scale = abs(q.selftest_shape_host(inp, idx, np.zeros((1, 3)))[0])
        assert np.abs(got - ref).max() <= 1e-12 * scale, (name, idx, np.abs(got - ref).max() / scale)
