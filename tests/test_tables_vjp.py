"""The reverse pass through the FP64 table builder (qd_tables_vjp) against central differences of
the whole builder (qd_tables_derivative): dL/dtheta for L = Re sum W T with random W, for the
dot's height, radius, lengths, centre and composition and the field, on lens, cone and prism
geometries, first- and second-order piezoelectricity, identical and different strain cells,
isotropic and anisotropic strain, free and commensurate strain cells.  Also: the restructured builder (shape sums before the
convolutions) still equals the oracle's tables."""
import os

import numpy as np
import pytest

import qudos_amd as q
from helpers import GOLDEN, ROOT, oracle_problem, oracle_tables, prism_paths
from oracle import qd_oracle as qo


def _cfg(name):
    d = os.path.join(ROOT, "qudos_amd", "configs", name)
    if not os.path.isdir(d):
        d = os.path.join(GOLDEN, name)
    return os.path.join(d, "parameters.in"), os.path.join(d, "supercell.in")


SMALL = ("-M_Gx=2", "-M_Gy=2", "-M_Gz=2", "-M_S_Gx=2", "-M_S_Gy=2", "-M_S_Gz=2")
CASES = [
    ("lens_9x9x9", SMALL, 2, 1),
    ("lens_9x9x9", SMALL + ("-Piezo=1", "-Fz=7"), 2, 1),
    ("cone_19x19x15", SMALL, 2, 1),
    ("lens_9x9x9", SMALL + ("-Isotropic=1",), 2, 1),
    ("lens_9x9x9", ("-M_Gx=2", "-M_Gy=2", "-M_Gz=2", "-M_S_Gx=3", "-M_S_Gy=3", "-M_S_Gz=2", "-Strain_supercell_type=0"), 2, 1),
    # commensurate strain cell (3 x 3 x 4 the electronic cell, FFT downsampling), second-order
    # piezoelectricity, 100 pyramid slices; differentiate slice 5 (and slice 6's composition)
    ("truncated_pyramid", ("-M_Gx=2", "-M_Gy=2", "-M_Gz=2", "-M_S_Gx=3", "-M_S_Gy=3", "-M_S_Gz=4"), 5, 6),
]


def _fd(inp, param, mask, step):
    return q.tables_derivative(inp, param, mask, step)


@pytest.mark.parametrize("name,over,dot,wl", CASES)
def test_vjp_matches_central_differences(name, over, dot, wl):
    pp, sp = _cfg(name)
    inp = q.read_inputs(pp, sp, over)
    t = q.build_tables(inp)
    NT = tuple(t.NT)
    rng = np.random.default_rng(11)
    W = rng.standard_normal((q.QD_NTAB,) + NT) + 1j * rng.standard_normal((q.QD_NTAB,) + NT)
    md, mx = 1 << dot, (1 << dot) | (1 << wl)
    params = [(q.QD_DP_L3, md), (q.QD_DP_L12, md), (q.QD_DP_L1, md), (q.QD_DP_L2, md), (q.QD_DP_X, mx),
              (q.QD_DP_OZ, md), (q.QD_DP_OX, md), (q.QD_DP_FX, 0), (q.QD_DP_FZ, 0)]
    if inp.shapes[dot].shape == 6:   # pyramids: base lengths together only
        params = [pm for pm in params if pm[0] not in (q.QD_DP_L1, q.QD_DP_L2)] + [(q.QD_DP_L4, md)]
    steps = {q.QD_DP_X: 1e-5, q.QD_DP_OZ: 1e-3, q.QD_DP_OX: 1e-3}
    g = q.tables_vjp(inp, W, params)
    for (param, mask), got in zip(params, g):
        dT = _fd(inp, param, mask, steps.get(param, 1e-6))
        want = float(np.sum(W.real * dT.real - W.imag * dT.imag))
        assert abs(got - want) <= 2e-7 * max(1.0, abs(want)), (name, over, param, got, want)


def test_pyramid_base_length_alone_is_rejected():
    pp, sp = _cfg("truncated_pyramid")
    inp = q.read_inputs(pp, sp, ("-M_Gx=1", "-M_Gy=1", "-M_Gz=1", "-M_S_Gx=2", "-M_S_Gy=2", "-M_S_Gz=2"))
    t = q.build_tables(inp)
    W = np.zeros((q.QD_NTAB,) + tuple(t.NT), complex)
    for p in (q.QD_DP_L1, q.QD_DP_L2):
        with pytest.raises(q.QdError) as e:
            q.tables_vjp(inp, W, [(p, 1 << 3)])
        assert e.value.code == q.QD_E_UNSUPPORTED


def test_x_derivative_at_the_binary_end_point():
    """The wetting layer of the lens configurations is pure InAs (x = 0): the derivative there is
    the one-sided limit of the interpolant, which the central difference also sees."""
    pp, sp = _cfg("lens_9x9x9")
    inp = q.read_inputs(pp, sp, SMALL)
    assert inp.shapes[1].x == 0.0
    t = q.build_tables(inp)
    rng = np.random.default_rng(3)
    NT = tuple(t.NT)
    W = rng.standard_normal((q.QD_NTAB,) + NT) + 1j * rng.standard_normal((q.QD_NTAB,) + NT)
    g = q.tables_vjp(inp, W, [(q.QD_DP_X, 1 << 1)])[0]
    dT = _fd(inp, q.QD_DP_X, 1 << 1, 1e-6)
    want = float(np.sum(W.real * dT.real - W.imag * dT.imag))
    assert abs(g - want) <= 2e-6 * max(1.0, abs(want))


@pytest.mark.parametrize("overrides", [SMALL, SMALL + ("-Piezo=1",), SMALL + ("-Fz=5", "-Kx=0.02")])
def test_weighted_shape_sums_match_the_oracle(overrides):
    pp, sp = _cfg("lens_9x9x9")
    t = q.build_tables(q.read_inputs(pp, sp, overrides))
    tab, _ = t.arrays()
    p, sh = oracle_problem(pp, sp, overrides)
    ref = oracle_tables(qo.coefficient_tables(p, sh))
    assert np.abs(tab - ref).max() <= 1e-13 * np.abs(ref).max()


def test_molecule_separation_and_field():
    """The QD molecule's gradient specs (qudos_amd/autograd.py MoleculeProblem): the upper dot's seven
    shapes moved rigidly along z (one mask), and the field, against central differences."""
    pp, sp = _cfg("molecule_17x17x21")
    over = ("-M_Gx=2", "-M_Gy=2", "-M_Gz=2", "-M_S_Gx=2", "-M_S_Gy=2", "-M_S_Gz=2")
    inp = q.read_inputs(pp, sp, over)
    t = q.build_tables(inp)
    NT = tuple(t.NT)
    rng = np.random.default_rng(17)
    W = rng.standard_normal((q.QD_NTAB,) + NT) + 1j * rng.standard_normal((q.QD_NTAB,) + NT)
    upper = 0
    for s in range(8, 15):
        upper |= 1 << s
    g = q.tables_vjp(inp, W, [(q.QD_DP_OZ, upper), (q.QD_DP_FZ, 0), (q.QD_DP_X, upper)])
    for (param, mask, step), got in zip([(q.QD_DP_OZ, upper, 1e-3), (q.QD_DP_FZ, 0, 1e-6), (q.QD_DP_X, upper, 1e-5)], g):
        dT = q.tables_derivative(inp, param, mask, step)
        want = float(np.sum(W.real * dT.real - W.imag * dT.imag))
        assert abs(got - want) <= 2e-7 * max(1.0, abs(want)), (param, got, want)


def test_retained_state_gives_the_same_gradient():
    """qd_tables_vjp_retained reuses the forward intermediates kept by qd_tables_build_retain; with
    the inputs changed after the build it runs the forward pass itself."""
    pp, sp = _cfg("lens_9x9x9")
    inp = q.read_inputs(pp, sp, SMALL + ("-Piezo=2", "-Fz=4"))
    t = q.build_tables(inp, retain=True)
    assert t.state
    plain = q.build_tables(inp)
    assert not plain.state
    assert np.array_equal(t.arrays()[0], plain.arrays()[0])
    NT = tuple(t.NT)
    rng = np.random.default_rng(23)
    W = rng.standard_normal((q.QD_NTAB,) + NT) + 1j * rng.standard_normal((q.QD_NTAB,) + NT)
    params = [(q.QD_DP_L3, 1 << 2), (q.QD_DP_L12, 1 << 2), (q.QD_DP_X, 6), (q.QD_DP_OZ, 1 << 2), (q.QD_DP_FZ, 0)]
    g0 = q.tables_vjp(inp, W, params)
    g1 = q.tables_vjp(inp, W, params, tables=t)
    assert np.array_equal(g0, g1)
    inp.shapes[2].L3 *= 1.01                      # the kept state no longer applies
    g2 = q.tables_vjp(inp, W, params, tables=t)
    g3 = q.tables_vjp(inp, W, params)
    assert np.array_equal(g2, g3) and not np.array_equal(g2, g0)
