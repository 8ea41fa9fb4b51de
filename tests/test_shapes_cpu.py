"""The library's shape transforms (host FP64, shapes.cpp) against the oracle's restatement for the
shapes 1, 2, 3, 5, 7-11, 13-15 and 18 (the oracle's closed forms are pinned by brute-force real-space
quadrature in test_oracle_pins.py), on every branch of the closed forms (G = 0, G along z, G in the
x-y plane, general, rotated shapes), and the exact derivatives (dual numbers) against central
differences of the library's own transforms."""
import numpy as np
import pytest

import qudos_amd as q
from helpers import prism_paths
from oracle import qd_oracle as qo

CELL = """QD shape: 0
Material: 0
Compos x: 1.0
Compos y: 0.0
Length 1: 20.0
Length 2: 20.0
Length 3: 16.0
Strain length 1: 1
Strain length 2: 1
Strain length 3: 1
Coulomb length 1: 20.0
Coulomb length 2: 20.0
Coulomb length 3: 16.0
"""

# (shape, lengths L1..L4 in nm); the pyramids' base lengths are equal (the reference's segment
# rotation tiles square bases)
SHAPES = {
    "rect_pyramid": (5, (5.5, 5.5, 4.0, 0.0)),
    "hex_pyramid": (7, (4.0, 5.0, 6.0, 0.0)),
    "oct_trunc_pyramid": (8, (4.0, 2.0, 3.0, 8.0)),
    "cylinder": (9, (5.0, 4.0, 0.0, 0.0)),
    "ring": (10, (6.0, 3.0, 4.0, 0.0)),
    "cone": (11, (5.0, 4.0, 0.0, 0.0)),
    "conic_ring": (13, (6.0, 3.0, 1.5, 4.0)),
    "elliptic_cone": (14, (6.0, 4.0, 4.5, 0.0)),
    "elliptic_trunc_cone": (15, (6.0, 4.0, 3.0, 4.0)),
    "wire_z": (18, (2.0, 4.1, 3.3, 0.0)),
    "wire_x": (18, (0.0, 3.7, 2.9, 0.0)),
    "sphere": (1, (5.0, 0.0, 0.0, 0.0)),
    "ellipsoid": (2, (6.0, 5.0, 4.0, 0.0)),
    "hemisphere": (3, (5.0, 0.0, 0.0, 0.0)),
}


def _files(tmp_path, shape, Ls, angles=(0.0, 0.0, 0.0), centre=(0.3, -0.2, -1.0)):
    text = CELL + f"""Embedded: 0
QD shape: {shape}
Material: 0
Compos x: 0.0
Compos y: 0.0
Centre x: {centre[0]}
Centre y: {centre[1]}
Centre z: {centre[2]}
Length 1: {Ls[0]}
Length 2: {Ls[1]}
Length 3: {Ls[2]}
Length 4: {Ls[3]}
Length 5: 0.0
Euler alpha angle: {angles[0]}
Euler beta angle: {angles[1]}
Euler gamma angle: {angles[2]}
"""
    sp = tmp_path / "supercell.in"
    sp.write_text(text)
    return str(sp), text


def _gvecs():
    k = 2 * np.pi / np.array([200.0, 200.0, 160.0])
    rng = np.random.default_rng(1)
    G = [np.zeros(3), k * np.array([0, 0, 2]), k * np.array([0, 0, -3]), k * np.array([3, 0, 0]),
         k * np.array([1, -2, 0]), k * np.array([0, 1, 1]), k * np.array([2, 0, -1])]
    G += [k * rng.integers(-4, 5, size=3) for _ in range(8)]
    return np.array(G)


@pytest.mark.parametrize("name", list(SHAPES))
@pytest.mark.parametrize("angles", [(0.0, 0.0, 0.0), (30.0, 20.0, 10.0)])
def test_library_shapes_match_oracle(tmp_path, name, angles):
    shape, Ls = SHAPES[name]
    sp, text = _files(tmp_path, shape, Ls, angles)
    pp, _ = prism_paths()
    inp = q.read_inputs(pp, sp, ("-N_shape=2",))
    G = _gvecs()
    got = q.selftest_shape_host(inp, 1, G)
    p = qo.read_parameters(open(pp).read() + "\nN_shape: 2\n")
    shapes = qo.read_supercell(text, p)
    ref = np.array([qo.chi(shapes, 1, g[0], g[1], g[2], 0)[0] for g in G])
    scale = abs(ref[0])
    assert scale > 0
    assert np.abs(got - ref).max() <= 1e-11 * scale, (name, np.abs(got - ref) / scale)


DERIV = [(q.QD_DP_L1, 1e-6), (q.QD_DP_L2, 1e-6), (q.QD_DP_L3, 1e-6), (q.QD_DP_L4, 1e-6), (q.QD_DP_L12, 1e-6),
         (q.QD_DP_OZ, 1e-4), (q.QD_DP_OX, 1e-4)]


@pytest.mark.parametrize("name", list(SHAPES))
def test_shape_derivatives_match_central_differences(tmp_path, name):
    shape, Ls = SHAPES[name]
    pp, _ = prism_paths()
    G = _gvecs()
    base_sp, _ = _files(tmp_path, shape, Ls)
    inp = q.read_inputs(pp, base_sp, ("-N_shape=2",))
    used = {1: (1,), 2: (1, 2, 3), 3: (1,), 5: (1, 2, 3), 7: (1, 2), 8: (1, 2, 3), 9: (1, 2), 10: (1, 2, 3),
            11: (1, 2), 13: (1, 2, 3, 4), 14: (1, 2, 3), 15: (1, 2, 3, 4), 18: (2, 3)}[shape]
    for param, rel in DERIV:
        if param in (q.QD_DP_L1, q.QD_DP_L2, q.QD_DP_L3, q.QD_DP_L4) and (param - q.QD_DP_L1 + 1 if param != q.QD_DP_L4 else 4) not in used:
            continue
        if shape == 5 and param in (q.QD_DP_L1, q.QD_DP_L2):
            continue                      # the pyramid's segment rotation depends on L2 / L1 (see qd_tables_vjp)
        if shape in (7, 8) and param == q.QD_DP_L12:
            continue                      # Length 2 is a height or top apothem there, not a base length
        if shape in (1, 3, 18) and param == q.QD_DP_L12:
            continue                      # one length (the radius), or a direction code in Length 1
        d = q.selftest_shape_host(inp, 1, G, param=param)
        fd = np.zeros(len(G), complex)
        for sgn in (1.0, -1.0):
            L2s = list(Ls)
            c = [0.3, -0.2, -1.0]
            if param == q.QD_DP_OZ:
                c[2] += sgn * rel * 10.0 / 10.0
                h = rel * 10.0
            elif param == q.QD_DP_OX:
                c[0] += sgn * rel * 10.0 / 10.0
                h = rel * 10.0
            else:
                idx = {q.QD_DP_L1: [0], q.QD_DP_L2: [1], q.QD_DP_L3: [2], q.QD_DP_L4: [3], q.QD_DP_L12: [0, 1]}[param]
                h = rel * 10.0 * Ls[idx[0]]
                for i in idx:
                    L2s[i] += sgn * rel * Ls[idx[0]]
            sp, _ = _files(tmp_path, shape, L2s, centre=c)
            v = q.selftest_shape_host(q.read_inputs(pp, sp, ("-N_shape=2",)), 1, G)
            fd += sgn * v
        fd /= 2 * h
        scale = max(np.abs(fd).max(), 1e-30)
        assert np.abs(d - fd).max() <= 1e-5 * scale, (name, param, np.abs(d - fd).max() / scale)
