"""Carrier localisation and overlaps on the GPU (qd_localisation, qd_overlaps): against the oracle's
restatement of carrier_localisation.c on the same eigenvectors, and against the Truncated_pyramid
example's carrier_localisation_*.dat; the overlaps of the orthonormal window states vanish."""
import os

import numpy as np
import pytest

import qudos_amd as q
from helpers import GOLDEN
from oracle import qd_oracle as qo

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def solver():
    s = q.Solver(0)
    yield s
    s.close()


def _oracle_loc(inp, sp, pp, Z):
    p = qo.read_parameters(open(pp).read())
    shapes = qo.read_supercell(open(sp).read(), p)
    G = [qo.basis_1d(inp.N[a], [shapes[0].L1, shapes[0].L2, shapes[0].L3][a]) for a in range(3)]
    chis = [None] + [qo.chi(shapes, s, *qo.grid3(G), 0) for s in range(1, len(shapes))]
    return qo.carrier_localisation(chis, shapes, tuple(inp.N), Z)


def test_truncated_pyramid_localisation_and_overlaps(solver):
    d = os.path.join(GOLDEN, "truncated_pyramid")
    pp, sp = os.path.join(d, "parameters.in"), os.path.join(d, "supercell.in")
    inp = q.read_inputs(pp, sp)
    solver.upload(q.build_tables(inp))
    il, iu = q.window(inp)
    w, Z, _ = solver.solve(il, iu, q.QD_FP64)
    vb, cb, secs = solver.localisation(inp)
    nvb = 2 * inp.N_vb
    ref = _oracle_loc(inp, sp, pp, Z)
    assert np.abs(vb - ref[nvb - 1::-1]).max() < 1e-10
    assert np.abs(cb - ref[nvb:]).max() < 1e-10
    for got, name in ((100 * vb, "carrier_localisation_vb.dat"), (100 * cb, "carrier_localisation_cb.dat")):
        r = np.loadtxt(os.path.join(d, name), ndmin=2)
        assert np.max(np.abs(got - r) / np.maximum(1.0, np.abs(r))) < 1e-5, name
    ov = solver.overlaps(inp)
    assert ov.shape == (2 * inp.N_cb, nvb)
    assert np.abs(ov).max() < 1e-10
    ovr = Z[:, nvb:].conj().T @ Z[:, nvb - 1::-1]
    assert np.abs(ov - ovr).max() < 1e-13
