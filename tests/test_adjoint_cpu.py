"""Host table derivatives used by the adjoint chain (qd_tables_derivative): the field derivative
is exact (tables are linear in F), the shape and composition derivatives converge with the step
(central differences, O(h^2))."""
import os

import numpy as np
import pytest

import qudos_amd as q

CFG = os.path.join(os.path.dirname(q.lib_path()), "..", "configs", "lens_11x11x9")
SMALL = ("-M_Gx=2", "-M_Gy=2", "-M_Gz=2", "-M_S_Gx=2", "-M_S_Gy=2", "-M_S_Gz=2")


def _inp(extra=()):
    return q.read_inputs(os.path.join(CFG, "parameters.in"), os.path.join(CFG, "supercell.in"), SMALL + tuple(extra))


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_field_derivative_is_exact(axis):
    key = "xyz"[axis]
    inp = _inp((f"-F{key}=7",))
    dT = q.tables_derivative(inp, q.QD_DP_FX + axis)
    tp = q.build_tables(_inp((f"-F{key}=9",))).arrays()[0]
    tm = q.build_tables(_inp((f"-F{key}=5",))).arrays()[0]
    fd = (tp - tm) / (4 * 1e-5)      # kV/cm -> V/Angstrom
    assert np.abs(fd - dT).max() <= 1e-9 * np.abs(dT).max()
    # only the four band-edge tables depend on the field
    assert np.abs(dT[4:]).max() == 0.0


@pytest.mark.parametrize("param,mask", [(q.QD_DP_L3, 1 << 2), (q.QD_DP_L12, 1 << 2), (q.QD_DP_X, (1 << 1) | (1 << 2))])
def test_shape_derivatives_converge(param, mask):
    inp = _inp()
    d1 = q.tables_derivative(inp, param, mask, 1e-5)
    d2 = q.tables_derivative(inp, param, mask, 1e-3)
    scale = np.abs(d1).max()
    assert scale > 0
    assert np.abs(d1 - d2).max() <= 1e-4 * scale


def test_derivative_argument_errors():
    inp = _inp()
    with pytest.raises(q.QdError):
        q.tables_derivative(inp, q.QD_DP_L3, 0, 1e-5)          # no shape selected
    with pytest.raises(q.QdError):
        q.tables_derivative(inp, q.QD_DP_L3, 1 << 2, -1.0)     # bad step
