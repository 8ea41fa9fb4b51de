"""CPU-side checks of the C ABI: exports, input parsing, error paths, FP64 coefficient tables."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

import qudos_amd as q
from helpers import ROOT, oracle_problem, oracle_tables, prism_paths
from oracle import qd_oracle as qo


def test_library_exports_every_declared_symbol():
    hdr = open(os.path.join(ROOT, "include", "qudos_gpu.h")).read()
    declared = sorted(set(re.findall(r"\b(qd_[a-z_0-9]+)\s*\(", hdr)))
    assert declared, "no declarations parsed"
    lib = ctypes.CDLL(q.lib_path())
    missing = [s for s in declared if not hasattr(lib, s)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", q.lib_path()], capture_output=True, text=True).stdout
    for s in declared:
        assert re.search(rf"\bT {s}\b", out), s


def test_device_entry_points_check_arguments_without_a_gpu():
    """The device coefficient passes reject a missing handle, inputs or outputs before any HIP call."""
    L = q.load_library()
    pp, sp = prism_paths()
    inp = q.read_inputs(pp, sp)
    t = q.Tables()
    W = np.zeros(2 * q.QD_NTAB, dtype=np.float64)
    params = (ctypes.c_int * 1)(q.QD_DP_FZ)
    masks = (ctypes.c_uint * 1)(0)
    grad = (ctypes.c_double * 1)()
    G = np.zeros(3)
    out = np.zeros(2)
    dp = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    assert L.qd_tables_build_device(None, ctypes.byref(inp), ctypes.byref(t)) == q.QD_E_ARG
    assert L.qd_tables_vjp_device(None, ctypes.byref(inp), dp(W), 1, params, masks, grad) == q.QD_E_ARG
    assert L.qd_selftest_shape_device(None, ctypes.byref(inp), 1, 0, -1, 1, dp(G), dp(out)) == q.QD_E_ARG
    assert "null" in L.qd_last_error().decode() or "bad argument" in L.qd_last_error().decode()


def test_read_prism_inputs():
    pp, sp = prism_paths()
    inp = q.read_inputs(pp, sp)
    assert inp.N_shape == 2 and list(inp.N) == [11, 11, 11]
    assert inp.piezo == 2 and inp.strain == 1
    s1 = inp.shapes[1]
    assert s1.shape == 16 and abs(s1.L1 - 43.30127) < 1e-12 and abs(s1.beta - 54.7356103172) < 1e-12
    assert list(inp.SL) == [200.0, 200.0, 200.0]
    assert q.window(inp) == (6 * 1331 - 8 + 1, 6 * 1331 + 4)


def test_cli_overrides_and_errors(tmp_path):
    pp, sp = prism_paths()
    inp = q.read_inputs(pp, sp, ["-M_Gx=3", "-Fz=10"])
    assert inp.N[0] == 7 and abs(inp.F[2] - 1e-4) < 1e-18
    with pytest.raises(q.QdError) as e:
        q.read_inputs(pp, sp, ["M_Gx=3"])
    assert e.value.code == -2
    with pytest.raises(q.QdError):
        q.read_inputs(pp, sp, ["-M_Gx"])
    with pytest.raises(q.QdError):
        q.read_inputs(str(tmp_path / "missing.in"), sp)
    with pytest.raises(q.QdError):   # more shapes than N_shape
        q.read_inputs(pp, sp, ["-N_shape=1"])
    bad = tmp_path / "supercell.in"
    bad.write_text(open(sp).read().replace("Material: 0 \nCompos x: 0.000000", "Material: 19 \nCompos x: 0.000000"))
    inp2 = q.read_inputs(pp, str(bad))
    with pytest.raises(q.QdError):
        q.build_tables(inp2)


@pytest.mark.parametrize("overrides", [
    (),
    ("-M_Gx=3", "-M_Gy=3", "-M_Gz=2", "-M_S_Gx=3", "-M_S_Gy=3", "-M_S_Gz=2", "-Fx=15", "-Fz=-7", "-Kx=0.01"),
    ("-M_Gx=2", "-M_Gy=2", "-M_Gz=2", "-M_S_Gx=2", "-M_S_Gy=2", "-M_S_Gz=2", "-Piezo=1", "-Isotropic=1"),
])
def test_tables_match_oracle(overrides):
    pp, sp = prism_paths()
    t = q.build_tables(q.read_inputs(pp, sp, overrides))
    tab, G = t.arrays()
    p, sh = oracle_problem(pp, sp, overrides)
    ref = oracle_tables(qo.coefficient_tables(p, sh))
    scale = np.abs(ref).max(axis=(1, 2, 3), keepdims=True)
    rel = np.abs(tab - ref) / np.maximum(scale, 1e-300)
    assert rel.max() < 1e-13, rel.max(axis=(1, 2, 3))
    assert np.allclose(G[0], qo.basis_1d(p.N[0], sh[0].L1), rtol=0, atol=1e-15)


@pytest.mark.parametrize("cfg,overrides", [
    ("lens_9x9x9", ()),
    ("cone_19x19x15", ("-M_Gx=4", "-M_Gy=4", "-M_Gz=3", "-M_S_Gx=4", "-M_S_Gy=4", "-M_S_Gz=3")),
])
def test_curved_shape_tables_match_oracle(cfg, overrides):
    """lens / truncated cone + wetting layer: quadrature-based transforms, strain, piezo"""
    d = os.path.join(ROOT, "qudos_amd", "configs", cfg)
    pp, sp = os.path.join(d, "parameters.in"), os.path.join(d, "supercell.in")
    tab, _ = q.build_tables(q.read_inputs(pp, sp, overrides)).arrays()
    p, sh = oracle_problem(pp, sp, overrides)
    ref = oracle_tables(qo.coefficient_tables(p, sh))
    scale = np.abs(ref).max(axis=(1, 2, 3), keepdims=True)
    assert (np.abs(tab - ref) / np.maximum(scale, 1e-300)).max() < 1e-13
