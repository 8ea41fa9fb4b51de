"""The oracle's C restatement of the reference CPU path (oracle/qd_cpu_path.c: per-element H fill
with the reference's loop structure, MKL zheevr) against the numpy oracle: it is the CPU baseline
of bench.py and the entry evaluator of the full-size GPU residual tests, so it must build the same
H and solve it the same way."""
import os

import numpy as np
import pytest

from helpers import ROOT, oracle_problem, prism_paths
from oracle import qd_cpu
from oracle import qd_oracle as qo

SMALL = ("-M_Gx=2", "-M_Gy=2", "-M_Gz=2", "-M_S_Gx=2", "-M_S_Gy=2", "-M_S_Gz=2")
LENS = (os.path.join(ROOT, "qudos_amd", "configs", "lens_9x9x9", "parameters.in"),
        os.path.join(ROOT, "qudos_amd", "configs", "lens_9x9x9", "supercell.in"))


def _problem(paths, overrides):
    p, sh = oracle_problem(*paths, overrides)
    t = qo.coefficient_tables(p, sh)
    return p, t, qo.build_hamiltonian(t), qd_cpu.CpuProblem(t)


@pytest.mark.parametrize("paths,overrides", [
    (None, SMALL + ("-Fx=0.001", "-Fy=-0.002", "-Fz=0.0015", "-Kx=0.02", "-Ky=-0.01", "-Kz=0.03")),
    (LENS, SMALL + ("-Fz=0.001", "-Kx=0.01")),
    (LENS, SMALL + ("-Piezo=0",)),
])
def test_c_fill_matches_numpy_oracle(paths, overrides):
    p, t, H, cp = _problem(paths or prism_paths(), overrides)
    L = np.tril(H)
    Hc = cp.hamiltonian(threads=4)
    assert np.abs(np.tril(Hc) - L).max() <= 1e-14 * np.abs(H).max()
    assert np.all(np.triu(Hc, 1) == 0)      # only the lower triangle is written, as the reference does
    rows = [0, 3, cp.NH // 2, cp.NH - 1]
    assert np.abs(cp.rows(rows, threads=4) - H[rows]).max() <= 1e-14 * np.abs(H).max()
    cols = np.array([1, cp.NH // 3, cp.NH - 2], dtype=np.int32)
    Hs = cp.hamiltonian_columns(cols, threads=4)
    for j, c in enumerate(cols):
        assert np.abs(Hs[c:, j] - H[c:, c]).max() <= 1e-14 * np.abs(H).max()
        assert np.all(Hs[:c, j] == 0)


def test_mkl_zheevr_window():
    if not qd_cpu.mkl_available():
        pytest.skip("Intel MKL (libmkl_rt.so) is not loadable on this host")
    p, t, H, cp = _problem(prism_paths(), SMALL + ("-Fz=0.002",))
    il, iu = qo.window(p)
    w, Z = qd_cpu.zheevr(cp.hamiltonian(threads=4), il, iu, threads=4)
    ref = np.linalg.eigvalsh(H)[il - 1:iu]
    assert np.abs(w - ref).max() <= 1e-12 * np.abs(ref).max()
    R = H @ Z - Z * w
    assert np.abs(R).max() <= 1e-11 * np.abs(H).max()
    assert np.abs(Z.conj().T @ Z - np.eye(iu - il + 1)).max() <= 1e-12
