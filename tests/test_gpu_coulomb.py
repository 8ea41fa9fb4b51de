"""GPU parity of the Coulomb matrix elements (qd_coulomb) against the FP64 oracle
(oracle/qd_coulomb.py) on the same coefficient vectors, and against the reference's Prism_111
fixture (coulomb_matrix_elements_*.dat) through gauge-invariant block sums.

Tolerance: 1e-6 relative to the largest element of each array (the contract for the Coulomb
elements); the FP64 device path reaches ~1e-13."""
import ctypes
import os

import numpy as np
import pytest

import qudos_amd as q
from helpers import GOLDEN, prism_paths
from oracle import qd_coulomb as qc
from test_coulomb_oracle import block_norms

pytestmark = pytest.mark.gpu

REL = 1e-6


@pytest.fixture(scope="module")
def solver():
    s = q.Solver(0)
    yield s
    s.close()


def _synthetic(N, L, NC, CL, ctype, cme, nr0=3.4, ranks=4, RCL=(1, 1, 1)):
    inp = q.Inputs()
    t = q.Tables()
    for a in range(3):
        t.N[a] = N[a]
        t.L[a] = L[a]
        inp.N[a] = N[a]
        inp.N_C[a] = NC[a]
        inp.CL[a] = CL[a]
        inp.RCL[a] = RCL[a]
    t.nr0 = nr0
    inp.coulomb_type = ctype
    inp.cme = cme
    inp.madelung_ranks = ranks
    return inp, t


def _rand_states(seed, N, S):
    rng = np.random.default_rng(seed)
    NG = N[0] * N[1] * N[2]
    C = rng.standard_normal((8 * NG, S)) + 1j * rng.standard_normal((8 * NG, S))
    return C / np.linalg.norm(C, axis=0)


def _assert_close(got, want):
    for key in ("hh", "ee", "eh_direct", "eh_exchange"):
        scale = np.abs(want[key]).max()
        err = np.abs(got[key] - want[key]).max()
        assert err <= REL * scale, f"{key}: max err {err:.3e} vs scale {scale:.3e}"


CASES = [
    # (N, L, NC, CL, type, cme, nh, ne)
    ((4, 3, 3), (120.0, 110.0, 90.0), (4, 3, 3), (120.0, 110.0, 90.0), 1, 1, 3, 2),
    ((5, 5, 3), (150.0, 150.0, 100.0), (5, 5, 3), (150.0, 150.0, 100.0), 0, 2, 2, 2),
    ((3, 4, 3), (100.0, 120.0, 80.0), (5, 6, 4), (160.0, 180.0, 130.0), 0, 4, 2, 3),
    ((5, 3, 4), (140.0, 100.0, 110.0), (5, 3, 6), (140.0, 100.0, 170.0), 0, 3, 4, 2),
    # commensurate Coulomb cells (B_coulomb_supercell_fft.c): relative lengths (1, 1, 2) and (2, 1, 3)
    ((3, 3, 4), (100.0, 100.0, 80.0), (3, 3, 7), (100.0, 100.0, 160.0), 1, 4, 2, 2, (1, 1, 2)),
    ((3, 4, 3), (90.0, 100.0, 70.0), (5, 4, 7), (180.0, 100.0, 210.0), 1, 2, 2, 3, (2, 1, 3)),
]


@pytest.mark.parametrize("case", CASES, ids=[f"cme{c[5]}_type{c[4]}" + ("_rcl" if len(c) > 8 else "") for c in CASES])
def test_coulomb_random_states_match_oracle(solver, case):
    N, L, NC, CL, ctype, cme, nh, ne = case[:8]
    RCL = case[8] if len(case) > 8 else (1, 1, 1)
    inp, t = _synthetic(N, L, NC, CL, ctype, cme, RCL=RCL)
    C = _rand_states(sum(N) + cme, N, nh + ne)
    solver.n = C.shape[0]
    got, st = solver.coulomb(inp, t, nh, ne, Z=C)
    want = qc.coulomb_elements(C, nh, ne, N, NC, L, CL, ctype, t.nr0, cme=cme, nprocs=4, RCL=RCL)
    _assert_close(got, want)
    assert st["n_pairs"] == (nh + ne) * (nh + ne + 1) // 2
    assert all(f >= 2 * n - 1 for f, n in zip(st["fft"], N))


def test_coulomb_rejects_unsupported_commensurate_cell(solver):
    """A commensurate Coulomb cell (relative lengths 2, 1, 1 are supported) around an electronic-
    structure cell that is displaced (shape 0 with a centre) is rejected with QD_E_UNSUPPORTED."""
    inp, t = _synthetic((3, 3, 3), (90.0,) * 3, (3, 3, 3), (180.0, 90.0, 90.0), 1, 1, RCL=(2, 1, 1))
    cell = (q._lib.QdShape * 1)()
    cell[0].Oz = 5.0
    inp.N_shape = 1
    inp.shapes = ctypes.cast(cell, ctypes.POINTER(q._lib.QdShape))
    solver.n = 8 * 27
    try:
        with pytest.raises(q.QdError) as e:
            solver.coulomb(inp, t, 1, 1, Z=_rand_states(0, (3, 3, 3), 2))
        assert e.value.code == q.QD_E_UNSUPPORTED
    finally:
        inp.shapes = None   # the array belongs to Python, not to qd_free_inputs
        inp.N_shape = 0


def _gauge_invariant_blocks(Z, nh, ne, p, sh, nr0):
    """Blocks (Kramers level per printed index) whose |V|^2 sum does not change under random
    U(2) rotations within each Kramers pair, found with the oracle; their sums are comparable
    across eigensolvers."""
    L = (sh[0].L1, sh[0].L2, sh[0].L3)
    CL = (sh[0].CL1, sh[0].CL2, sh[0].CL3)
    base = qc.write_coulomb_text(qc.coulomb_elements(Z, nh, ne, p.N, p.N_C, L, CL, p.coulomb_type, nr0, cme=p.cme))
    rng = np.random.default_rng(11)
    Zr = Z.copy()
    for lv in range(Z.shape[1] // 2):
        U, _ = np.linalg.qr(rng.standard_normal((2, 2)) + 1j * rng.standard_normal((2, 2)))
        Zr[:, 2 * lv:2 * lv + 2] = Z[:, 2 * lv:2 * lv + 2] @ U
    rot = qc.write_coulomb_text(qc.coulomb_elements(Zr, nh, ne, p.N, p.N_C, L, CL, p.coulomb_type, nr0, cme=p.cme))
    out = {}
    for name in base:
        a = block_norms(_parse(base[name]))
        b = block_norms(_parse(rot[name]))
        out[name] = [k for k in a if abs(a[k] - b[k]) <= 1e-9 * max(a[k], 1e-300)]
    return out


def _parse(text):
    rows = []
    for line in text.splitlines():
        tok = line.split()
        if tok:
            rows.append((tuple(int(x) for x in tok[:4]), complex(float(tok[4]), float(tok[5])) if len(tok) >= 6 else None))
    return rows


@pytest.mark.parametrize("prec", [q.QD_FP64, q.QD_FP32_REFINE])
def test_prism_fixture_coulomb(solver, tmp_path, prec):
    """Prism_111 (11x11x11 plane waves, CME 4, free-length Coulomb cell): device solve, device
    Coulomb on the device-resident eigenvectors, written with qd_write_coulomb."""
    from oracle import qd_oracle as qo
    pp, sp = prism_paths()
    inp = q.read_inputs(pp, sp)
    t = q.build_tables(inp)
    solver.upload(t)
    il, iu = q.window(inp)
    w, Z, _ = solver.solve(il, iu, prec)
    nh, ne = 2 * inp.N_vb, 2 * inp.N_cb
    got, st = solver.coulomb(inp, t, nh, ne)
    # same vectors through the oracle: element-wise parity
    p = qo.read_parameters(open(pp).read())
    sh = qo.read_supercell(open(sp).read(), p)
    L = (sh[0].L1, sh[0].L2, sh[0].L3)
    CL = (sh[0].CL1, sh[0].CL2, sh[0].CL3)
    want = qc.coulomb_elements(Z, nh, ne, p.N, p.N_C, L, CL, p.coulomb_type, t.nr0, cme=p.cme)
    _assert_close(got, want)
    # gauge-invariant comparison with the reference's files
    q.write_coulomb(tmp_path, got)
    inv = _gauge_invariant_blocks(Z, nh, ne, p, sh, t.nr0)
    for name, keys in inv.items():
        mine = block_norms(qc.read_coulomb_file(os.path.join(tmp_path, name)))
        ref = block_norms(qc.read_coulomb_file(os.path.join(GOLDEN, "prism_111", name)))
        assert len(keys) >= len(ref) // 2
        for k in keys:
            assert abs(mine[k] - ref[k]) <= REL * ref[k], (name, k, mine[k], ref[k])
    # direct densities: V_iikk real, V_iiii > 0 (size-independent identities)
    ed = got["eh_direct"]
    for i in range(nh):
        for k in range(ne):
            v = ed[i, i, k, k]
            assert abs(v.imag) <= 1e-9 * abs(v)
