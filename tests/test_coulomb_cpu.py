"""CPU checks of the library's Coulomb host code: the writer against the oracle's restatement of
write_coulomb_matrix_elements.c (byte-identical text) and argument errors."""
import os

import numpy as np
import pytest

import qudos_amd as q
from oracle import qd_coulomb as qc


def _rand_cme(rng, nv, nc):
    def r(shape):
        return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    cme = {"hh": r((nv,) * 4), "ee": r((nc,) * 4), "eh_direct": r((nv, nv, nc, nc)), "eh_exchange": r((nv, nc, nc, nv))}
    cme["hh"][0, 0, 0, 0] = complex(-0.0, 0.0)   # signed zeros print as the C library prints them
    return cme


@pytest.mark.parametrize("nv,nc", [(4, 2), (2, 6), (8, 4)])
def test_writer_matches_reference_format(tmp_path, nv, nc):
    cme = _rand_cme(np.random.default_rng(nv * 10 + nc), nv, nc)
    q.write_coulomb(tmp_path, cme)
    want = qc.write_coulomb_text(cme)
    for name, text in want.items():
        got = open(os.path.join(tmp_path, name)).read()
        assert got == text, name


def test_writer_rejects_bad_directory(tmp_path):
    cme = _rand_cme(np.random.default_rng(0), 2, 2)
    with pytest.raises(q.QdError):
        q.write_coulomb(os.path.join(tmp_path, "missing", "dir"), cme)


def _synthetic(N, L, NC, CL, ctype, RCL=(1, 1, 1)):
    inp = q.Inputs()
    t = q.Tables()
    for a in range(3):
        t.N[a] = N[a]
        t.NT[a] = 2 * N[a] - 1
        t.L[a] = L[a]
        inp.N[a] = N[a]
        inp.N_C[a] = NC[a]
        inp.CL[a] = CL[a]
        inp.RCL[a] = RCL[a]
    t.nr0 = 3.4
    inp.coulomb_type = ctype
    inp.cme = 1
    return inp, t


@pytest.mark.parametrize("N,L,NC,RCL", [
    ((9, 9, 11), (600.0, 600.0, 350.0), (9, 9, 21), (1, 1, 2)),     # the Truncated_pyramid cells
    ((3, 4, 3), (90.0, 100.0, 70.0), (5, 4, 7), (2, 1, 3)),
])
def test_commensurate_upsampling_matrices_match_oracle(N, L, NC, RCL):
    """The library's Coulomb-cell upsampling (host matrices, applied in the device's order) against
    the oracle's FFT restatement of B_coulomb_supercell_fft.c on random pair grids."""
    CL = tuple(RCL[a] * L[a] for a in range(3))
    inp, t = _synthetic(N, L, NC, CL, 1, RCL)
    rng = np.random.default_rng(2)
    B = rng.standard_normal((3,) + tuple(2 * n - 1 for n in N)) + 1j * rng.standard_normal((3,) + tuple(2 * n - 1 for n in N))
    got = q.selftest_coulomb_upsample_host(inp, t, B)
    want = qc.upsample_commensurate(B, N, NC, L, CL, RCL)
    assert np.abs(got - want).max() <= 1e-13 * np.abs(want).max()


def test_free_upsampling_matrices_match_oracle():
    N, L, NC, CL = (3, 4, 3), (100.0, 120.0, 80.0), (5, 6, 4), (160.0, 180.0, 130.0)
    inp, t = _synthetic(N, L, NC, CL, 0)
    rng = np.random.default_rng(4)
    B = rng.standard_normal((2,) + tuple(2 * n - 1 for n in N)) + 1j * rng.standard_normal((2,) + tuple(2 * n - 1 for n in N))
    got = q.selftest_coulomb_upsample_host(inp, t, B)
    want = qc.upsample_free(B, N, NC, L, CL)
    assert np.abs(got - want).max() <= 1e-13 * np.abs(want).max()
