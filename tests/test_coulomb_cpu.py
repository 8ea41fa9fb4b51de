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
