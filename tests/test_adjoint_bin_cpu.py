"""CPU check of the adjoint's FFT-correlation binning: the binning kernels' own per-element code
(zero padding with the u = (1, k) weights, the spectral accumulation over the bilinear element
coefficients, the extraction of the Delta-G offsets), run on the host with a host FFT, against
the binning by its definition (every lower band block, every plane-wave pair, the element
formula's coefficient with a one-hot table)."""
import numpy as np
import pytest

import qudos_amd as q
from helpers import prism_paths


@pytest.mark.parametrize("overrides", [
    ("-M_Gx=1", "-M_Gy=1", "-M_Gz=1", "-M_S_Gx=1", "-M_S_Gy=1", "-M_S_Gz=1"),
    ("-M_Gx=1", "-M_Gy=2", "-M_Gz=1", "-M_S_Gx=1", "-M_S_Gy=2", "-M_S_Gz=1", "-Kx=0.21", "-Ky=-0.13", "-Kz=0.08",
     "-Fz=15"),
])
def test_fft_binning_equals_the_definition(overrides):
    pp, sp = prism_paths()
    inp = q.read_inputs(pp, sp, overrides)
    t = q.build_tables(inp)
    nh = 8 * inp.N[0] * inp.N[1] * inp.N[2]
    rng = np.random.default_rng(5)
    V = rng.standard_normal((nh, 3)) + 1j * rng.standard_normal((nh, 3))
    terms = [(0, 0, 0.7), (1, 2, -0.3), (2, 1, 0.25), (1, 1, 1.5)]
    Wf, Wd = q.selftest_adjoint_bin_host(t, V, terms)
    scale = np.abs(Wd).max()
    assert scale > 0
    assert np.abs(Wf - Wd).max() <= 1e-12 * scale, np.abs(Wf - Wd).max() / scale
    # every table that the element formula uses receives weight
    assert all(np.abs(Wd[k]).max() > 0 for k in range(q.QD_NTAB))
