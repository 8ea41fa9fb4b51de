"""The qd_pwem drop-in driver: input errors on the CPU; on the GPU, the Prism_111 example run the
way the reference runs it (parameters.in + supercell.in in the working directory, -key=value
overrides) against the reference's PWEM_test_ref output files."""
import os
import shutil
import subprocess

import numpy as np
import pytest

import qudos_amd as q
from helpers import GOLDEN, momentum_pair_sums, read_momentum_file
from oracle import qd_coulomb as qc

DRIVER = os.path.join(os.path.dirname(q.lib_path()), "qd_pwem")


def _run(cwd, *args, timeout=300):
    return subprocess.run([DRIVER, *args], cwd=cwd, capture_output=True, text=True, timeout=timeout)


def test_driver_reports_missing_inputs(tmp_path):
    r = _run(tmp_path)
    assert r.returncode != 0
    assert "parameters.in" in r.stderr


def test_driver_rejects_malformed_override(tmp_path):
    for name in ("parameters.in", "supercell.in"):
        shutil.copy(os.path.join(GOLDEN, "prism_111", name), tmp_path)
    r = _run(tmp_path, "M_Gx=3")
    assert r.returncode != 0 and "must begin with '-'" in r.stderr
    r = _run(tmp_path, "-M_Gx=three")
    assert r.returncode != 0 and "non-numeric" in r.stderr


def test_driver_help():
    r = subprocess.run([DRIVER, "-h"], capture_output=True, text=True, timeout=30)
    assert r.returncode == 0 and "usage" in r.stdout


@pytest.mark.gpu
@pytest.mark.parametrize("fixture,precision", [("prism_111", "0"), ("prism_111", "1"), ("truncated_pyramid", "1")])
def test_driver_fixture(tmp_path, fixture, precision):
    """Prism_111 (free strain and Coulomb cells) and Truncated_pyramid (commensurate strain cell
    3 x 3 x 4 and Coulomb cell 1 x 1 x 2, 100 graded slices)."""
    for name in ("parameters.in", "supercell.in"):
        shutil.copy(os.path.join(GOLDEN, fixture, name), tmp_path)
    r = _run(tmp_path, "-Eigensolver=1", f"-Precision={precision}")
    assert r.returncode == 0, r.stderr
    for name in ("energies_vb.dat", "energies_cb.dat"):
        got = np.loadtxt(os.path.join(tmp_path, name), ndmin=2)
        ref = np.loadtxt(os.path.join(GOLDEN, fixture, name), ndmin=2)
        assert got.shape == ref.shape
        assert np.max(np.abs(got[:, 0] - ref[:, 0])) < 1e-9
        # Bloch characters per Kramers pair (gauge invariant)
        g2 = got[:, 1:].reshape(-1, 2, 4).sum(axis=1)
        r2 = ref[:, 1:].reshape(-1, 2, 4).sum(axis=1)
        assert np.max(np.abs(g2 - r2)) < 1e-6
    if fixture == "prism_111":
        # real-space envelope outputs: linescans, cross-sections and the 3-D grid against the
        # example's probability-density files (Kramers-pair sums, 7 significant digits)
        from test_envelopes_cpu import driver_file_errors
        errs = driver_file_errors(str(tmp_path))
        assert max(errs.values()) < 5e-6, errs
    times = np.loadtxt(os.path.join(tmp_path, "calculation_times.dat"))
    assert times.shape == (21,) and times[-1] > 0 and times[19] > 0 and times[16] > 0
    assert (times[15] > 0) == (fixture == "truncated_pyramid")   # Localisation: 1 only there
    assert times[11] > 0 and times[12] > 0   # both examples request linescans and cross-sections
    # carrier localisation (per state and shape, invariant within Kramers pairs) and overlaps
    if fixture == "truncated_pyramid":
        for name in ("carrier_localisation_vb.dat", "carrier_localisation_cb.dat"):
            got = np.loadtxt(os.path.join(tmp_path, name), ndmin=2)
            ref = np.loadtxt(os.path.join(GOLDEN, fixture, name), ndmin=2)
            assert got.shape == ref.shape
            assert np.max(np.abs(got - ref) / np.maximum(1.0, np.abs(ref))) < 1e-5, name
        got = np.loadtxt(os.path.join(tmp_path, "overlaps.dat"), ndmin=2)
        ref = np.loadtxt(os.path.join(GOLDEN, fixture, "overlaps.dat"), ndmin=2)
        assert got.shape == ref.shape and np.array_equal(got[:, :2], ref[:, :2])
        assert np.abs(got[:, 2:]).max() < 1e-10
    # momentum matrix elements: same rows; Kramers-pair block sums of |p|^2 per polarisation
    fn = "momentum_matrix_elements.dat"
    got_p, got_m = read_momentum_file(os.path.join(tmp_path, fn))
    ref_p, ref_m = read_momentum_file(os.path.join(GOLDEN, fixture, fn))
    assert got_p.shape == ref_p.shape
    gs, rs = momentum_pair_sums(got_m), momentum_pair_sums(ref_m)
    assert np.abs(gs - rs).max() <= 1e-6 * rs.max(), np.abs(gs - rs).max()
    # Coulomb files: same line layout as the reference; direct hh/ee/eh diagonal blocks
    # (Kramers-pair density sums, gauge invariant) agree to 1e-6
    for name in ("hh", "ee", "eh_direct", "eh_exchange"):
        fn = f"coulomb_matrix_elements_{name}.dat"
        got = open(os.path.join(tmp_path, fn)).read().splitlines()
        ref = open(os.path.join(GOLDEN, fixture, fn)).read().splitlines()
        assert len(got) == len(ref)
        assert [g.split()[:4] for g in got] == [x.split()[:4] for x in ref]
    for name in ("ee", "eh_exchange"):
        fn = f"coulomb_matrix_elements_{name}.dat"
        from test_coulomb_oracle import block_norms
        mine = block_norms(qc.read_coulomb_file(os.path.join(tmp_path, fn)))
        ref = block_norms(qc.read_coulomb_file(os.path.join(GOLDEN, fixture, fn)))
        assert set(mine) == set(ref)
        # 1e-6 relative for every block: the driver refines the vectors that feed the Coulomb
        # elements to ||H x - E x|| <= 5e-12 ||T||_1 (two sweeps), which keeps even the 3e-12 ee
        # blocks of this fixture, ten orders below the largest, within 1e-6 of themselves
        for k in ref:
            assert abs(mine[k] - ref[k]) <= 1e-6 * ref[k], (fn, k, mine[k], ref[k])
