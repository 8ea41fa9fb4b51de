"""The library's HIP kernels executed on the CPU under the host emulation (tests/emulated.py).

The same .hip sources that hipcc builds for gfx950 run here with work-items as fibers and the
wavefront instructions modelled lane by lane, through the same C ABI and the same test code as
the GPU tests, at sizes the emulation finishes in seconds.

  * Calibration: the instruction models (the MFMA operand and accumulator layouts, the DPP lane
    reductions) and the eigensolver, both of which pass their GPU tests on the MI355X, give the
    same results under the emulation.
  * Kernels whose GPU tests have not run on an MI355X: momentum matrix elements, carrier
    localisation and overlaps, the FFT-correlation binning of the adjoint, the commensurate
    Coulomb cells, the adjoint chained through the table builder's reverse pass (the QD dot and
    the QD molecule) and the status reporting, each against the oracle or central differences
    exactly as its GPU test checks it.
  * Coefficient generation on the device (shape transforms, strain, piezoelectric fields, the
    transfer to the electronic grid, convolutions, tables) and its reverse pass against the host
    FP64 builder and reverse pass, in every cell mode, and gradients with both passes on the device.
  * Memory safety: a child process runs the same paths on the AddressSanitizer build, with
    array-bounds checks on every fixed-size array (work-group shared arrays included).

What the emulation cannot show: code generation for gfx950, timing, and memory-ordering effects
between work-groups beyond sequential consistency.
"""
import os

import numpy as np
import pytest

import emulated
import qudos_amd as q
import test_gpu_adjoint as gadj
import test_gpu_coulomb as gcoul
import test_gpu_diag as gdiag
import test_gpu_fields as gfld
import test_gpu_momentum as gmom
import test_gpu_solve as gsolve
import test_gpu_twostage as gtwo
from helpers import GOLDEN, prism_paths
from qudos_amd.autograd import DotProblem, MoleculeProblem

TINY = ("-M_Gx=1", "-M_Gy=1", "-M_Gz=1", "-M_S_Gx=1", "-M_S_Gy=1", "-M_S_Gz=1")
TINY_FIELD = TINY + ("-Fx=12", "-Fy=-5", "-Fz=20", "-Kz=0.02")
TINY_K = ("-M_Gx=1", "-M_Gy=1", "-M_Gz=2", "-M_S_Gx=1", "-M_S_Gy=1", "-M_S_Gz=2", "-Kx=0.13", "-Kz=-0.2", "-Fz=12")
LENS = os.path.join(os.path.dirname(q.lib_path()), "..", "configs", "lens_11x11x9")
LENS_TINY = ("-M_Gx=1", "-M_Gy=1", "-M_Gz=2", "-M_S_Gx=1", "-M_S_Gy=1", "-M_S_Gz=2", "-M_C_Gx=1", "-M_C_Gy=1",
             "-M_C_Gz=2", "-N_VB=2", "-N_CB=2")


@pytest.fixture(scope="module", autouse=True)
def _emulation():
    with emulated.emulated():
        yield


@pytest.fixture(scope="module")
def solver(_emulation):
    s = q.Solver(0)
    yield s
    s.close()


# ---- calibration: code that passes on the MI355X ------------------------------------------------
def test_instruction_models():
    gdiag.test_mfma_layout()
    gdiag.test_dpp_lane_reductions()


@pytest.mark.parametrize("n", [2, 33, 65, 97, 130])
@pytest.mark.parametrize("prec", [q.QD_FP64, q.QD_FP32_REFINE])
def test_two_stage_eigensolver(solver, n, prec):
    if prec == q.QD_FP32_REFINE and n < 31:
        pytest.skip("the refinement's Rayleigh-Ritz space [X E] needs n >= 2k")
    gtwo.test_two_stage_sizes(solver, n, prec)


@pytest.mark.parametrize("prec", [q.QD_FP64, q.QD_FP32_REFINE])
def test_random_matrix_window(solver, prec):
    gsolve.test_eigensolver_random_matrix(solver, 300, 101, 20, prec)


def test_one_stage_matches_two_stage(solver):
    # FP32 + refinement exercises both reductions' FP32 kernels and the FP64 refinement
    gtwo.test_two_stage_matches_one_stage(solver, q.QD_FP32_REFINE)


@pytest.mark.parametrize("overrides", [TINY, TINY_FIELD])
def test_hamiltonian_matches_oracle(solver, overrides):
    gsolve.test_hamiltonian_matches_oracle(solver, overrides)


@pytest.mark.parametrize("prec", [q.QD_FP64, q.QD_FP32_REFINE])
@pytest.mark.parametrize("overrides", [TINY, TINY_FIELD])
def test_prism_end_to_end(solver, prec, overrides):
    gsolve.test_small_prism_end_to_end(solver, prec, overrides)


# ---- kernels not yet run on an MI355X -------------------------------------------------------------
@pytest.mark.parametrize("overrides", [TINY, TINY_K])
def test_momentum_matches_oracle(solver, overrides):
    gmom.test_momentum_matches_oracle(solver, overrides)


def test_localisation_and_overlaps_match_oracle(solver):
    import test_gpu_outputs as gout
    d = os.path.join(GOLDEN, "truncated_pyramid")
    pp, sp = os.path.join(d, "parameters.in"), os.path.join(d, "supercell.in")
    # the example's commensurate strain cell (3 x 3 x 4 electronic cells) at the smallest sizes
    over = ("-M_Gx=1", "-M_Gy=1", "-M_Gz=2", "-M_S_Gx=3", "-M_S_Gy=3", "-M_S_Gz=5")
    inp = q.read_inputs(pp, sp, over)
    solver.upload(q.build_tables(inp))
    il, iu = q.window(inp)
    w, Z, _ = solver.solve(il, iu, q.QD_FP64)
    vb, cb, _ = solver.localisation(inp)
    nvb = 2 * inp.N_vb
    ref = gout._oracle_loc(inp, sp, pp, Z)
    assert np.abs(vb - ref[nvb - 1::-1]).max() < 1e-10
    assert np.abs(cb - ref[nvb:]).max() < 1e-10
    ov = solver.overlaps(inp)
    assert ov.shape == (2 * inp.N_cb, nvb)
    assert np.abs(ov - Z[:, nvb:].conj().T @ Z[:, nvb - 1::-1]).max() < 1e-13


def test_envelopes_match_oracle(solver):
    from oracle import qd_oracle as qo
    pp, sp = prism_paths()
    inp = q.read_inputs(pp, sp, TINY_K)
    solver.upload(q.build_tables(inp))
    il, iu = q.window(inp)
    w, Z, _ = solver.solve(il, iu, q.QD_FP64)
    rng = np.random.default_rng(7)
    x, y, z = (rng.uniform(-120.0, 120.0, 300) for _ in range(3))
    L = (inp.shapes[0].L1, inp.shapes[0].L2, inp.shapes[0].L3)
    ref = qo.envelopes(Z, tuple(inp.N), L, tuple(inp.K), x, y, z)
    F = solver.envelopes(inp, x, y, z)
    assert F.shape == ref.shape
    assert np.abs(F - ref).max() <= 1e-12 * np.abs(ref).max()
    # a sub-window that ends inside a group of states, at fewer points than a work-group
    F2 = solver.envelopes(inp, x[:7], y[:7], z[:7], first=3, count=6)
    assert np.abs(F2 - ref[3:9, :, :7]).max() <= 1e-12 * np.abs(ref).max()


def test_driver_writes_envelope_files(tmp_path):
    """qd_pwem on the emulated library with the example's linescans, cross-sections and 3-D grid
    at small sizes: file names, positions against the oracle's grids, and Kramers-pair sums of the
    probability densities against the oracle's envelopes of its own solve of the same problem."""
    import shutil
    import subprocess
    import test_envelopes_cpu as tenv
    from helpers import oracle_problem
    from oracle import qd_oracle as qo
    pp, sp = prism_paths()
    for f in (pp, sp):
        shutil.copy(f, tmp_path)
    over = TINY + ("-N_line=9", "-N_cross=5", "-N_3D=4", "-CME=0", "-MME=0")
    emulated.ensure_built(emulated.DRIVER)
    r = subprocess.run([emulated.DRIVER, *over, "-Precision=0"], cwd=tmp_path, capture_output=True, text=True,
                       timeout=600)
    assert r.returncode == 0, r.stderr
    p, sh = oracle_problem(pp, sp, over)
    w, Z = qo.solve(qo.build_hamiltonian(qo.coefficient_tables(p, sh)), p)
    nvb, ncb = 2 * p.N_vb, 2 * p.N_cb
    L = (sh[0].L1, sh[0].L2, sh[0].L3)
    names = [f"vb_{i}" for i in range(1, nvb + 1)] + [f"cb_{i}" for i in range(1, ncb + 1)]
    grids = [("", f"_{d}.dat", qo.linescan_grid(i, 9, L)) for i, d in enumerate(qo.LINESCAN_NAMES)
             if d in ("100", "010", "001", "110", "111")]
    grids += [("_cross_section", f"_{d}.dat", qo.cross_section_grid(i, 5, L)) for i, d in enumerate(qo.CROSS_NAMES)
              if d in ("100", "010", "001", "-110", "111")]
    grids.append(("_3D", ".dat", qo.grid_3d(4, L)))
    for kind, suffix, (x, y, z) in grids:
        dens = np.stack([np.loadtxt(tmp_path / f"probability_density{kind}_{n}{suffix}", ndmin=2) for n in names])
        env = np.stack([np.loadtxt(tmp_path / f"envelope_function{kind}_{n}{suffix}", ndmin=2) for n in names])
        assert dens.shape == (nvb + ncb, len(x), 9) and env.shape == (nvb + ncb, len(x), 20), (kind, suffix)
        assert np.abs(dens[0, :, :3] - 0.1 * np.stack([x, y, z], axis=1)).max() <= 1e-6 * max(L)
        assert np.abs(dens[:, 0, 3] - np.concatenate([w[nvb - 1::-1], w[nvb:]])).max() < 1e-6
        F = qo.envelopes(Z, tuple(p.N), L, tuple(p.K), x, y, z)
        got = tenv.pair_sums(dens[:, :, 4:])
        want = tenv.pair_sums(tenv.fixture_order(F, nvb, grid3d=kind == "_3D"))
        scale = want.max(axis=(1, 2))[:, None, None]
        assert (np.abs(got - want) / scale).max() < 2e-6, (kind, suffix)
        # the envelope files' components: |F|^2 summed per band pair reproduces the densities
        c = env[:, :, 4:].reshape(len(names), len(x), 8, 2)
        band_dens = (c ** 2).sum(axis=3)   # sqrt(1000) F: SO1, SO2, LH1, LH2, HH1, HH2, CB1, CB2 in nm^-3
        per_band = band_dens.reshape(len(names), len(x), 4, 2).sum(axis=3)
        assert np.abs(per_band - dens[:, :, 5:]).max() <= 1e-5 * dens[:, :, 5:].max()


def test_adjoint_bin_equals_trace_with_kernel_dH(solver):
    gadj.test_adjoint_bin_equals_trace_with_kernel_dH(solver)


@pytest.mark.parametrize("case", gcoul.CASES, ids=[f"cme{c[5]}_type{c[4]}" + ("_rcl" if len(c) > 8 else "")
                                                   for c in gcoul.CASES])
def test_coulomb_matches_oracle(solver, case):
    gcoul.test_coulomb_random_states_match_oracle(solver, case)


def test_coulomb_rejects_unsupported_commensurate_cell(solver):
    gcoul.test_coulomb_rejects_unsupported_commensurate_cell(solver)


@pytest.mark.parametrize("prec", [q.QD_FP64, q.QD_FP32_REFINE])
def test_dot_gradient_matches_finite_differences(solver, prec):
    pp, sp = os.path.join(LENS, "parameters.in"), os.path.join(LENS, "supercell.in")
    prob = DotProblem(pp, sp, overrides=LENS_TINY, precision=prec, solver=solver)
    try:
        theta = gadj._draw(2)
        val, info = prob.forward(theta)
        g, ginfo = prob.gradient()
        assert abs(ginfo["J"] - info["J_eh"]) <= 1e-10 * abs(info["J_eh"])
        fd = gadj._fd_gradient(prob, theta)
    finally:
        prob.close()
    scale = np.abs(fd).max()
    for j in range(6):
        assert abs(g[j] - fd[j]) <= 2e-5 * scale + 1e-6 * abs(fd[j]), (j, g, fd)


def test_gradient_with_device_tables(solver):
    import test_gpu_molecule as gmol
    pp, sp = os.path.join(LENS, "parameters.in"), os.path.join(LENS, "supercell.in")
    gadj.check_device_tables_gradient(
        lambda dev: DotProblem(pp, sp, overrides=LENS_TINY, precision=q.QD_FP64, solver=solver, device_tables=dev),
        gadj._draw(2))
    pm, sm = os.path.join(gmol.CFG, "parameters.in"), os.path.join(gmol.CFG, "supercell.in")
    over = ("-M_Gx=1", "-M_Gy=1", "-M_Gz=2", "-M_S_Gx=1", "-M_S_Gy=1", "-M_S_Gz=2", "-N_VB=2", "-N_CB=2")
    gadj.check_device_tables_gradient(
        lambda dev: MoleculeProblem(pm, sm, gmol.LOWER, gmol.UPPER, overrides=over, precision=q.QD_FP64, solver=solver,
                                    device_tables=dev), gmol.draw(1))


@pytest.mark.parametrize("prec", [q.QD_FP64, q.QD_FP32_REFINE])
def test_adjoint_z_matches_finite_differences(solver, prec):
    gadj.check_adjoint_z(solver, LENS_TINY, prec)


def test_molecule_gradient_matches_finite_differences(solver):
    import test_gpu_molecule as gmol
    pp, sp = os.path.join(gmol.CFG, "parameters.in"), os.path.join(gmol.CFG, "supercell.in")
    over = ("-M_Gx=1", "-M_Gy=1", "-M_Gz=2", "-M_S_Gx=1", "-M_S_Gy=1", "-M_S_Gz=2", "-M_C_Gx=1", "-M_C_Gy=1",
            "-M_C_Gz=2", "-N_VB=2", "-N_CB=2")
    prob = MoleculeProblem(pp, sp, gmol.LOWER, gmol.UPPER, overrides=over, precision=q.QD_FP64, solver=solver)
    theta = gmol.draw(1)
    try:
        val, _ = prob.forward(theta)
        g, _ = prob.gradient()
        fd = np.zeros(4)
        for j in range(4):
            h = 1e-4 * theta[0] if j == 0 else 1e-4 * 30.0
            tp, tm = theta.copy(), theta.copy()
            tp[j] += h
            tm[j] -= h
            fd[j] = (prob.forward(tp)[0] - prob.forward(tm)[0]) / (2 * h)
    finally:
        prob.close()
    scale = np.abs(fd).max()
    assert val > 0
    for j in range(4):
        assert abs(g[j] - fd[j]) <= 2e-5 * scale + 1e-6 * abs(fd[j]), (j, g, fd)


@pytest.mark.parametrize("limits,prec", [
    (dict(refine_sweeps=0), q.QD_FP32_REFINE),
    (dict(invit_sweeps=0), q.QD_FP64),
    (dict(bisect_rounds=1), q.QD_FP64),
])
def test_limit_reports_noconv(limits, prec):
    import test_gpu_status as gstat
    pp, sp = prism_paths()
    inp = q.read_inputs(pp, sp, TINY)
    s = q.Solver(0)
    try:
        s.upload(q.build_tables(inp))
        il, iu = q.window(inp)
        w_ok, _, _ = s.solve(il, iu, q.QD_FP64)
        gstat.test_limit_reports_noconv((s, il, iu, w_ok), limits, prec)
        gstat.test_marked_wait_failure_fails_later_uses((s, il, iu, w_ok))
    finally:
        s.close()


def test_adjoint_noconv_and_wait_failure():
    pp, sp = os.path.join(LENS, "parameters.in"), os.path.join(LENS, "supercell.in")
    s = q.Solver(0)
    prob = DotProblem(pp, sp, overrides=LENS_TINY, precision=q.QD_FP32_REFINE, solver=s)
    try:
        theta = np.array([5.0, 10.0, 0.9, 3.0, -2.0, 8.0])
        prob.forward(theta)
        s.set_limits(sternheimer_iters=0)
        with pytest.raises(q.QdError) as e:
            prob.gradient()
        assert e.value.code == q.QD_E_NOCONV, str(e.value)
        s.set_limits()
        g, info = prob.gradient()
        assert info["sternheimer_residual"] < 1e-12
        s.mark_wait_failure()
        with pytest.raises(q.QdError) as e:
            prob.gradient()
        assert e.value.code == q.QD_E_DEVICE, str(e.value)
        prob.forward(theta)
        g2, _ = prob.gradient()
        assert np.abs(g2 - g).max() <= 1e-9 * np.abs(g).max()
    finally:
        prob.close()
        s.close()


# ---- coefficient generation on the device ---------------------------------------------------------
@pytest.mark.parametrize("name,over", gfld.CASES)
def test_device_tables_match_host_builder(solver, name, over):
    gfld.check_device_tables(solver, name, over)


# full size: the bench's lens and the QD molecule's 12 slices (the Truncated_pyramid at full size, 29 s
# here, runs in the GPU tests; its commensurate cell and 100 slices are covered above at small size)
@pytest.mark.parametrize("name,over", [c for c in gfld.FULL if c[0] != "truncated_pyramid"])
def test_device_tables_full_size(solver, name, over):
    gfld.check_device_tables(solver, name, over)


@pytest.mark.parametrize("name,over,dot,wl", gfld.VJP_CASES)
def test_device_vjp_matches_host(solver, name, over, dot, wl):
    gfld.check_device_vjp(solver, name, over, dot, wl)


def test_bench_device_tables_fields(solver):
    """bench.py --device-tables: the timing helper's fields and its comparison with the host tables."""
    import sys
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    import bench
    pp, sp = gfld.cfg("lens_9x9x9")
    inp = q.read_inputs(pp, sp, TINY)
    t = q.build_tables(inp)
    solver.upload(t)
    out = bench.device_tables_timing(q, solver, inp, t, reps=1)
    assert set(out) == {"seconds", "max_rel_diff_vs_host", "reps"} and out["seconds"] > 0
    assert out["max_rel_diff_vs_host"] < 1e-12


def test_device_shapes_match_host(solver, tmp_path):
    gfld.check_device_shapes(solver, tmp_path)


def test_device_tables_resident_and_input_errors(solver):
    gfld.check_resident(solver, "lens_9x9x9", TINY)
    gfld.check_input_errors(solver)


# ---- memory safety: the same paths on the AddressSanitizer build ------------------------------------
SANITIZED = r"""
import os
import numpy as np
import qudos_amd as q
import test_emu_kernels as T
import test_gpu_coulomb as gcoul
import test_gpu_fields as gfld
s = q.Solver(0)
for name, over in gfld.CASES[5:7]:
    gfld.check_device_tables(s, name, over)
for case in gfld.VJP_CASES[5:7]:
    gfld.check_device_vjp(s, *case)
T.test_two_stage_eigensolver(s, 65, q.QD_FP32_REFINE)
T.test_two_stage_eigensolver(s, 33, q.QD_FP64)
T.test_momentum_matches_oracle(s, T.TINY_K)
T.test_localisation_and_overlaps_match_oracle(s)
for case in gcoul.CASES[-2:]:
    T.test_coulomb_matches_oracle(s, case)
pp, sp = os.path.join(T.LENS, "parameters.in"), os.path.join(T.LENS, "supercell.in")
prob = T.DotProblem(pp, sp, overrides=T.LENS_TINY, precision=q.QD_FP32_REFINE, solver=s)
prob.forward(T.gadj._draw(2))
g, info = prob.gradient()
prob.close()
s.close()
print("sanitized run complete", g)
"""


def test_address_sanitizer_run():
    out = emulated.run_sanitized(SANITIZED)
    assert "sanitized run complete" in out
