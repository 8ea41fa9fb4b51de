"""GPU checks of the adjoint (qd_adjoint / qd_adjoint_bin / qudos_amd.autograd):
  * the Delta-G binning equals Re tr(D dH) with dH formed by the Hamiltonian kernel itself;
  * J_eh of the adjoint equals the direct Coulomb elements of qd_coulomb;
  * dL/dtheta for L = (E_e1 - E_h1) + J_eh matches central finite differences of L (step 1e-4
    relative; field step 1e-4 x 20 kV/cm) on seeded random lens parameters, for the FP64 and
    the FP32 + refinement paths."""
import ctypes as C
import os

import numpy as np
import pytest

import qudos_amd as q
from qudos_amd.autograd import DotProblem

pytestmark = pytest.mark.gpu

CFG = os.path.join(os.path.dirname(q.lib_path()), "..", "configs", "lens_11x11x9")
SMALL = ("-M_Gx=2", "-M_Gy=2", "-M_Gz=2", "-M_S_Gx=2", "-M_S_Gy=2", "-M_S_Gz=2", "-M_C_Gx=2", "-M_C_Gy=2",
         "-M_C_Gz=2", "-N_VB=4", "-N_CB=4")
MED = ("-M_Gx=3", "-M_Gy=3", "-M_Gz=3", "-M_S_Gx=3", "-M_S_Gy=3", "-M_S_Gz=3", "-M_C_Gx=3", "-M_C_Gy=3",
       "-M_C_Gz=3", "-N_VB=4", "-N_CB=4")


@pytest.fixture(scope="module")
def solver():
    s = q.Solver(0)
    yield s
    s.close()


def _paths():
    return os.path.join(CFG, "parameters.in"), os.path.join(CFG, "supercell.in")


def _tab_view(t):
    nt = t.NT[0] * t.NT[1] * t.NT[2]
    raw = np.ctypeslib.as_array(t.tab, shape=(q.QD_NTAB * nt * 2,))
    return raw.view(np.complex128).reshape(q.QD_NTAB, t.NT[0], t.NT[1], t.NT[2])


def test_adjoint_bin_equals_trace_with_kernel_dH(solver):
    pp, sp = _paths()
    inp = q.read_inputs(pp, sp, SMALL + ("-Fx=5", "-Fz=-12"))
    t = q.build_tables(inp)
    solver.upload(t)
    n = solver.n
    NT = tuple(t.NT)
    rng = np.random.default_rng(4)
    V = rng.standard_normal((n, 3)) + 1j * rng.standard_normal((n, 3))
    terms = [(0, 0, 0.7), (1, 1, -1.3), (2, 0, 0.4), (0, 2, 0.4)]   # Hermitian D
    W = solver.adjoint_bin(V, terms, NT)
    D = sum(g * np.outer(V[:, a], V[:, b].conj()) for a, b, g in terms)
    view = _tab_view(t)
    base = view.copy()
    for param, mask in [(q.QD_DP_FZ, 0), (q.QD_DP_L3, 1 << 2), (q.QD_DP_X, (1 << 1) | (1 << 2))]:
        dT = q.tables_derivative(inp, param, mask, 1e-5)
        view[...] = base + dT
        solver.upload(t)
        Hp = solver.hamiltonian(q.QD_FP64)
        view[...] = base - dT
        solver.upload(t)
        Hm = solver.hamiltonian(q.QD_FP64)
        view[...] = base
        dHl = 0.5 * (np.tril(Hp) - np.tril(Hm))
        dH = dHl + np.tril(dHl, -1).conj().T
        dH[np.diag_indices(n)] = dH[np.diag_indices(n)].real
        want = np.real(np.sum(D.T * dH))           # Re tr(D dH) = Re sum_{r,c} D[c,r] dH[r,c]
        got = float(np.sum(W.real * dT.real - W.imag * dT.imag))
        assert abs(got - want) <= 1e-9 * max(1.0, abs(want)), (param, got, want)
    solver.upload(t)


def _fd_gradient(problem, theta, rel=1e-4):
    g = np.zeros(6)
    for j in range(6):
        h = rel * abs(theta[j]) if j < 3 else rel * 20.0
        tp = theta.copy(); tp[j] += h
        tm = theta.copy(); tm[j] -= h
        g[j] = (problem.forward(tp)[0] - problem.forward(tm)[0]) / (2 * h)
    return g


def _draw(seed):
    rng = np.random.default_rng(seed)
    h = rng.uniform(4, 6)
    R = rng.uniform(8, 12)
    f = rng.uniform(0.6, 1.0)
    u = rng.standard_normal(3)
    u /= np.linalg.norm(u)
    F = rng.uniform(0, 20) * u
    return np.array([h, R, f, *F])


@pytest.mark.parametrize("seed", [0, 1])
def test_gradient_matches_finite_differences(solver, seed):
    pp, sp = _paths()
    prob = DotProblem(pp, sp, overrides=MED, precision=q.QD_FP64, solver=solver)
    theta = _draw(seed)
    val, info = prob.forward(theta)
    g, ginfo = prob.gradient()
    assert abs(ginfo["J"] - info["J_eh"]) <= 1e-10 * abs(info["J_eh"])
    assert ginfo["sternheimer_residual"] < 1e-10
    fd = _fd_gradient(prob, theta)
    scale = np.abs(fd).max()
    for j in range(6):
        assert abs(g[j] - fd[j]) <= 2e-5 * scale + 1e-6 * abs(fd[j]), (j, g, fd)
    prob.close()


def test_fp32_refine_gradient_matches_fp64(solver):
    pp, sp = _paths()
    theta = _draw(3)
    out = {}
    for prec in (q.QD_FP64, q.QD_FP32_REFINE):
        prob = DotProblem(pp, sp, overrides=MED, precision=prec, solver=solver)
        prob.forward(theta)
        out[prec] = prob.gradient()[0]
        prob.close()
    a, b = out[q.QD_FP64], out[q.QD_FP32_REFINE]
    assert np.abs(a - b).max() <= 1e-6 * np.abs(a).max()


def check_device_tables_gradient(make_problem, theta):
    """The same objective and gradient with both table passes on the device (qd_tables_build_device,
    qd_tables_vjp_device) as with the host builder and its reverse pass."""
    out = {}
    for dev in (False, True):
        prob = make_problem(dev)
        try:
            val, _ = prob.forward(theta)
            g, info = prob.gradient()
        finally:
            prob.close()
        assert ("device" in info["table_derivatives"]) == dev
        out[dev] = (val, g)
    (v0, g0), (v1, g1) = out[False], out[True]
    assert abs(v1 - v0) <= 1e-12 * abs(v0), (v0, v1)
    assert np.abs(g1 - g0).max() <= 1e-9 * np.abs(g0).max(), (g0, g1)


def test_gradient_with_device_tables(solver):
    pp, sp = _paths()
    check_device_tables_gradient(
        lambda dev: DotProblem(pp, sp, overrides=MED, precision=q.QD_FP64, solver=solver, device_tables=dev), _draw(4))


def _projector_objective(Z, w, levels, vecs, gE):
    """L = sum_p sum_{n in level p} |a_p^H psi_n|^2 + sum gE E: invariant under unitary mixing within
    each level; returns L and the columns and gradients dL/dpsi_n^* = a_p a_p^H psi_n."""
    L = float(np.dot(gE, w))
    cols, gZ = [], []
    for (first, count), a in zip(levels, vecs):
        for n in range(first, first + count):
            c = np.vdot(a, Z[:, n])
            L += abs(c) ** 2
            cols.append(n)
            gZ.append(a * c)
    return L, cols, np.stack(gZ, axis=1)


def check_adjoint_z(solver, overrides, prec, rel_tol=2e-5):
    """qd_adjoint_z against central differences of a gauge-invariant objective of two levels' eigenvectors
    and two energies, for the field F_z and the dot height (tables_derivative gives dT/dtheta)."""
    pp, sp = _paths()
    inp = q.read_inputs(pp, sp, overrides + ("-Fx=4", "-Fz=-9"))
    il, iu = q.window(inp)
    k = iu - il + 1
    nvb = 2 * inp.N_vb
    levels = [(nvb - 2, 2), (nvb, 2)]            # the highest hole level and the lowest electron level
    rng = np.random.default_rng(12)
    n = 8 * inp.N[0] * inp.N[1] * inp.N[2]
    vecs = [v / np.linalg.norm(v) for v in (rng.standard_normal((2, n)) + 1j * rng.standard_normal((2, n)))]
    gE = np.zeros(k)
    gE[nvb - 2:nvb] = -0.5
    gE[nvb:nvb + 2] = 0.5

    def objective():
        solver.upload(q.build_tables(inp))
        w, Z, _ = solver.solve(il, iu, prec)
        return _projector_objective(Z, w, levels, vecs, gE)

    L0, cols, gZ = objective()
    t = q.build_tables(inp)
    W, st = solver.adjoint_z(tuple(t.NT), gE, cols, gZ)
    assert st["sternheimer_residual"] < 1e-12
    for param, mask, ref in [(q.QD_DP_FZ, 0, lambda: inp.F[2]), (q.QD_DP_L3, 1 << 2, lambda: inp.shapes[2].L3)]:
        dT = q.tables_derivative(inp, param, mask, 1e-5)
        grad = float(np.sum(W.real * dT.real - W.imag * dT.imag))
        if param == q.QD_DP_FZ:
            h = 1e-4 * 20.0 * 1e-5                   # 1e-4 x 20 kV/cm in V/Angstrom
            orig = inp.F[2]
            inp.F[2] = orig + h
            Lp = objective()[0]
            inp.F[2] = orig - h
            Lm = objective()[0]
            inp.F[2] = orig
        else:
            orig = inp.shapes[2].L3
            h = 1e-4 * orig
            inp.shapes[2].L3 = orig + h
            Lp = objective()[0]
            inp.shapes[2].L3 = orig - h
            Lm = objective()[0]
            inp.shapes[2].L3 = orig
        fd = (Lp - Lm) / (2 * h)
        assert abs(grad - fd) <= rel_tol * max(abs(fd), 1e-3 * abs(L0)), (param, grad, fd)


@pytest.mark.parametrize("prec", [q.QD_FP64, q.QD_FP32_REFINE])
def test_adjoint_z_matches_finite_differences(solver, prec):
    check_adjoint_z(solver, SMALL, prec)


def test_torch_autograd_function(solver):
    torch = pytest.importorskip("torch")
    from qudos_amd.autograd import gap_coulomb_objective
    pp, sp = _paths()
    prob = DotProblem(pp, sp, overrides=SMALL, precision=q.QD_FP64, solver=solver)
    theta = torch.tensor(_draw(5), dtype=torch.float64, requires_grad=True)
    L = gap_coulomb_objective(theta, prob)
    L.backward()
    g_np, _ = prob.gradient()
    assert torch.allclose(theta.grad, torch.as_tensor(g_np), rtol=1e-12, atol=0)
    prob.close()
