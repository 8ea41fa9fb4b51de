"""Differentiable QUDOS objectives: a thin ``torch.autograd.Function`` over the C ABI.

The objective is L = (E_e1 - E_h1) + J_eh, with E_h1 the highest hole level, E_e1 the lowest
electron level and J_eh = 1/4 sum_{i in h1, k in e1} V^eh_direct_iikk, the direct Coulomb energy of
the two Kramers levels.  Parameters theta = (dot height [nm], dot base radius [nm], In fraction of
dot and wetting layer, Fx, Fy, Fz [kV/cm]).

forward : qd_tables_build (host FP64) -> qd_upload_tables -> qd_solve_device -> qd_coulomb.
backward: qd_adjoint gives W = dL/dT over the Delta-G tables (Hellmann-Feynman for the energies,
          Sternheimer solutions for the eigenvector part of J_eh, both binned over G - G' on the
          GPU); dL/dtheta = Re sum W * dT/dtheta with qd_tables_derivative (analytic for the field,
          central differences of the FP64 table builder for the shape lengths and composition).
"""
from __future__ import annotations

import time

import numpy as np

from . import _lib as L

PARAM_NAMES = ("height_nm", "radius_nm", "in_fraction", "Fx_kV_cm", "Fy_kV_cm", "Fz_kV_cm")


class DotProblem:
    """One dot configuration on one GPU handle: parameters.in + supercell.in with the dot at shape
    index ``dot`` (base radius = Length 1 = Length 2, height = Length 3) and the wetting layer at
    ``wl`` (its composition follows the dot's)."""

    def __init__(self, parameters_path, supercell_path, dot=2, wl=1, overrides=(), precision=L.QD_FP64,
                 solver=None, fd_rel_step=1e-5):
        self.inp = L.read_inputs(parameters_path, supercell_path, overrides)
        self.dot, self.wl = dot, wl
        self.precision = precision
        self.solver = solver or L.Solver(0)
        self.fd_rel_step = fd_rel_step
        self.tables = None
        self.timings = {}

    # ---- parameters ----
    def get_theta(self):
        d = self.inp.shapes[self.dot]
        return np.array([d.L3 / 10.0, d.L1 / 10.0, 1.0 - d.x] + [1e5 * self.inp.F[a] for a in range(3)])

    def set_theta(self, theta):
        th = [float(v) for v in theta]
        d = self.inp.shapes[self.dot]
        d.L3 = 10.0 * th[0]
        d.L1 = d.L2 = 10.0 * th[1]
        for s in (self.dot, self.wl):
            if s is not None and s >= 1:
                self.inp.shapes[s].x = 1.0 - th[2]
        for a in range(3):
            self.inp.F[a] = 1e-5 * th[3 + a]

    def levels(self):
        """0-based window indices of the first state of the h1 and e1 Kramers levels."""
        nvb = 2 * self.inp.N_vb
        return nvb - 2, nvb

    # ---- forward ----
    def forward(self, theta):
        self.set_theta(theta)
        t0 = time.perf_counter()
        if self.tables is not None:
            L.load_library().qd_tables_free(L.C.byref(self.tables))
        self.tables = L.build_tables(self.inp)
        t1 = time.perf_counter()
        self.solver.upload(self.tables)
        il, iu = L.window(self.inp)
        st = self.solver.solve_device(il, iu, self.precision)
        w, _ = self.solver.fetch(iu - il + 1, vectors=False)
        cme, cst = self.solver.coulomb(self.inp, self.tables, 2, 2)
        ed = cme["eh_direct"]
        J = 0.25 * sum(ed[i, i, k, k].real for i in range(2) for k in range(2))
        h1, e1 = self.levels()
        t2 = time.perf_counter()
        self.w = w
        self.timings = {"tables_s": t1 - t0, "forward_s": t2 - t1, "solve_s": st["t_total"],
                        "coulomb_s": cst["t_total"]}
        return (w[e1] - w[h1]) + J, {"E_h1": w[h1], "E_e1": w[e1], "J_eh": J}

    # ---- adjoint ----
    def gradient(self):
        """dL/dtheta after forward(); returns (grad, info)."""
        t0 = time.perf_counter()
        k = len(self.w)
        h1, e1 = self.levels()
        gE = np.zeros(k)
        gE[e1] = gE[e1 + 1] = 0.5
        gE[h1] = gE[h1 + 1] = -0.5
        W, st = self.solver.adjoint(self.inp, self.tables, gE, h1, e1, 1.0)
        t1 = time.perf_counter()
        mask_dot = 1 << self.dot
        mask_x = mask_dot | ((1 << self.wl) if self.wl is not None and self.wl >= 1 else 0)
        specs = [(L.QD_DP_L3, mask_dot, 10.0), (L.QD_DP_L12, mask_dot, 10.0), (L.QD_DP_X, mask_x, -1.0),
                 (L.QD_DP_FX, 0, 1e-5), (L.QD_DP_FY, 0, 1e-5), (L.QD_DP_FZ, 0, 1e-5)]
        grad = np.zeros(6)
        for j, (param, mask, scale) in enumerate(specs):
            dT = L.tables_derivative(self.inp, param, mask, self.fd_rel_step)
            grad[j] = scale * float(np.sum(W.real * dT.real - W.imag * dT.imag))
        t2 = time.perf_counter()
        info = dict(st)
        info.update({"adjoint_s": t1 - t0, "table_derivatives_s": t2 - t1})
        return grad, info

    def close(self):
        if self.tables is not None:
            L.load_library().qd_tables_free(L.C.byref(self.tables))
            self.tables = None


def _torch_function():
    import torch

    class GapCoulombObjective(torch.autograd.Function):
        """L(theta) = (E_e1 - E_h1) + J_eh; backward through qd_adjoint (see module docstring)."""

        @staticmethod
        def forward(ctx, theta, problem):
            val, _ = problem.forward(theta.detach().cpu().double().numpy())
            grad, _ = problem.gradient()
            ctx.save_for_backward(torch.as_tensor(grad, dtype=theta.dtype, device=theta.device))
            return theta.new_tensor(val)

        @staticmethod
        def backward(ctx, gout):
            (grad,) = ctx.saved_tensors
            return gout * grad, None

    return GapCoulombObjective


def gap_coulomb_objective(theta, problem: DotProblem):
    """torch entry point: L(theta) with autograd support (theta: 6-element float64 tensor)."""
    return _torch_function().apply(theta, problem)
