"""Differentiable QUDOS objectives: a thin ``torch.autograd.Function`` over the C ABI.

forward : qd_tables_build (host FP64) -> qd_upload_tables -> qd_solve_device [-> qd_coulomb].
backward: qd_adjoint gives W = dL/dT over the Delta-G tables (Hellmann-Feynman for the energies,
          Sternheimer solutions for the eigenvector part of a Coulomb term, both binned over
          G - G' on the GPU); qd_tables_vjp then carries W back through the FP64 table builder
          (strain, piezoelectric field, shape transforms, alloy interpolation) to every parameter
          in one reverse pass, with the exact derivatives of the shape transforms; the forward
          pass's intermediates are kept with the tables (build_tables(..., retain=True)).
With device_tables=True both table passes run on the GPU instead: qd_tables_build_device in the
forward (the tables are left resident) and qd_tables_vjp_device in the backward (a forward pass of
the fields on the device, then the reverse pass there, contracted with the shapes' exact
derivatives there as well; only the alloy interpolation's derivatives are applied on the host).

Two problems are provided:
  DotProblem      one lens dot on a wetting layer; theta = (height [nm], base radius [nm], In
                  fraction of dot and wetting layer, Fx, Fy, Fz [kV/cm]);
                  L = (E_e1 - E_h1) + J_eh, J_eh = 1/4 sum_{i in h1, k in e1} V^eh_direct_iikk.
  MoleculeProblem two stacked dots, each a stack of shapes (graded slices and a wetting layer);
                  theta = (separation [nm], Fx, Fy, Fz [kV/cm]);
                  L = E_e2 - E_e1, the bonding-antibonding tunnel splitting of the electrons.
"""
from __future__ import annotations

import time

import numpy as np

from . import _lib as L

PARAM_NAMES = ("height_nm", "radius_nm", "in_fraction", "Fx_kV_cm", "Fy_kV_cm", "Fz_kV_cm")
MOLECULE_PARAM_NAMES = ("separation_nm", "Fx_kV_cm", "Fy_kV_cm", "Fz_kV_cm")


class _Problem:
    """Inputs on one GPU handle; subclasses map theta onto the inputs and name the objective."""

    coulomb = False

    def __init__(self, parameters_path, supercell_path, overrides=(), precision=L.QD_FP64, solver=None,
                 refine_tolerance=None, device_tables=False):
        self.inp = L.read_inputs(parameters_path, supercell_path, overrides)
        self.precision = precision
        self.device_tables = device_tables
        self.solver = solver or L.Solver(0)
        self.refine_tolerance = refine_tolerance
        self.tables = None
        self.timings = {}

    # subclass interface
    def set_theta(self, theta):
        raise NotImplementedError

    def objective(self, w, cme):
        """(L, info) from the window eigenvalues w and the Coulomb elements (or None)."""
        raise NotImplementedError

    def weights(self, k):
        """(gE, hole_level, electron_level, wJ) of qd_adjoint for this objective."""
        raise NotImplementedError

    def theta_specs(self):
        """[(QD_DP_*, shape mask, d(parameter)/d(theta))] per component of theta."""
        raise NotImplementedError

    # ---- forward ----
    def forward(self, theta):
        self.set_theta(theta)
        t0 = time.perf_counter()
        if self.tables is not None:
            L.load_library().qd_tables_free(L.C.byref(self.tables))
        if self.device_tables:
            self.tables = self.solver.build_tables(self.inp)   # built on the device, left resident
            t1 = time.perf_counter()
        else:
            self.tables = L.build_tables(self.inp, retain=True)
            t1 = time.perf_counter()
            self.solver.upload(self.tables)
        if self.refine_tolerance is not None:
            self.solver.set_refine_tolerance(self.refine_tolerance)
        il, iu = L.window(self.inp)
        st = self.solver.solve_device(il, iu, self.precision)
        w, _ = self.solver.fetch(iu - il + 1, vectors=False)
        cme, cst = (self.solver.coulomb(self.inp, self.tables, 2, 2) if self.coulomb else (None, {"t_total": 0.0}))
        val, info = self.objective(w, cme)
        t2 = time.perf_counter()
        self.w = w
        self.timings = {"tables_s": t1 - t0, "forward_s": t2 - t1, "solve_s": st["t_total"],
                        "coulomb_s": cst["t_total"]}
        return val, info

    # ---- adjoint ----
    def gradient(self):
        """dL/dtheta after forward(); returns (grad, info)."""
        t0 = time.perf_counter()
        gE, hl, el, wJ = self.weights(len(self.w))
        W, st = self.solver.adjoint(self.inp, self.tables, gE, hl, el, wJ)
        t1 = time.perf_counter()
        specs = self.theta_specs()
        pm = [(p, m) for p, m, _ in specs]
        if self.device_tables:
            g = self.solver.tables_vjp(self.inp, W, pm)
            how = "qd_tables_vjp_device (reverse pass through the table builder on the device)"
        else:
            g = L.tables_vjp(self.inp, W, pm, tables=self.tables)
            how = "qd_tables_vjp (reverse pass through the FP64 table builder)"
        grad = np.array([scale * gj for (_, _, scale), gj in zip(specs, g)])
        t2 = time.perf_counter()
        info = dict(st)
        info.update({"adjoint_s": t1 - t0, "table_derivatives_s": t2 - t1, "table_derivatives": how})
        return grad, info

    def close(self):
        if self.tables is not None:
            L.load_library().qd_tables_free(L.C.byref(self.tables))
            self.tables = None


class DotProblem(_Problem):
    """One dot configuration: parameters.in + supercell.in with the dot at shape index ``dot``
    (base radius = Length 1 = Length 2, height = Length 3) and the wetting layer at ``wl`` (its
    composition follows the dot's)."""

    coulomb = True

    def __init__(self, parameters_path, supercell_path, dot=2, wl=1, overrides=(), precision=L.QD_FP64, solver=None,
                 refine_tolerance=None, device_tables=False):
        super().__init__(parameters_path, supercell_path, overrides, precision, solver, refine_tolerance, device_tables)
        self.dot, self.wl = dot, wl

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

    def objective(self, w, cme):
        ed = cme["eh_direct"]
        J = 0.25 * sum(ed[i, i, k, k].real for i in range(2) for k in range(2))
        h1, e1 = self.levels()
        return (w[e1] - w[h1]) + J, {"E_h1": w[h1], "E_e1": w[e1], "J_eh": J}

    def weights(self, k):
        h1, e1 = self.levels()
        gE = np.zeros(k)
        gE[e1] = gE[e1 + 1] = 0.5
        gE[h1] = gE[h1 + 1] = -0.5
        return gE, h1, e1, 1.0

    def theta_specs(self):
        mask_dot = 1 << self.dot
        mask_x = mask_dot | ((1 << self.wl) if self.wl is not None and self.wl >= 1 else 0)
        # Angstrom per nm for the lengths, x = 1 - f for the composition, V/Angstrom per kV/cm
        return [(L.QD_DP_L3, mask_dot, 10.0), (L.QD_DP_L12, mask_dot, 10.0), (L.QD_DP_X, mask_x, -1.0),
                (L.QD_DP_FX, 0, 1e-5), (L.QD_DP_FY, 0, 1e-5), (L.QD_DP_FZ, 0, 1e-5)]


class MoleculeProblem(_Problem):
    """Two stacked dots: shapes ``lower`` and ``upper`` (lists of supercell.in shape indices, each a
    dot's slices and wetting layer); the separation moves the upper dot's shapes rigidly along z."""

    def __init__(self, parameters_path, supercell_path, lower, upper, overrides=(), precision=L.QD_FP32_REFINE,
                 solver=None, refine_tolerance=None, device_tables=False):
        super().__init__(parameters_path, supercell_path, overrides, precision, solver, refine_tolerance, device_tables)
        self.lower, self.upper = list(lower), list(upper)
        base = min(self.inp.shapes[s].Oz for s in self.lower)
        self._offsets = {s: self.inp.shapes[s].Oz - base for s in self.upper}
        self._base_upper = min(self.inp.shapes[s].Oz for s in self.upper)

    def get_theta(self):
        sep = (self._base_upper - min(self.inp.shapes[s].Oz for s in self.lower)) / 10.0
        return np.array([sep] + [1e5 * self.inp.F[a] for a in range(3)])

    def set_theta(self, theta):
        th = [float(v) for v in theta]
        base = min(self.inp.shapes[s].Oz for s in self.lower)
        shift = base + 10.0 * th[0] - self._base_upper
        for s in self.upper:
            self.inp.shapes[s].Oz += shift
        self._base_upper += shift
        for a in range(3):
            self.inp.F[a] = 1e-5 * th[1 + a]

    def levels(self):
        nvb = 2 * self.inp.N_vb
        return nvb, nvb + 2            # first states of e1 and e2

    def objective(self, w, cme):
        e1, e2 = self.levels()
        return w[e2] - w[e1], {"E_e1": w[e1], "E_e2": w[e2]}

    def weights(self, k):
        e1, e2 = self.levels()
        gE = np.zeros(k)
        gE[e2] = gE[e2 + 1] = 0.5
        gE[e1] = gE[e1 + 1] = -0.5
        return gE, -1, -1, 0.0

    def theta_specs(self):
        mask_up = 0
        for s in self.upper:
            mask_up |= 1 << s
        return [(L.QD_DP_OZ, mask_up, 10.0), (L.QD_DP_FX, 0, 1e-5), (L.QD_DP_FY, 0, 1e-5), (L.QD_DP_FZ, 0, 1e-5)]


def _torch_function():
    import torch

    class SpectralObjective(torch.autograd.Function):
        """L(theta) of a problem; backward through qd_adjoint + qd_tables_vjp (module docstring)."""

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

    return SpectralObjective


def gap_coulomb_objective(theta, problem: DotProblem):
    """torch entry point: L(theta) = (E_e1 - E_h1) + J_eh with autograd support (6-element tensor)."""
    return _torch_function().apply(theta, problem)


def tunnel_splitting_objective(theta, problem: MoleculeProblem):
    """torch entry point: L(theta) = E_e2 - E_e1 of a dot molecule (4-element tensor)."""
    return _torch_function().apply(theta, problem)
