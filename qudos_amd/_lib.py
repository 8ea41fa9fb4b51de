"""ctypes binding of include/qudos_gpu.h (the library is built in-tree under qudos_amd/lib)."""
from __future__ import annotations

import ctypes as C
import sys
import os

import numpy as np

QD_OK = 0
QD_E_ARG, QD_E_INPUT, QD_E_ALLOC, QD_E_DEVICE, QD_E_NOCONV, QD_E_UNSUPPORTED, QD_E_NAN = -1, -2, -3, -4, -5, -6, -7
QD_FP64 = 0
QD_FP32_REFINE = 1
QD_REDUCE_TWO_STAGE = 0
QD_REDUCE_ONE_STAGE = 1
QD_NTAB = 20
_ERRS = {-1: "QD_E_ARG", -2: "QD_E_INPUT", -3: "QD_E_ALLOC", -4: "QD_E_DEVICE", -5: "QD_E_NOCONV",
         -6: "QD_E_UNSUPPORTED", -7: "QD_E_NAN"}

EXPORTED_SYMBOLS = ["qd_last_error", "qd_read_inputs", "qd_free_inputs", "qd_tables_build", "qd_tables_free",
                    "qd_window", "qd_create", "qd_destroy", "qd_upload_tables", "qd_hamiltonian", "qd_solve",
                    "qd_solve_device", "qd_fetch_results", "qd_solve_matrix", "qd_profile_enable",
                    "qd_profile_read", "qd_debug_tridiagonal", "qd_debug_apply_q", "qd_selftest_quad_sum", "qd_selftest_mfma", "qd_coulomb",
                    "qd_write_coulomb", "qd_adjoint", "qd_tables_derivative", "qd_tables_vjp", "qd_adjoint_bin",
                    "qd_set_reduction", "qd_set_refine_tolerance", "qd_debug_set_limits",
                    "qd_debug_mark_wait_failure", "qd_momentum", "qd_selftest_adjoint_bin_host", "qd_selftest_momentum_host",
                    "qd_selftest_coulomb_upsample_host", "qd_selftest_hamiltonian_host",
                    "qd_selftest_shape_host", "qd_localisation", "qd_overlaps",
                    "qd_tables_build_retain", "qd_tables_vjp_retained", "qd_envelopes", "qd_adjoint_z",
                    "qd_tables_build_device", "qd_selftest_shape_device", "qd_tables_vjp_device"]

QD_DP_FX, QD_DP_FY, QD_DP_FZ, QD_DP_L1, QD_DP_L2, QD_DP_L3, QD_DP_L12, QD_DP_X, QD_DP_OX, QD_DP_OY, QD_DP_OZ, QD_DP_L4 = range(12)


class QdShape(C.Structure):
    _fields_ = [("shape", C.c_int), ("embedded", C.c_int), ("material", C.c_int), ("x", C.c_double),
                ("y", C.c_double), ("Ox", C.c_double), ("Oy", C.c_double), ("Oz", C.c_double),
                ("L1", C.c_double), ("L2", C.c_double), ("L3", C.c_double), ("L4", C.c_double),
                ("L5", C.c_double), ("alpha", C.c_double), ("beta", C.c_double), ("gamma", C.c_double)]


class Inputs(C.Structure):
    """qd_inputs; the shapes array of an instance filled by qd_read_inputs is freed with it."""
    _fields_ = [("N_shape", C.c_int), ("shapes", C.POINTER(QdShape)), ("N", C.c_int * 3), ("N_S", C.c_int * 3),
                ("N_C", C.c_int * 3), ("strain_type", C.c_int), ("coulomb_type", C.c_int), ("RL", C.c_int * 3),
                ("RCL", C.c_int * 3), ("SL", C.c_double * 3), ("CL", C.c_double * 3), ("N_vb", C.c_int),
                ("N_cb", C.c_int), ("T", C.c_double), ("F", C.c_double * 3), ("K", C.c_double * 3),
                ("strain", C.c_int), ("isotropic", C.c_int), ("piezo", C.c_int), ("cme", C.c_int),
                ("softening", C.c_int), ("sigma_softening", C.c_double), ("madelung_ranks", C.c_int),
                ("mme", C.c_int), ("overlaps", C.c_int), ("localisation", C.c_int), ("N_line", C.c_int),
                ("linescan", C.c_int * 7), ("cross_sections", C.c_int), ("N_cross", C.c_int),
                ("cross_section", C.c_int * 8), ("out3d", C.c_int), ("N_3D", C.c_int)]

    def __del__(self):
        try:
            if self.shapes:
                load_library().qd_free_inputs(C.byref(self))
        except Exception:
            pass


class Tables(C.Structure):
    """qd_tables; the host arrays of an instance filled by qd_tables_build are freed with it
    (qd_tables_free is idempotent, so an explicit free before is harmless)."""
    _fields_ = [("N", C.c_int * 3), ("NT", C.c_int * 3), ("K", C.c_double * 3), ("L", C.c_double * 3),
                ("G", C.POINTER(C.c_double) * 3), ("tab", C.POINTER(C.c_double)), ("N_vb", C.c_int),
                ("N_cb", C.c_int), ("nr0", C.c_double), ("state", C.c_void_p)]

    def arrays(self):
        """(QD_NTAB, NTx, NTy, NTz) complex table copy and the three G axes."""
        nt = self.NT[0] * self.NT[1] * self.NT[2]
        raw = np.ctypeslib.as_array(self.tab, shape=(QD_NTAB * nt * 2,)).copy()
        tab = raw.view(np.complex128).reshape(QD_NTAB, self.NT[0], self.NT[1], self.NT[2])
        G = [np.ctypeslib.as_array(self.G[a], shape=(self.NT[a],)).copy() for a in range(3)]
        return tab, G

    def __del__(self):
        try:
            if self.tab:
                load_library().qd_tables_free(C.byref(self))
        except Exception:
            pass


class Stats(C.Structure):
    _fields_ = [("t_hamiltonian", C.c_double), ("t_reduce", C.c_double), ("t_tridiag", C.c_double),
                ("t_backtransform", C.c_double), ("t_refine", C.c_double), ("t_total", C.c_double),
                ("refine_iters", C.c_int), ("max_residual", C.c_double), ("max_orth", C.c_double),
                ("t_band", C.c_double), ("tri_residual", C.c_double)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class CoulombStats(C.Structure):
    _fields_ = [("t_pairs", C.c_double), ("t_contract", C.c_double), ("t_total", C.c_double),
                ("n_pairs", C.c_int), ("fft", C.c_int * 3), ("nq", C.c_long)]

    def as_dict(self):
        d = {f: getattr(self, f) for f, _ in self._fields_ if f != "fft"}
        d["fft"] = tuple(self.fft)
        return d


class MomentumStats(C.Structure):
    _fields_ = [("t_total", C.c_double)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class AdjointStats(C.Structure):
    _fields_ = [("t_density", C.c_double), ("t_sternheimer", C.c_double), ("t_bin", C.c_double),
                ("t_total", C.c_double), ("sternheimer_iters", C.c_int), ("sternheimer_residual", C.c_double),
                ("J", C.c_double)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class QdError(RuntimeError):
    def __init__(self, code, msg, stats=None):
        super().__init__(f"{_ERRS.get(code, code)}: {msg}")
        self.code = code
        self.stats = stats      # solve statistics when a solve returned QD_E_NOCONV


_LIB = None


def lib_path():
    return os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib", "libqdgpu.so")


def load_library():
    """Load the in-tree libqdgpu.so; raises if it has not been built (no fallback)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    path = lib_path()
    if not os.path.exists(path):
        raise OSError(f"{path} is missing: build it with __graft_entry__.build() or make -C qudos_amd/csrc")
    _LIB = bind(C.CDLL(path))
    return _LIB


def bind(L):
    """Declare the argument and result types of the C ABI on a loaded library object."""
    P = C.POINTER
    L.qd_last_error.restype = C.c_char_p
    L.qd_read_inputs.argtypes = [C.c_char_p, C.c_char_p, C.c_int, P(C.c_char_p), P(Inputs)]
    L.qd_free_inputs.argtypes = [P(Inputs)]
    L.qd_free_inputs.restype = None
    L.qd_tables_build.argtypes = [P(Inputs), P(Tables)]
    L.qd_tables_free.argtypes = [P(Tables)]
    L.qd_tables_free.restype = None
    L.qd_window.argtypes = [P(Inputs), P(C.c_int), P(C.c_int)]
    L.qd_create.argtypes = [C.c_int, P(C.c_void_p)]
    L.qd_destroy.argtypes = [C.c_void_p]
    L.qd_upload_tables.argtypes = [C.c_void_p, P(Tables)]
    L.qd_tables_build_device.argtypes = [C.c_void_p, P(Inputs), P(Tables)]
    L.qd_tables_vjp_device.argtypes = [C.c_void_p, P(Inputs), P(C.c_double), C.c_int, P(C.c_int), P(C.c_uint),
                                       P(C.c_double)]
    L.qd_selftest_shape_device.argtypes = [C.c_void_p, P(Inputs), C.c_int, C.c_int, C.c_int, C.c_int,
                                           P(C.c_double), P(C.c_double)]
    L.qd_hamiltonian.argtypes = [C.c_void_p, C.c_int, P(C.c_double)]
    L.qd_solve.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, P(C.c_double), P(C.c_double), P(Stats)]
    L.qd_solve_device.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, P(Stats)]
    L.qd_fetch_results.argtypes = [C.c_void_p, P(C.c_double), P(C.c_double)]
    L.qd_solve_matrix.argtypes = [C.c_void_p, C.c_int, P(C.c_double), C.c_int, C.c_int, C.c_int, P(C.c_double),
                                  P(C.c_double), P(Stats)]
    L.qd_debug_tridiagonal.argtypes = [C.c_void_p, P(C.c_double), P(C.c_double)]
    L.qd_debug_apply_q.argtypes = [C.c_void_p, C.c_int, C.c_int, P(C.c_double)]
    L.qd_selftest_mfma.argtypes = [P(C.c_double)]
    L.qd_selftest_quad_sum.argtypes = [P(C.c_double), P(C.c_double)]
    L.qd_profile_enable.argtypes = [C.c_void_p, C.c_int]
    L.qd_set_reduction.argtypes = [C.c_void_p, C.c_int]
    L.qd_set_refine_tolerance.argtypes = [C.c_void_p, C.c_double]
    L.qd_debug_set_limits.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int]
    L.qd_debug_mark_wait_failure.argtypes = [C.c_void_p]
    L.qd_profile_read.argtypes = [C.c_void_p, P(C.c_int), P(C.c_double), P(C.c_double)]
    L.qd_coulomb.argtypes = [C.c_void_p, P(Inputs), P(Tables), C.c_int, C.c_int, P(C.c_double), P(C.c_double),
                             P(C.c_double), P(C.c_double), P(C.c_double), P(CoulombStats)]
    L.qd_selftest_adjoint_bin_host.argtypes = [P(Tables), C.c_int, P(C.c_double), C.c_int, P(C.c_int), P(C.c_int),
                                               P(C.c_double), P(C.c_double), P(C.c_double)]
    L.qd_selftest_momentum_host.argtypes = [P(Tables), C.c_int, P(C.c_double), C.c_int, P(C.c_double), P(C.c_double)]
    L.qd_selftest_coulomb_upsample_host.argtypes = [P(Inputs), P(Tables), C.c_int, P(C.c_double), P(C.c_double)]
    L.qd_selftest_hamiltonian_host.argtypes = [P(Tables), P(C.c_double)]
    L.qd_selftest_shape_host.argtypes = [P(Inputs), C.c_int, C.c_int, C.c_int, C.c_int, P(C.c_double), P(C.c_double)]
    L.qd_localisation.argtypes = [C.c_void_p, P(Inputs), P(C.c_double), P(C.c_double), P(C.c_double)]
    L.qd_overlaps.argtypes = [C.c_void_p, P(Inputs), P(C.c_double)]
    L.qd_envelopes.argtypes = [C.c_void_p, P(Inputs), C.c_int, C.c_int, C.c_int, P(C.c_double), P(C.c_double),
                               P(C.c_double), P(C.c_double)]
    L.qd_momentum.argtypes = [C.c_void_p, P(Inputs), C.c_int, C.c_int, P(C.c_double), P(MomentumStats)]
    L.qd_adjoint.argtypes = [C.c_void_p, P(Inputs), P(Tables), P(C.c_double), C.c_int, C.c_int, C.c_double,
                             P(C.c_double), P(AdjointStats)]
    L.qd_tables_derivative.argtypes = [P(Inputs), C.c_int, C.c_uint, C.c_double, P(C.c_double)]
    L.qd_tables_build_retain.argtypes = [P(Inputs), P(Tables)]
    L.qd_tables_vjp_retained.argtypes = [P(Tables), P(Inputs), P(C.c_double), C.c_int, P(C.c_int), P(C.c_uint),
                                         P(C.c_double)]
    L.qd_tables_vjp.argtypes = [P(Inputs), P(C.c_double), C.c_int, P(C.c_int), P(C.c_uint), P(C.c_double)]
    L.qd_adjoint_z.argtypes = [C.c_void_p, P(C.c_double), C.c_int, P(C.c_int), P(C.c_double), P(C.c_double),
                               P(AdjointStats)]
    L.qd_adjoint_bin.argtypes = [C.c_void_p, C.c_int, P(C.c_double), C.c_int, P(C.c_int), P(C.c_int),
                                 P(C.c_double), P(C.c_double)]
    L.qd_write_coulomb.argtypes = [C.c_char_p, C.c_int, C.c_int, P(C.c_double), P(C.c_double), P(C.c_double),
                                   P(C.c_double)]
    return L


def _check(rc, stats=None):
    if rc != QD_OK:
        raise QdError(rc, load_library().qd_last_error().decode(), stats.as_dict() if stats is not None else None)


def _dptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def read_inputs(parameters_path, supercell_path, overrides=()):
    """parameters.in + supercell.in (+ '-key=value' overrides) -> Inputs."""
    L = load_library()
    inp = Inputs()
    argv = (C.c_char_p * max(1, len(overrides)))(*[o.encode() for o in overrides])
    _check(L.qd_read_inputs(str(parameters_path).encode(), str(supercell_path).encode(), len(overrides), argv,
                            C.byref(inp)))
    return inp


def build_tables(inp: Inputs, retain=False) -> Tables:
    """qd_tables_build, or qd_tables_build_retain (retain=True: the builder's intermediates stay with
    the tables for tables_vjp(..., tables=t))."""
    L = load_library()
    t = Tables()
    _check((L.qd_tables_build_retain if retain else L.qd_tables_build)(C.byref(inp), C.byref(t)))
    return t


def window(inp: Inputs):
    il, iu = C.c_int(), C.c_int()
    _check(load_library().qd_window(C.byref(inp), C.byref(il), C.byref(iu)))
    return il.value, iu.value


class Solver:
    """One GPU handle (qd_create / qd_destroy)."""

    def __init__(self, device=0):
        L = load_library()
        h = C.c_void_p()
        _check(L.qd_create(device, C.byref(h)))
        self._h = h
        self.n = None

    def close(self):
        if self._h:
            load_library().qd_destroy(self._h)
            self._h = None

    def __del__(self):
        # no HIP calls while the interpreter shuts down (the runtime may already be finalising):
        # a handle still open then is left to the process exit
        if sys.is_finalizing():
            return
        try:
            self.close()
        except Exception:
            pass

    def upload(self, tables: Tables):
        _check(load_library().qd_upload_tables(self._h, C.byref(tables)))
        self.n = 8 * tables.N[0] * tables.N[1] * tables.N[2]

    def tables_vjp(self, inp: Inputs, W, params):
        """qd_tables_vjp_device: tables_vjp(inp, W, params) with the reverse pass on this handle's device."""
        Wc = np.ascontiguousarray(W, dtype=np.complex128).reshape(-1)
        pa = np.array([p for p, _ in params], dtype=np.int32)
        ma = np.array([m for _, m in params], dtype=np.uint32)
        g = np.zeros(len(params))
        _check(load_library().qd_tables_vjp_device(self._h, C.byref(inp), _dptr(Wc.view(np.float64)), len(params),
                                                   pa.ctypes.data_as(C.POINTER(C.c_int)),
                                                   ma.ctypes.data_as(C.POINTER(C.c_uint)), _dptr(g)))
        return g

    def shape_transform(self, inp: Inputs, idx, G, cell=0, param=-1):
        """qd_selftest_shape_device: chi_idx at the rows of G (n x 3) on this handle's device, or its
        derivative with respect to param."""
        Gc = np.ascontiguousarray(G, dtype=np.float64).reshape(-1, 3)
        out = np.zeros(Gc.shape[0], dtype=np.complex128)
        _check(load_library().qd_selftest_shape_device(self._h, C.byref(inp), idx, cell, param, Gc.shape[0],
                                                       _dptr(Gc), _dptr(out.view(np.float64))))
        return out

    def build_tables(self, inp: Inputs) -> Tables:
        """qd_tables_build_device: the tables built on this handle's device (fields, convolutions,
        tables; shape transforms on the host), returned as a host copy and left uploaded."""
        t = Tables()
        _check(load_library().qd_tables_build_device(self._h, C.byref(inp), C.byref(t)))
        self.n = 8 * t.N[0] * t.N[1] * t.N[2]
        return t

    def hamiltonian(self, precision=QD_FP64):
        n = self.n
        H = np.zeros((n, n), dtype=np.complex128)   # row-major view of column-major data
        _check(load_library().qd_hamiltonian(self._h, precision, _dptr(H.view(np.float64))))
        Hl = H.T.copy()   # column-major -> H[row, col]
        return Hl

    def solve(self, il, iu, precision=QD_FP64, vectors=True):
        k = iu - il + 1
        w = np.zeros(k)
        Z = np.zeros((k, self.n), dtype=np.complex128) if vectors else None
        st = Stats()
        _check(load_library().qd_solve(self._h, il, iu, precision, _dptr(w),
                                       _dptr(Z.view(np.float64)) if vectors else None, C.byref(st)), st)
        return w, (Z.T.copy() if vectors else None), st.as_dict()

    def solve_device(self, il, iu, precision=QD_FP64):
        st = Stats()
        _check(load_library().qd_solve_device(self._h, il, iu, precision, C.byref(st)), st)
        return st.as_dict()

    def fetch(self, k, vectors=True):
        w = np.zeros(k)
        Z = np.zeros((k, self.n), dtype=np.complex128) if vectors else None
        _check(load_library().qd_fetch_results(self._h, _dptr(w), _dptr(Z.view(np.float64)) if vectors else None))
        return w, (Z.T.copy() if vectors else None)

    def tridiagonal(self, n):
        d, e = np.zeros(n), np.zeros(n - 1)
        _check(load_library().qd_debug_tridiagonal(self._h, _dptr(d), _dptr(e)))
        return d, e

    def apply_q(self, X, adjoint=False):
        """Q X (or Q^H X) with the unitary factor of the last reduction (test support)."""
        Xt = np.ascontiguousarray(np.asarray(X, dtype=np.complex128).T)   # column-major n x k
        _check(load_library().qd_debug_apply_q(self._h, Xt.shape[0], 1 if adjoint else 0, _dptr(Xt.view(np.float64))))
        return Xt.T.copy()

    def set_refine_tolerance(self, rel_residual):
        """Stop of the FP64 refinement: every ||H x - w x|| <= rel_residual ||T||_1 (default 1e-10)."""
        _check(load_library().qd_set_refine_tolerance(self._h, float(rel_residual)))

    def set_limits(self, refine_sweeps=-1, sternheimer_iters=-1, invit_sweeps=-1, bisect_rounds=-1):
        """Iteration limits of later solves (test support); -1 keeps the default."""
        _check(load_library().qd_debug_set_limits(self._h, refine_sweeps, sternheimer_iters, invit_sweeps,
                                                  bisect_rounds))

    def mark_wait_failure(self):
        """Record a bounded-wait failure for the last factorisation (test support)."""
        _check(load_library().qd_debug_mark_wait_failure(self._h))

    def set_reduction(self, method):
        """QD_REDUCE_TWO_STAGE (default) or QD_REDUCE_ONE_STAGE for later solves."""
        _check(load_library().qd_set_reduction(self._h, int(method)))

    def profile(self, enable=True):
        _check(load_library().qd_profile_enable(self._h, 1 if enable else 0))

    def profile_read(self):
        n, t, b = C.c_int(), C.c_double(), C.c_double()
        _check(load_library().qd_profile_read(self._h, C.byref(n), C.byref(t), C.byref(b)))
        return n.value, t.value, b.value

    def solve_matrix(self, A, il, iu, precision=QD_FP64):
        """Eigenpairs il..iu of a Hermitian matrix given by its lower triangle (numpy, any layout)."""
        n = A.shape[0]
        Acm = np.ascontiguousarray(np.tril(A).T)   # column-major storage as row-major of A^T
        k = iu - il + 1
        w = np.zeros(k)
        Z = np.zeros((k, n), dtype=np.complex128)
        st = Stats()
        _check(load_library().qd_solve_matrix(self._h, n, _dptr(Acm.view(np.float64)), il, iu, precision, _dptr(w),
                                              _dptr(Z.view(np.float64)), C.byref(st)))
        return w, Z.T.copy(), st.as_dict()


    def coulomb(self, inp: Inputs, tables: Tables, nh, ne, Z=None):
        """Coulomb matrix elements of nh hole + ne electron states (qd_coulomb).  Z: N_H x (nh+ne)
        complex (holes first) or None for the device-resident eigenvectors of the last solve.
        Returns ({hh, ee, eh_direct, eh_exchange} arrays in the reference's index order, stats)."""
        hh = np.zeros((nh,) * 4, dtype=np.complex128)
        ee = np.zeros((ne,) * 4, dtype=np.complex128)
        ed = np.zeros((nh, nh, ne, ne), dtype=np.complex128)
        ex = np.zeros((nh, ne, ne, nh), dtype=np.complex128)
        zp = None
        if Z is not None:
            Zc = np.asfortranarray(np.asarray(Z, dtype=np.complex128))
            if Zc.shape != (self.n, nh + ne):
                raise ValueError(f"Z must be {self.n} x {nh + ne}")
            zp = Zc.T.copy()   # row-major of Z^T == column-major Z
        st = CoulombStats()
        _check(load_library().qd_coulomb(self._h, C.byref(inp), C.byref(tables), nh, ne,
                                         _dptr(zp.view(np.float64)) if zp is not None else None,
                                         _dptr(hh.view(np.float64)), _dptr(ee.view(np.float64)),
                                         _dptr(ed.view(np.float64)), _dptr(ex.view(np.float64)), C.byref(st)))
        return {"hh": hh, "ee": ee, "eh_direct": ed, "eh_exchange": ex}, st.as_dict()


    def momentum(self, inp: Inputs, nv, nc):
        """qd_momentum: p[a, c, v] = <psi_v| dH/dk_a |psi_c> (eV Angstrom) for a = x, y, z, the nc
        lowest conduction states c and the nv highest valence states v (v = 0 highest) of the last
        solve."""
        p = np.zeros((3, nc, nv), dtype=np.complex128)
        st = MomentumStats()
        _check(load_library().qd_momentum(self._h, C.byref(inp), nv, nc, _dptr(p.view(np.float64)), C.byref(st)))
        return p, st.as_dict()

    def localisation(self, inp: Inputs):
        """qd_localisation: (vb, cb) fractions per shape, rows as in carrier_localisation_*.dat."""
        ns = inp.N_shape
        vb = np.zeros((2 * inp.N_vb, ns))
        cb = np.zeros((2 * inp.N_cb, ns))
        t = C.c_double(0.0)
        _check(load_library().qd_localisation(self._h, C.byref(inp), _dptr(vb), _dptr(cb), C.byref(t)))
        return vb, cb, t.value

    def overlaps(self, inp: Inputs):
        """qd_overlaps: (2 N_cb, 2 N_vb) complex <cb|vb>, valence index 0 the highest state."""
        ov = np.zeros((2 * inp.N_cb, 2 * inp.N_vb), dtype=np.complex128)
        _check(load_library().qd_overlaps(self._h, C.byref(inp), _dptr(ov.view(np.float64))))
        return ov

    def envelopes(self, inp: Inputs, x, y, z, first=0, count=None):
        """qd_envelopes: F[s, b, p] (complex, Angstrom^-3/2) of the window states first ..
        first + count - 1 (ascending energy) at the points (x, y, z) [Angstrom]."""
        x, y, z = (np.ascontiguousarray(v, dtype=np.float64) for v in (x, y, z))
        if count is None:
            count = 2 * (inp.N_vb + inp.N_cb) - first
        F = np.zeros((count, 8, len(x)), dtype=np.complex128)
        _check(load_library().qd_envelopes(self._h, C.byref(inp), first, count, len(x), _dptr(x), _dptr(y), _dptr(z),
                                           _dptr(F.view(np.float64))))
        return F

    def adjoint(self, inp: Inputs, tables: Tables, gE, hole_level=-1, electron_level=-1, wJ=0.0):
        """qd_adjoint: W = dL/dT (QD_NTAB, NTx, NTy, NTz) complex for L = sum gE E_n + wJ J_eh."""
        nt = tables.NT[0] * tables.NT[1] * tables.NT[2]
        W = np.zeros(QD_NTAB * nt, dtype=np.complex128)
        g = np.ascontiguousarray(gE, dtype=np.float64)
        st = AdjointStats()
        _check(load_library().qd_adjoint(self._h, C.byref(inp), C.byref(tables), _dptr(g), hole_level,
                                         electron_level, wJ, _dptr(W.view(np.float64)), C.byref(st)))
        return W.reshape(QD_NTAB, tables.NT[0], tables.NT[1], tables.NT[2]), st.as_dict()

    def adjoint_z(self, NT, gE=None, cols=(), gZ=None):
        """qd_adjoint_z: W = dL/dT (QD_NTAB, NTx, NTy, NTz) complex for dL = sum gE dE_n + 2 Re sum_j
        <gZ[:, j], d psi_{cols[j]}> (gZ: N_H x len(cols) complex, dL/dpsi^*), and the statistics."""
        nt = NT[0] * NT[1] * NT[2]
        W = np.zeros(QD_NTAB * nt, dtype=np.complex128)
        g = None if gE is None else np.ascontiguousarray(gE, dtype=np.float64)
        c = np.ascontiguousarray(cols, dtype=np.int32)
        Zg = None if gZ is None else np.asfortranarray(np.asarray(gZ, dtype=np.complex128))
        st = AdjointStats()
        _check(load_library().qd_adjoint_z(self._h, _dptr(g) if g is not None else None, len(c),
                                           c.ctypes.data_as(C.POINTER(C.c_int)) if len(c) else None,
                                           _dptr(Zg.T.copy().view(np.float64)) if Zg is not None else None,
                                           _dptr(W.view(np.float64)), C.byref(st)), st)
        return W.reshape(QD_NTAB, *NT), st.as_dict()

    def adjoint_bin(self, V, terms, NT):
        """qd_adjoint_bin: W for D = sum gamma V[:, a] V[:, b]^H (terms: list of (a, b, gamma))."""
        Vc = np.asfortranarray(np.asarray(V, dtype=np.complex128))
        vp = Vc.T.copy()
        a = np.array([t[0] for t in terms], dtype=np.int32)
        b = np.array([t[1] for t in terms], dtype=np.int32)
        g = np.array([t[2] for t in terms], dtype=np.float64)
        nt = NT[0] * NT[1] * NT[2]
        W = np.zeros(QD_NTAB * nt, dtype=np.complex128)
        _check(load_library().qd_adjoint_bin(self._h, Vc.shape[1], _dptr(vp.view(np.float64)), len(terms),
                                             a.ctypes.data_as(C.POINTER(C.c_int)), b.ctypes.data_as(C.POINTER(C.c_int)),
                                             _dptr(g), _dptr(W.view(np.float64))))
        return W.reshape(QD_NTAB, *NT)


def tables_derivative(inp: Inputs, param, shape_mask=0, rel_step=1e-5):
    """qd_tables_derivative: dT/dtheta as (QD_NTAB, NTx, NTy, NTz) complex."""
    NT = [2 * inp.N[a] - 1 for a in range(3)]
    dT = np.zeros(QD_NTAB * NT[0] * NT[1] * NT[2], dtype=np.complex128)
    _check(load_library().qd_tables_derivative(C.byref(inp), param, shape_mask, rel_step, _dptr(dT.view(np.float64))))
    return dT.reshape(QD_NTAB, *NT)


def selftest_adjoint_bin_host(tables: Tables, V, terms):
    """qd_selftest_adjoint_bin_host: (W_fft, W_direct), each (QD_NTAB, NTx, NTy, NTz) complex, for
    D = sum gamma V[:, a] V[:, b]^H; terms = [(a, b, gamma), ...].  Host only."""
    Vc = np.asfortranarray(np.asarray(V, dtype=np.complex128))
    nv = Vc.shape[1]
    vp = Vc.T.copy()
    a = np.array([t[0] for t in terms], dtype=np.int32)
    b = np.array([t[1] for t in terms], dtype=np.int32)
    g = np.array([t[2] for t in terms], dtype=np.float64)
    NT = tuple(tables.NT)
    Wf = np.zeros((QD_NTAB,) + NT, dtype=np.complex128)
    Wd = np.zeros((QD_NTAB,) + NT, dtype=np.complex128)
    lib = load_library()
    _check(lib.qd_selftest_adjoint_bin_host(C.byref(tables), nv, _dptr(vp.view(np.float64)), len(terms),
                                            a.ctypes.data_as(C.POINTER(C.c_int)), b.ctypes.data_as(C.POINTER(C.c_int)),
                                            _dptr(g), _dptr(Wf.view(np.float64)), _dptr(Wd.view(np.float64))))
    return Wf, Wd


def selftest_momentum_host(tables: Tables, V, Cc):
    """qd_selftest_momentum_host: p[a, c, v] for valence states V (ascending, v = 0 is the last
    column) and conduction states Cc, on the host with the momentum kernel's element code."""
    Vf = np.asfortranarray(np.asarray(V, dtype=np.complex128))
    Cf = np.asfortranarray(np.asarray(Cc, dtype=np.complex128))
    nv, nc = Vf.shape[1], Cf.shape[1]
    vp, cp = Vf.T.copy(), Cf.T.copy()
    p = np.zeros((3, nc, nv), dtype=np.complex128)
    _check(load_library().qd_selftest_momentum_host(C.byref(tables), nv, _dptr(vp.view(np.float64)), nc,
                                                    _dptr(cp.view(np.float64)), _dptr(p.view(np.float64))))
    return p


def selftest_coulomb_upsample_host(inp: Inputs, tables: Tables, B):
    """qd_selftest_coulomb_upsample_host: B (P, NTx, NTy, NTz) complex -> (P, NTCx, NTCy, NTCz)."""
    Bc = np.ascontiguousarray(B, dtype=np.complex128)
    P = Bc.shape[0]
    NTC = tuple(2 * inp.N_C[a] - 1 for a in range(3))
    out = np.zeros((P,) + NTC, dtype=np.complex128)
    _check(load_library().qd_selftest_coulomb_upsample_host(C.byref(inp), C.byref(tables), P, _dptr(Bc.view(np.float64)),
                                                            _dptr(out.view(np.float64))))
    return out


def selftest_hamiltonian_host(tables: Tables):
    """qd_selftest_hamiltonian_host: the full Hermitian H (N_H x N_H complex) built on the host with
    the assembly kernels' element code."""
    n = 8 * tables.N[0] * tables.N[1] * tables.N[2]
    H = np.zeros((n, n), dtype=np.complex128, order="F")
    _check(load_library().qd_selftest_hamiltonian_host(C.byref(tables), _dptr(H.reshape(-1, order="F").view(np.float64))))
    return H


def selftest_shape_host(inp: Inputs, idx, G, cell=0, param=-1):
    """qd_selftest_shape_host: chi_idx at the rows of G (n x 3), or d chi / d param."""
    Gc = np.ascontiguousarray(G, dtype=np.float64).reshape(-1, 3)
    out = np.zeros(Gc.shape[0], dtype=np.complex128)
    _check(load_library().qd_selftest_shape_host(C.byref(inp), idx, cell, param, Gc.shape[0], _dptr(Gc),
                                                 _dptr(out.view(np.float64))))
    return out


def tables_vjp(inp: Inputs, W, params, tables: Tables = None):
    """qd_tables_vjp: dL/dtheta for L = Re sum W T; params: list of (QD_DP_*, shape_mask).  With
    tables built by build_tables(inp, retain=True), qd_tables_vjp_retained reuses their intermediates."""
    Wc = np.ascontiguousarray(W, dtype=np.complex128).reshape(-1)
    pa = np.array([p for p, _ in params], dtype=np.int32)
    ma = np.array([m for _, m in params], dtype=np.uint32)
    g = np.zeros(len(params))
    L = load_library()
    args = (C.byref(inp), _dptr(Wc.view(np.float64)), len(params), pa.ctypes.data_as(C.POINTER(C.c_int)),
            ma.ctypes.data_as(C.POINTER(C.c_uint)), _dptr(g))
    _check(L.qd_tables_vjp_retained(C.byref(tables), *args) if tables is not None else L.qd_tables_vjp(*args))
    return g


def write_coulomb(directory, cme):
    """qd_write_coulomb: the four coulomb_matrix_elements_*.dat files."""
    hh, ee = np.ascontiguousarray(cme["hh"]), np.ascontiguousarray(cme["ee"])
    ed, ex = np.ascontiguousarray(cme["eh_direct"]), np.ascontiguousarray(cme["eh_exchange"])
    nh, ne = hh.shape[0], ee.shape[0]
    _check(load_library().qd_write_coulomb(str(directory).encode(), nh, ne, _dptr(hh.view(np.float64)),
                                           _dptr(ee.view(np.float64)), _dptr(ed.view(np.float64)),
                                           _dptr(ex.view(np.float64))))


def selftest_mfma():
    out = np.zeros(2 + 512)
    _check(load_library().qd_selftest_mfma(_dptr(out)))
    return out


def selftest_quad_sum(x):
    """DPP reductions of 64 doubles on the device (test support): 4-, 8-, 16-lane group sums, the
    lane-0 broadcast and the 64-lane sum, as a (5, 64) array."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.zeros(320)
    _check(load_library().qd_selftest_quad_sum(_dptr(x), _dptr(out)))
    return out.reshape(5, 64)
